// rt_fused.hip -- fused sweep-block kernel (K colour steps per launch,
// cells in LDS, TinyMT states in VGPRs) and its host driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "rt_fused.hpp"

namespace rt {

constexpr int kLW = 64;        // loaded tile width: one column per lane
constexpr int kKMax = 64;      // steps per window

template <int MODEL>
struct CellOf {
  using type = uint8_t;
};
template <>
struct CellOf<M_RECTTRI> {
  using type = int32_t;
};

template <int MODEL>
__device__ __forceinline__ bool has_stream(int i, int j, int rows, int cols) {
  if (MODEL == M_DOMINO) return i >= 1 && i <= rows - 2 && j >= 1 && j <= cols - 2;
  return i >= 0 && i < rows && j >= 0 && j < cols;
}

// TinyMT stream index = the reference's NDRange sequential id
// (tinymt.clh:23-30): domino NDRange (N-2, N/2-2) offset by one,
// lozenge / rect-triangle NDRange (N, N/3).
template <int MODEL>
__device__ __forceinline__ long stream_id(int i, int j, int rows, int cols) {
  if (MODEL == M_DOMINO) return (long)(j - 1) + (long)(cols - 2) * (i - 1);
  return (long)j + (long)cols * i;
}

// Per-lane constants of a tile, fixed for a whole window.  Row r of a
// thread is global row gr0 + W*r and LDS index li0 + W*r*RS, so every LDS
// address of the step is li0 plus a compile-time offset.
struct Lanes {
  int li0;          // LDS index of the thread's row-0 cell
  int gr0;          // its global row
  int im0;          // gr0 mod 3
  uint32_t hsm;     // bit r: row r owns a TinyMT stream (and moves)
  uint32_t interm;  // bit r: row r is interior (pulls from neighbours)
};

__device__ __forceinline__ uint32_t pick(uint32_t mask_bits, int r, uint32_t yes, uint32_t no) {
  const uint32_t m = 0u - ((mask_bits >> r) & 1u);
  return (yes & m) | (no & ~m);
}

// One colour step on the LDS tile, specialised on the (workgroup-uniform)
// pull pattern so the body is branch-free.  All reads of the step (own cell
// and neighbours, every row and chain) come before any write: the planes
// read and the plane written differ, but the compiler cannot see that from
// runtime colour offsets, and stores issued between reads would serialise
// every row behind an LDS round trip.
// Lozenge rows of up to 4 cells per lane turn the draw into the move's
// target once for both chains; 8-row lanes keep the draw (the targets would
// push the pair kernel past 128 VGPRs).
template <int MODEL, int R>
constexpr bool kTarget = MODEL == M_LOZENGE && R <= 4;

// The next step's draw of every row (its move target for kTarget).
template <int MODEL, int R>
__device__ __forceinline__ void step_draws(Tiny* st, const Lanes& L, uint32_t* b) {
#pragma unroll
  for (int r = 0; r < R; r++) {
    b[r] = tiny_bits(st[r]);
    if (kTarget<MODEL, R>) b[r] = lozenge_target(b[r], (L.hsm >> r) & 1u);
  }
}

template <int MODEL, int NC, int W, int R, int PULL, class Cell, int RS, int PL>
__device__ __forceinline__ void sweep_step(Cell* lds, const Lanes& L, const uint32_t* b,
                                           uint32_t op, const UnitDesc& U, const ModelParams& mp) {
  constexpr int NCOL = MODEL == M_DOMINO ? 2 : 3;
  const int c = op & OP_COLOUR;
  uint32_t out[NC][R];
#pragma unroll
  for (int ch = 0; ch < NC; ch++) {
    const Cell* P = lds + ch * NCOL * PL + L.li0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int off = r * W * RS;
      const int gr = L.gr0 + W * r;
      uint32_t v = (uint32_t)P[c * PL + off];
      uint32_t nv;
      if (MODEL == M_DOMINO) {
        if (PULL & 1) {
          const uint32_t pv = domino_rebuild(P + (1 - c) * PL + off, RS, gr, c);
          v = pick(L.interm, r, pv, v);
        }
        nv = domino_move(v, b[r], U.rule, gr, 1 - c, U.dw);
      } else if (MODEL == M_LOZENGE) {
        // Both pulls, then one select on interm (a cell on the array's
        // border keeps its bits).
        const int im = (L.im0 + (W % 3) * r) % 3;
        if (PULL) {
          uint32_t pv = v;
          if (PULL & 1) pv = lozenge_pull_prev_m(pv, P + ((c + 2) % 3) * PL + off, RS, im, (c + 2) % 3);
          if (PULL & 2) pv = lozenge_pull_prev2_m(pv, P + ((c + 1) % 3) * PL + off, RS, im, (c + 1) % 3);
          v = pick(L.interm, r, pv, v);
        }
        if (kTarget<MODEL, R>) {  // the target already holds the has-stream test
          out[ch][r] = lozenge_apply(v, b[r]);
          continue;
        }
        nv = lozenge_move(v, b[r]);
      } else {
        if (PULL & 1) {
          const int pv = (op >> 4) & 3;
          const int32_t* src = (const int32_t*)(P + pv * PL + off);
          const uint32_t w = c == (pv + 1) % 3 ? recttri_rebuild1(src, RS, gr, pv)
                                               : recttri_rebuild2(src, RS, gr, pv);
          v = pick(L.interm, r, w, v);
        }
        nv = ((L.hsm >> r) & 1u) ? recttri_move(v, b[r], mp.rw) : v;
      }
      out[ch][r] = pick(L.hsm, r, nv, v);
    }
  }
#pragma unroll
  for (int ch = 0; ch < NC; ch++)
#pragma unroll
    for (int r = 0; r < R; r++)
      lds[ch * NCOL * PL + L.li0 + c * PL + r * W * RS] = (Cell)out[ch][r];
}

// Work item of this workgroup.  Workgroups go to the 8 XCDs round robin
// (block b to XCD b % 8); with xcd set, XCD x takes the contiguous x-th
// eighth of the work list, so tiles that share halo rows share an L2.  The
// grid is then rounded up to a multiple of 8 and surplus blocks get -1.
__device__ __forceinline__ int work_index(int nwork, int xcd) {
  unsigned b = blockIdx.x;
  if (xcd) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
  return b < (unsigned)nwork ? (int)b : -1;
}

// The K colour steps of a window on the LDS tile, NCR chains.  PIPE: the
// next step's op byte and draws are taken before the barrier, so a step
// starts with its LDS loads instead of an op read and the TinyMT
// arithmetic.  Measured per launch on the side-300 hexagon CFTP: 16x4
// pairs at eight waves per SIMD (32 samples) 151 -> 139 us, but 16x4 pairs
// at four waves (4 samples) 28.0 -> 29.6 and 16x4 split pairs (32 samples)
// 155 -> 160, so only the eight-wave kernel takes it.
template <int MODEL, int NCR, int W, int R, bool PIPE, class Cell, int RS, int PL>
__device__ __forceinline__ void run_window(Cell* lds, const Lanes& L, Tiny* st,
                                           const uint8_t* sops, int K, const UnitDesc& U,
                                           const ModelParams& mp) {
  if (K <= 0) return;
  uint32_t b[R];
  if (PIPE) step_draws<MODEL, R>(st, L, b);
  uint32_t op = sops[0];
  for (int s = 0; s < K; s++) {
    if (!PIPE) {
      op = sops[s];
      step_draws<MODEL, R>(st, L, b);
    }
    const int pull = (op >> 2) & 3;
    if (MODEL == M_LOZENGE) {
      switch (pull) {
        case 0: sweep_step<MODEL, NCR, W, R, 0, Cell, RS, PL>(lds, L, b, op, U, mp); break;
        case 1: sweep_step<MODEL, NCR, W, R, 1, Cell, RS, PL>(lds, L, b, op, U, mp); break;
        case 2: sweep_step<MODEL, NCR, W, R, 2, Cell, RS, PL>(lds, L, b, op, U, mp); break;
        default: sweep_step<MODEL, NCR, W, R, 3, Cell, RS, PL>(lds, L, b, op, U, mp); break;
      }
    } else {
      if (pull & 1) sweep_step<MODEL, NCR, W, R, 1, Cell, RS, PL>(lds, L, b, op, U, mp);
      else sweep_step<MODEL, NCR, W, R, 0, Cell, RS, PL>(lds, L, b, op, U, mp);
    }
    if (PIPE && s + 1 < K) {
      op = sops[s + 1];
      step_draws<MODEL, R>(st, L, b);
    }
    __syncthreads();
  }
}

// W wave64s per workgroup; each lane owns one tile column and R rows
// (rows wv, wv+W, ...), so the loaded tile is (W*R) x 64 cells.
template <int MODEL, int NC, int W, int R, bool PIPE>
__device__ __forceinline__ void fused_body(const UnitDesc* __restrict__ units,
                                           const WorkItem* __restrict__ work, int nwork, int xcd,
                                           int H, const ModelParams& mp,
                                           const WindowRec* __restrict__ wins) {
  using Cell = typename CellOf<MODEL>::type;
  constexpr int NCOL = MODEL == M_DOMINO ? 2 : 3;
  constexpr int kThreads = 64 * W;
  constexpr int LH = W * R;
  constexpr int RS = kLW + 2;
  constexpr int PL = (LH + 2) * RS;
  __shared__ __align__(16) Cell lds[NC * NCOL * PL];
  static_assert((NCOL * PL * sizeof(Cell)) % 4 == 0, "chain planes must be whole words");
  __shared__ uint8_t sops[kKMax];

  const int b = work_index(nwork, xcd);
  if (b < 0) return;
  const WorkItem wi = work[b];
  const WindowRec w = wins[wi.unit];
  const long s0 = w.s0;
  const int K = w.K, init = w.init;
  const uint32_t seed = w.seed;
  if (K <= 0 && !init) return;  // unit idle in this launch
  UnitDesc U = units[wi.unit];
  window_buffers(U, w.flags);
  const int rows = U.rows, cols = U.cols;
  const int HC = col_halo(MODEL, H);
  const int TH = LH - 2 * H, TW = kLW - 2 * HC;
  const int tr = wi.tile / U.tiles_c, tc = wi.tile % U.tiles_c;
  const int r0 = tr * TH - H, c0 = tc * TW - HC + wi.cs;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform row offset
  const int gc = c0 + lane;
  const bool col_in = gc >= 0 && gc < cols;

  for (int k = threadIdx.x; k < NC * NCOL * PL; k += kThreads) lds[k] = 0;
  if (threadIdx.x < K) sops[threadIdx.x] = U.ops[s0 + threadIdx.x];
  __syncthreads();

  Tiny st[R];
  Lanes L;
  L.li0 = (wv + 1) * RS + lane + 1;
  L.gr0 = r0 + wv;
  L.im0 = ((L.gr0 % 3) + 3) % 3;
  L.hsm = 0;
  L.interm = 0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int lr = wv + W * r, gr = r0 + lr;
    const bool in = col_in && gr >= 0 && gr < rows;
    const int li = L.li0 + r * W * RS;
    if (col_in && gr >= 1 && gr <= rows - 2 && gc >= 1 && gc <= cols - 2) L.interm |= 1u << r;
    if (in) {
#pragma unroll
      for (int ch = 0; ch < NC; ch++)
#pragma unroll
        for (int k = 0; k < NCOL; k++)
          lds[(ch * NCOL + k) * PL + li] = ((const Cell*)U.plane[ch][k])[(long)gr * cols + gc];
    }
    const bool hs = in && has_stream<MODEL>(gr, gc, rows, cols);
    if (hs) L.hsm |= 1u << r;
    st[r] = Tiny{0, 0, 0, 0, 0, 0, 0};
    if (hs) {
      const long id = stream_id<MODEL>(gr, gc, rows, cols);
      st[r].mat1 = U.param[id];
      st[r].mat2 = U.param[U.pstride + id];
      st[r].tmat = U.param[2 * U.pstride + id];
      if (init) {
        tiny_init(st[r], seed + (uint32_t)id);
      } else {
        st[r].s0 = U.state[id];
        st[r].s1 = U.state[U.nstream + id];
        st[r].s2 = U.state[2 * U.nstream + id];
        st[r].s3 = U.state[3 * U.nstream + id];
      }
    }
  }
  __syncthreads();

  // A CFTP pair whose two chains hold the same cells over the whole loaded
  // region (coalesced there, or outside the domain) ends the window with
  // the same cells too: every step is a function of the loaded cells and
  // the shared draws.  Such a tile runs one chain and writes it to both.
  // (Not in the 8-row shape: the second loop body would push it past 128
  // VGPRs, i.e. below two workgroups per CU.)
  constexpr bool kOne = NC == 2 && R <= 4;
  bool one = false;
  if (kOne) {
    constexpr int kWords = NCOL * PL * (int)sizeof(Cell) / 4;
    const uint32_t* w32 = (const uint32_t*)lds;
    int diff = 0;
    for (int k = threadIdx.x; k < kWords; k += kThreads) diff |= w32[k] != w32[kWords + k];
    one = !__syncthreads_or(diff);
  }
  if (kOne && one) run_window<MODEL, 1, W, R, PIPE, Cell, RS, PL>(lds, L, st, sops, K, U, mp);
  else run_window<MODEL, NC, W, R, PIPE, Cell, RS, PL>(lds, L, st, sops, K, U, mp);

#pragma unroll
  for (int r = 0; r < R; r++) {
    const int lr = wv + W * r, gr = r0 + lr;
    const bool inner = col_in && gr >= 0 && gr < rows && lr >= H && lr < H + TH && lane >= HC &&
                       lane < HC + TW;
    if (!inner) continue;
#pragma unroll
    for (int ch = 0; ch < NC; ch++)
#pragma unroll
      for (int k = 0; k < NCOL; k++)
        ((Cell*)U.plane_b[ch][k])[(long)gr * cols + gc] =
            lds[((one ? 0 : ch) * NCOL + k) * PL + L.li0 + r * W * RS];
    if (((L.hsm >> r) & 1u) && !U.no_state_write) {
      const long id = stream_id<MODEL>(gr, gc, rows, cols);
      U.state_b[id] = st[r].s0;
      U.state_b[U.nstream + id] = st[r].s1;
      U.state_b[2 * U.nstream + id] = st[r].s2;
      U.state_b[3 * U.nstream + id] = st[r].s3;
    }
  }
}

// Four waves per SIMD: a 16-wave workgroup alone on its CU, up to 128
// VGPRs per lane.
template <int MODEL, int NC, int W, int R>
__global__ __launch_bounds__(64 * W, 4) void k_fused(const UnitDesc* __restrict__ units,
                                                    const WorkItem* __restrict__ work, int nwork,
                                                    int xcd, int H, ModelParams mp,
                                                    const WindowRec* __restrict__ wins) {
  fused_body<MODEL, NC, W, R, false>(units, work, nwork, xcd, H, mp, wins);
}

// 48-row chain-pair tiles (16 waves x 3 rows) at eight waves per SIMD: two
// workgroups per CU within 64 VGPRs, with no spills in the step loop (4
// rows per lane would spill ~25 VGPRs to scratch every step).  Pays once a
// launch has several tiles per CU.
template <int MODEL>
__global__ __launch_bounds__(1024, 8) void k_fused_o8(const UnitDesc* __restrict__ units,
                                                     const WorkItem* __restrict__ work, int nwork,
                                                     int xcd, int H, ModelParams mp,
                                                     const WindowRec* __restrict__ wins) {
  fused_body<MODEL, 2, 16, 3, true>(units, work, nwork, xcd, H, mp, wins);
}

// The same tiles for one-chain jobs (walks, batched walks).
template <int MODEL>
__global__ __launch_bounds__(1024, 8) void k_walk_o8(const UnitDesc* __restrict__ units,
                                                    const WorkItem* __restrict__ work, int nwork,
                                                    int xcd, int H, ModelParams mp,
                                                    const WindowRec* __restrict__ wins) {
  fused_body<MODEL, 1, 16, 3, true>(units, work, nwork, xcd, H, mp, wins);
}

}  // namespace rt

#include "swar_lozenge.hpp"
#include "pair_lozenge.hpp"
#include "pair_domino.hpp"

namespace rt {

// ------------------------------------------------------------- host side
// Lozenge pull state: two bits per colour y, bit 2y set when colour y-1
// (mod 3) rotated since y's last rotation (y refreshes bits {2, 4, 32} from
// it: OP_PULL_A), bit 2y+1 when colour y+1 did (OP_PULL_B); the table maps
// (state, colour) at index 4 * state + colour to the op byte and the next
// state.  Equal to comparing the colours' last rotation steps, which is
// what `last` keeps between calls.
namespace {
struct LozengeOps {
  uint8_t op[256], next[256];
  LozengeOps() {
    for (int st = 0; st < 64; st++)
      for (int c = 0; c < 3; c++) {
        const int k = 4 * st + c, b = st >> (2 * c);
        op[k] = (uint8_t)(c | ((b & 1) ? OP_PULL_A : 0) | ((b & 2) ? OP_PULL_B : 0));
        int ns = st & ~(3 << (2 * c));
        ns |= 1 << (2 * ((c + 1) % 3));
        ns |= 2 << (2 * ((c + 2) % 3));
        next[k] = (uint8_t)ns;
      }
  }
};
}  // namespace

void OpBuilder::append(const std::vector<uint8_t>& colours, std::vector<uint8_t>& ops) {
  const size_t n = colours.size(), base = ops.size();
  if (n == 0) return;
  ops.resize(base + n);
  uint8_t* o = ops.data() + base;
  const uint8_t* c = colours.data();
  if (model == M_LOZENGE) {
    static const LozengeOps T;
    int st = 0;
    for (int y = 0; y < 3; y++) {
      if (last[(y + 2) % 3] > last[y]) st |= 1 << (2 * y);
      if (last[(y + 1) % 3] > last[y]) st |= 2 << (2 * y);
    }
    for (size_t i = 0; i < n; i++) {
      const int k = 4 * st + c[i];
      o[i] = T.op[k];
      st = T.next[k];
    }
    int found = 0;  // each colour's last rotation in this segment
    for (size_t i = n; i-- > 0 && found != 7;)
      if (!(found & (1 << c[i]))) {
        found |= 1 << c[i];
        last[c[i]] = g + (long)i;
      }
  } else {
    // domino: pull whenever the colour changes; rectangle-triangle also
    // records the previous colour in the op's high bits
    int p = prev;
    for (size_t i = 0; i < n; i++) {
      const int ci = c[i];
      uint8_t op = (uint8_t)ci;
      if (p >= 0 && p != ci) op |= model == M_DOMINO ? OP_PULL_A : (uint8_t)(OP_PULL_A | (p << 4));
      o[i] = op;
      p = ci;
    }
  }
  prev = c[n - 1];
  g += (long)n;
}

// Tile shapes: waves W, rows per lane R (tile height W*R), halo H.  Byte
// cells use 64-row tiles (8 x 8 or 16 x 4, same tiles, same halo) or 32-row
// tiles (16 x 2); int32 rect-triangle pairs use 32-row tiles (8 x 4, H = 4)
// to keep two chains' three planes within 64 KB of LDS.
struct Geo {
  int W, R, H;
};
// SH_16x3_O8: 48-row tiles at eight waves per SIMD (k_fused_o8 for chain
// pairs, k_walk_o8 for one chain).
enum TileShape { SH_AUTO = 0, SH_8x8 = 1, SH_16x4 = 2, SH_16x2 = 3, SH_RT8x4 = 4, SH_16x3_O8 = 7 };

static Geo shape_geo(int shape) {
  switch (shape) {
    case SH_16x4: return Geo{16, 4, 6};
    case SH_16x3_O8: return Geo{16, 3, 6};
    case SH_16x2: return Geo{16, 2, 6};
    case SH_RT8x4: return Geo{8, 4, 4};
    default: return Geo{8, 8, 6};
  }
}

// Loaded tile of a launch: LH x LW cells, a halo of Hr rows and Hc columns.
struct TileGeo {
  int LH, LW, Hr, Hc;
};

// Lozenge chain pairs on tile shape 7 run k_pair_lozenge<32>: 96 x 32
// tiles (mirrored row bands), a halo of H rows and H / 3 columns, H = 9
// unless rt_set_halo chose another.  Per launch window set on the side-300
// hexagon CFTP (6 doubling rounds, kernel time), against the best of the
// other shapes at H = 6 (autotuned): 32 samples 516 against 632 ms, 16
// samples 316 / 430, 8 samples 179 / 229, 4 samples 108 / 135, one sample
// 93 / 88; the 48 x 64 form of the same kernel (k_pair_lozenge<64>, same
// column halo) 643 ms at 32 samples and H = 9, 671 at H = 6.  A taller
// tile needs fewer of its rows as halo and, with the narrow column halo,
// a deeper row halo costs little, so windows are half again as long and
// the per-tile-load cost (HBM traffic, tile set-up) falls by a third.
constexpr int kPairLW = 32;
// Default halo of the packed-pair kernels: lozenge 9 (above), domino 8
// (k_pair_domino, Aztec diamond of order 300, 5 doubling rounds, kernel ms
// at H = 6 / 8 / 10: 8 samples 88 / 83 / 81, 32 samples 309 / 304 / 308;
// against 620 ms for the byte-cell k_fused_o8<0> at 32 samples and 140
// against 69 at 4).
static int pair_halo(int model) { return model == M_DOMINO ? 8 : 9; }

// Launches that run the packed kernels on 96 x 32 tiles, on tile shape 7:
// lozenge units (chain pairs and one-chain walks) with a halo divisible by
// 3 on the byte-cell engine (k_pair_lozenge); domino units with an even
// halo (k_pair_domino), chain pairs (CFTP pairs: every CFTP entry point and
// the server take heat-bath rules only) and one-chain walks whose rule is a
// heat-bath one, the only moves that kernel has (rule: the job's, or -1
// when its units differ).
// The tile geometry and the kernel a launch function picks both follow it.
static bool packed_pair(const rt_ctx* ctx, int model, int nc, int shape, int H, int rule) {
  if (shape != SH_16x3_O8) return false;
  if (model == M_LOZENGE) return H % 3 == 0 && ctx->engine != RT_ENGINE_FUSED_SWAR;
  if (model != M_DOMINO || H % 2 != 0) return false;
  return nc == 2 || (nc == 1 && (rule == D_HEATBATH || rule == D_HEATBATH2P));
}

// A lone small lozenge job (64-row tiles for under a third of the CUs)
// does better on the rule of thumb / autotuner at H = 6 (one hexagon-300
// sample: 88 against 93 ms; hexagon 64..160, 1 or 3 samples, 4 doubling
// rounds from 1024: 20 / 22 ms), so k_pair_lozenge takes over from that
// size up.  Domino units take k_pair_domino at H = 8 at every size:
// one-chain heat-bath walks (us per step against the tuned choice at
// H = 6: Aztec order 100 1.16 / 1.31, 300 1.17 / 1.31, 512 1.79 / 2.26)
// and chain pairs (one Aztec sample, 4 doubling rounds from 1024, ms
// against the same kernel at H = 6: order 64 18.1 / 20.5, 256 19.0 / 21.7).
static bool pair_job_is_large(const rt_ctx* ctx, long tiles64) { return 3 * tiles64 > ctx->num_cu; }

static TileGeo tile_geo(const rt_ctx* ctx, int model, int nc, int shape, int H, int rule) {
  if (packed_pair(ctx, model, nc, shape, H, rule))
    return TileGeo{3072 / kPairLW, kPairLW, H, col_halo(model, H)};
  const Geo g = shape_geo(shape);
  if (model == M_LOZENGE && ctx->engine == RT_ENGINE_FUSED_SWAR) return TileGeo{g.W * g.R, kLW, H, H};
  return TileGeo{g.W * g.R, kLW, H, col_halo(model, H)};
}

static void tile_counts(const UnitDesc& u, const TileGeo& g, int* tr, int* tc) {
  const int TH = g.LH - 2 * g.Hr, TW = g.LW - 2 * g.Hc;
  *tr = (u.rows + TH - 1) / TH;
  *tc = (u.cols + TW - 1) / TW;
}

static long tiles64(const FusedJob& job) {
  long t64 = 0;
  for (const UnitDesc& u : job.units) {
    int tr, tc;
    tile_counts(u, TileGeo{64, kLW, 6, 6}, &tr, &tc);
    t64 += (long)tr * tc;
  }
  return t64;
}

// Rule of thumb for a walk / batch job's shape.  Few tiles (a sample or
// chain running alone) means idle CUs and a launch bound by each wave's
// chain of dependent LDS round trips: 16 waves halve that chain (-30 % per
// launch for one side-300 hexagon CFTP sample), and 32-row tiles halve it
// again while putting 2.6x more workgroups on the chip, at a higher halo
// share.  With several tiles per CU the 8-wave shape's ILP per wave wins
// (-12 % per launch at 32 samples).
static int rule_of_thumb(const rt_ctx* ctx, long t64) {
  if (3 * t64 <= ctx->num_cu) return SH_16x2;
  if (t64 <= ctx->num_cu) return SH_16x4;
  return SH_8x8;
}

// Jobs of one kind share a tuned shape: model, chains per unit and the
// 64-row tile count to a factor of two.
static uint32_t tune_key(const FusedJob& job, long t64) {
  int lg = 0;
  while (lg < 40 && (2L << lg) <= t64) lg++;
  return (uint32_t)job.model | ((uint32_t)job.nc << 4) | ((uint32_t)lg << 8);
}

constexpr int kTrialWindows = 4;     // launches per timed trial
// Trials per candidate shape, interleaved; the first round of trials is
// not counted: a kernel's first launch can pay one-off costs (the scratch
// setup of a kernel that spills), which would bias a 4-launch trial.
constexpr int kTrialReps = 4;
struct LaunchShape {
  int shape;
  bool split;
};
// Candidates: the first tune_count(job) of the job's list.
static const LaunchShape kTunePairs[6] = {{SH_8x8, false},  {SH_16x4, false},
                                          {SH_16x2, false}, {SH_16x4, true},
                                          {SH_16x2, true},  {SH_16x3_O8, false}};
static const LaunchShape kTuneOne[4] = {
    {SH_8x8, false}, {SH_16x4, false}, {SH_16x2, false}, {SH_16x3_O8, false}};
static int tune_count(const FusedJob& job) { return job.nc == 2 ? 6 : 4; }
static LaunchShape tune_shape(const FusedJob& job, int c) {
  return job.nc == 2 ? kTunePairs[c] : kTuneOne[c];
}
// A job is tuned when at least half its launches run on the chosen shape
// (else the rule of thumb): 194 launches for chain pairs, 98 for one chain.
static size_t tune_min_windows(const FusedJob& job) {
  return 2 * (1 + (size_t)kTrialReps * tune_count(job) * kTrialWindows);
}

static bool tunable(const rt_ctx* ctx, const FusedJob& job) {
  return ctx->autotune && ctx->tile_shape == SH_AUTO && job.model != M_RECTTRI &&
         !(job.model == M_LOZENGE && ctx->engine == RT_ENGINE_FUSED_SWAR);
}

static void choose_shape(const rt_ctx* ctx, FusedJob& job) {
  job.tune = false;
  job.split = false;
  job.shape_source = 0;
  if (job.model == M_RECTTRI) {
    job.shape = job.nc == 2 ? SH_RT8x4 : SH_8x8;
    return;
  }
  if (ctx->tile_shape == SH_16x3_O8) {  // byte-cell kernels only (no 48-row SWAR kernel)
    const bool swar = job.model == M_LOZENGE && ctx->engine == RT_ENGINE_FUSED_SWAR;
    job.shape = swar ? SH_16x4 : SH_16x3_O8;
    return;
  }
  if (ctx->tile_shape == SH_AUTO && job.halo == pair_halo(job.model) &&
      packed_pair(ctx, job.model, job.nc, SH_16x3_O8, job.halo, job.mp.rule)) {
    job.shape = SH_16x3_O8;  // fastest at every batch size measured: no trials
    return;
  }
  if (ctx->tile_shape != SH_AUTO) {  // 1..3, or 4..6 = the same with pairs split
    job.shape = ctx->tile_shape > 3 ? ctx->tile_shape - 3 : ctx->tile_shape;
    job.split = ctx->tile_shape > 3 && job.nc == 2;
    return;
  }
  const long t64 = tiles64(job);
  job.shape = rule_of_thumb(ctx, t64);
  if (!tunable(ctx, job)) return;
  auto it = ctx->tuned.find(tune_key(job, t64));
  if (it != ctx->tuned.end()) {
    job.shape = it->second & 15;
    job.split = job.nc == 2 && (it->second >> 4) != 0;
    job.shape_source = 2;
    return;
  }
  job.tune = job.nlaunch >= tune_min_windows(job);
}

// Launch windows of each unit: at most kKMax steps and at most H of the
// unit's pull steps each, a new window at every segment start (TinyMT
// re-seeding).
static void cut_windows(FusedJob& job, int H) {
  long T = 0;
  for (long s : job.seg_steps) T += s;
  job.windows.assign(job.units.size(), {});
  job.nlaunch = 0;
  if (job.units.empty() || T == 0) return;
  parallel_for((int)job.units.size(), [&](int u) {
    const std::vector<uint8_t>& o = job.ops[u];
    std::vector<FusedWindow>& win = job.windows[u];
    long s0 = 0;
    for (size_t k = 0; k < job.seg_steps.size(); k++) {
      const long end = std::min<long>(s0 + job.seg_steps[k], (long)o.size());
      bool first = true;
      while (s0 < end) {
        int K = 0, pulls = 0;
        while (s0 + K < end && K < kKMax) {
          const int p1 = (o[s0 + K] & (OP_PULL_A | OP_PULL_B)) ? 1 : 0;
          if (pulls + p1 > H) break;
          pulls += p1;
          K++;
        }
        win.push_back(FusedWindow{s0, K, first ? 1 : 0, (int)k});
        first = false;
        s0 += K;
      }
    }
  });
  for (const std::vector<FusedWindow>& win : job.windows)
    job.nlaunch = std::max(job.nlaunch, win.size());
}

static size_t tiles_bytes(const rt_ctx* ctx, const FusedJob& job, int shape, bool split) {
  const TileGeo g = tile_geo(ctx, job.model, split ? 1 : job.nc, shape, job.halo, job.mp.rule);
  const size_t per = split ? 2 : 1;
  size_t work = 0;
  for (const UnitDesc& u : job.units) {
    int tr, tc;
    tile_counts(u, g, &tr, &tc);
    work += (size_t)tr * tc * per;
  }
  size_t b = ((job.units.size() * per * sizeof(UnitDesc) + 255) & ~(size_t)255) +
             ((work * sizeof(WorkItem) + 255) & ~(size_t)255);
  if (split) b += (job.nlaunch * job.units.size() * 2 * sizeof(WindowRec) + 255) & ~(size_t)255;
  return b;
}

// Second buffer set of a unit: every colour plane of every chain, and the
// TinyMT states.
static size_t plane_bytes(const FusedJob& job, const UnitDesc& u) {
  return (size_t)u.rows * u.cols * (job.model == M_RECTTRI ? 4 : 1);
}
static size_t second_set_bytes(const FusedJob& job, const UnitDesc& u) {
  const int ncol = job.model == M_DOMINO ? 2 : 3;
  return (size_t)job.nc * ncol * ((plane_bytes(job, u) + 255) & ~(size_t)255) +
         ((16 * (size_t)u.nstream + 255) & ~(size_t)255);
}

// Copy-back of a unit whose last window wrote the second set.
struct CopyDesc {
  const void* src;
  void* dst;
  long bytes;
};

__global__ void k_copy_list(const CopyDesc* __restrict__ list, unsigned n) {
  const long stride = (long)gridDim.x * blockDim.x;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (unsigned k = blockIdx.y; k < n; k += gridDim.y) {
    const CopyDesc d = list[k];
    if ((((uintptr_t)d.src | (uintptr_t)d.dst | (uintptr_t)d.bytes) & 15) == 0) {
      const uint4* s = (const uint4*)d.src;
      uint4* o = (uint4*)d.dst;
      for (long i = t; i < d.bytes / 16; i += stride) o[i] = s[i];
    } else {
      const uint8_t* s = (const uint8_t*)d.src;
      uint8_t* o = (uint8_t*)d.dst;
      for (long i = t; i < d.bytes; i += stride) o[i] = s[i];
    }
  }
}

// Halo of a job's tiles: the context's setting, except rect-triangle pairs
// (32-row int32 tiles, H = 4) and, unless rt_set_halo was called, lozenge
// and domino chain pairs that may run the packed-pair kernels (H = 9 and
// 8).  Every shape of a job has the same halo, so one window list serves
// all.
static void plan_windows(rt_ctx* ctx, FusedJob& job) {
  if (job.model == M_RECTTRI && job.nc == 2)
    job.halo = shape_geo(SH_RT8x4).H;
  else if (!ctx->halo_set && (ctx->tile_shape == SH_AUTO || ctx->tile_shape == SH_16x3_O8) &&
           packed_pair(ctx, job.model, job.nc, SH_16x3_O8, pair_halo(job.model), job.mp.rule) &&
           (ctx->tile_shape == SH_16x3_O8 || pair_job_is_large(ctx, tiles64(job)) ||
            job.model == M_DOMINO))
    job.halo = pair_halo(job.model);
  else
    job.halo = ctx->halo;
  cut_windows(job, job.halo);
  choose_shape(ctx, job);
}

size_t fused_scratch_bytes(rt_ctx* ctx, FusedJob& job) {
  plan_windows(ctx, job);
  long T = 0;
  for (long s : job.seg_steps) T += s;
  const size_t nu = job.units.size();
  size_t bytes = ((nu * (size_t)T + 255) & ~(size_t)255) +
                 ((job.nlaunch * nu * sizeof(WindowRec) + 255) & ~(size_t)255) + 1024;
  if (job.tune)
    for (int c = 0; c < tune_count(job); c++)
      bytes += tiles_bytes(ctx, job, tune_shape(job, c).shape, tune_shape(job, c).split);
  else
    bytes += tiles_bytes(ctx, job, job.shape, job.split);
  for (const UnitDesc& u : job.units) bytes += second_set_bytes(job, u);
  bytes += 2 * (((nu * 7 * sizeof(CopyDesc)) + 255) & ~(size_t)255);
  return bytes;
}

// XCD order for launches with no more workgroups than CUs (one workgroup
// per CU): per launch on the side-300 hexagon CFTP it takes a lone sample
// from 15.3 to 13.3 us and 4 samples (64-row tiles) from 33.2 to 30.1 us;
// with more workgroups than CUs it was as often slower (+2..5 %) as faster.
static int xcd_of(const rt_ctx* ctx, unsigned nwork) {
  return ctx->xcd_order == 2 || (ctx->xcd_order == 1 && nwork <= (unsigned)ctx->num_cu) ? 1 : 0;
}
static unsigned grid_of(int xcd, unsigned nwork) { return xcd ? (nwork + 7u) & ~7u : nwork; }

template <int MODEL, int NC, int W, int R>
static void launch(rt_ctx* ctx, hipStream_t st, unsigned nblocks, const UnitDesc* units,
                   const WorkItem* work, int H, const ModelParams& mp, const WindowRec* wins) {
  const int xcd = xcd_of(ctx, nblocks);
  hipLaunchKernelGGL((k_fused<MODEL, NC, W, R>), dim3(grid_of(xcd, nblocks)), dim3(64 * W), 0,
                     st, units, work, (int)nblocks, xcd, H, mp, wins);
}

typedef void (*LaunchFn)(rt_ctx*, hipStream_t, unsigned, const UnitDesc*, const WorkItem*, int,
                         const ModelParams&, const WindowRec*);

template <int MODEL, int NC>
static void launch_o8(rt_ctx* ctx, hipStream_t st, unsigned nblocks, const UnitDesc* units,
                      const WorkItem* work, int H, const ModelParams& mp, const WindowRec* wins) {
  const int xcd = xcd_of(ctx, nblocks);
  if (NC == 2)
    hipLaunchKernelGGL((k_fused_o8<MODEL>), dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
  else
    hipLaunchKernelGGL((k_walk_o8<MODEL>), dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
}

// Lozenge chain pairs on tile shape 7: both chains packed into one LDS word
// per cell (k_pair_lozenge) when the halo is a multiple of 3, which that
// kernel's compile-time row residues need; k_fused_o8<1> otherwise.
// XCD order pays for every grid of k_pair_lozenge, whose tile loads
// re-read their neighbours' halo: 6 doubling rounds of the side-300 hexagon
// CFTP, kernel ms plain -> XCD order, one stream: 8 samples 179 -> 171, 16
// samples 316 -> 290, 32 samples 516 -> 516; two streams: 8 samples 140 ->
// 139, 16 samples 247 -> 231, 32 samples 471 -> 451.  (The 48 x 64
// byte-cell kernel k_fused_o8<1> keeps the rule of the other shapes above
// four workgroups per CU, where it was measured.)
template <int NC>
static void launch_pair_lozenge(rt_ctx* ctx, hipStream_t st, unsigned nblocks,
                                const UnitDesc* units, const WorkItem* work, int H,
                                const ModelParams& mp, const WindowRec* wins) {
  const bool pair = packed_pair(ctx, M_LOZENGE, NC, SH_16x3_O8, H, mp.rule);
  const int xcd =
      xcd_of(ctx, nblocks) ||
              (ctx->xcd_order == 1 && (pair || (NC == 2 && nblocks > 4u * (unsigned)ctx->num_cu)))
          ? 1
          : 0;
  if (pair)
    hipLaunchKernelGGL(k_pair_lozenge<kPairLW>, dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
  else if (NC == 2)
    hipLaunchKernelGGL((k_fused_o8<M_LOZENGE>), dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
  else
    hipLaunchKernelGGL((k_walk_o8<M_LOZENGE>), dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
}

// Domino chain pairs on tile shape 7: k_pair_domino when the halo is even
// (its row parities), k_fused_o8<0> otherwise (the same test as
// packed_pair, which fixes the tiles).
template <int NC>
static void launch_pair_domino(rt_ctx* ctx, hipStream_t st, unsigned nblocks,
                               const UnitDesc* units, const WorkItem* work, int H,
                               const ModelParams& mp, const WindowRec* wins) {
  const bool pair = packed_pair(ctx, M_DOMINO, NC, SH_16x3_O8, H, mp.rule);
  const int xcd = xcd_of(ctx, nblocks) || (ctx->xcd_order == 1 && pair) ? 1 : 0;
  if (pair)
    hipLaunchKernelGGL(k_pair_domino, dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st, units,
                       work, (int)nblocks, xcd, H, mp, wins);
  else if (NC == 2)
    hipLaunchKernelGGL((k_fused_o8<M_DOMINO>), dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
  else
    hipLaunchKernelGGL((k_walk_o8<M_DOMINO>), dim3(grid_of(xcd, nblocks)), dim3(1024), 0, st,
                       units, work, (int)nblocks, xcd, H, mp, wins);
}

template <int MODEL, int NC>
static LaunchFn pick_byte(int shape) {
  if (shape == SH_16x3_O8 && MODEL == M_LOZENGE) return launch_pair_lozenge<NC>;
  if (shape == SH_16x3_O8 && MODEL == M_DOMINO) return launch_pair_domino<NC>;
  if (shape == SH_16x3_O8) return launch_o8<MODEL, NC>;
  if (shape == SH_16x4) return launch<MODEL, NC, 16, 4>;
  if (shape == SH_16x2) return launch<MODEL, NC, 16, 2>;
  return launch<MODEL, NC, 8, 8>;
}

template <int NC, int W, int R>
static void launch_swar(rt_ctx* ctx, hipStream_t st, unsigned nblocks, const UnitDesc* units,
                        const WorkItem* work, int H, const ModelParams& mp,
                        const WindowRec* wins) {
  const int xcd = xcd_of(ctx, nblocks);
  hipLaunchKernelGGL((k_swar_lozenge<NC, W, R>), dim3(grid_of(xcd, nblocks)), dim3(64 * W), 0,
                     st, units, work, (int)nblocks, xcd, H, mp, wins);
}

// Lozenge: four cells per lane-word (k_swar_lozenge), same tiles: 64-row
// tiles by 16 waves x 1 word row, 32-row tiles by 8 waves x 1.
template <int NC>
static LaunchFn pick_swar(int shape) {
  if (shape == SH_16x2) return launch_swar<NC, 8, 1>;
  if (shape == SH_16x3_O8) return nullptr;  // no 48-row SWAR kernel (choose_shape never asks)
  return launch_swar<NC, 16, 1>;
}

static LaunchFn pick(const rt_ctx* ctx, int model, int nc, int shape) {
  if (model == M_DOMINO) return nc == 1 ? pick_byte<M_DOMINO, 1>(shape) : pick_byte<M_DOMINO, 2>(shape);
  if (model == M_LOZENGE && ctx->engine == RT_ENGINE_FUSED_SWAR)
    return nc == 1 ? pick_swar<1>(shape) : pick_swar<2>(shape);
  if (model == M_LOZENGE)
    return nc == 1 ? pick_byte<M_LOZENGE, 1>(shape) : pick_byte<M_LOZENGE, 2>(shape);
  return nc == 1 ? launch<M_RECTTRI, 1, 8, 8> : launch<M_RECTTRI, 2, 8, 4>;
}

std::vector<int> fused_live_spans(const uint8_t* const* planes, int nplanes, int rows,
                                  int cols) {
  std::vector<int> sp(2 * (size_t)rows);
  for (int r = 0; r < rows; r++) {
    int lo = cols, hi = -1;
    for (int k = 0; k < nplanes; k++) {
      const uint8_t* row = planes[k] + (size_t)r * cols;
      int a = 0, b = cols - 1;
      while (a < cols && !row[a]) a++;
      while (b > a && !row[b]) b--;
      if (a < cols) {
        lo = std::min(lo, a);
        hi = std::max(hi, b);
      }
    }
    sp[2 * r] = lo;
    sp[2 * r + 1] = hi;
  }
  return sp;
}

// Work items of one unit under geometry g (fused_unit_work), `copies`
// consecutive unit indices per tile (split pairs); cells += the cells the
// items' interiors cover inside the grid.
static void unit_work(const std::vector<int>* sp, const UnitDesc& U, const TileGeo& g, int unit,
                      int copies, std::vector<WorkItem>& out, long* cells) {
  const int TH = g.LH - 2 * g.Hr, TW = g.LW - 2 * g.Hc;
  int tr, tc;
  tile_counts(U, g, &tr, &tc);
  for (int r = 0; r < tr; r++) {
    const int ra = r * TH, rb = std::min(ra + TH, U.rows);  // interior rows [ra, rb)
    int lo = 0, hi = U.cols - 1;
    if (sp) {
      lo = U.cols;
      hi = -1;
      for (int i = std::max(ra - 1, 0); i <= std::min(rb, U.rows - 1); i++)
        if ((*sp)[2 * i] <= (*sp)[2 * i + 1]) {
          lo = std::min(lo, (*sp)[2 * i] - 1);
          hi = std::max(hi, (*sp)[2 * i + 1] + 1);
        }
      lo = std::max(lo, 0);
      hi = std::min(hi, U.cols - 1);
      if (lo > hi) continue;  // no cell of this row band can turn nonzero
    }
    const int n = (hi - lo + TW) / TW;  // <= tc
    for (int k = 0; k < n; k++) {
      for (int c = 0; c < copies; c++) out.push_back(WorkItem{unit + c, r * tc + k, lo});
      *cells += (long)(rb - ra) * (std::min(lo + (k + 1) * TW, U.cols) - (lo + k * TW));
    }
  }
}

// Tiles may be left out (FusedJob::spans): domino and byte-cell lozenge
// units that carry spans.
static bool skips_tiles(const rt_ctx* ctx, const FusedJob& job, size_t u) {
  return (job.model == M_DOMINO ||
          (job.model == M_LOZENGE && ctx->engine != RT_ENGINE_FUSED_SWAR)) &&
         u < job.spans.size() && (int)job.spans[u].size() == 2 * job.units[u].rows;
}

// Per launch, the sites its windows compute under the tiles t.
static void set_launch_sites(const FusedJob& job, const FusedPlan::Tiles& t,
                             std::vector<long>& out) {
  out.assign(job.nlaunch, 0);
  for (size_t u = 0; u < job.units.size(); u++)
    for (size_t l = 0; l < job.windows[u].size(); l++) out[l] += t.unit_sites[u];
}
static long sites_of(const FusedJob& job, const FusedPlan::Tiles& t, size_t l0, size_t l1) {
  long n = 0;
  for (size_t u = 0; u < job.units.size(); u++) {
    const size_t e = std::min(l1, job.windows[u].size());
    if (e > l0) n += (long)(e - l0) * t.unit_sites[u];
  }
  return n;
}

// Unit table, work list (and, split, window records) of one launch shape,
// at p (advanced past them).  `wins` are the job's [launch][unit] records.
static int upload_tiles(rt_ctx* ctx, const FusedJob& job, LaunchShape ls, uint8_t*& p,
                        const WindowRec* d_wins, const std::vector<WindowRec>& wins,
                        FusedPlan::Tiles* out) {
  const TileGeo geo =
      tile_geo(ctx, job.model, ls.split ? 1 : job.nc, ls.shape, job.halo, job.mp.rule);
  const size_t nu = job.units.size();
  std::vector<UnitDesc> units;
  if (ls.split) {
    // unit 2u = chain 0 of pair u (writes the streams), 2u+1 = chain 1
    for (size_t u = 0; u < nu; u++)
      for (int ch = 0; ch < 2; ch++) {
        UnitDesc v = job.units[u];
        for (int k = 0; k < 3; k++) {
          v.plane[0][k] = job.units[u].plane[ch][k];
          v.plane_b[0][k] = job.units[u].plane_b[ch][k];
          v.plane[1][k] = v.plane_b[1][k] = nullptr;
        }
        v.no_state_write = ch;
        units.push_back(v);
      }
  } else {
    units = job.units;
  }
  std::vector<WorkItem> work;
  std::vector<size_t> unit_end;  // work-list size after each (pair) unit
  std::vector<long> unit_sites(nu);
  for (size_t u = 0; u < units.size(); u += ls.split ? 2 : 1) {
    const size_t ju = ls.split ? u / 2 : u;  // the job's unit
    int tr, tc;
    tile_counts(units[u], geo, &tr, &tc);
    units[u].tiles_c = tc;
    if (ls.split) units[u + 1].tiles_c = tc;
    const bool skip = skips_tiles(ctx, job, ju);
    long cells = 0;
    // Split pairs: the two chains' workgroups of a tile sit next to each
    // other in the list, so they land in the same XCD share and the second
    // one finds the shared TinyMT states in L2.
    unit_work(skip ? &job.spans[ju] : nullptr, units[u], geo, (int)u, ls.split ? 2 : 1, work,
              &cells);
    const long all = (long)units[u].rows * units[u].cols;
    unit_sites[ju] = skip ? (long)((double)job.units[ju].nstream * cells / all + 0.5)
                          : job.units[ju].nstream;
    unit_end.push_back(work.size());
  }

  UnitDesc* d_units = (UnitDesc*)p;
  p += (units.size() * sizeof(UnitDesc) + 255) & ~(size_t)255;
  WorkItem* d_work = (WorkItem*)p;
  p += (work.size() * sizeof(WorkItem) + 255) & ~(size_t)255;
  HIPCHK_RET(ctx, hipMemcpyAsync(d_units, units.data(), units.size() * sizeof(UnitDesc),
                                 hipMemcpyHostToDevice, ctx->stream));
  HIPCHK_RET(ctx, hipMemcpyAsync(d_work, work.data(), work.size() * sizeof(WorkItem),
                                 hipMemcpyHostToDevice, ctx->stream));
  std::vector<WindowRec> w2;
  if (ls.split) {
    const size_t L = job.nlaunch;
    w2.resize(L * 2 * nu);
    for (size_t l = 0; l < L; l++)
      for (size_t u = 0; u < nu; u++) w2[(l * nu + u) * 2] = w2[(l * nu + u) * 2 + 1] = wins[l * nu + u];
    WindowRec* d_w2 = (WindowRec*)p;
    p += (w2.size() * sizeof(WindowRec) + 255) & ~(size_t)255;
    HIPCHK_RET(ctx, hipMemcpyAsync(d_w2, w2.data(), w2.size() * sizeof(WindowRec),
                                   hipMemcpyHostToDevice, ctx->stream));
    d_wins = d_w2;
  }
  // Pageable host sources: the copies may still read them after the calls
  // return, so wait before the vectors go out of scope.
  HIPCHK_RET(ctx, hipStreamSynchronize(ctx->stream));
  *out = FusedPlan::Tiles{d_units, d_work, (unsigned)work.size(), ls.shape, ls.split, d_wins,
                          units.size(), std::vector<unsigned>(unit_end.begin(), unit_end.end()),
                          unit_sites};
  return RT_OK;
}

int fused_upload(rt_ctx* ctx, FusedJob& job, void* scratch, FusedPlan* plan) {
  if (job.shape == SH_AUTO) plan_windows(ctx, job);
  const size_t nu = job.units.size();
  *plan = FusedPlan();
  long T = 0;
  for (long s : job.seg_steps) T += s;
  plan->total_steps = T;
  for (const UnitDesc& u : job.units) plan->streams += u.nstream;
  if (nu == 0 || T == 0) return RT_OK;
  // Device layout of the scratch block: ops | window records | second
  // buffer sets | copy-back list | first-set copy list | (units | work) per
  // tile shape.
  uint8_t* p = (uint8_t*)scratch;
  uint8_t* d_ops = p;
  p += (nu * (size_t)T + 255) & ~(size_t)255;
  WindowRec* d_wins = (WindowRec*)p;
  const size_t L = job.nlaunch;
  p += (L * nu * sizeof(WindowRec) + 255) & ~(size_t)255;
  const int ncol = job.model == M_DOMINO ? 2 : 3;
  std::vector<CopyDesc> copies, seed_b;
  for (size_t u = 0; u < nu; u++) {
    UnitDesc& U = job.units[u];
    const size_t pb = plane_bytes(job, U);
    for (int ch = 0; ch < job.nc; ch++)
      for (int k = 0; k < ncol; k++) {
        U.plane_b[ch][k] = p;
        p += (pb + 255) & ~(size_t)255;
        // the cells of left-out tiles are never written: both sets start
        // as the start planes (zero there)
        if (skips_tiles(ctx, job, u))
          seed_b.push_back(CopyDesc{U.plane[ch][k], U.plane_b[ch][k], (long)pb});
      }
    U.state_b = (uint32_t*)p;
    p += (16 * (size_t)U.nstream + 255) & ~(size_t)255;
    if (job.windows[u].size() % 2) {  // the last window writes the second set
      for (int ch = 0; ch < job.nc; ch++)
        for (int k = 0; k < ncol; k++)
          copies.push_back(CopyDesc{U.plane_b[ch][k], U.plane[ch][k], (long)pb});
      copies.push_back(CopyDesc{U.state_b, U.state, 16 * U.nstream});
    }
  }
  CopyDesc* d_copies = (CopyDesc*)p;
  p += ((nu * 7 * sizeof(CopyDesc)) + 255) & ~(size_t)255;
  CopyDesc* d_seed_b = (CopyDesc*)p;
  p += ((nu * 7 * sizeof(CopyDesc)) + 255) & ~(size_t)255;
  if (!seed_b.empty()) {
    HIPCHK_RET(ctx, hipMemcpyAsync(d_seed_b, seed_b.data(), seed_b.size() * sizeof(CopyDesc),
                                   hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_copy_list, dim3(64, std::min((unsigned)seed_b.size(), 4096u)), dim3(256),
                       0, ctx->stream, d_seed_b, (unsigned)seed_b.size());
    HIPCHK_RET(ctx, hipGetLastError());
  }
  std::vector<WindowRec> wins(L * nu, WindowRec{0, 0, 0, 0, 0});
  plan->launch_work.assign(L, 0);
  for (size_t u = 0; u < nu; u++) {
    job.units[u].ops = d_ops + u * (size_t)T;
    job.units[u].rule = job.mp.rule;
    job.units[u].dw = job.mp.dw;
    const std::vector<uint8_t>& o = job.ops[u];
    if ((long)o.size() != T) {
      ctx->err = "fused job: op stream length differs from the segment total";
      return RT_EINVAL;
    }
    HIPCHK_RET(ctx, hipMemcpyAsync(d_ops + u * (size_t)T, o.data(), T, hipMemcpyHostToDevice,
                                   ctx->stream));
    const std::vector<FusedWindow>& win = job.windows[u];
    for (size_t l = 0; l < win.size(); l++) {
      wins[l * nu + u] = WindowRec{win[l].s0, win[l].K, win[l].init,
                                   job.seg_seeds[u][win[l].seg], (int)(l & 1)};
      plan->launch_work[l] += (long)win[l].K * job.units[u].nstream;
    }
  }
  HIPCHK_RET(ctx, hipMemcpyAsync(d_wins, wins.data(), wins.size() * sizeof(WindowRec),
                                 hipMemcpyHostToDevice, ctx->stream));
  if (!copies.empty())
    HIPCHK_RET(ctx, hipMemcpyAsync(d_copies, copies.data(), copies.size() * sizeof(CopyDesc),
                                   hipMemcpyHostToDevice, ctx->stream));
  plan->d_copies = d_copies;
  plan->ncopies = (unsigned)copies.size();
  FusedPlan::Tiles t;
  const LaunchShape first{job.shape, job.split};
  int rc = upload_tiles(ctx, job, first, p, d_wins, wins, &t);  // waits for the copies above
  if (rc) return rc;
  plan->cand.push_back(t);
  if (job.tune)
    for (int c = 0; c < tune_count(job); c++) {
      const LaunchShape ls = tune_shape(job, c);
      if (ls.shape == first.shape && ls.split == first.split) continue;
      rc = upload_tiles(ctx, job, ls, p, d_wins, wins, &t);
      if (rc) return rc;
      plan->cand.push_back(t);
    }
  plan->d_units = plan->cand[0].d_units;
  plan->d_work = plan->cand[0].d_work;
  plan->nwork = plan->cand[0].nwork;
  plan->d_wins = plan->cand[0].d_wins;
  plan->nlaunch = L;
  set_launch_sites(job, plan->cand[0], plan->launch_sites);
  return RT_OK;
}

static void count_launches(rt_ctx* ctx, const FusedPlan& plan, size_t l0, size_t l1) {
  long work = 0;
  for (size_t l = l0; l < l1; l++) work += plan.launch_work[l];
  ctx->last.launches += (long)(l1 - l0);
  ctx->last.attempts += work;
  if (l1 == plan.nlaunch) ctx->last.steps += plan.total_steps;
}

int fused_tune(rt_ctx* ctx, FusedJob& job, FusedPlan* plan) {
  const size_t nc = plan->cand.size();
  const size_t need = 1 + (size_t)kTrialReps * nc * kTrialWindows;
  if (!job.tune || nc < 2 || plan->nlaunch < need) return RT_OK;
  for (const FusedPlan::Tiles& t : plan->cand)
    if (t.nwork == 0) return RT_OK;
  const int H = job.halo;
  // Launch 0 (TinyMT seeding) under the rule-of-thumb shape, then
  // kTrialReps rounds of kTrialWindows launches under each candidate;
  // event k + 1 closes launch group k.
  struct Trial {
    int cand;
    size_t l0, l1;
  };
  std::vector<Trial> trials{{0, 0, 1}};
  size_t l = 1;
  for (int r = 0; r < kTrialReps; r++)
    for (size_t c = 0; c < nc; c++, l += kTrialWindows)
      trials.push_back(Trial{(int)c, l, l + kTrialWindows});
  if (trials.size() + 1 > sizeof(ctx->evt) / sizeof(ctx->evt[0])) return RT_OK;
  HIPCHK_RET(ctx, hipEventRecord(ctx->evt[0], ctx->stream));
  for (size_t k = 0; k < trials.size(); k++) {
    const FusedPlan::Tiles& t = plan->cand[trials[k].cand];
    LaunchFn fn = pick(ctx, job.model, t.split ? 1 : job.nc, t.shape);
    if (!fn) {
      ctx->err = "no sweep-block kernel for this tile shape";
      return RT_EINVAL;
    }
    for (size_t x = trials[k].l0; x < trials[k].l1; x++)
      fn(ctx, ctx->stream, t.nwork, t.d_units, t.d_work, H, job.mp, t.d_wins + x * t.nunits);
    HIPCHK_RET(ctx, hipGetLastError());
    HIPCHK_RET(ctx, hipEventRecord(ctx->evt[k + 1], ctx->stream));
  }
  HIPCHK_RET(ctx, hipEventSynchronize(ctx->evt[trials.size()]));
  std::vector<double> ms(nc, 0.0), work(nc, 0.0);
  for (size_t k = 0; k < trials.size(); k++) {
    float e = 0;
    HIPCHK_RET(ctx, hipEventElapsedTime(&e, ctx->evt[k], ctx->evt[k + 1]));
    const int c = trials[k].cand, sl = shape_slot(plan->cand[c].shape, plan->cand[c].split);
    plan->trial_ms[sl] += e;
    plan->trial_launches[sl] += (long)(trials[k].l1 - trials[k].l0);
    plan->trial_sites[sl] += sites_of(job, plan->cand[c], trials[k].l0, trials[k].l1);
    if (k <= nc) continue;  // the seeding launch and the first round are not fair trials
    ms[c] += e;
    for (size_t x = trials[k].l0; x < trials[k].l1; x++) work[c] += (double)plan->launch_work[x];
  }
  size_t best = 0;
  std::vector<double> cost(nc);
  for (size_t c = 0; c < nc; c++) {
    cost[c] = work[c] > 0 ? ms[c] / work[c] : 1e30;
    if (cost[c] < cost[best]) best = c;
  }
  count_launches(ctx, *plan, 0, l);
  plan->first = l;
  job.shape = plan->cand[best].shape;
  job.split = plan->cand[best].split;
  job.shape_source = 1;
  std::swap(plan->cand[0], plan->cand[best]);
  plan->d_units = plan->cand[0].d_units;
  plan->d_work = plan->cand[0].d_work;
  plan->nwork = plan->cand[0].nwork;
  plan->d_wins = plan->cand[0].d_wins;
  set_launch_sites(job, plan->cand[0], plan->launch_sites);
  ctx->tuned[tune_key(job, tiles64(job))] = job.shape | (job.split ? 16 : 0);
  return RT_OK;
}

void fused_add_trials(const FusedPlan& plan, rt_timing* acc) {
  for (int s = 0; s < 8; s++) {
    acc->kernel_ms += plan.trial_ms[s];
    acc->kernel_launches += plan.trial_launches[s];
    acc->kernel_sites += plan.trial_sites[s];
    acc->shape_ms[s] += plan.trial_ms[s];
    acc->shape_launches[s] += plan.trial_launches[s];
    acc->shape_sites[s] += plan.trial_sites[s];
  }
}

int fused_launch(rt_ctx* ctx, const FusedJob& job, const FusedPlan& plan, bool split_streams) {
  ctx->last.shape = job.shape;
  ctx->last.shape_source = job.shape_source;
  ctx->last.split = job.split ? 1 : 0;
  if (plan.cand.empty()) return RT_OK;  // nothing to run
  const FusedPlan::Tiles& t = plan.cand[0];
  LaunchFn fn = pick(ctx, job.model, t.split ? 1 : job.nc, t.shape);
  if (!fn) {
    ctx->err = "no sweep-block kernel for this tile shape";
    return RT_EINVAL;
  }
  const int H = job.halo;
  // The units in two halves on two streams: a half's next window waits only
  // for its own last one, so one half's tile loads and stores (HBM-bound,
  // a third of a launch) overlap the other half's steps (VALU-bound) and
  // launch boundaries.  Kernel ms, 6 doubling rounds of the side-300
  // hexagon CFTP on k_pair_lozenge (XCD order), streams 1 / 2 / 3 / 4:
  // 8 samples 171 / 139 / 150 / 234, 16 samples 290 / 231 / 234 / 293, 32
  // samples 516 / 451 / 456 / 524.
  // parts: contiguous work-list ranges at unit boundaries, about equal
  std::vector<unsigned> cut{0};
  // (only grids of two workgroups per CU or more: below that a part's
  // launches are latency-bound and splitting only doubles the dispatches,
  // 8 order-24 two-periodic domino samples 2.2 -> 1.2 samples/s)
  const int want = split_streams && plan.nlaunch > plan.first + 1 &&
                           t.nwork >= 2u * (unsigned)ctx->num_cu
                       ? ctx->nstreams
                       : 1;
  for (int k = 1; k < want; k++) {
    const unsigned target = (unsigned)((size_t)t.nwork * k / want);
    unsigned best = 0;
    for (unsigned e : t.ends)
      if (e > cut.back() && e < t.nwork &&
          (best == 0 || std::labs((long)e - (long)target) < std::labs((long)best - (long)target)))
        best = e;
    if (best) cut.push_back(best);
  }
  cut.push_back(t.nwork);
  const int parts = (int)cut.size() - 1;
  if (parts > 1) {
    int rc = ctx->ensure_streams(parts);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->evj[0], ctx->stream));
    for (int k = 1; k < parts; k++)
      HIPCHK_RET(ctx, hipStreamWaitEvent(ctx->sx[k], ctx->evj[0], 0));
  }
  for (size_t l = plan.first; l < plan.nlaunch; l++) {
    const WindowRec* w = t.d_wins + l * t.nunits;
    for (int k = 0; k < parts; k++)
      if (cut[k + 1] > cut[k])  // no work items: every tile of the job stays zero
        fn(ctx, k ? ctx->sx[k] : ctx->stream, cut[k + 1] - cut[k], t.d_units, t.d_work + cut[k],
           H, job.mp, w);
  }
  for (int k = 1; k < parts; k++) {
    HIPCHK_RET(ctx, hipEventRecord(ctx->evj[k], ctx->sx[k]));
    HIPCHK_RET(ctx, hipStreamWaitEvent(ctx->stream, ctx->evj[k], 0));
  }
  if (plan.ncopies)
    hipLaunchKernelGGL(k_copy_list, dim3(64, std::min(plan.ncopies, 4096u)), dim3(256), 0,
                       ctx->stream, plan.d_copies, plan.ncopies);
  HIPCHK_RET(ctx, hipGetLastError());
  count_launches(ctx, plan, plan.first, plan.nlaunch);
  return RT_OK;
}

int fused_serve_shape(const rt_ctx* ctx, int model, int nc, int rule, long tiles64) {
  if (model == M_RECTTRI) return nc == 2 ? SH_RT8x4 : SH_8x8;
  // The packed kernels at every size: a lone small lozenge job is faster on
  // 64-row tiles (plan_windows), but the server's groups share the chip
  // with each other and run many windows of many requests (the default
  // served stream's burst capacity 3.31 -> 4.04 requests/s, p99 17.6 ->
  // 11.5 s, with lozenge pairs packed at every size too).
  if (packed_pair(ctx, model, nc, SH_16x3_O8, pair_halo(model), rule)) return SH_16x3_O8;
  // Few tiles: one 16-wave workgroup per CU on 64-row tiles; more tiles
  // than CUs: the 48-row tiles at eight waves per SIMD (two workgroups per
  // CU), which the batch tuner picks from about 16 hexagon-300 samples up.
  return tiles64 <= ctx->num_cu ? SH_16x4 : SH_16x3_O8;
}

int fused_launch_windows(rt_ctx* ctx, hipStream_t st, int model, int nc, int shape,
                         const ModelParams& mp, const UnitDesc* d_units, const WorkItem* d_work,
                         unsigned nwork, const WindowRec* d_wins) {
  if (nwork == 0) return RT_OK;
  LaunchFn fn = pick(ctx, model, nc, shape);
  if (!fn) {
    ctx->err = "no sweep-block kernel for this tile shape";
    return RT_EINVAL;
  }
  fn(ctx, st, nwork, d_units, d_work, fused_shape_halo(ctx, model, nc, mp.rule, shape), mp,
     d_wins);
  HIPCHK_RET(ctx, hipGetLastError());
  ctx->last.launches += 1;
  return RT_OK;
}

void fused_unit_work(const rt_ctx* ctx, int model, int nc, int rule, int shape,
                     const UnitDesc& u, const std::vector<int>& spans, int unit,
                     std::vector<WorkItem>& out) {
  const bool skip = (model == M_DOMINO ||
                     (model == M_LOZENGE && ctx->engine != RT_ENGINE_FUSED_SWAR)) &&
                    (int)spans.size() == 2 * u.rows;
  long cells = 0;
  unit_work(skip ? &spans : nullptr, u,
            tile_geo(ctx, model, nc, shape, fused_shape_halo(ctx, model, nc, rule, shape), rule),
            unit, 1, out, &cells);
}

int fused_shape_halo(const rt_ctx* ctx, int model, int nc, int rule, int shape) {
  return packed_pair(ctx, model, nc, shape, pair_halo(model), rule) ? pair_halo(model)
                                                                    : shape_geo(shape).H;
}

void fused_tiles(const rt_ctx* ctx, int model, int nc, int rule, int shape, int rows, int cols,
                 int* tiles_r, int* tiles_c) {
  UnitDesc u{};
  u.rows = rows;
  u.cols = cols;
  tile_counts(u,
              tile_geo(ctx, model, nc, shape, fused_shape_halo(ctx, model, nc, rule, shape), rule),
              tiles_r, tiles_c);
}

int fused_run(rt_ctx* ctx, FusedJob& job, void* scratch) {
  FusedPlan plan;  // job.shape was fixed when the scratch was sized
  int rc = fused_upload(ctx, job, scratch, &plan);
  if (!rc) rc = fused_tune(ctx, job, &plan);
  if (rc) return rc;
  return fused_launch(ctx, job, plan, false);
}

int fused_sync(rt_ctx* ctx, const FusedJob& job) {
  for (size_t u = 0; u < job.units.size(); u++) {
    const UnitDesc& U = job.units[u];
    const OpBuilder& h = job.hist[u];
    if (h.prev < 0) continue;  // no step ran: planes are as loaded
    int rc = RT_OK;
    if (job.model == M_DOMINO) {
      DominoArrays a[2];
      for (int ch = 0; ch < 2; ch++)
        for (int k = 0; k < 2; k++) a[ch].col[k] = (uint8_t*)U.plane[ch < job.nc ? ch : 0][k];
      rc = launch_domino_sync(ctx, a, job.nc, U.rows, h.prev);
    } else if (job.model == M_LOZENGE) {
      TriArrays8 a[2];
      for (int ch = 0; ch < 2; ch++)
        for (int k = 0; k < 3; k++) a[ch].col[k] = (uint8_t*)U.plane[ch < job.nc ? ch : 0][k];
      rc = lozenge_device_sync(ctx, a, job.nc, U.rows, h.last);
    } else {
      TriArrays32 a[2];
      for (int ch = 0; ch < 2; ch++)
        for (int k = 0; k < 3; k++) a[ch].col[k] = (int32_t*)U.plane[ch < job.nc ? ch : 0][k];
      rc = launch_recttri_sync(ctx, a, job.nc, U.rows, h.prev);
    }
    if (rc) return rc;
  }
  return RT_OK;
}

}  // namespace rt

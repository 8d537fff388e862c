// pair_lozenge.hpp -- lozenge CFTP chain pairs at eight waves per SIMD,
// the two chains of a pair packed into one LDS word per cell (included by
// rt_fused.hip).
//
// Windows, op bytes, buffer sets and TinyMT handling are those of k_fused
// (one tile per workgroup, one lane per column).  What differs:
//
// * Cell layout in LDS: cell (i, j) of colour c is one 32-bit word holding
//   the top chain's byte in bits 0-7 and the bottom chain's in bits 16-23.
//   The lozenge pulls (lozengekernel.cl:49-52) move single bits inside a
//   cell and a shift moves both halves alike, so one load, one shift and
//   one bit-field insert per neighbour refresh both chains; the cube move
//   (lozengekernel.cl:20-25: 25 -> 38 when rd < .4, 38 -> 25 when
//   .4 <= rd < .8) tests both halves against the draw's target at once.
// * Tiles of 3072 cells, 96 rows x 32 columns (LW = 32, the one the host
//   launches) or 48 x 64, each lane 3 rows of one column: row bands of 3.
//   With the halo H a multiple of 3 every tile starts on a global row that
//   is a multiple of 3 (tiles step by LH - 2H rows from -H), so the row's
//   residue mod 3, which selects the pull's neighbour columns, is the row's
//   index in the band: a compile-time constant.  With the colour as a
//   template parameter too (twelve step bodies: 3 colours x 4 pull
//   patterns), every LDS address is the lane's base plus an immediate
//   offset.
// * Halo of H rows but only H / 3 columns (kernel comment below).
// * Stale bands: after k pulling steps of a window the top and bottom k
//   rows of the loaded tile are stale (their neighbours outside the tile
//   were not loaded); halo rows are never written back, so a wave whose
//   rows are all stale skips the rest of the window, draws included.
//
// The host uses this kernel on tile shape 7 when H % 3 == 0, for lozenge
// chain pairs and for one-chain walks (the chain in the low halves, the
// high halves zero: a packed operation costs the same for one live half);
// k_fused_o8<1> / k_walk_o8<1> otherwise; results are identical.
#pragma once

namespace rt {

constexpr int kPkRows = 3;  // rows per lane (one band per wave)

// (x & m) | (y & ~m), the mask a constant held in an SGPR (a literal
// cannot be a VOP3 operand on this chip).
__device__ __forceinline__ uint32_t pk_bfi(uint32_t m, uint32_t x, uint32_t y) {
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "s"(m), "v"(x), "v"(y));
  return r;
}

// Two 16-bit halves of a word, for the per-half shifts that keep one
// chain's bits out of the other's.
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u32(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }

// The same with a per-lane mask.
__device__ __forceinline__ uint32_t pk_bfi_v(uint32_t m, uint32_t x, uint32_t y) {
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(x), "v"(y));
  return r;
}

// Target of a draw in both halves: the cell value the move changes (25 for
// rd < .4, 38 for .4 <= rd < .8), else 0x100, which no cell equals.
__device__ __forceinline__ uint32_t pk_target(uint32_t b) {
  return b < kBelow04 ? 0x00190019u : (b < kBelow08 ? 0x00260026u : 0x01000100u);
}

// Both halves of v (cells <= 63) with the target's value flipped 25 <-> 38:
// d = v ^ t2 has zero halves where a cell equals the target; the packed
// unsigned minimum with 1 turns each half into 0 (equal) or 1, times 63 is
// 0 or 63 per half (below 2^24, so the 24-bit multiply is exact), and
// v ^ that ^ 0x003F003F (one three-input XOR, the constant in an SGPR)
// flips exactly the equal halves.
__device__ __forceinline__ uint32_t pk_apply(uint32_t v, uint32_t t2) {
  const uint32_t d = v ^ t2;
  uint32_t m, r;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(d), "s"(0x00010001u));
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(v), "v"(__umul24(m, 63u)),
      "s"(0x003F003Fu));
  return r;
}

// One colour step C with pull pattern PULL on the packed tile.  `p` is the
// lane's row-0 word of plane 0; row r of the band has residue r mod 3.
// Reads of all rows precede the writes (as in sweep_step).
//   pull A (PULL & 1), bits {2, 4, 32} from colour a = C-1
//   (lozenge_pull_prev_m): (i-1, j+p2) bit 16 -> 2, (i, j-p4) bit 8 -> 4,
//   (i+1, j) bit 1 -> 32, p2 = [i%3 == (2a+1)%3], p4 = [i%3 == (2a+2)%3];
//   pull B (PULL & 2), bits {1, 8, 16} from colour b = C-2
//   (lozenge_pull_prev2_m): (i-1, j) bit 32 -> 1, (i, j+p8) bit 4 -> 8,
//   (i+1, j-p16) bit 2 -> 16, p8 = [i%3 == 2b%3], p16 = [i%3 == (2b+1)%3].
template <int C, int PULL, int RS, int PL>
__device__ __forceinline__ void pk_step(uint32_t* p, const uint32_t* interior,
                                        const uint32_t* t2) {
  constexpr int A = (C + 2) % 3, B = (C + 1) % 3;
  constexpr uint32_t MA = 0x00260026u, MB = 0x00190019u;
  uint32_t out[kPkRows];
#pragma unroll
  for (int r = 0; r < kPkRows; r++) {
    const int im = r;  // global row mod 3
    const int o = r * RS;
    uint32_t v = p[C * PL + o];
    if (PULL) {
      uint32_t bits = 0;
      if (PULL & 1) {
        const int p2 = im == (2 * A + 1) % 3, p4 = im == (2 * A + 2) % 3;
        const uint32_t* q = p + A * PL + o;
        const uint32_t a = q[-RS + p2], b = q[-p4], c = q[RS];
        bits = pk_bfi(0x00200020u, c << 5, pk_bfi(0x00040004u, b >> 1, a >> 3));
      }
      if (PULL & 2) {
        const int p8 = im == (2 * B) % 3, p16 = im == (2 * B + 1) % 3;
        const uint32_t* q = p + B * PL + o;
        const uint32_t a = q[-RS], b = q[p8], c = q[RS - p16];
        const uint32_t bb = pk_bfi(0x00100010u, c << 3, pk_bfi(0x00080008u, b << 1, a >> 5));
        bits = (PULL & 1) ? pk_bfi(MB, bb, bits) : bb;
      }
      constexpr uint32_t M = (PULL & 1 ? MA : 0u) | (PULL & 2 ? MB : 0u);
      v = pk_bfi_v(M & interior[r], bits, v);
    }
    out[r] = pk_apply(v, t2[r]);
  }
#pragma unroll
  for (int r = 0; r < kPkRows; r++) p[C * PL + r * RS] = out[r];
}

// Kernel arguments, work items, windows and buffer sets as k_fused; mp is
// unused (the lozenge move has no parameters).  Needs H % 3 == 0.
//
// LW = 64: 48 x 64 tiles, wave wv owns rows 3wv .. 3wv+2.  LW = 32: 96 x 32
// tiles, lanes 0-31 of wave wv own rows 3wv .. 3wv+2 and lanes 32-63 rows
// 93-3wv .. 95-3wv, so the two bands of a wave go stale together (both at
// distance 3wv from a tile edge).  Both load a 3072-cell tile.
//
// Halo: H rows, H / 3 columns.  A pulling step reads rows i-1 and i+1, so
// rows go stale one deeper per pull; its column offsets (p2, p4, p8, p16)
// are 1 only for one row residue each, and over every colour sequence the
// stale columns at either side number at most ceil(pulls / 3) (checked
// exhaustively by the host tests on the lozenge pull rule), so a window of
// at most H pulls leaves columns H / 3 .. LW - 1 - H / 3 exact.
template <int LW>
__global__ __launch_bounds__(1024, 8) void k_pair_lozenge(const UnitDesc* __restrict__ units,
                                                          const WorkItem* __restrict__ work,
                                                          int nwork, int xcd, int H,
                                                          ModelParams mp,
                                                          const WindowRec* __restrict__ wins) {
  constexpr int W = 16, R = kPkRows;
  constexpr int kThreads = 64 * W;
  constexpr int HALVES = 64 / LW;
  constexpr int LH = W * R * HALVES;
  constexpr int RS = LW + 2;
  constexpr int PL = (LH + 2) * RS;
  __shared__ __align__(16) uint32_t lds[3 * PL];
  __shared__ uint8_t sops[kKMax];

  const int blk = work_index(nwork, xcd);
  if (blk < 0) return;
  const WorkItem wi = work[blk];
  const WindowRec w = wins[wi.unit];
  const long s0 = w.s0;
  const int K = w.K, init = w.init;
  const uint32_t seed = w.seed;
  if (K <= 0 && !init) return;  // unit idle in this launch
  UnitDesc U = units[wi.unit];
  window_buffers(U, w.flags);
  const int rows = U.rows, cols = U.cols;
  const int HC = H / 3;
  const int TH = LH - 2 * H, TW = LW - 2 * HC;
  const int tr = wi.tile / U.tiles_c, tc = wi.tile % U.tiles_c;
  const int r0 = tr * TH - H, c0 = tc * TW - HC + wi.cs;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & (LW - 1);
  const int bw = R * wv;  // the wave's band row at the top of the tile
  const int band = (HALVES == 2 && lane >= LW) ? LH - R - bw : bw;  // tile row of the lane's row 0
  const int gc = c0 + col;
  const bool col_in = gc >= 0 && gc < cols;
  // a CFTP pair, or one chain (a walk) in the low halves, the high halves
  // then zero: never a target, never stored
  const bool pair = U.plane[1][0] != nullptr;

  for (int k = threadIdx.x; k < 3 * PL; k += kThreads) lds[k] = 0;
  if (threadIdx.x < K) sops[threadIdx.x] = U.ops[s0 + threadIdx.x];
  __syncthreads();

  Tiny st[R];
  uint32_t interior[R];
  uint32_t* const p = lds + (band + 1) * RS + col + 1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int gr = r0 + band + r;
    const bool in = col_in && gr >= 0 && gr < rows;
    interior[r] = (col_in && gr >= 1 && gr <= rows - 2 && gc >= 1 && gc <= cols - 2) ? ~0u : 0u;
    st[r] = Tiny{0, 0, 0, 0, 0, 0, 0};
    if (in) {
      const long idx = (long)gr * cols + gc;
#pragma unroll
      for (int k = 0; k < 3; k++)
        p[k * PL + r * RS] =
            (uint32_t)((const uint8_t*)U.plane[0][k])[idx] |
            (pair ? (uint32_t)((const uint8_t*)U.plane[1][k])[idx] << 16 : 0u);
      // every lozenge cell of the grid owns a stream: id = j + (N/3) i
      st[r].mat1 = U.param[idx];
      st[r].mat2 = U.param[U.pstride + idx];
      st[r].tmat = U.param[2 * U.pstride + idx];
      if (init) {
        tiny_init(st[r], seed + (uint32_t)idx);
      } else {
        st[r].s0 = U.state[idx];
        st[r].s1 = U.state[U.nstream + idx];
        st[r].s2 = U.state[2 * U.nstream + idx];
        st[r].s3 = U.state[3 * U.nstream + idx];
      }
    }
  }
  __syncthreads();

  // Cells outside the grid hold 0 in both halves and never pull, so they
  // never equal a target: their (zero) streams need no masking.
  if (K > 0) {
    uint32_t t2[R];
#pragma unroll
    for (int r = 0; r < R; r++) t2[r] = pk_target(tiny_bits(st[r]));
    uint32_t op = sops[0];
    int pulls = 0;  // pulling steps of the window so far, this one included
    for (int s = 0; s < K; s++) {
      pulls += (op & (OP_PULL_A | OP_PULL_B)) ? 1 : 0;
      // After k pulls, tile rows < k and > LH - 1 - k are stale: a row that
      // pulls is exact when the rows above and below it were.  A wave whose
      // rows are all stale after this step skips it; every row is still
      // computed for as long as it is exact, so the rows that exact rows
      // read are too.  (With LW = 32 the mirrored band is stale exactly
      // when the top one is.)
      const bool stale = bw + R - 1 < pulls || (HALVES == 1 && bw > LH - 1 - pulls);
      if (!stale) {
        switch (op & 15) {
#define RT_PK_CASE(c, pl) \
  case (c) | ((pl) << 2): pk_step<c, pl, RS, PL>(p, interior, t2); break;
          RT_PK_CASE(0, 0) RT_PK_CASE(1, 0) RT_PK_CASE(2, 0)
          RT_PK_CASE(0, 1) RT_PK_CASE(1, 1) RT_PK_CASE(2, 1)
          RT_PK_CASE(0, 2) RT_PK_CASE(1, 2) RT_PK_CASE(2, 2)
          RT_PK_CASE(0, 3) RT_PK_CASE(1, 3) RT_PK_CASE(2, 3)
#undef RT_PK_CASE
          default: break;
        }
      }
      // the next step's op byte and draws before the barrier
      if (s + 1 < K) {
        op = sops[s + 1];
        if (!stale) {
#pragma unroll
          for (int r = 0; r < R; r++) t2[r] = pk_target(tiny_bits(st[r]));
        }
      }
      __syncthreads();
    }
  }

#pragma unroll
  for (int r = 0; r < R; r++) {
    const int lr = band + r, gr = r0 + lr;
    const bool inner = col_in && gr >= 0 && gr < rows && lr >= H && lr < H + TH && col >= HC &&
                       col < HC + TW;
    if (!inner) continue;
    const long idx = (long)gr * cols + gc;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const uint32_t x = p[k * PL + r * RS];
      ((uint8_t*)U.plane_b[0][k])[idx] = (uint8_t)x;
      if (pair) ((uint8_t*)U.plane_b[1][k])[idx] = (uint8_t)(x >> 16);
    }
    if (!U.no_state_write) {
      U.state_b[idx] = st[r].s0;
      U.state_b[U.nstream + idx] = st[r].s1;
      U.state_b[2 * U.nstream + idx] = st[r].s2;
      U.state_b[3 * U.nstream + idx] = st[r].s3;
    }
  }
}

}  // namespace rt

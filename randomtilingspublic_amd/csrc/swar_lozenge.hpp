// swar_lozenge.hpp -- lozenge sweep-block kernel with four byte cells per
// 32-bit word (included by rt_fused.hip).
//
// Same tiles, halo, windows and op bytes as k_fused (one 64-column tile
// per workgroup, H = 6), but each lane owns one word = 4 consecutive cells
// of a row, and a wave covers 4 rows x 16 words.  The lozenge pulls
// (lozengekernel.cl:49-52, see lozenge_pull_prev_m / _prev2_m) only move
// single bits inside a cell, so they are whole-word mask-and-shift
// operations; the +-1 column neighbours are byte funnel shifts of two
// adjacent words by a per-lane amount (no divergent branches).  The cube move (lozengekernel.cl:20-25:
// 25 -> 38 when rd < .4, 38 -> 25 when .4 <= rd < .8) is a per-byte
// equality test and an XOR with 25 ^ 38 = 0x3F under byte masks built from
// the four cells' own TinyMT draws.  Every cell still draws from its own
// stream, so results are those of the one-cell-per-lane kernel.
#pragma once

namespace rt {

// 0x80 in every byte of x that is zero, 0x00 elsewhere (exact: no carries
// cross byte lanes since (x & 0x7F) + 0x7F <= 0xFE).
__device__ __forceinline__ uint32_t swar_zero80(uint32_t x) {
  return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// 0x80 in byte k when b[k] < t.
__device__ __forceinline__ uint32_t swar_below80(const uint32_t* b, uint32_t t) {
  return (b[0] < t ? 0x00000080u : 0u) | (b[1] < t ? 0x00008000u : 0u) |
         (b[2] < t ? 0x00800000u : 0u) | (b[3] < t ? 0x80000000u : 0u);
}

// Cells one column to the right (shift 8: byte k <- cell k+1) or unshifted
// (shift 0), from the lane's word `cur` and the next word `next`; and one
// column to the left (shift 24: byte k <- cell k-1) or unshifted (shift 32)
// from `cur` and the previous word.  A 64-bit funnel shift by a per-lane
// amount keeps both words in use, so no lane branches on its row.
__device__ __forceinline__ uint32_t swar_funnel(uint32_t hi, uint32_t lo, uint32_t sh) {
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
}

struct SwarRow {
  int li;             // LDS word index of the lane's word in this row slot
  int im;             // global row mod 3
  uint32_t in;        // 0x3F in bytes of cells inside the domain (they move)
  uint32_t interior;  // 0xFF in bytes of cells that pull from neighbours
};

// One colour step.  All reads precede all writes (see sweep_step).
template <int NC, int R, int PULL, int PLW>
__device__ __forceinline__ void swar_step(uint32_t* lds, const SwarRow* row, Tiny (*st)[4],
                                          uint32_t op) {
  const int c = op & OP_COLOUR;
  const int ca = (c + 2) % 3, cb = (c + 1) % 3;
  const int pa2 = (2 * ca + 1) % 3, pa4 = (2 * ca + 2) % 3;  // rows where prev shifts
  const int pb8 = (2 * cb) % 3, pb16 = (2 * cb + 1) % 3;      // rows where prev2 shifts
  uint32_t up[R], down[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    uint32_t b[4];
#pragma unroll
    for (int k = 0; k < 4; k++) b[k] = tiny_bits(st[r][k]);
    up[r] = swar_below80(b, kBelow04);
    down[r] = swar_below80(b, kBelow08) & ~up[r];
  }
  uint32_t out[NC][R];
#pragma unroll
  for (int ch = 0; ch < NC; ch++) {
    const uint32_t* P = lds + ch * 3 * PLW;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = row[r].li, im = row[r].im;
      uint32_t v = P[c * PLW + li];
      if (PULL & 1) {  // bits {2,4,32} from colour c-1: (i-1, j+p2), (i, j-p4), (i+1, j)
        const uint32_t* Q = P + ca * PLW;
        const uint32_t a = swar_funnel(Q[li - 15], Q[li - 16], im == pa2 ? 8u : 0u);
        const uint32_t b = swar_funnel(Q[li], Q[li - 1], im == pa4 ? 24u : 32u);
        const uint32_t cc = Q[li + 16];
        const uint32_t pv = (v & 0xD9D9D9D9u) | ((a & 0x10101010u) >> 3) |
                            ((b & 0x08080808u) >> 1) | ((cc & 0x01010101u) << 5);
        v = (pv & row[r].interior) | (v & ~row[r].interior);
      }
      if (PULL & 2) {  // bits {1,8,16} from colour c-2: (i-1, j), (i, j+p8), (i+1, j-p16)
        const uint32_t* Q = P + cb * PLW;
        const uint32_t a = Q[li - 16];
        const uint32_t b = swar_funnel(Q[li + 1], Q[li], im == pb8 ? 8u : 0u);
        const uint32_t cc = swar_funnel(Q[li + 16], Q[li + 15], im == pb16 ? 24u : 32u);
        const uint32_t pv = (v & 0xE6E6E6E6u) | ((a & 0x20202020u) >> 5) |
                            ((b & 0x04040404u) << 1) | ((cc & 0x02020202u) << 3);
        v = (pv & row[r].interior) | (v & ~row[r].interior);
      }
      const uint32_t x = v ^ 0x19191919u;  // zero bytes: cells equal to 25
      const uint32_t sel = (swar_zero80(x) & up[r]) | (swar_zero80(x ^ 0x3F3F3F3Fu) & down[r]);
      out[ch][r] = v ^ (((sel >> 1) - (sel >> 7)) & row[r].in);  // 0x3F in flagged bytes
    }
  }
#pragma unroll
  for (int ch = 0; ch < NC; ch++)
#pragma unroll
    for (int r = 0; r < R; r++) lds[ch * 3 * PLW + c * PLW + row[r].li] = out[ch][r];
}

// Tile of 4*W*R rows x 64 columns; wave wv, row slot r, lane quarter q
// (lane >> 4) own tile row 4*(wv + W*r) + q, word w = lane & 15 (cells
// 4w .. 4w+3).  Kernel arguments and buffer sets as k_fused.
//
// TinyMT states and parameters travel through an LDS staging array, one
// 32-bit word of every tile cell at a time: the workgroup loads (and
// stores) it with one lane per column, as k_fused does, and each lane
// moves its four cells' words with one 16-byte LDS access.  Loading them
// straight from a lane's four consecutive cells would spread every load
// instruction over four times as many cache lines.
template <int NC, int W, int R>
__global__ __launch_bounds__(64 * W, W >= 16 ? 1 : 2) void k_swar_lozenge(
    const UnitDesc* __restrict__ units, const WorkItem* __restrict__ work, int nwork, int xcd,
    int H, ModelParams mp, const WindowRec* __restrict__ wins) {
  constexpr int kThreads = 64 * W;
  constexpr int LH = 4 * W * R;
  constexpr int PLW = (LH + 2) * 16;  // words per plane, with a zero margin row each side
  __shared__ uint32_t lds[NC * 3 * PLW];
  __shared__ __align__(16) uint32_t stage[LH * kLW];
  __shared__ uint8_t sops[kKMax];

  const int b = work_index(nwork, xcd);
  if (b < 0) return;
  const WorkItem wi = work[b];
  const WindowRec wr = wins[wi.unit];
  const long s0 = wr.s0;
  const int K = wr.K, init = wr.init;
  const uint32_t seed = wr.seed;
  if (K <= 0 && !init) return;  // unit idle in this launch
  UnitDesc U = units[wi.unit];
  window_buffers(U, wr.flags);
  const int rows = U.rows, cols = U.cols;
  const int TH = LH - 2 * H, TW = kLW - 2 * H;
  const int tr = wi.tile / U.tiles_c, tc = wi.tile % U.tiles_c;
  const int r0 = tr * TH - H, c0 = tc * TW - H + wi.cs;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = lane >> 4, w = lane & 15;
  const int gc0 = c0 + 4 * w;

  for (int k = threadIdx.x; k < NC * 3 * PLW; k += kThreads) lds[k] = 0;
  if (threadIdx.x < K) sops[threadIdx.x] = U.ops[s0 + threadIdx.x];
  __syncthreads();

  SwarRow row[R];
  Tiny st[R][4];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int lr = 4 * (wv + W * r) + q, gr = r0 + lr;
    row[r].li = (lr + 1) * 16 + w;
    row[r].im = ((gr % 3) + 3) % 3;
    const bool row_in = gr >= 0 && gr < rows;
    const bool row_int = gr >= 1 && gr <= rows - 2;
    uint32_t in = 0, inter = 0;
    uint32_t cell[NC][3];
#pragma unroll
    for (int ch = 0; ch < NC; ch++)
#pragma unroll
      for (int k = 0; k < 3; k++) cell[ch][k] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int gc = gc0 + k;
      const bool hs = row_in && gc >= 0 && gc < cols;
      if (!hs) continue;
      in |= 0xFFu << (8 * k);
      if (row_int && gc >= 1 && gc <= cols - 2) inter |= 0xFFu << (8 * k);
      const long cidx = (long)gr * cols + gc;
#pragma unroll
      for (int ch = 0; ch < NC; ch++)
#pragma unroll
        for (int pl = 0; pl < 3; pl++)
          cell[ch][pl] |= (uint32_t)((const uint8_t*)U.plane[ch][pl])[cidx] << (8 * k);
    }
    row[r].in = in & 0x3F3F3F3Fu;
    row[r].interior = inter;
#pragma unroll
    for (int ch = 0; ch < NC; ch++)
#pragma unroll
      for (int pl = 0; pl < 3; pl++) lds[(ch * 3 + pl) * PLW + row[r].li] = cell[ch][pl];
  }

  // Stream words (lozenge stream id = cell index j + (N/3) i): parameters,
  // then the state, or the seeding.
  auto stage_in = [&](const uint32_t* a, int field) {
    for (int t = threadIdx.x; t < LH * kLW; t += kThreads) {
      const int gr = r0 + (t >> 6), gc = c0 + (t & 63);
      stage[t] = (gr >= 0 && gr < rows && gc >= 0 && gc < cols) ? a[(long)gr * cols + gc] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int lr = 4 * (wv + W * r) + q;
      const uint4 v = *(const uint4*)&stage[lr * kLW + 4 * w];
      const uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        uint32_t* f = field == 0 ? &st[r][k].mat1 : field == 1 ? &st[r][k].mat2
                    : field == 2 ? &st[r][k].tmat : field == 3 ? &st[r][k].s0
                    : field == 4 ? &st[r][k].s1 : field == 5 ? &st[r][k].s2 : &st[r][k].s3;
        *f = x[k];
      }
    }
    __syncthreads();
  };
  stage_in(U.param, 0);
  stage_in(U.param + U.pstride, 1);
  stage_in(U.param + 2 * U.pstride, 2);
  if (init) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int gr = r0 + 4 * (wv + W * r) + q;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int gc = gc0 + k;
        st[r][k].s0 = st[r][k].s1 = st[r][k].s2 = st[r][k].s3 = 0;
        if (gr >= 0 && gr < rows && gc >= 0 && gc < cols)
          tiny_init(st[r][k], seed + (uint32_t)((long)gr * cols + gc));
      }
    }
  } else {
    stage_in(U.state, 3);
    stage_in(U.state + U.nstream, 4);
    stage_in(U.state + 2 * U.nstream, 5);
    stage_in(U.state + 3 * U.nstream, 6);
  }

  for (int s = 0; s < K; s++) {
    const uint32_t op = sops[s];
    switch ((op >> 2) & 3) {
      case 0: swar_step<NC, R, 0, PLW>(lds, row, st, op); break;
      case 1: swar_step<NC, R, 1, PLW>(lds, row, st, op); break;
      case 2: swar_step<NC, R, 2, PLW>(lds, row, st, op); break;
      default: swar_step<NC, R, 3, PLW>(lds, row, st, op); break;
    }
    __syncthreads();
  }

  // Cells of the tile interior back to the destination planes.
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int lr = 4 * (wv + W * r) + q, gr = r0 + lr;
    if (lr < H || lr >= H + TH) continue;
    uint32_t words[NC][3];
#pragma unroll
    for (int ch = 0; ch < NC; ch++)
#pragma unroll
      for (int pl = 0; pl < 3; pl++) words[ch][pl] = lds[(ch * 3 + pl) * PLW + row[r].li];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int lc = 4 * w + k, gc = gc0 + k;
      const bool hs = gr >= 0 && gr < rows && gc >= 0 && gc < cols;
      if (lc < H || lc >= H + TW || !hs) continue;
      const long cidx = (long)gr * cols + gc;
#pragma unroll
      for (int ch = 0; ch < NC; ch++)
#pragma unroll
        for (int pl = 0; pl < 3; pl++)
          ((uint8_t*)U.plane_b[ch][pl])[cidx] = (uint8_t)(words[ch][pl] >> (8 * k));
    }
  }
  if (U.no_state_write) return;
  // States of the tile interior, through the staging array.
#pragma unroll
  for (int field = 0; field < 4; field++) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int lr = 4 * (wv + W * r) + q;
      uint4 v;
      if (field == 0) v = make_uint4(st[r][0].s0, st[r][1].s0, st[r][2].s0, st[r][3].s0);
      else if (field == 1) v = make_uint4(st[r][0].s1, st[r][1].s1, st[r][2].s1, st[r][3].s1);
      else if (field == 2) v = make_uint4(st[r][0].s2, st[r][1].s2, st[r][2].s2, st[r][3].s2);
      else v = make_uint4(st[r][0].s3, st[r][1].s3, st[r][2].s3, st[r][3].s3);
      *(uint4*)&stage[lr * kLW + 4 * w] = v;
    }
    __syncthreads();
    uint32_t* dst = U.state_b + (long)field * U.nstream;
    for (int t = threadIdx.x; t < LH * kLW; t += kThreads) {
      const int lr = t >> 6, lc = t & 63, gr = r0 + lr, gc = c0 + lc;
      if (lr >= H && lr < H + TH && lc >= H && lc < H + TW && gr >= 0 && gr < rows && gc >= 0 &&
          gc < cols)
        dst[(long)gr * cols + gc] = stage[t];
    }
    __syncthreads();
  }
}

}  // namespace rt

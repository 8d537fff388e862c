// pair_domino.hpp -- domino CFTP chain pairs at eight waves per SIMD, the
// two chains of a pair packed into one LDS word per cell (included by
// rt_fused.hip).  The domino counterpart of k_pair_lozenge
// (pair_lozenge.hpp): same windows, op bytes, buffer sets, 96 x 32 tiles
// with mirrored 3-row bands, stale bands and TinyMT handling.  What differs:
//
// * Two colour planes (the reference's black and white vertex arrays,
//   DominoTiler.cpp:24-33); a pull rebuilds the whole 4-bit cell from the
//   other colour's four neighbours (dominokernel.cl:22-31): bit 1 from the
//   cell above (its bit 2), bit 2 from the cell below (its bit 1), bit 4
//   from the left neighbour (its bit 8), bit 8 from the right one (its bit
//   4).  Above and below share one shift, left and right the other, so
//   both chains' cells come from two bit-field inserts, two shifts and one
//   and-or.
// * Which columns are "left" and "right" depends on the row's parity and
//   the colour: the pair (j + pi - 1, j + pi), pi = (i + c) & 1.  With an
//   even halo every tile starts on an even global row, so a lane's row
//   parity is its band's start parity q (3wv or 93 - 3wv) plus the row's
//   index in the band; the lane keeps two base pointers, p + q and
//   p + 1 - q, and each step body (colour and row fixed at compile time)
//   reads through one of them.
// * Halo of H rows (H even) and H / 2 columns: the pull reads one column to
//   the left or to the right by row parity, so columns go stale at most
//   ceil(pulls / 2) per side (host tests, exhaustively and on the oracle).
// * Moves: the heat-bath rule (dominokernelCFTP.cl:21-30) flips 3 -> 12
//   when rd < 0.5 and 12 -> 3 otherwise, so the cell value that flips is
//   3 or 12 by the draw's top bit.  The two-periodic heat bath
//   (dominokernelCFTPtwoperiodic.cl:15-44) compares 2 rd (or 2 rd - 1) with
//   a ratio of face weights that is 1 on colour-1 steps (the same rule as
//   the uniform one there) and, on colour-0 steps, wa^2 / wb^2 or its
//   inverse by row parity; with rd = u 2^-23 and u = b >> 9, "ratio > th"
//   is "(u mod 2^22) < ceil(min(ratio, 1) 2^22)", two integers per
//   workgroup (exact: every float in the comparison is a multiple of 2^-22
//   or the ratio itself, scaled by a power of two).  Cells without a TinyMT
//   stream (the border ring) get a target no cell equals.
//
// The host uses this kernel on tile shape 7 when H is even, for domino
// chain pairs (heat-bath rules, which CFTP requires) and for one-chain walks
// with a heat-bath rule (the chain in the low halves; a packed operation
// costs the same for one live half).
#pragma once

namespace rt {

// Both halves of v (cells <= 15) with the target's value flipped 3 <-> 12
// (pk_apply with the 4-bit mask).
__device__ __forceinline__ uint32_t pkd_apply(uint32_t v, uint32_t t2) {
  const uint32_t d = v ^ t2;
  uint32_t m, r;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(d), "s"(0x00010001u));
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(v), "v"(__umul24(m, 15u)),
      "s"(0x000F000Fu));
  return r;
}

constexpr uint32_t kPkdNone = 0x01000100u;  // a target no cell equals

// Flipping value of a heat-bath draw in both halves: 12 for b >= 2^31
// (12 -> 3), else 3 (3 -> 12).
__device__ __forceinline__ uint32_t pkd_target_hb(uint32_t b) {
  const uint32_t hi = (uint32_t)((int32_t)b >> 31);
  return pk_bfi_v(hi, 0x000C000Cu, 0x00030003u);
}

// The same for a two-periodic colour-0 step: the flip also needs
// (u mod 2^22) < K.  Odd rows compare wa^2 / wb^2 (ka) when rd < 0.5 and
// wb^2 / wa^2 (kb) otherwise, even rows the other way round, so K = ka
// where the draw's half (hi, all ones for rd >= 0.5) differs from the
// row's parity (odd, all ones for odd rows), kb where they agree.
__device__ __forceinline__ uint32_t pkd_target_2p(uint32_t b, uint32_t odd, uint32_t ka,
                                                  uint32_t kb) {
  const uint32_t hi = (uint32_t)((int32_t)b >> 31);
  const uint32_t u2 = (b >> 9) & 0x3fffffu;
  const uint32_t k = pk_bfi_v(hi ^ odd, ka, kb);
  const uint32_t pass = (uint32_t)((int32_t)(u2 - k) >> 31);
  return pk_bfi_v(pass, pk_bfi_v(hi, 0x000C000Cu, 0x00030003u), kPkdNone);
}

// One colour step C (the plane rotated; the pull, when PULL, rebuilds it
// from plane 1 - C) on the packed tile.  pq = p + q, pn = p + 1 - q.
template <int C, int PULL, int RS, int PL>
__device__ __forceinline__ void pkd_step(uint32_t* p, const uint32_t* pq, const uint32_t* pn,
                                         const uint32_t* interior, const uint32_t* t2) {
  constexpr int S = 1 - C;  // source plane of the pull
  uint32_t out[kPkRows];
#pragma unroll
  for (int r = 0; r < kPkRows; r++) {
    const int o = r * RS;
    uint32_t v = p[C * PL + o];
    if (PULL) {
      // left / right neighbours at j + pi - 1, j + pi, pi = q ^ ((r + C) & 1)
      const uint32_t* P = ((r + C) & 1) ? pn : pq;
      const uint32_t up = p[S * PL + o - RS], dn = p[S * PL + o + RS];
      const uint32_t lf = P[S * PL + o - 1], rg = P[S * PL + o];
      // up bit 2 -> 1 and left bit 8 -> 4 (shift right, per half: a holds
      // nothing but those bits and bit 2 of each half); down bit 1 -> 2 and
      // right bit 4 -> 8 (shift left), of which bits 1 and 3 are taken
      const uint32_t a = as_u32(as_u16x2(pk_bfi(0x00080008u, lf, up)) >> (u16x2)(1));
      const uint32_t b = pk_bfi(0x00040004u, rg, dn) << 1;
      v = pk_bfi_v(interior[r], pk_bfi(0x000A000Au, b, a), v);
    }
    out[r] = pkd_apply(v, t2[r]);
  }
#pragma unroll
  for (int r = 0; r < kPkRows; r++) p[C * PL + r * RS] = out[r];
}

// Kernel arguments, work items, windows and buffer sets as k_fused; the
// rule and weights come from each unit.  Needs H even.
__global__ __launch_bounds__(1024, 8) void k_pair_domino(const UnitDesc* __restrict__ units,
                                                         const WorkItem* __restrict__ work,
                                                         int nwork, int xcd, int H,
                                                         ModelParams mp,
                                                         const WindowRec* __restrict__ wins) {
  constexpr int W = 16, R = kPkRows, LW = 32;
  constexpr int kThreads = 64 * W;
  constexpr int LH = 2 * W * R;  // 96: two mirrored bands per wave
  constexpr int RS = LW + 2;
  constexpr int PL = (LH + 2) * RS;
  __shared__ __align__(16) uint32_t lds[2 * PL];
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
  const int HC = col_halo(M_DOMINO, H);
  const int TH = LH - 2 * H, TW = LW - 2 * HC;
  const int tr = wi.tile / U.tiles_c, tc = wi.tile % U.tiles_c;
  const int r0 = tr * TH - H, c0 = tc * TW - HC + wi.cs;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & (LW - 1);
  const int bw = R * wv;  // the wave's band row at the top of the tile
  const int half = lane >= LW ? 1 : 0;
  const int band = half ? LH - R - bw : bw;  // tile row of the lane's row 0
  const int q = band & 1;                    // its parity (r0 is even)
  const int gc = c0 + col;
  const bool col_in = gc >= 0 && gc < cols;
  // a CFTP pair, or one chain (a walk) in the low halves, the high halves
  // then zero: never a target, never stored
  const bool pair = U.plane[1][0] != nullptr;

  for (int k = threadIdx.x; k < 2 * PL; k += kThreads) lds[k] = 0;
  if (threadIdx.x < K) sops[threadIdx.x] = U.ops[s0 + threadIdx.x];
  __syncthreads();

  Tiny st[R];
  uint32_t interior[R];
  uint32_t* const p = lds + (band + 1) * RS + col + 1;
  const uint32_t* const pq = p + q;
  const uint32_t* const pn = p + 1 - q;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int gr = r0 + band + r;
    const bool in = col_in && gr >= 0 && gr < rows;
    // cells that pull and move: the interior (dominokernel*.cl NDRanges
    // start at 1 and end before the last row and column)
    const bool hs = col_in && gr >= 1 && gr <= rows - 2 && gc >= 1 && gc <= cols - 2;
    interior[r] = hs ? ~0u : 0u;
    st[r] = Tiny{0, 0, 0, 0, 0, 0, 0};
    if (in) {
      const long idx = (long)gr * cols + gc;
#pragma unroll
      for (int k = 0; k < 2; k++)
        p[k * PL + r * RS] =
            (uint32_t)((const uint8_t*)U.plane[0][k])[idx] |
            (pair ? (uint32_t)((const uint8_t*)U.plane[1][k])[idx] << 16 : 0u);
    }
    if (hs) {
      const long id = stream_id<M_DOMINO>(gr, gc, rows, cols);
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
  // two-periodic thresholds (colour-0 steps): ka for ratio wa^2 / wb^2,
  // kb for its inverse, as in domino_move's float comparisons
  const bool two_p = U.rule == D_HEATBATH2P;
  uint32_t ka = 0, kb = 0;
  if (two_p) {
    const float ra = U.dw.wa2 / U.dw.wb2, rb = U.dw.wb2 / U.dw.wa2;
    ka = ra >= 1.0f ? 0x400000u : (uint32_t)ceilf(ra * 4194304.0f);
    kb = rb >= 1.0f ? 0x400000u : (uint32_t)ceilf(rb * 4194304.0f);
  }
  const uint32_t qodd = q ? ~0u : 0u;  // odd global row for the band's rows 0 and 2
  __syncthreads();

  if (K > 0) {
    uint32_t t2[R];
    uint32_t op = sops[0];
    // the next step's flipping values (its colour decides the two-periodic
    // rule); streams of cells outside the interior are zero and masked
#define RT_PKD_DRAWS()                                                               \
  do {                                                                               \
    if (two_p && (op & OP_COLOUR) == 0) {                                            \
      _Pragma("unroll") for (int r = 0; r < R; r++) t2[r] = pk_bfi_v(                \
          interior[r], pkd_target_2p(tiny_bits(st[r]), (r & 1) ? ~qodd : qodd, ka, kb), \
          kPkdNone);                                                                 \
    } else {                                                                         \
      _Pragma("unroll") for (int r = 0; r < R; r++) t2[r] =                          \
          pk_bfi_v(interior[r], pkd_target_hb(tiny_bits(st[r])), kPkdNone);          \
    }                                                                                \
  } while (0)
    RT_PKD_DRAWS();
    int pulls = 0;  // pulling steps of the window so far, this one included
    for (int s = 0; s < K; s++) {
      pulls += (op & OP_PULL_A) ? 1 : 0;
      // stale rows as in k_pair_lozenge: a wave whose rows are all stale
      // after this step skips it (the mirrored band is stale exactly when
      // the top one is)
      const bool stale = bw + R - 1 < pulls;
      if (!stale) {
        switch (op & 7) {
          case 0: pkd_step<0, 0, RS, PL>(p, pq, pn, interior, t2); break;
          case 1: pkd_step<1, 0, RS, PL>(p, pq, pn, interior, t2); break;
          case 4: pkd_step<0, 1, RS, PL>(p, pq, pn, interior, t2); break;
          case 5: pkd_step<1, 1, RS, PL>(p, pq, pn, interior, t2); break;
          default: break;
        }
      }
      if (s + 1 < K) {
        op = sops[s + 1];
        if (!stale) RT_PKD_DRAWS();
      }
      __syncthreads();
    }
#undef RT_PKD_DRAWS
  }

#pragma unroll
  for (int r = 0; r < R; r++) {
    const int lr = band + r, gr = r0 + lr;
    const bool inner = col_in && gr >= 0 && gr < rows && lr >= H && lr < H + TH && col >= HC &&
                       col < HC + TW;
    if (!inner) continue;
    const long idx = (long)gr * cols + gc;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t x = p[k * PL + r * RS];
      ((uint8_t*)U.plane_b[0][k])[idx] = (uint8_t)x;
      if (pair) ((uint8_t*)U.plane_b[1][k])[idx] = (uint8_t)(x >> 16);
    }
    if (((interior[r] & 1u) != 0) && !U.no_state_write) {
      const long id = stream_id<M_DOMINO>(gr, gc, rows, cols);
      U.state_b[id] = st[r].s0;
      U.state_b[U.nstream + id] = st[r].s1;
      U.state_b[2 * U.nstream + id] = st[r].s2;
      U.state_b[3 * U.nstream + id] = st[r].s3;
    }
  }
}

}  // namespace rt

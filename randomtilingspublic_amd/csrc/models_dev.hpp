// models_dev.hpp -- per-cell local moves and neighbour rebuilds of the four
// reference models, as device functions shared by every kernel shape
// (one-step, fused LDS sweep block, CFTP pair, batched chains).
//
// Cell states and bit layouts are the reference's:
//   domino vertex bits 1 up, 2 down, 4 left, 8 right (DominoTiler.h:28-32,
//     DominoTiler.cpp:337-341: pws = {2,1,8,4} for down, up, right, left);
//   lozenge vertex bits 1,2,4,8,16,32 (LozengeTiler.h:16-19);
//   rectangle-triangle: six 4-bit face codes d5..d0 (RectTriangleTiler.h:19-24).
//
// All comparisons against the reference's double literals (.8, .4, 0.5,
// 0.75, 0.25) are done in float: the draws lie on the 2^-23 grid and no
// grid value falls between such a literal and its float rounding (checked
// exhaustively in tests/test_oracle.py), and 0.5/0.75/0.25 are exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tinymt_dev.hpp"

namespace rt {

// ------------------------------------------------------------ domino ----
enum DominoRule { D_TOGGLE08 = 0, D_HEATBATH = 1, D_METRO2P = 2, D_HEATBATH2P = 3 };

struct DominoWeights {
  float wa2;  // (float)(a*a), a*a evaluated in double like the reference's
  float wb2;  // (float)(b*b)  `#define a 0.7` literals
};

// Local move at a vertex of colour t in compressed row i
// (dominokernel.cl:14-17, dominokernelCFTP.cl:65-69,
//  dominokerneltwoperiodic.cl:108-123, dominokernelCFTPtwoperiodic.cl:175-190).
// `b` is the tempered draw; rd = (b>>9)*2^-23.
__device__ __forceinline__ uint32_t domino_move(uint32_t v, uint32_t b, int rule,
                                                int i, int t, DominoWeights w) {
  if (rule == D_TOGGLE08) {
    const uint32_t t = (v == 3u) ? 12u : ((v == 12u) ? 3u : v);
    v = b < kBelow08 ? t : v;
  } else if (rule == D_HEATBATH) {
    // a flippable vertex ends in 12 if b < 2^31, else in 3
    const bool flippable = v == 3u || v == 12u;
    v = flippable ? (b < kBelow05 ? 12u : 3u) : v;
  } else {
    const float rd = tiny_to_float01(b);
    // Horizontal pair weight: row parity i; vertical pair weight: (i+t).
    const float wrow = (i & 1) ? w.wa2 : w.wb2;
    const float wcol = ((i + t) & 1) ? w.wa2 : w.wb2;
    if (rule == D_METRO2P) {
      if (rd < 0.8f) {
        const float th = rd / 0.8f;
        if (v == 3u) {
          if (wrow / wcol > th) v = 12u;
        } else if (v == 12u) {
          if (wcol / wrow > th) v = 3u;
        }
      }
    } else {  // D_HEATBATH2P: thresholds rd/0.5 and (rd-0.5)/0.5 are exact
      if (rd < 0.5f) {
        const float th = rd * 2.0f;
        if (v == 3u && wrow / wcol > th) v = 12u;
      } else {
        const float th = (rd - 0.5f) * 2.0f;
        if (v == 12u && wcol / wrow > th) v = 3u;
      }
    }
  }
  return v;
}

// Rebuild of a vertex from its four neighbours of the other colour, where
// `tr` is the colour that was rotated (dominokernel.cl:27-30).  `p` points
// at the rotated array's cell with the same compressed index (i,j), `rs` is
// that array's row stride; i is the global row (it fixes the parities).
template <class P>
__device__ __forceinline__ uint32_t domino_rebuild(const P* p, int rs, int i, int tr) {
  const uint32_t up = (uint32_t)p[-rs] & 0xffu;
  const uint32_t dn = (uint32_t)p[rs] & 0xffu;
  const uint32_t lf = (uint32_t)p[-((i + tr + 1) & 1)] & 0xffu;
  const uint32_t rg = (uint32_t)p[(i + tr) & 1] & 0xffu;
  return ((up >> 1) & 1u) | ((dn & 1u) << 1) | ((lf >> 1) & 4u) | ((rg & 4u) << 1);
}

// ----------------------------------------------------------- lozenge ----
// Cube add (25 -> 38) with rd < .4, remove (38 -> 25) with .4 <= rd < .8
// (lozengekernel.cl:20-25).  `b` is the tempered draw (rd = (b>>9)*2^-23).
__device__ __forceinline__ uint32_t lozenge_move(uint32_t v, uint32_t b) {
  const bool up = b < kBelow04;
  const bool down = !up && b < kBelow08;
  return (up && v == 25u) ? 38u : ((down && v == 38u) ? 25u : v);
}

// Bits of a colour-x vertex that face colour x-1 ("next" of the rotated
// colour in lozengekernel.cl:49-50) are {2,4,32}; bits facing colour x+1
// ("next-next", 51-52) are {1,8,16}.  `p` points at cell (i,j) of the
// source colour's array (row stride rs); i is the global row.
// Refresh the {2,4,32} bits of v from the array of colour tr = x-1 (mod 3).
template <class P>
__device__ __forceinline__ uint32_t lozenge_pull_prev(uint32_t v, const P* p, int rs, int i,
                                                      int tr) {
  const int im = i % 3;
  const int p2 = im == ((2 * tr + 1) % 3);
  const int p4 = im == ((2 * tr + 2) % 3);
  v &= ~(2u | 4u | 32u);
  v |= ((uint32_t)p[-rs + p2] & 16u) >> 3;
  v |= ((uint32_t)p[-p4] & 8u) >> 1;
  v |= ((uint32_t)p[rs] & 1u) << 5;
  return v & 0xffu;
}

// Bit-field insert: the bits of x under mask M, the rest from y.  Written
// as the instruction itself: the compiler turns (x & M) | (y & ~M) chains
// into separate and / or ops.
template <uint32_t M>
__device__ __forceinline__ uint32_t bfi(uint32_t x, uint32_t y) {
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "n"(M), "v"(x), "v"(y));
  return r;
}

// Same, with im = i % 3 precomputed; v is a cell byte.
template <class P>
__device__ __forceinline__ uint32_t lozenge_pull_prev_m(uint32_t v, const P* p, int rs, int im,
                                                        int tr) {
  const int p2 = im == ((2 * tr + 1) % 3);
  const int p4 = im == ((2 * tr + 2) % 3);
  const uint32_t a = p[-rs + p2], b = p[-p4], c = p[rs];
  return bfi<32u>(c << 5, bfi<4u>(b >> 1, bfi<2u>(a >> 3, v)));
}

template <class P>
__device__ __forceinline__ uint32_t lozenge_pull_prev2_m(uint32_t v, const P* p, int rs, int im,
                                                         int tr) {
  const int p8 = im == ((2 * tr) % 3);
  const int p16 = im == ((2 * tr + 1) % 3);
  const uint32_t a = p[-rs], b = p[p8], c = p[rs - p16];
  return bfi<16u>(c << 3, bfi<8u>(b << 1, bfi<1u>(a >> 5, v)));
}

// lozenge_move split in two: the draw of a cell decides which byte value
// (25 or 38) a move would change, or none (0x100, no byte equals it); the
// pair's two chains then apply it with one compare and one select each.
// hs = 0 (a cell without a TinyMT stream) never moves.
__device__ __forceinline__ uint32_t lozenge_target(uint32_t b, bool hs) {
  const uint32_t t = b < kBelow04 ? 25u : (b < kBelow08 ? 38u : 0x100u);
  return hs ? t : 0x100u;
}
__device__ __forceinline__ uint32_t lozenge_apply(uint32_t v, uint32_t t) {
  return v == t ? (t ^ 0x3Fu) : v;  // 25 <-> 38
}

// Refresh the {1,8,16} bits of v from colour tr = x-2 (mod 3), i.e. x is
// the "next-next" colour of tr.
template <class P>
__device__ __forceinline__ uint32_t lozenge_pull_prev2(uint32_t v, const P* p, int rs, int i,
                                                       int tr) {
  const int im = i % 3;
  const int p8 = im == ((2 * tr) % 3);
  const int p16 = im == ((2 * tr + 1) % 3);
  v &= ~(1u | 8u | 16u);
  v |= ((uint32_t)p[-rs] & 32u) >> 5;
  v |= ((uint32_t)p[p8] & 4u) << 1;
  v |= ((uint32_t)p[rs - p16] & 2u) << 3;
  return v & 0xffu;
}

// ------------------------------------------------- rectangle-triangle ----
struct RTWeights {
  float w[16];  // recttrianglekernel.cl:13: {0, wt1..3, wr1..3, wr1..3, wm1..3, wm1..3}
};

__device__ __forceinline__ float rt_weight(const RTWeights& W, uint32_t h) {
  return W.w[(h >> 20) & 15u] * W.w[(h >> 16) & 15u] * W.w[(h >> 12) & 15u] *
         W.w[(h >> 8) & 15u] * W.w[(h >> 4) & 15u] * W.w[h & 15u];
}

// Hexagon move at one vertex (recttrianglekernel.cl:37-117).  Returns the
// new 24-bit state.  The five "all-triangle / single-rectangle" states form
// a 4-way choice by quarter of rd/0.9; the twelve "rectangle pair" states
// form fixed 2-cycles.
__device__ __forceinline__ uint32_t recttri_move(uint32_t h, uint32_t b, const RTWeights& W) {
  float rd = tiny_to_float01(b);
  if (!(rd < 0.9f)) return h;
  rd = rd / 0.9f;
  const float q = (rd < 0.25f) ? 0.0f : (rd < 0.5f) ? 0.25f : (rd < 0.75f) ? 0.5f : 0.75f;
  // Targets for quarters [0,.25), [.25,.5), [.5,.75), [.75,1).
  uint32_t nt = h;
  bool pair = false;
  switch (h) {
    case 0x122133u: nt = rd < 0.25f ? 0xa7dd7au : rd < 0.5f ? 0xfc99cfu : rd < 0.75f ? 0x8ebbe8u : 0x331221u; break;
    case 0x331221u: nt = rd < 0.25f ? 0xa7dd7au : rd < 0.5f ? 0xfc99cfu : rd < 0.75f ? 0x8ebbe8u : 0x122133u; break;
    case 0x8ebbe8u: nt = rd < 0.25f ? 0xa7dd7au : rd < 0.5f ? 0xfc99cfu : rd < 0.75f ? 0x122133u : 0x331221u; break;
    case 0xfc99cfu: nt = rd < 0.25f ? 0xa7dd7au : rd < 0.5f ? 0x122133u : rd < 0.75f ? 0x8ebbe8u : 0x331221u; break;
    case 0xa7dd7au: nt = rd < 0.25f ? 0x122133u : rd < 0.5f ? 0xfc99cfu : rd < 0.75f ? 0x8ebbe8u : 0x331221u; break;
    case 0x47d47au: nt = 0xd22a33u; pair = true; break;
    case 0xd22a33u: nt = 0x47d47au; pair = true; break;
    case 0xa74d74u: nt = 0x33a22du; pair = true; break;
    case 0x33a22du: nt = 0xa74d74u; pair = true; break;
    case 0x8eb558u: nt = 0x331eb1u; pair = true; break;
    case 0x331eb1u: nt = 0x8eb558u; pair = true; break;
    case 0x855be8u: nt = 0x1be133u; pair = true; break;
    case 0x1be133u: nt = 0x855be8u; pair = true; break;
    case 0x6699cfu: nt = 0xcf1221u; pair = true; break;
    case 0xcf1221u: nt = 0x6699cfu; pair = true; break;
    case 0xfc9966u: nt = 0x1221fcu; pair = true; break;
    case 0x1221fcu: nt = 0xfc9966u; pair = true; break;
    default: return h;
  }
  // Accept if w(new)/w(old) >= rescaled rd; 2-cycles compare against rd itself.
  const float r2 = pair ? rd : (rd - q) * 4.0f;
  const float ratio = rt_weight(W, nt) / rt_weight(W, h);
  return (ratio >= r2) ? nt : h;
}

// updateTiles1 (recttrianglekernel.cl:120-132): rebuild a vertex of colour
// tr+1 from the rotated colour tr; p points at cell (i,j) of colour tr.
__device__ __forceinline__ uint32_t recttri_rebuild1(const int32_t* p, int rs, int i, int tr) {
  const int p1 = (i % 3) == tr;
  const int p2 = (i % 3) == ((tr + 1) % 3);
  const uint32_t u = (uint32_t)p[-rs];
  const uint32_t r = (uint32_t)p[p1];
  const uint32_t d = (uint32_t)p[rs - p2];
  return ((u & 0xffu) << 16) | ((r & 0xf00000u) >> 8) | ((d & 0x0ff000u) >> 8) |
         ((r & 0x000f00u) >> 8);
}

// updateTiles2 (recttrianglekernel.cl:134-146): rebuild a vertex of colour
// tr+2 from the rotated colour tr.
__device__ __forceinline__ uint32_t recttri_rebuild2(const int32_t* p, int rs, int i, int tr) {
  const int p1 = (i % 3) == ((tr + 2) % 3);
  const int p2 = (i % 3) == ((tr + 1) % 3);
  const uint32_t l = (uint32_t)p[-p1];
  const uint32_t u = (uint32_t)p[-rs + p2];
  const uint32_t d = (uint32_t)p[rs];
  return ((l & 0x00f000u) << 8) | ((u & 0x000ff0u) << 8) | ((l & 0x00000fu) << 8) |
         ((d & 0xff0000u) >> 16);
}

}  // namespace rt

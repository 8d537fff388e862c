// rt_bibone.hip -- the bibone (triangular-lattice dimer) walk on gfx950.
//
// State (TriangleDimerTiler.cpp:101-160): eight M x M int32 planes, M = N/2:
// lozenge-type cells eH, eL, eR (4 dimer bits), triangle-type cells tU, tD
// (three 2-bit fields) and butterfly-type cells bH, bL, bR (four 2-bit wing
// fields + bit 8), kept mutually consistent after every move.  One TinyMT
// stream per plane cell, id = j + M*i (the reference's (M, M) NDRange id),
// advanced once per step by whichever plane rotates.
//
// A step is one of three move families picked by the host schedule
// (TriangleDimerTiler.cpp:291-298, 447-448):
//   lozenge   rotate(e_o) -> refresh e_o -> other two e -> {tU, b*} -> tD
//   triangle  rotate(t)   -> refresh t   -> other t    -> e* -> b*
//   butterfly rotate(b_o) -> four refresh stages of b_o -> e* -> {tU, other b} -> tD
// Every kernel writes only cells it does not read, so one launch per phase
// is exact; phases that only read the lozenge planes share a launch.
#include <hip/hip_runtime.h>

#include <random>
#include <vector>

#include "rt_kernels.hpp"

namespace rt {
namespace {

enum Plane { EH = 0, EL, ER, TU, TD, BH, BL, BR, NPLANE };

struct Bi {
  int32_t* p[NPLANE];
  int M;
};

struct Cell {
  int k, i, j;
};

__device__ __forceinline__ bool cell_of(const Bi& B, Cell& c) {
  c.k = blockIdx.x * blockDim.x + threadIdx.x;
  if (c.k >= B.M * B.M) return false;
  c.i = c.k / B.M;
  c.j = c.k - c.i * B.M;
  return true;
}

__device__ __forceinline__ bool interior(const Bi& B, const Cell& c) {
  return c.i > 0 && c.j > 0 && c.i < B.M - 1 && c.j < B.M - 1;
}

__device__ __forceinline__ int at(const Bi& B, int pl, int i, int j) { return B.p[pl][i * B.M + j]; }
__device__ __forceinline__ int bt(int x, int k) { return (x >> k) & 1; }
__device__ __forceinline__ int fld(int x, int k) { return (x >> (2 * k)) & 3; }
__device__ __forceinline__ int mod3(int x) { return ((x % 3) + 3) % 3; }

// One draw of stream k; true when rd < 0.5, i.e. the tempered word < 2^31.
__device__ __forceinline__ bool coin(const StreamTable& S, int k) {
  Tiny t = tiny_load(S, k);
  const uint32_t b = tiny_bits(t);
  tiny_store(S, k, t);
  return b < kBelow05;
}

__global__ void k_bi_init(StreamTable S, int n, uint32_t seed) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  Tiny t;
  t.mat1 = S.param[k];
  t.mat2 = S.param[S.pstride + k];
  t.tmat = S.param[2 * S.pstride + k];
  tiny_init(t, seed + (uint32_t)k);
  tiny_store(S, k, t);
}

// ---------------------------------------------------------------- rotations
// RotateLozenges (triangledimerkernel.cl:4-22): sublattice parity of
// orientation o is i (H), j (L) or i+j (R); 5 <-> 10 with probability 1/2.
__global__ void k_bi_rot_loz(Bi B, StreamTable S, int o, int c) {
  Cell x;
  if (!cell_of(B, x)) return;
  const bool heads = coin(S, x.k);
  const int par = (o == 0 ? x.i : (o == 1 ? x.j : x.i + x.j)) & 1;
  int32_t* p = B.p[EH + o] + x.k;
  const int v = *p;
  if (heads && par == c && (v == 5 || v == 10)) *p = 15 - v;
}

// RotateTriangles (190-211): class (j-i) mod 3; 45 <-> 54.
__global__ void k_bi_rot_tri(Bi B, StreamTable S, int pl, int c) {
  Cell x;
  if (!cell_of(B, x)) return;
  const bool heads = coin(S, x.k);
  int32_t* p = B.p[pl] + x.k;
  const int v = *p;
  if (heads && mod3(x.j - x.i) == c && (v == 45 || v == 54)) *p = 99 - v;
}

// RotateButterflys (330-352): slice i, j or i+j (mod 3) == p1; the flip
// 85 <-> 170 happens on the cells whose parity (j for H and R, i for L)
// equals p2.
__global__ void k_bi_rot_bf(Bi B, StreamTable S, int o, int p1, int p2) {
  Cell x;
  if (!cell_of(B, x)) return;
  const bool heads = coin(S, x.k);
  const int slice = (o == 0 ? x.i : (o == 1 ? x.j : x.i + x.j)) % 3;
  const int q = (o == 1 ? x.i : x.j) & 1;
  int32_t* p = B.p[BH + o] + x.k;
  const int v = *p;
  if (heads && slice == p1 && q == p2 && (v == 85 || v == 170)) *p = 255 - v;
}

// ------------------------------------------------------------ lozenge moves
// UpdateLozengesFlipped (24-34): cells of the other parity rebuilt from
// their four neighbours a, b, c, d (bit 2 of a, bit 3 of b, bit 0 of c and
// bit 1 of d, in that order).
__global__ void k_bi_loz_refresh(Bi B, int o, int c) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int par = (o == 0 ? x.i : (o == 1 ? x.j : x.i + x.j)) & 1;
  if (par == c) return;
  const int pl = EH + o, i = x.i, j = x.j;
  const int s = o & 1;            // column shift of the a / c neighbours
  const int r = o == 0 ? 1 : 0;   // row shift of the b / d neighbours
  const int a = at(B, pl, i - 1, j + s), b = at(B, pl, i - r, j + 1);
  const int cc = at(B, pl, i + 1, j - s), d = at(B, pl, i + r, j - 1);
  B.p[pl][x.k] = bt(a, 2) | bt(b, 3) << 1 | bt(cc, 0) << 2 | bt(d, 1) << 3;
}

// UpdateLozenges0/1/2 (36-79): the two other lozenge orientations take the
// dimer bits they share with the rotated orientation o.
__global__ void k_bi_loz_share(Bi B, int o) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int i = x.i, j = x.j, k = x.k;
  if (o == 0) {
    const int dn = at(B, EH, i + 1, j);
    B.p[EL][k] = (B.p[EL][k] & ~10) | bt(dn, 0) << 1 | bt(at(B, EH, i + 1, j - 1), 0) << 3;
    B.p[ER][k] = (B.p[ER][k] & ~10) | (dn & 2) | (at(B, EH, i, j) & 8);
  } else if (o == 1) {
    const int me = at(B, EL, i, j);
    B.p[ER][k] = (B.p[ER][k] & ~5) | (me & 1) | bt(at(B, EL, i + 1, j), 0) << 2;
    B.p[EH][k] = (B.p[EH][k] & ~5) | bt(at(B, EL, i - 1, j), 1) | bt(me, 1) << 2;
  } else {
    const int lf = at(B, ER, i, j - 1);
    B.p[EH][k] = (B.p[EH][k] & ~10) | (at(B, ER, i - 1, j) & 2) | bt(lf, 1) << 3;
    B.p[EL][k] = (B.p[EL][k] & ~5) | (at(B, ER, i, j) & 1) | (lf & 4);
  }
}

// wing(x, a, b, own, k) = bit a of x + 2 * bit b of x + 3 * bit k of own:
// which of a butterfly wing's three slots holds the dimer.
__device__ __forceinline__ int wing(int x, int a, int b, int own, int k) {
  return bt(x, a) + 2 * bt(x, b) + 3 * bt(own, k);
}
__device__ __forceinline__ int pack4(int f0, int f1, int f2, int f3, int c) {
  return f0 + 4 * f1 + 16 * f2 + 64 * f3 + 256 * c;
}

// Everything rebuilt from the lozenge planes alone, selected by `mask`:
// bit 0 tU (UpdateTriangleUFromLozenges, 97-119), bits 1..3 bH, bL, bR
// (UpdateButterflys{H,L,R}FromLozenge, 121-188).
__global__ void k_bi_from_loz(Bi B, int mask) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int i = x.i, j = x.j, k = x.k;
  const int h = at(B, EH, i, j), l = at(B, EL, i, j), r = at(B, ER, i, j);
  const int hd = at(B, EH, i + 1, j), lr = at(B, EL, i, j + 1);
  if (mask & 1) {
    const int L = bt(l, 0) + 2 * bt(h, 3) + 3 * bt(hd, 0);
    const int R = bt(lr, 0) + 2 * bt(hd, 1) + 3 * bt(lr, 1);
    const int V = bt(lr, 2) + 2 * bt(hd, 3) + 3 * bt(hd, 2);
    B.p[TU][k] = L + 4 * R + 16 * V;
  }
  if (mask & 2) {
    const int ru = at(B, ER, i - 1, j), lu = at(B, EL, i - 1, j + 1);
    B.p[BH][k] = pack4(wing(ru, 3, 0, h, 0), wing(lu, 0, 1, h, 1), wing(r, 1, 2, h, 2),
                       wing(l, 2, 3, h, 3), bt(lu, 2));
  }
  if (mask & 4) {
    const int hdl = at(B, EH, i + 1, j - 1), rl = at(B, ER, i, j - 1);
    B.p[BL][k] = pack4(wing(h, 0, 1, l, 0), wing(r, 1, 2, l, 1), wing(hdl, 2, 3, l, 2),
                       wing(rl, 3, 0, l, 3), bt(r, 3));
  }
  if (mask & 8) {
    B.p[BR][k] = pack4(wing(h, 0, 1, r, 0), wing(lr, 0, 1, r, 1), wing(hd, 2, 3, r, 2),
                       wing(l, 2, 3, r, 3), bt(lr, 3));
  }
}

// ----------------------------------------------------------- triangle moves
// tri_code(x) = 1, 2 or 3 when x's fields say the shared slot is used in
// the way UpdateTriangles (257-278) tests, summed as the reference sums.
__device__ __forceinline__ int tri_code(int x) {
  return (fld(x, 2) == 1) + 2 * (fld(x, 0) == 3) + 3 * (fld(x, 1) == 2);
}

// UpdateTriangles: dir 0 rebuilds tD from tU (vertical neighbour above),
// dir 1 rebuilds tU from tD (neighbour below).
__global__ void k_bi_tri_pair(Bi B, int dir) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int src = dir ? TD : TU, dst = dir ? TU : TD;
  const int i = x.i, j = x.j;
  const int a = at(B, src, i, j - 1 + dir), b = at(B, src, i, j + dir);
  const int v = at(B, src, i + (dir ? 1 : -1), j);
  B.p[dst][x.k] = tri_code(a) + 4 * tri_code(b) + 16 * tri_code(v);
}

// UpdateTrianglesFlipped0/1 (213-255): the two classes that did not rotate
// take their fields from three class-c neighbours.  For class c+1 (p = 1)
// the fields come from (f1, f2, f0) of the neighbours, for class c+2 from
// (f2, f0, f1).
__global__ void k_bi_tri_refresh(Bi B, int down, int c) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int cls = mod3(x.j - x.i);
  if (cls == c) return;
  const int p = cls == (c + 1) % 3;
  const int pl = down ? TD : TU, i = x.i, j = x.j;
  int a, b, d;
  if (!down) {
    a = at(B, pl, i - 1 + p, j - p);
    b = at(B, pl, i - p, j + 1);
    d = at(B, pl, i + 1, j - 1 + p);
  } else {
    a = at(B, pl, i + 1 - p, j - 1);
    b = at(B, pl, i + p, j + 1 - p);
    d = at(B, pl, i - 1, j + p);
  }
  const int g0 = p ? fld(a, 1) : fld(a, 2);
  const int g1 = p ? fld(b, 2) : fld(b, 0);
  const int g2 = p ? fld(d, 0) : fld(d, 1);
  B.p[pl][x.k] = g0 + 4 * g1 + 16 * g2;
}

// UpdateLozenge{H,L,R}FromTriangles (280-328).
__global__ void k_bi_loz_from_tri(Bi B) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int i = x.i, j = x.j, k = x.k;
  const int d = at(B, TD, i, j);
  const int uu = at(B, TU, i - 1, j), ul = at(B, TU, i, j - 1), u = at(B, TU, i, j);
  B.p[EH][k] = (fld(d, 2) == 2) | (fld(d, 2) == 3) << 1 | (fld(uu, 2) == 3) << 2 |
               (fld(uu, 2) == 2) << 3;
  B.p[EL][k] = (fld(ul, 1) == 1) | (fld(ul, 1) == 3) << 1 | (fld(d, 0) == 1) << 2 |
               (fld(d, 0) == 2) << 3;
  B.p[ER][k] = (fld(u, 0) == 1) | (fld(d, 1) == 3) << 1 | (fld(d, 1) == 1) << 2 |
               (fld(u, 0) == 2) << 3;
}

// ---------------------------------------------------------- butterfly moves
// Slice and per-orientation neighbour offsets of the butterfly refresh
// kernels (355-649).  Orientation o: 0 H (slice i mod 3), 1 L (j mod 3),
// 2 R ((i+j) mod 3).
__device__ __forceinline__ int bf_slice(int o, int i, int j) {
  return (o == 0 ? i : (o == 1 ? j : i + j)) % 3;
}
// rot1: 1->2->3->1, rot2: 1->3->2->1; both keep 0.
__device__ __forceinline__ int rot1(int f) { return f ? f % 3 + 1 : 0; }
__device__ __forceinline__ int rot2(int f) { return f ? (f + 1) % 3 + 1 : 0; }
__device__ __forceinline__ int is3(int x, int k) { return fld(x, k) == 3; }
__device__ __forceinline__ int flag(int x) { return bt(x, 8); }

// Same slice, other parity (UpdateButterflysFlipped{H,L,R}1): the wings
// turn with their two in-slice neighbours n1 (before) and n2 (after).
__global__ void k_bi_bf_same(Bi B, int o, int p1, int p2) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int i = x.i, j = x.j, pl = BH + o;
  const int q = (o == 1 ? i : j) & 1;
  if (bf_slice(o, i, j) != p1 || q == p2) return;
  const int c = flag(B.p[pl][x.k]);
  int v;
  if (o == 2) {
    const int n1 = at(B, pl, i + 1, j - 1), n2 = at(B, pl, i - 1, j + 1);
    v = pack4(rot2(fld(n2, 3)), rot1(fld(n2, 2)), rot2(fld(n1, 1)), rot1(fld(n1, 0)), c);
  } else {
    const int n1 = o == 0 ? at(B, pl, i, j - 1) : at(B, pl, i - 1, j);
    const int n2 = o == 0 ? at(B, pl, i, j + 1) : at(B, pl, i + 1, j);
    v = pack4(rot1(fld(n1, 1)), rot2(fld(n2, 0)), rot1(fld(n2, 3)), rot2(fld(n1, 2)), c);
  }
  B.p[pl][x.k] = v;
}

// Adjacent slices (UpdateButterflysFlipped{H,L,R}2{1,2,3}).  Stage 1 and 3
// work on slice p1+1 (H) / p1+2 (L, R), stage 2 on the remaining one.
// H and L read the same fields from differently placed neighbours.
__global__ void k_bi_bf_stage(Bi B, int o, int p1, int stage) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int i = x.i, j = x.j, pl = BH + o;
  const int s13 = o == 0 ? (p1 + 1) % 3 : (p1 + 2) % 3;
  const int s2 = o == 0 ? (p1 + 2) % 3 : (p1 + 1) % 3;
  const int sl = bf_slice(o, i, j);
  if (sl != (stage == 2 ? s2 : s13)) return;
  const int v = B.p[pl][x.k];
  int f0 = fld(v, 0), f1 = fld(v, 1), f2 = fld(v, 2), f3 = fld(v, 3), c = flag(v);
  if (o != 2) {
    // neighbours: u1, u2 above (H: (i-1,j),(i-1,j+1); L: (i-1,j+1),(i,j+1)),
    //             d1, d2 below (H: (i+1,j),(i+1,j-1); L: (i+1,j-1),(i,j-1)).
    const bool H = o == 0;
    if (stage == 1) {
      const int u1 = H ? at(B, pl, i - 1, j) : at(B, pl, i - 1, j + 1);
      const int u2 = H ? at(B, pl, i - 1, j + 1) : at(B, pl, i, j + 1);
      c = fld(u1, 2) == 2;
      if (f0 != 2) f0 = is3(u1, 3) + 3 * is3(u1, 2);
      if (f1 != 1) f1 = 3 * is3(u2, 3) + 2 * is3(u2, 2);
    } else if (stage == 2) {
      const int u1 = H ? at(B, pl, i - 1, j) : at(B, pl, i - 1, j + 1);
      const int u2 = H ? at(B, pl, i - 1, j + 1) : at(B, pl, i, j + 1);
      const int d1 = H ? at(B, pl, i + 1, j) : at(B, pl, i + 1, j - 1);
      const int d2 = H ? at(B, pl, i + 1, j - 1) : at(B, pl, i, j - 1);
      c = fld(d1, 0) == 2;
      const int n0 = flag(u1) ? 2 : (f0 != 2) * f0;
      const int n1 = flag(u2) ? 1 : (f1 != 1) * f1;
      if (f2 != 2) f2 = 3 * is3(d1, 0) + is3(d1, 1);
      if (f3 != 1) f3 = 2 * is3(d2, 0) + 3 * is3(d2, 1);
      f0 = n0;
      f1 = n1;
    } else {
      const int d1 = H ? at(B, pl, i + 1, j) : at(B, pl, i + 1, j - 1);
      const int d2 = H ? at(B, pl, i + 1, j - 1) : at(B, pl, i, j - 1);
      f2 = flag(d1) ? 2 : (f2 != 2) * f2;
      f3 = flag(d2) ? 1 : (f3 != 1) * f3;
    }
  } else {
    // R: neighbours up (i-1,j), right (i,j+1), down (i+1,j), left (i,j-1).
    if (stage == 1) {
      const int rt = at(B, pl, i, j + 1), dn = at(B, pl, i + 1, j);
      c = fld(dn, 0) == 1;
      if (f1 != 2) f1 = is3(rt, 0) + 3 * is3(rt, 3);
      if (f2 != 1) f2 = 3 * is3(dn, 0) + 2 * is3(dn, 3);
    } else if (stage == 2) {
      const int up = at(B, pl, i - 1, j), rt = at(B, pl, i, j + 1);
      const int dn = at(B, pl, i + 1, j), lf = at(B, pl, i, j - 1);
      c = fld(up, 2) == 1;
      if (f0 != 1) f0 = 3 * is3(up, 2) + 2 * is3(up, 1);
      const int n1 = flag(rt) ? 2 : (f1 != 2) * f1;
      const int n2 = flag(dn) ? 1 : (f2 != 1) * f2;
      if (f3 != 2) f3 = is3(lf, 2) + 3 * is3(lf, 1);
      f1 = n1;
      f2 = n2;
    } else {
      const int up = at(B, pl, i - 1, j), lf = at(B, pl, i, j - 1);
      f0 = flag(up) ? 1 : (f0 != 1) * f0;
      f3 = flag(lf) ? 2 : (f3 != 2) * f3;
    }
  }
  B.p[pl][x.k] = pack4(f0, f1, f2, f3, c);
}

// UpdateLozengeFromButterflys{H,L,R} (652-749): the three lozenge planes at
// a butterfly cell of orientation o.  `full` is the lozenge of the
// butterfly's own orientation, the two others read one field pair + flag.
__global__ void k_bi_loz_from_bf(Bi B, int o) {
  Cell x;
  if (!cell_of(B, x) || !interior(B, x)) return;
  const int v = B.p[BH + o][x.k];
  const int f0 = fld(v, 0), f1 = fld(v, 1), f2 = fld(v, 2), f3 = fld(v, 3), c = flag(v);
  const int full = (f0 == 3) | (f1 == 3) << 1 | (f2 == 3) << 2 | (f3 == 3) << 3;
  int h, l, r;
  if (o == 0) {
    h = full;
    l = c | (f2 == 3) << 1 | (f3 == 1) << 2 | (f3 == 2) << 3;
    r = c | (f2 == 1) << 1 | (f2 == 2) << 2 | (f3 == 3) << 3;
  } else if (o == 1) {
    l = full;
    r = (f0 == 3) | (f1 == 1) << 1 | (f1 == 2) << 2 | c << 3;
    h = (f0 == 1) | (f0 == 2) << 1 | (f1 == 3) << 2 | c << 3;
  } else {
    r = full;
    h = (f0 == 1) | (f0 == 2) << 1 | c << 2 | (f3 == 3) << 3;
    l = (f0 == 3) | c << 1 | (f3 == 1) << 2 | (f3 == 2) << 3;
  }
  B.p[EH][x.k] = h;
  B.p[EL][x.k] = l;
  B.p[ER][x.k] = r;
}

// ------------------------------------------------------------------ host
struct Move {
  uint8_t type, r1, r2, p1, p2;
};

// The reference's schedule (TriangleDimerTiler.cpp:272-298, 447-448): one
// default_random_engine seeded with `seed`; per step rFlipType, r1 ~ U{0,2},
// r2 ~ U{0,1}; butterfly steps then draw p1 ~ U{0,2}, p2 ~ U{0,1}.
std::vector<Move> bibone_schedule(long seed, long steps) {
  std::default_random_engine g;
  std::uniform_int_distribution<int> d3(0, 2), d2(0, 1);
  g.seed(seed);
  std::vector<Move> out((size_t)steps);
  for (Move& m : out) {
    m.type = (uint8_t)d3(g);
    m.r1 = (uint8_t)d3(g);
    m.r2 = (uint8_t)d2(g);
    m.p1 = m.p2 = 0;
    if (m.type == 2) {
      m.p1 = (uint8_t)d3(g);
      m.p2 = (uint8_t)d2(g);
    }
  }
  return out;
}

#define BI_LAUNCH(kern, ...)                                                                \
  do {                                                                                      \
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kBs), 0, ctx->stream, __VA_ARGS__);             \
    launches++;                                                                             \
  } while (0)

constexpr int kBs = 256;

int bibone_run(rt_ctx* ctx, const Bi& B, const StreamTable& S, const std::vector<Move>& sched,
               long* launches_out) {
  const int n = B.M * B.M;
  const unsigned nb = (unsigned)((n + kBs - 1) / kBs);
  long launches = 0;
  for (const Move& m : sched) {
    if (m.type == 0) {
      BI_LAUNCH(k_bi_rot_loz, B, S, (int)m.r1, (int)m.r2);
      BI_LAUNCH(k_bi_loz_refresh, B, (int)m.r1, (int)m.r2);
      BI_LAUNCH(k_bi_loz_share, B, (int)m.r1);
      BI_LAUNCH(k_bi_from_loz, B, 15);
      BI_LAUNCH(k_bi_tri_pair, B, 0);
    } else if (m.type == 1) {
      const int down = m.r2 == 1;
      BI_LAUNCH(k_bi_rot_tri, B, S, down ? (int)TD : (int)TU, (int)m.r1);
      BI_LAUNCH(k_bi_tri_refresh, B, down, (int)m.r1);
      BI_LAUNCH(k_bi_tri_pair, B, down);
      BI_LAUNCH(k_bi_loz_from_tri, B);
      BI_LAUNCH(k_bi_from_loz, B, 14);
    } else {
      const int o = m.r1;
      BI_LAUNCH(k_bi_rot_bf, B, S, o, (int)m.p1, (int)m.p2);
      BI_LAUNCH(k_bi_bf_same, B, o, (int)m.p1, (int)m.p2);
      BI_LAUNCH(k_bi_bf_stage, B, o, (int)m.p1, 1);
      BI_LAUNCH(k_bi_bf_stage, B, o, (int)m.p1, 2);
      BI_LAUNCH(k_bi_bf_stage, B, o, (int)m.p1, 3);
      BI_LAUNCH(k_bi_loz_from_bf, B, o);
      BI_LAUNCH(k_bi_from_loz, B, 1 | (14 & ~(2 << o)));
      BI_LAUNCH(k_bi_tri_pair, B, 0);
    }
  }
  HIPCHK_RET(ctx, hipGetLastError());
  *launches_out = launches;
  return RT_OK;
}

}  // namespace
}  // namespace rt

using namespace rt;

// TriangleDimerTiler::Walk(t, steps, seed) (TriangleDimerTiler.cpp:101-695).
extern "C" int rt_bibone_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps, long seed) {
  if (!ctx) return RT_EINVAL;
  if (!tiling || N < 6 || N % 2) {
    ctx->err = "bibone tiling must be N*N with N even and N >= 6";
    return RT_EINVAL;
  }
  if (steps < 0) {
    ctx->err = "negative step count";
    return RT_EINVAL;
  }
  const int M = N / 2;
  const size_t n = (size_t)M * M;
  int rc = ctx->check_params(n);
  if (rc) return rc;
  HIPCHK_RET(ctx, hipSetDevice(ctx->device));
  const size_t plane_bytes = n * sizeof(int32_t);
  rc = ctx->reserve(NPLANE * plane_bytes + 4 * n * sizeof(uint32_t) + 256);
  if (rc) return rc;

  // Split into the eight planes on the host (TriangleDimerTiler.cpp:122-148).
  std::vector<int> t(tiling, tiling + (size_t)N * N);
  const std::vector<int> tri = bibone_triangle_flips(t, N);
  const std::vector<int> bf = bibone_butterfly_flips(t, N);
  std::vector<int32_t> host(NPLANE * n, 0);
  auto P = [&](int pl, int i, int j) -> int32_t& { return host[pl * n + (size_t)i * M + j]; };
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int pl = (i % 2 == 0) ? (j % 2 ? 0 : -1) : (j % 2 ? 2 : 1);
      if (pl < 0) continue;
      P(EH + pl, i / 2, j / 2) = t[(size_t)i * N + j];
      P(BH + pl, i / 2, j / 2) = bf[(size_t)i * N + j];
    }
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++) P(j % 2 ? TU : TD, i, j / 2) = tri[(size_t)i * N + j];

  Bi B;
  B.M = M;
  int32_t* base = (int32_t*)ctx->arena;
  for (int pl = 0; pl < NPLANE; pl++) B.p[pl] = base + pl * n;
  uint32_t* state = (uint32_t*)(base + NPLANE * n);
  StreamTable S;
  S.state = state;
  S.param = ctx->d_param;
  S.stride = (long)n;
  S.pstride = (long)ctx->nparam;

  const std::vector<Move> sched = bibone_schedule(seed, steps);
  ctx->last = rt_timing{};
  HIPCHK_RET(ctx, hipMemcpyAsync(base, host.data(), NPLANE * plane_bytes, hipMemcpyHostToDevice,
                                 ctx->stream));
  HIPCHK_RET(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_bi_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, S,
                     (int)n, (uint32_t)seed);
  long launches = 0;
  rc = bibone_run(ctx, B, S, sched, &launches);
  if (rc) return rc;
  HIPCHK_RET(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK_RET(ctx, hipMemcpyAsync(host.data(), base, 3 * plane_bytes, hipMemcpyDeviceToHost,
                                 ctx->stream));
  HIPCHK_RET(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK_RET(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last.device_ms = ms;
  ctx->last.launches = launches + 1;
  ctx->last.steps = steps;
  ctx->last.attempts = steps * (long)n;

  // Merge the lozenge-type planes back (TriangleDimerTiler.cpp:563-577).
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int pl = (i % 2 == 0) ? (j % 2 ? 0 : -1) : (j % 2 ? 2 : 1);
      if (pl >= 0) tiling[(size_t)i * N + j] = P(EH + pl, i / 2, j / 2);
    }
  return RT_OK;
}

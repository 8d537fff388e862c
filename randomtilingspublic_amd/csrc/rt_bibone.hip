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
// Three engines: one launch per phase (ENGINE_STEP), all phases of a few
// steps per launch on LDS tiles of the eight planes (ENGINE_FUSED, and
// starts the edge form does not take), and the edge form below, one bit per
// lattice edge and one cycle flip per move (the default).
#include <hip/hip_runtime.h>

#include <random>
#include <vector>

#include "bibone_edges.hpp"
#include "rt_internal.hpp"
#include "rt_kernels.hpp"

namespace rt {
namespace {

struct Bi {
  int32_t* p[NPLANE];
  int M;
};

// Cell accessor of the one-launch-per-phase kernels: the planes in HBM.
struct GAcc {
  const Bi& B;
  __device__ int M() const { return B.M; }
  __device__ int get(int pl, int i, int j) const { return B.p[pl][i * B.M + j]; }
  __device__ void set(int pl, int i, int j, int v) const { B.p[pl][i * B.M + j] = v; }
};

template <class A>
__device__ __forceinline__ bool interior(const A& a, int i, int j) {
  return i > 0 && j > 0 && i < a.M() - 1 && j < a.M() - 1;
}

__device__ __forceinline__ int bt(int x, int k) { return (x >> k) & 1; }
__device__ __forceinline__ int fld(int x, int k) { return (x >> (2 * k)) & 3; }
__device__ __forceinline__ int mod3(int x) { return ((x % 3) + 3) % 3; }

// One draw of stream k; true when rd < 0.5, i.e. the tempered word < 2^31.
__device__ __forceinline__ bool coin(Tiny& t) { return tiny_bits(t) < kBelow05; }

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

// ------------------------------------------------------------- rotations
// RotateLozenges (triangledimerkernel.cl:4-22): sublattice parity of
// orientation o is i (H), j (L) or i+j (R); 5 <-> 10 with probability 1/2.
template <class A>
__device__ __forceinline__ void ph_rot_loz(const A& a, int i, int j, bool heads, int o, int c) {
  const int par = (o == 0 ? i : (o == 1 ? j : i + j)) & 1;
  const int v = a.get(EH + o, i, j);
  if (heads && par == c && (v == 5 || v == 10)) a.set(EH + o, i, j, 15 - v);
}

// RotateTriangles (190-211): class (j-i) mod 3; 45 <-> 54.
template <class A>
__device__ __forceinline__ void ph_rot_tri(const A& a, int i, int j, bool heads, int pl, int c) {
  const int v = a.get(pl, i, j);
  if (heads && mod3(j - i) == c && (v == 45 || v == 54)) a.set(pl, i, j, 99 - v);
}

// RotateButterflys (330-352): slice i, j or i+j (mod 3) == p1; the flip
// 85 <-> 170 happens on the cells whose parity (j for H and R, i for L)
// equals p2.
template <class A>
__device__ __forceinline__ void ph_rot_bf(const A& a, int i, int j, bool heads, int o, int p1,
                                          int p2) {
  const int slice = (o == 0 ? i : (o == 1 ? j : i + j)) % 3;
  const int q = (o == 1 ? i : j) & 1;
  const int v = a.get(BH + o, i, j);
  if (heads && slice == p1 && q == p2 && (v == 85 || v == 170)) a.set(BH + o, i, j, 255 - v);
}

// ---------------------------------------------------------- lozenge moves
// UpdateLozengesFlipped (24-34): cells of the other parity rebuilt from
// their four neighbours a, b, c, d (bit 2 of a, bit 3 of b, bit 0 of c and
// bit 1 of d, in that order).
template <class A>
__device__ __forceinline__ void ph_loz_refresh(const A& a, int i, int j, int o, int c) {
  if (!interior(a, i, j)) return;
  const int par = (o == 0 ? i : (o == 1 ? j : i + j)) & 1;
  if (par == c) return;
  const int pl = EH + o;
  const int s = o & 1;            // column shift of the a / c neighbours
  const int r = o == 0 ? 1 : 0;   // row shift of the b / d neighbours
  const int na = a.get(pl, i - 1, j + s), nb = a.get(pl, i - r, j + 1);
  const int nc = a.get(pl, i + 1, j - s), nd = a.get(pl, i + r, j - 1);
  a.set(pl, i, j, bt(na, 2) | bt(nb, 3) << 1 | bt(nc, 0) << 2 | bt(nd, 1) << 3);
}

// UpdateLozenges0/1/2 (36-79): the two other lozenge orientations take the
// dimer bits they share with the rotated orientation o.
template <class A>
__device__ __forceinline__ void ph_loz_share(const A& a, int i, int j, int o) {
  if (!interior(a, i, j)) return;
  if (o == 0) {
    const int dn = a.get(EH, i + 1, j);
    a.set(EL, i, j, (a.get(EL, i, j) & ~10) | bt(dn, 0) << 1 | bt(a.get(EH, i + 1, j - 1), 0) << 3);
    a.set(ER, i, j, (a.get(ER, i, j) & ~10) | (dn & 2) | (a.get(EH, i, j) & 8));
  } else if (o == 1) {
    const int me = a.get(EL, i, j);
    a.set(ER, i, j, (a.get(ER, i, j) & ~5) | (me & 1) | bt(a.get(EL, i + 1, j), 0) << 2);
    a.set(EH, i, j, (a.get(EH, i, j) & ~5) | bt(a.get(EL, i - 1, j), 1) | bt(me, 1) << 2);
  } else {
    const int lf = a.get(ER, i, j - 1);
    a.set(EH, i, j, (a.get(EH, i, j) & ~10) | (a.get(ER, i - 1, j) & 2) | bt(lf, 1) << 3);
    a.set(EL, i, j, (a.get(EL, i, j) & ~5) | (a.get(ER, i, j) & 1) | (lf & 4));
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
template <class A>
__device__ __forceinline__ void ph_from_loz(const A& a, int i, int j, int mask) {
  if (!interior(a, i, j)) return;
  const int h = a.get(EH, i, j), l = a.get(EL, i, j), r = a.get(ER, i, j);
  const int hd = a.get(EH, i + 1, j), lr = a.get(EL, i, j + 1);
  if (mask & 1) {
    const int L = bt(l, 0) + 2 * bt(h, 3) + 3 * bt(hd, 0);
    const int R = bt(lr, 0) + 2 * bt(hd, 1) + 3 * bt(lr, 1);
    const int V = bt(lr, 2) + 2 * bt(hd, 3) + 3 * bt(hd, 2);
    a.set(TU, i, j, L + 4 * R + 16 * V);
  }
  if (mask & 2) {
    const int ru = a.get(ER, i - 1, j), lu = a.get(EL, i - 1, j + 1);
    a.set(BH, i, j, pack4(wing(ru, 3, 0, h, 0), wing(lu, 0, 1, h, 1), wing(r, 1, 2, h, 2),
                          wing(l, 2, 3, h, 3), bt(lu, 2)));
  }
  if (mask & 4) {
    const int hdl = a.get(EH, i + 1, j - 1), rl = a.get(ER, i, j - 1);
    a.set(BL, i, j, pack4(wing(h, 0, 1, l, 0), wing(r, 1, 2, l, 1), wing(hdl, 2, 3, l, 2),
                          wing(rl, 3, 0, l, 3), bt(r, 3)));
  }
  if (mask & 8) {
    a.set(BR, i, j, pack4(wing(h, 0, 1, r, 0), wing(lr, 0, 1, r, 1), wing(hd, 2, 3, r, 2),
                          wing(l, 2, 3, r, 3), bt(lr, 3)));
  }
}

// --------------------------------------------------------- triangle moves
// tri_code(x) = 1, 2 or 3 when x's fields say the shared slot is used in
// the way UpdateTriangles (257-278) tests, summed as the reference sums.
__device__ __forceinline__ int tri_code(int x) {
  return (fld(x, 2) == 1) + 2 * (fld(x, 0) == 3) + 3 * (fld(x, 1) == 2);
}

// UpdateTriangles: dir 0 rebuilds tD from tU (vertical neighbour above),
// dir 1 rebuilds tU from tD (neighbour below).
template <class A>
__device__ __forceinline__ void ph_tri_pair(const A& a, int i, int j, int dir) {
  if (!interior(a, i, j)) return;
  const int src = dir ? TD : TU, dst = dir ? TU : TD;
  const int na = a.get(src, i, j - 1 + dir), nb = a.get(src, i, j + dir);
  const int v = a.get(src, i + (dir ? 1 : -1), j);
  a.set(dst, i, j, tri_code(na) + 4 * tri_code(nb) + 16 * tri_code(v));
}

// UpdateTrianglesFlipped0/1 (213-255): the two classes that did not rotate
// take their fields from three class-c neighbours.  For class c+1 (p = 1)
// the fields come from (f1, f2, f0) of the neighbours, for class c+2 from
// (f2, f0, f1).
template <class A>
__device__ __forceinline__ void ph_tri_refresh(const A& a, int i, int j, int down, int c) {
  if (!interior(a, i, j)) return;
  const int cls = mod3(j - i);
  if (cls == c) return;
  const int p = cls == (c + 1) % 3;
  const int pl = down ? TD : TU;
  int na, nb, nd;
  if (!down) {
    na = a.get(pl, i - 1 + p, j - p);
    nb = a.get(pl, i - p, j + 1);
    nd = a.get(pl, i + 1, j - 1 + p);
  } else {
    na = a.get(pl, i + 1 - p, j - 1);
    nb = a.get(pl, i + p, j + 1 - p);
    nd = a.get(pl, i - 1, j + p);
  }
  const int g0 = p ? fld(na, 1) : fld(na, 2);
  const int g1 = p ? fld(nb, 2) : fld(nb, 0);
  const int g2 = p ? fld(nd, 0) : fld(nd, 1);
  a.set(pl, i, j, g0 + 4 * g1 + 16 * g2);
}

// UpdateLozenge{H,L,R}FromTriangles (280-328).
template <class A>
__device__ __forceinline__ void ph_loz_from_tri(const A& a, int i, int j) {
  if (!interior(a, i, j)) return;
  const int d = a.get(TD, i, j);
  const int uu = a.get(TU, i - 1, j), ul = a.get(TU, i, j - 1), u = a.get(TU, i, j);
  a.set(EH, i, j, (fld(d, 2) == 2) | (fld(d, 2) == 3) << 1 | (fld(uu, 2) == 3) << 2 |
                      (fld(uu, 2) == 2) << 3);
  a.set(EL, i, j, (fld(ul, 1) == 1) | (fld(ul, 1) == 3) << 1 | (fld(d, 0) == 1) << 2 |
                      (fld(d, 0) == 2) << 3);
  a.set(ER, i, j, (fld(u, 0) == 1) | (fld(d, 1) == 3) << 1 | (fld(d, 1) == 1) << 2 |
                      (fld(u, 0) == 2) << 3);
}

// -------------------------------------------------------- butterfly moves
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
template <class A>
__device__ __forceinline__ void ph_bf_same(const A& a, int i, int j, int o, int p1, int p2) {
  if (!interior(a, i, j)) return;
  const int pl = BH + o;
  const int q = (o == 1 ? i : j) & 1;
  if (bf_slice(o, i, j) != p1 || q == p2) return;
  const int c = flag(a.get(pl, i, j));
  int v;
  if (o == 2) {
    const int n1 = a.get(pl, i + 1, j - 1), n2 = a.get(pl, i - 1, j + 1);
    v = pack4(rot2(fld(n2, 3)), rot1(fld(n2, 2)), rot2(fld(n1, 1)), rot1(fld(n1, 0)), c);
  } else {
    const int n1 = o == 0 ? a.get(pl, i, j - 1) : a.get(pl, i - 1, j);
    const int n2 = o == 0 ? a.get(pl, i, j + 1) : a.get(pl, i + 1, j);
    v = pack4(rot1(fld(n1, 1)), rot2(fld(n2, 0)), rot1(fld(n2, 3)), rot2(fld(n1, 2)), c);
  }
  a.set(pl, i, j, v);
}

// Adjacent slices (UpdateButterflysFlipped{H,L,R}2{1,2,3}).  Stage 1 and 3
// work on slice p1+1 (H) / p1+2 (L, R), stage 2 on the remaining one.
// H and L read the same fields from differently placed neighbours.
template <class A>
__device__ __forceinline__ void ph_bf_stage(const A& a, int i, int j, int o, int p1, int stage) {
  if (!interior(a, i, j)) return;
  const int pl = BH + o;
  const int s13 = o == 0 ? (p1 + 1) % 3 : (p1 + 2) % 3;
  const int s2 = o == 0 ? (p1 + 2) % 3 : (p1 + 1) % 3;
  const int sl = bf_slice(o, i, j);
  if (sl != (stage == 2 ? s2 : s13)) return;
  const int v = a.get(pl, i, j);
  int f0 = fld(v, 0), f1 = fld(v, 1), f2 = fld(v, 2), f3 = fld(v, 3), c = flag(v);
  if (o != 2) {
    // neighbours: u1, u2 above (H: (i-1,j),(i-1,j+1); L: (i-1,j+1),(i,j+1)),
    //             d1, d2 below (H: (i+1,j),(i+1,j-1); L: (i+1,j-1),(i,j-1)).
    const bool H = o == 0;
    if (stage == 1) {
      const int u1 = H ? a.get(pl, i - 1, j) : a.get(pl, i - 1, j + 1);
      const int u2 = H ? a.get(pl, i - 1, j + 1) : a.get(pl, i, j + 1);
      c = fld(u1, 2) == 2;
      if (f0 != 2) f0 = is3(u1, 3) + 3 * is3(u1, 2);
      if (f1 != 1) f1 = 3 * is3(u2, 3) + 2 * is3(u2, 2);
    } else if (stage == 2) {
      const int u1 = H ? a.get(pl, i - 1, j) : a.get(pl, i - 1, j + 1);
      const int u2 = H ? a.get(pl, i - 1, j + 1) : a.get(pl, i, j + 1);
      const int d1 = H ? a.get(pl, i + 1, j) : a.get(pl, i + 1, j - 1);
      const int d2 = H ? a.get(pl, i + 1, j - 1) : a.get(pl, i, j - 1);
      c = fld(d1, 0) == 2;
      const int n0 = flag(u1) ? 2 : (f0 != 2) * f0;
      const int n1 = flag(u2) ? 1 : (f1 != 1) * f1;
      if (f2 != 2) f2 = 3 * is3(d1, 0) + is3(d1, 1);
      if (f3 != 1) f3 = 2 * is3(d2, 0) + 3 * is3(d2, 1);
      f0 = n0;
      f1 = n1;
    } else {
      const int d1 = H ? a.get(pl, i + 1, j) : a.get(pl, i + 1, j - 1);
      const int d2 = H ? a.get(pl, i + 1, j - 1) : a.get(pl, i, j - 1);
      f2 = flag(d1) ? 2 : (f2 != 2) * f2;
      f3 = flag(d2) ? 1 : (f3 != 1) * f3;
    }
  } else {
    // R: neighbours up (i-1,j), right (i,j+1), down (i+1,j), left (i,j-1).
    if (stage == 1) {
      const int rt = a.get(pl, i, j + 1), dn = a.get(pl, i + 1, j);
      c = fld(dn, 0) == 1;
      if (f1 != 2) f1 = is3(rt, 0) + 3 * is3(rt, 3);
      if (f2 != 1) f2 = 3 * is3(dn, 0) + 2 * is3(dn, 3);
    } else if (stage == 2) {
      const int up = a.get(pl, i - 1, j), rt = a.get(pl, i, j + 1);
      const int dn = a.get(pl, i + 1, j), lf = a.get(pl, i, j - 1);
      c = fld(up, 2) == 1;
      if (f0 != 1) f0 = 3 * is3(up, 2) + 2 * is3(up, 1);
      const int n1 = flag(rt) ? 2 : (f1 != 2) * f1;
      const int n2 = flag(dn) ? 1 : (f2 != 1) * f2;
      if (f3 != 2) f3 = is3(lf, 2) + 3 * is3(lf, 1);
      f1 = n1;
      f2 = n2;
    } else {
      const int up = a.get(pl, i - 1, j), lf = a.get(pl, i, j - 1);
      f0 = flag(up) ? 1 : (f0 != 1) * f0;
      f3 = flag(lf) ? 2 : (f3 != 2) * f3;
    }
  }
  a.set(pl, i, j, pack4(f0, f1, f2, f3, c));
}

// UpdateLozengeFromButterflys{H,L,R} (652-749): the three lozenge planes at
// a butterfly cell of orientation o.  `full` is the lozenge of the
// butterfly's own orientation, the two others read one field pair + flag.
template <class A>
__device__ __forceinline__ void ph_loz_from_bf(const A& a, int i, int j, int o) {
  if (!interior(a, i, j)) return;
  const int v = a.get(BH + o, i, j);
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
  a.set(EH, i, j, h);
  a.set(EL, i, j, l);
  a.set(ER, i, j, r);
}

// ------------------------------------- one launch per phase (engine 1)
// Each kernel: one work item per plane cell, the phase above on it.
__device__ __forceinline__ bool cell_of(const Bi& B, int& i, int& j) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B.M * B.M) return false;
  i = k / B.M;
  j = k - i * B.M;
  return true;
}

#define BI_KERNEL(name, phase, params, ...)  \
  __global__ void name params {              \
    int i, j;                                \
    if (!cell_of(B, i, j)) return;           \
    phase(GAcc{B}, i, j, __VA_ARGS__);       \
  }

__device__ __forceinline__ bool gcoin(const StreamTable& S, int k) {
  Tiny t = tiny_load(S, k);
  const bool h = coin(t);
  tiny_store(S, k, t);
  return h;
}

__global__ void k_bi_rot_loz(Bi B, StreamTable S, int o, int c) {
  int i, j;
  if (!cell_of(B, i, j)) return;
  ph_rot_loz(GAcc{B}, i, j, gcoin(S, i * B.M + j), o, c);
}
__global__ void k_bi_rot_tri(Bi B, StreamTable S, int pl, int c) {
  int i, j;
  if (!cell_of(B, i, j)) return;
  ph_rot_tri(GAcc{B}, i, j, gcoin(S, i * B.M + j), pl, c);
}
__global__ void k_bi_rot_bf(Bi B, StreamTable S, int o, int p1, int p2) {
  int i, j;
  if (!cell_of(B, i, j)) return;
  ph_rot_bf(GAcc{B}, i, j, gcoin(S, i * B.M + j), o, p1, p2);
}
BI_KERNEL(k_bi_loz_refresh, ph_loz_refresh, (Bi B, int o, int c), o, c)
BI_KERNEL(k_bi_loz_share, ph_loz_share, (Bi B, int o), o)
BI_KERNEL(k_bi_from_loz, ph_from_loz, (Bi B, int mask), mask)
BI_KERNEL(k_bi_tri_pair, ph_tri_pair, (Bi B, int dir), dir)
BI_KERNEL(k_bi_tri_refresh, ph_tri_refresh, (Bi B, int down, int c), down, c)
__global__ void k_bi_loz_from_tri(Bi B) {
  int i, j;
  if (!cell_of(B, i, j)) return;
  ph_loz_from_tri(GAcc{B}, i, j);
}
BI_KERNEL(k_bi_bf_same, ph_bf_same, (Bi B, int o, int p1, int p2), o, p1, p2)
BI_KERNEL(k_bi_bf_stage, ph_bf_stage, (Bi B, int o, int p1, int stage), o, p1, stage)
BI_KERNEL(k_bi_loz_from_bf, ph_loz_from_bf, (Bi B, int o), o)
#undef BI_KERNEL

// ---------------------------- eight-plane sweep blocks (engine 2, fallback)
// One workgroup owns a tile of LW x LW plane cells (an interior of side
// LW - 2H plus an H-cell halo), holds all eight planes of it in LDS as
// 16-bit values (every plane value is below 512) and the TinyMT states of
// its cells in VGPRs, and runs every phase of a window of steps with one
// workgroup barrier between phases.  Every phase reads neighbours at most
// one cell away (rotations and UpdateLozengeFromButterflys none), so a step
// makes the cells within its radius -- 4 for lozenge and triangle steps, 6
// for butterfly steps -- of the tile's edge stale; a window whose radii sum
// to at most H leaves the interior exact.  Halo cells draw from their
// owners' streams, redundantly, so every draw is the reference's.  Planes
// and states are double-buffered (read one set, write the other), as in
// k_fused.
constexpr int kBiMaxK = 16;  // steps per window
// Edge-form tile geometry used when rt_set_tile_shape leaves it at 0: 64 x
// 64 cells, halo 16 (8 steps per launch), 1024 threads.  Device time per
// step on the bibone hexagon 200, 201, 200 (406 x 406 plane cells): 2.43 us
// against 2.74 for 48 x 48 / halo 8 (shape 2), 2.75 for 48 x 48 / halo 12
// (6, 289 tiles against 169), 3.22 for 64 x 64 / halo 8 (3), 3.31 for
// 40 x 40 / halo 8 (5), 3.95 for 48 x 48 / halo 12 at 576 threads (7) and
// 4.39 for 32 x 32 / halo 4 (1), all in plain tile order; the XCD order
// of the default takes it to 2.28 us.  VALU issue is a third of the chip's
// SIMD cycles: one 1024-thread workgroup per CU, two barriers per step.
constexpr int kBiEdgeDefaultShape = 4;
// Tile geometry of the walk: 48 x 48 loaded cells, halo 8 (32 x 32
// interior), 768 threads (three cells each).  Measured on the bibone
// hexagon 200, 201, 200 (406 x 406 plane cells), device time per step:
// 13.7 us against 14.7 for 64 x 64 tiles with halo 12 (1024 threads),
// 13.8 with halo 16, 17.7 for 48 x 48 with halo 12, 17.7 for 32 x 32 with
// halo 8 (256 threads); 21.8 for one launch per phase.
constexpr int kBiLW = 48, kBiH = 8, kBiThreads = 768;

__host__ __device__ constexpr int bi_radius(int type) { return type == 2 ? 6 : 4; }

// Cell accessor over an LDS tile of LW x LW cells with a one-cell zero
// margin; (r0, c0) is the plane cell of tile position (0, 0).
template <int LW>
struct TAcc {
  static constexpr int RS = LW + 2, PL = RS * RS;
  uint16_t* l;
  int r0, c0, m;
  __device__ int M() const { return m; }
  __device__ int get(int pl, int i, int j) const {
    return l[pl * PL + (i - r0 + 1) * RS + (j - c0 + 1)];
  }
  __device__ void set(int pl, int i, int j, int v) const {
    l[pl * PL + (i - r0 + 1) * RS + (j - c0 + 1)] = (uint16_t)v;
  }
};

struct BiMove {
  uint8_t type, r1, r2, p1, p2;
};

// Tile geometry: LW x LW loaded cells, halo H, T threads; thread t owns
// tile cells (t / LW + (T / LW) q, t % LW), q < LW * LW / T: one column per
// lane, rows T / LW apart.  The four state words of its cells stay in
// VGPRs; the streams' parameters, read only by the rotation draws, sit in
// LDS.
template <int LW>
__device__ __forceinline__ bool bi_coin(uint32_t* s, const uint32_t* sp, int lc) {
  Tiny t{s[0], s[1], s[2], s[3], sp[lc], sp[LW * LW + lc], sp[2 * LW * LW + lc]};
  const bool h = coin(t);
  s[0] = t.s0;
  s[1] = t.s1;
  s[2] = t.s2;
  s[3] = t.s3;
  return h;
}

template <int LW, int H, int T>
__global__ __launch_bounds__(T) void k_bi_sweep(
    const int32_t* __restrict__ pin, int32_t* __restrict__ pout, const uint32_t* __restrict__ sin,
    uint32_t* __restrict__ sout, const uint32_t* __restrict__ param, long pstride, int M,
    int tiles_c, const BiMove* __restrict__ moves, int K) {
  constexpr int TH = LW - 2 * H;  // interior rows and columns
  constexpr int kCells = LW * LW / T, kRowStep = T / LW;
  constexpr int PLS = TAcc<LW>::PL;
  static_assert(T % LW == 0 && (LW * LW) % T == 0, "tile rows split evenly over the threads");
  __shared__ uint16_t lds[NPLANE * PLS];
  __shared__ uint32_t sp[3 * LW * LW];
  __shared__ BiMove smv[kBiMaxK];
  const long n = (long)M * M;
  const int tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
  const int r0 = tr * TH - H, c0 = tc * TH - H;
  for (int k = threadIdx.x; k < NPLANE * PLS; k += T) lds[k] = 0;
  if (threadIdx.x < K) smv[threadIdx.x] = moves[threadIdx.x];
  __syncthreads();
  const TAcc<LW> a{lds, r0, c0, M};
  const int lj = threadIdx.x % LW, li0 = threadIdx.x / LW;
  const int gj = c0 + lj;
  const bool col_in = gj >= 0 && gj < M;
  uint32_t st[kCells][4];
#pragma unroll
  for (int q = 0; q < kCells; q++) {
    const int gi = r0 + li0 + kRowStep * q, lc = (li0 + kRowStep * q) * LW + lj;
    st[q][0] = st[q][1] = st[q][2] = st[q][3] = 0;
    sp[lc] = sp[LW * LW + lc] = sp[2 * LW * LW + lc] = 0;
    if (col_in && gi >= 0 && gi < M) {
      const long k = (long)gi * M + gj;
#pragma unroll
      for (int pl = 0; pl < NPLANE; pl++) a.set(pl, gi, gj, pin[pl * n + k]);
#pragma unroll
      for (int w = 0; w < 4; w++) st[q][w] = sin[w * n + k];
      sp[lc] = param[k];
      sp[LW * LW + lc] = param[pstride + k];
      sp[2 * LW * LW + lc] = param[2 * pstride + k];
    }
  }
  __syncthreads();
#define BI_PHASE(call)                                                      \
  do {                                                                      \
    _Pragma("unroll") for (int q = 0; q < kCells; q++) {                  \
      const int gi = r0 + li0 + kRowStep * q;                                     \
      if (col_in && gi >= 0 && gi < M) call;                                \
    }                                                                       \
    __syncthreads();                                                        \
  } while (0)
#define BI_COIN bi_coin<LW>(st[q], sp, (li0 + kRowStep * q) * LW + lj)
  for (int s = 0; s < K; s++) {
    const BiMove m = smv[s];
    if (m.type == 0) {
      BI_PHASE(ph_rot_loz(a, gi, gj, BI_COIN, m.r1, m.r2));
      BI_PHASE(ph_loz_refresh(a, gi, gj, m.r1, m.r2));
      BI_PHASE(ph_loz_share(a, gi, gj, m.r1));
      BI_PHASE(ph_from_loz(a, gi, gj, 15));
      BI_PHASE(ph_tri_pair(a, gi, gj, 0));
    } else if (m.type == 1) {
      const int down = m.r2 == 1;
      BI_PHASE(ph_rot_tri(a, gi, gj, BI_COIN, down ? (int)TD : (int)TU, m.r1));
      BI_PHASE(ph_tri_refresh(a, gi, gj, down, m.r1));
      BI_PHASE(ph_tri_pair(a, gi, gj, down));
      BI_PHASE(ph_loz_from_tri(a, gi, gj));
      BI_PHASE(ph_from_loz(a, gi, gj, 14));
    } else {
      const int o = m.r1;
      BI_PHASE(ph_rot_bf(a, gi, gj, BI_COIN, o, m.p1, m.p2));
      BI_PHASE(ph_bf_same(a, gi, gj, o, m.p1, m.p2));
      BI_PHASE(ph_bf_stage(a, gi, gj, o, m.p1, 1));
      BI_PHASE(ph_bf_stage(a, gi, gj, o, m.p1, 2));
      BI_PHASE(ph_bf_stage(a, gi, gj, o, m.p1, 3));
      BI_PHASE(ph_loz_from_bf(a, gi, gj, o));
      BI_PHASE(ph_from_loz(a, gi, gj, 1 | (14 & ~(2 << o))));
      BI_PHASE(ph_tri_pair(a, gi, gj, 0));
    }
  }
#undef BI_COIN
#undef BI_PHASE
#pragma unroll
  for (int q = 0; q < kCells; q++) {
    const int li = li0 + kRowStep * q, gi = r0 + li;
    if (!col_in || gi < 0 || gi >= M || li < H || li >= H + TH || lj < H ||
        lj >= H + TH)
      continue;
    const long k = (long)gi * M + gj;
#pragma unroll
    for (int pl = 0; pl < NPLANE; pl++) pout[pl * n + k] = a.get(pl, gi, gj);
#pragma unroll
    for (int w = 0; w < 4; w++) sout[w * n + k] = st[q][w];
  }
}

// ------------------------------------------------ edge form (engine 0, 3)
// The eight planes above are redundant: on every tiling the walk reaches
// from a consistent start, each plane value is a function of the dimer
// cover, i.e. of one occupation bit per edge of the triangular lattice
// (bibone_edges.hpp), and each move of TriangleDimerTiler::Walk flips one
// cycle of that cover: a lozenge rotation (5 <-> 10) a 4-cycle, a triangle
// rotation (45 <-> 54) a 6-cycle, a butterfly rotation (85 <-> 170) an
// 8-cycle.  The reference's refresh kernels only copy those edge changes
// into the other planes.  The edge-form walk keeps the three edge families
// as three LDS byte planes of a tile (one byte per edge, 0 or 1) and runs a
// step as: one draw per cell of the array, the rotation test of the heads
// cells in the step's class (are the edges of cycle state A, or of state
// B, all set), a barrier, the flipped cycles' edges written with their new
// values (cycles of one class share no edge, so plain byte stores do not
// race), a barrier.  A cell reads and writes edges at most one cell away,
// so a step makes two cells of the tile's edge stale and a window of H / 2
// steps leaves the interior of an H-cell halo exact.  HBM keeps one byte
// per plane cell (bit 0 h, 1 l, 2 r) and the TinyMT state.  The walk takes
// this form when bibone_edge_form accepts the start (rt_bibone_host.cpp);
// other starts run on the eight planes.
// Per-cell lattice classes, computed once: bit 0 i & 1, bit 1 j & 1,
// bit 2 (i + j) & 1, bits 3-4 (j - i) mod 3, bits 5-6 i mod 3, bits 7-8
// j mod 3, bits 9-10 (i + j) mod 3.
__device__ __forceinline__ uint32_t edge_classes(int i, int j) {
  return (uint32_t)(i & 1) | (uint32_t)(j & 1) << 1 | (uint32_t)((i + j) & 1) << 2 |
         (uint32_t)mod3(j - i) << 3 | (uint32_t)(i % 3) << 5 | (uint32_t)(j % 3) << 7 |
         (uint32_t)((i + j) % 3) << 9;
}

// Code of a move: 0..2 lozenge of orientation r1, 3 / 4 triangle up / down,
// 5..7 butterfly of orientation r1.
__host__ __device__ inline int edge_code(const BiMove& m) {
  return m.type == 0 ? m.r1 : (m.type == 1 ? 3 + (m.r2 == 1) : 5 + m.r1);
}

// On a matching every rotation test is "these edges all hold a dimer": a
// lozenge is 5 when its bits 0 and 2 are set (bits 1 and 3 share their
// vertices, so they are clear), a triangle cell 45 when its fields' slots
// 1, 3, 2 are set, a butterfly 85 when every wing's slot 1 is set (the flag
// edge shares a vertex with a wing).  EdgeCycle lists the move's cycle:
// n / 2 edges of state A (5, 45, 85) then n / 2 of state B (10, 54, 170); a
// flip sets the other state's edges and clears its own.
struct EdgeCycle {
  int n;
  ESlot s[8];
};
__host__ __device__ constexpr EdgeCycle edge_cycle(int code) {
  if (code < 3) {
    const ESlot* l = kELoz[code];
    return {4, {l[0], l[2], l[1], l[3]}};
  }
  if (code < 5) {
    const auto& f = kETri[code - 3];
    return {6, {f[0][0], f[1][2], f[2][1], f[0][1], f[1][0], f[2][2]}};
  }
  const auto& f = kEBf[code - 5];
  return {8, {f[0][0], f[1][0], f[2][0], f[3][0], f[0][1], f[1][1], f[2][1], f[3][1]}};
}

// Whether a cell with classes `cl` is in the class of move m.
template <int CODE>
__device__ __forceinline__ bool edge_in_class(uint32_t cl, const BiMove& m) {
  if constexpr (CODE < 3) {
    return (int)((cl >> (CODE == 0 ? 0 : (CODE == 1 ? 1 : 2))) & 1) == m.r2;
  } else if constexpr (CODE < 5) {
    return (int)((cl >> 3) & 3) == m.r1;
  } else {
    constexpr int o = CODE - 5;
    const int slice = (cl >> (o == 0 ? 5 : (o == 1 ? 7 : 9))) & 3;
    const int q = (cl >> (o == 1 ? 0 : 1)) & 1;
    return slice == m.p1 && q == m.p2;
  }
}

template <int LW, int H, int T>
struct EdgeTile {
  static constexpr int RS = LW + 2;                 // LDS row: one margin cell each side
  static constexpr int PL = RS * RS;                // one edge family (h, l, r) per plane
  static constexpr int kBytes = (3 * PL + 3) & ~3;
  static constexpr int kCells = LW * LW / T, kRowStep = T / LW;
  static constexpr int TH = LW - 2 * H;             // interior side
  static constexpr int K = H / 2;                   // steps per window
  static_assert(T % LW == 0 && (LW * LW) % T == 0, "tile rows split evenly over the threads");
  static_assert(H % 2 == 0 && K <= kBiMaxK, "halo covers whole steps");
  // LDS offset of slot s from the byte of cell (li - 1, lj - 1), h plane
  static constexpr int off(const ESlot& s) { return s.bit * PL + (s.di + 1) * RS + (s.dj + 1); }
};

// Rotation tests of move class CODE on the thread's cells: bit q of the
// result is set when cell q flips (heads, in the class, cycle in state A
// or B); bit 16 + q says it was in state A.  `base` is the LDS byte of
// cell 0's up-left neighbour in the h plane.
template <int CODE, int LW, int H, int T>
__device__ __forceinline__ uint32_t edge_decide(const uint8_t* base, uint32_t heads,
                                                const uint32_t (&cls)[EdgeTile<LW, H, T>::kCells],
                                                const BiMove& m) {
  using G = EdgeTile<LW, H, T>;
  constexpr EdgeCycle c = edge_cycle(CODE);
  uint32_t flip = 0;
#pragma unroll
  for (int q = 0; q < G::kCells; q++) {
    const uint8_t* b = base + q * G::kRowStep * G::RS;
    uint32_t a = 1, bb = 1;
#pragma unroll
    for (int k = 0; k < c.n / 2; k++) {
      a &= b[G::off(c.s[k])];
      bb &= b[G::off(c.s[c.n / 2 + k])];
    }
    if (((heads >> q) & 1) && edge_in_class<CODE>(cls[q], m) && (a | bb))
      flip |= 1u << q | a << (16 + q);
  }
  return flip;
}

template <int CODE, int LW, int H, int T>
__device__ __forceinline__ void edge_apply(uint8_t* base, uint32_t flip) {
  using G = EdgeTile<LW, H, T>;
  constexpr EdgeCycle c = edge_cycle(CODE);
#pragma unroll
  for (int q = 0; q < G::kCells; q++) {
    if (!((flip >> q) & 1)) continue;
    uint8_t* b = base + q * G::kRowStep * G::RS;
    const uint8_t was_a = (uint8_t)((flip >> (16 + q)) & 1);
#pragma unroll
    for (int k = 0; k < c.n / 2; k++) {
      b[G::off(c.s[k])] = was_a ^ 1;
      b[G::off(c.s[c.n / 2 + k])] = was_a;
    }
  }
}

template <int LW, int H, int T>
__global__ __launch_bounds__(T) void k_bi_edge(const uint8_t* __restrict__ ein,
                                               uint8_t* __restrict__ eout,
                                               const uint32_t* __restrict__ sin,
                                               uint32_t* __restrict__ sout,
                                               const uint32_t* __restrict__ param, long pstride,
                                               long sstride, int M, int tiles_c,
                                               const BiMove* __restrict__ moves, int K, int xcd) {
  using G = EdgeTile<LW, H, T>;
  __shared__ uint32_t ew[G::kBytes / 4];
  __shared__ BiMove smv[kBiMaxK];
  uint8_t* e = (uint8_t*)ew;
  // XCD order: block b runs on XCD b % 8; XCD x takes the x-th contiguous
  // eighth of the tiles, so tiles that read each other's halo share an L2
  // (the grid is rounded up to a multiple of 8; surplus blocks leave).
  unsigned tile = blockIdx.x;
  if (xcd) tile = (tile & 7u) * (gridDim.x >> 3) + (tile >> 3);
  if (tile >= (unsigned)(tiles_c * tiles_c)) return;
  const int tr = (int)tile / tiles_c, tc = (int)tile % tiles_c;
  const int r0 = tr * G::TH - H, c0 = tc * G::TH - H;
  for (int k = threadIdx.x; k < G::kBytes / 4; k += T) ew[k] = 0;
  if (threadIdx.x < K) smv[threadIdx.x] = moves[threadIdx.x];
  __syncthreads();
  const int lj = threadIdx.x % LW, li0 = threadIdx.x / LW;
  const int gj = c0 + lj;
  // byte of cell (li0, lj) in the h plane, and of its up-left neighbour
  const int p0 = (li0 + 1) * G::RS + lj + 1;
  uint8_t* base = e + p0 - G::RS - 1;
  uint32_t st[G::kCells][4], pm[G::kCells][3], cls[G::kCells];
  uint32_t valid = 0;
#pragma unroll
  for (int q = 0; q < G::kCells; q++) {
    const int gi = r0 + li0 + G::kRowStep * q;
    st[q][0] = st[q][1] = st[q][2] = st[q][3] = 0;
    pm[q][0] = pm[q][1] = pm[q][2] = 0;
    cls[q] = 0;
    if (gj >= 0 && gj < M && gi >= 0 && gi < M) {
      valid |= 1u << q;
      const long k = (long)gi * M + gj;
      const uint32_t v = ein[k];
      const int p = p0 + q * G::kRowStep * G::RS;
      e[p] = v & 1;
      e[G::PL + p] = (v >> 1) & 1;
      e[2 * G::PL + p] = (v >> 2) & 1;
#pragma unroll
      for (int w = 0; w < 4; w++) st[q][w] = sin[w * sstride + k];
      pm[q][0] = param[k];
      pm[q][1] = param[pstride + k];
      pm[q][2] = param[2 * pstride + k];
      cls[q] = edge_classes(gi, gj);
    }
  }
  __syncthreads();
  for (int s = 0; s < K; s++) {
    const BiMove m = smv[s];
    // one draw per cell of the array, whatever the move
    uint32_t heads = 0;
#pragma unroll
    for (int q = 0; q < G::kCells; q++) {
      Tiny t{st[q][0], st[q][1], st[q][2], st[q][3], pm[q][0], pm[q][1], pm[q][2]};
      if (coin(t) && ((valid >> q) & 1)) heads |= 1u << q;
      st[q][0] = t.s0;
      st[q][1] = t.s1;
      st[q][2] = t.s2;
      st[q][3] = t.s3;
    }
    const int code = edge_code(m);
    uint32_t flip;
    switch (code) {
      case 0: flip = edge_decide<0, LW, H, T>(base, heads, cls, m); break;
      case 1: flip = edge_decide<1, LW, H, T>(base, heads, cls, m); break;
      case 2: flip = edge_decide<2, LW, H, T>(base, heads, cls, m); break;
      case 3: flip = edge_decide<3, LW, H, T>(base, heads, cls, m); break;
      case 4: flip = edge_decide<4, LW, H, T>(base, heads, cls, m); break;
      case 5: flip = edge_decide<5, LW, H, T>(base, heads, cls, m); break;
      case 6: flip = edge_decide<6, LW, H, T>(base, heads, cls, m); break;
      default: flip = edge_decide<7, LW, H, T>(base, heads, cls, m); break;
    }
    __syncthreads();
    switch (code) {
      case 0: edge_apply<0, LW, H, T>(base, flip); break;
      case 1: edge_apply<1, LW, H, T>(base, flip); break;
      case 2: edge_apply<2, LW, H, T>(base, flip); break;
      case 3: edge_apply<3, LW, H, T>(base, flip); break;
      case 4: edge_apply<4, LW, H, T>(base, flip); break;
      case 5: edge_apply<5, LW, H, T>(base, flip); break;
      case 6: edge_apply<6, LW, H, T>(base, flip); break;
      default: edge_apply<7, LW, H, T>(base, flip); break;
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < G::kCells; q++) {
    const int li = li0 + G::kRowStep * q, gi = r0 + li;
    if (!((valid >> q) & 1) || li < H || li >= H + G::TH || lj < H || lj >= H + G::TH) continue;
    const long k = (long)gi * M + gj;
    const int p = p0 + q * G::kRowStep * G::RS;
    eout[k] = (uint8_t)(e[p] | e[G::PL + p] << 1 | e[2 * G::PL + p] << 2);
#pragma unroll
    for (int w = 0; w < 4; w++) sout[w * sstride + k] = st[q][w];
  }
}

// ------------------------------------------------------------------ host
using Move = BiMove;

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

// Sweep-block walk: windows of consecutive steps whose radii sum to at most
// H (and at most kBiMaxK steps), one launch per window over all tiles,
// reading one buffer set and writing the other.  `a` and `b` are the two
// sets (eight planes then four state words per cell); returns the set that
// holds the result.
template <int LW, int H, int T>
int bibone_sweep(rt_ctx* ctx, int M, int32_t* a, int32_t* b, BiMove* d_moves,
                 const std::vector<Move>& sched, int32_t** result, long* launches_out) {
  constexpr int TH = LW - 2 * H;
  const long n = (long)M * M;
  const int tiles = (M + TH - 1) / TH;
  int32_t* cur = a;
  int32_t* nxt = b;
  long launches = 0;
  size_t s0 = 0;
  while (s0 < sched.size()) {
    int K = 0, rad = 0;
    while (s0 + K < sched.size() && K < kBiMaxK && rad + bi_radius(sched[s0 + K].type) <= H) {
      rad += bi_radius(sched[s0 + K].type);
      K++;
    }
    hipLaunchKernelGGL((k_bi_sweep<LW, H, T>), dim3((unsigned)(tiles * tiles)), dim3(T), 0,
                       ctx->stream, cur, nxt, (const uint32_t*)(cur + NPLANE * n),
                       (uint32_t*)(nxt + NPLANE * n), ctx->d_param, (long)ctx->nparam, M, tiles,
                       d_moves + s0, K);
    launches++;
    std::swap(cur, nxt);
    s0 += K;
  }
  HIPCHK_RET(ctx, hipGetLastError());
  *result = cur;
  *launches_out = launches;
  return RT_OK;
}

// Edge-form walk: windows of G::K steps, one launch per window over all
// tiles; `a` / `b` are the two buffer sets (edge bytes, then four TinyMT
// words per cell at offset `soff`).  Returns the set holding the result.
template <int LW, int H, int T>
int bibone_edge_sweep(rt_ctx* ctx, int M, uint8_t* a, uint8_t* b, size_t soff, BiMove* d_moves,
                      long steps, uint8_t** result, long* launches_out) {
  using G = EdgeTile<LW, H, T>;
  const long n = (long)M * M;
  const int tiles = (M + G::TH - 1) / G::TH;
  const unsigned ntiles = (unsigned)(tiles * tiles);
  const int xcd = ctx->xcd_order == 2 || (ctx->xcd_order == 1 && ntiles <= (unsigned)ctx->num_cu);
  const unsigned grid = xcd ? (ntiles + 7u) & ~7u : ntiles;
  uint8_t* cur = a;
  uint8_t* nxt = b;
  long launches = 0;
  for (long s0 = 0; s0 < steps; s0 += G::K) {
    const int K = (int)std::min<long>(G::K, steps - s0);
    hipLaunchKernelGGL((k_bi_edge<LW, H, T>), dim3(grid), dim3(T), 0, ctx->stream, cur, nxt,
                       (const uint32_t*)(cur + soff), (uint32_t*)(nxt + soff), ctx->d_param,
                       (long)ctx->nparam, n, M, tiles, d_moves + s0, K, xcd);
    launches++;
    std::swap(cur, nxt);
  }
  HIPCHK_RET(ctx, hipGetLastError());
  *result = cur;
  *launches_out = launches;
  return RT_OK;
}

// Tile geometries of the edge form, by rt_set_tile_shape value (0: the
// default).
int bibone_edge_run(rt_ctx* ctx, int shape, int M, uint8_t* a, uint8_t* b, size_t soff,
                    BiMove* d_moves, long steps, uint8_t** result, long* launches) {
  switch (shape) {
    case 1: return bibone_edge_sweep<32, 4, 256>(ctx, M, a, b, soff, d_moves, steps, result, launches);
    case 2: return bibone_edge_sweep<48, 8, 768>(ctx, M, a, b, soff, d_moves, steps, result, launches);
    case 3: return bibone_edge_sweep<64, 8, 1024>(ctx, M, a, b, soff, d_moves, steps, result, launches);
    case 5: return bibone_edge_sweep<40, 8, 320>(ctx, M, a, b, soff, d_moves, steps, result, launches);
    case 6: return bibone_edge_sweep<48, 12, 768>(ctx, M, a, b, soff, d_moves, steps, result, launches);
    case 7: return bibone_edge_sweep<48, 12, 576>(ctx, M, a, b, soff, d_moves, steps, result, launches);
    default: return bibone_edge_sweep<64, 16, 1024>(ctx, M, a, b, soff, d_moves, steps, result, launches);
  }
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
  // two buffer sets (the sweep blocks read one and write the other) of
  // eight planes + four TinyMT words per cell, then the move list; the
  // edge form uses the front of each set (edge bytes, then the state)
  const size_t set_bytes = (NPLANE + 4) * plane_bytes;
  const size_t move_bytes = ((size_t)steps * sizeof(Move) + 255) & ~(size_t)255;
  rc = ctx->reserve(2 * set_bytes + move_bytes + 512);
  if (rc) return rc;

  // The eight planes of the start (TriangleDimerTiler.cpp:122-148) and,
  // where it applies, its edge form.
  const std::vector<int> t(tiling, tiling + (size_t)N * N);
  std::vector<int32_t> host = bibone_planes(t, N);
  const bool edge_engine = ctx->engine == RT_ENGINE_AUTO || ctx->engine == RT_ENGINE_FUSED_SWAR;
  std::vector<uint8_t> edges;
  if (edge_engine) edges = bibone_edge_form(host, M);

  uint8_t* set_a = (uint8_t*)ctx->arena;
  uint8_t* set_b = set_a + set_bytes;
  const size_t soff = (n + 255) & ~(size_t)255;  // edge form: state after the edge bytes
  StreamTable S;
  S.state = (uint32_t*)(edges.empty() ? set_a + NPLANE * plane_bytes : set_a + soff);
  S.param = ctx->d_param;
  S.stride = (long)n;
  S.pstride = (long)ctx->nparam;

  const std::vector<Move> sched = bibone_schedule(seed, steps);
  ctx->last = rt_timing{};
  if (edges.empty())
    HIPCHK_RET(ctx, hipMemcpyAsync(set_a, host.data(), NPLANE * plane_bytes,
                                   hipMemcpyHostToDevice, ctx->stream));
  else
    HIPCHK_RET(ctx, hipMemcpyAsync(set_a, edges.data(), n, hipMemcpyHostToDevice, ctx->stream));
  BiMove* d_moves = (BiMove*)(set_a + 2 * set_bytes);
  if (steps > 0)
    HIPCHK_RET(ctx, hipMemcpyAsync(d_moves, sched.data(), (size_t)steps * sizeof(Move),
                                   hipMemcpyHostToDevice, ctx->stream));
  HIPCHK_RET(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_bi_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, S,
                     (int)n, (uint32_t)seed);
  long launches = 0;
  int shape = 0;
  if (!edges.empty()) {
    shape = ctx->tile_shape ? ctx->tile_shape : kBiEdgeDefaultShape;
    uint8_t* result = set_a;
    rc = bibone_edge_run(ctx, shape, M, set_a, set_b, soff, d_moves, steps, &result, &launches);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK_RET(ctx, hipMemcpyAsync(edges.data(), result, n, hipMemcpyDeviceToHost, ctx->stream));
  } else {
    Bi B;
    B.M = M;
    for (int pl = 0; pl < NPLANE; pl++) B.p[pl] = (int32_t*)set_a + pl * n;
    int32_t* result = (int32_t*)set_a;
    if (ctx->engine == RT_ENGINE_STEP)
      rc = bibone_run(ctx, B, S, sched, &launches);
    else
      rc = bibone_sweep<kBiLW, kBiH, kBiThreads>(ctx, M, (int32_t*)set_a, (int32_t*)set_b,
                                                 d_moves, sched, &result, &launches);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK_RET(ctx, hipMemcpyAsync(host.data(), result, 3 * plane_bytes, hipMemcpyDeviceToHost,
                                   ctx->stream));
  }
  HIPCHK_RET(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK_RET(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last.device_ms = ms;
  ctx->last.launches = launches + 1;
  ctx->last.steps = steps;
  ctx->last.attempts = steps * (long)n;
  ctx->last.shape = shape;

  // Merge the lozenge-type cells back (TriangleDimerTiler.cpp:563-577).
  if (!edges.empty()) {
    bibone_edge_merge(edges, M, tiling);
    return RT_OK;
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int o = (i % 2 == 0) ? (j % 2 ? 0 : -1) : (j % 2 ? 2 : 1);
      if (o >= 0) tiling[(size_t)i * N + j] = host[o * n + (size_t)(i / 2) * M + j / 2];
    }
  return RT_OK;
}

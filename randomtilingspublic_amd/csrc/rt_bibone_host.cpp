// rt_bibone_host.cpp -- host statics of the bibone (triangular-lattice
// dimer) model: start tilings, the derived flip-state arrays, vertices,
// domain and dimer indicators (TriangleDimerTiler.cpp:35-95, 697-982).
//
// Tiling format (the reference's): an N x N int array of 4-bit cell
// states.  Cells at (even row, odd col) are "H" lozenge cells, (odd, even)
// "L" cells and (odd, odd) "R" cells; bit k says whether the k-th of the
// cell's four dimer slots holds a dimer.  The derived arrays are
//   triangle flips: (N/2) x N, 6-bit states (three 2-bit fields);
//   butterfly flips: N x N, 9-bit states (four 2-bit fields + a flag).
#include <algorithm>
#include <cmath>
#include <vector>

#include "bibone_edges.hpp"
#include "rt_internal.hpp"

namespace rt {

namespace {

inline int bit(int x, int k) { return (x >> k) & 1; }

// Three dimer-edge families on an E x E grid: h (horizontal), l (up-left),
// r (up-right).  Vertex-state cells are read off the edges around them
// (TriangleDimerTiler.cpp:738-753, 838-852).
struct Edges {
  int E;
  std::vector<int> h, l, r;
  explicit Edges(int e) : E(e), h(e * e, 0), l(e * e, 0), r(e * e, 0) {}
  int H(int i, int j) const { return h[i * E + j]; }
  int L(int i, int j) const { return l[i * E + j]; }
  int R(int i, int j) const { return r[i * E + j]; }
};

tiling edges_to_tiling(const Edges& e) {
  const int T = 2 * e.E;
  tiling t((size_t)T * T, 0);
  for (int i = 3; i < T - 1; i++) {
    const int a = i / 2;
    for (int j = 3; j < T - 1; j++) {
      const int b = j / 2;
      int v = 0;
      if ((i & 1) == 0) {
        if (j & 1) v = e.R(a - 1, b) | e.L(a - 1, b + 1) << 1 | e.R(a, b) << 2 | e.L(a, b) << 3;
      } else if ((j & 1) == 0) {
        v = e.H(a, b) | e.R(a, b) << 1 | e.H(a + 1, b - 1) << 2 | e.R(a, b - 1) << 3;
      } else {
        v = e.H(a, b) | e.L(a, b + 1) << 1 | e.H(a + 1, b) << 2 | e.L(a, b) << 3;
      }
      t[(size_t)i * T + j] = v;
    }
  }
  return t;
}

}  // namespace

// IceCreamCone(N, M) (TriangleDimerTiler.cpp:697-757): downward triangle of
// side N with a side-M triangle cut from its top-left corner, covered by
// horizontal dimers on alternate vertices of each row and a fringe of l / r
// dimers along the right edge.  Untileable sizes are built the same way
// (the reference only prints a warning).
tiling bibone_ice_cream_cone(int N, int M) {
  const int E = N + 4;
  const int k = ((N + 1) * (N + 2) / 2) % 2;
  Edges e(E);
  for (int i = 2; i < E - 2; i++) {
    for (int j = 2; j < E - i; j++) {
      const bool right_edge = j == E - i - 1;
      const int at = i * E + j;
      if (i < M + 2) {
        if (M != 0 && (i + j) % 2 == M % 2 && j > M + 3 - i) e.h[at] = 1;
        if ((N - M) % 2 == 0 && i % 2 == 0 && right_edge) e.r[at] = 1;
      } else {
        if (j % 2 == 0) e.h[at] = 1;
        if (right_edge && i % 4 == (3 + N % 2 + 2 * k) % 4) {
          e.h[at] = 0;
          e.l[at] = 1;
        }
        if (right_edge && i % 4 == (2 + N % 2 + 2 * k) % 4) e.r[at] = 1;
      }
    }
  }
  return edges_to_tiling(e);
}

// Hexagon with sides a, c, b, a, c, b (clockwise from the top edge), with a
// dimer cover.  No reference counterpart (the reference builds only
// IceCreamCone).  In the vertex coordinates of TilingToVertices a vertex
// (i, j) neighbours (i, j+-1), (i+-1, j), (i-1, j+1) and (i+1, j-1); edge
// h(i, j) joins (i, j)-(i, j+1), l(i, j) joins (i, j)-(i+1, j) and r(i, j)
// joins (i, j+1)-(i+1, j).  Vertex rows i = 0..b+c hold the columns
// L(i)..R(i), L(i) = -min(i, b), R(i) = a - max(0, i - c), placed at offset
// (2, 2 + b) of the edge grid.  The row ends of consecutive rows are always
// adjacent (L and R move by 0 or -1 per row), so the cover is built
// top-down: a row with an odd number of unmatched vertices sends one free
// end down to the end of the next row, and the rest pairs horizontally.
// Returns an empty tiling when the vertex count ab + bc + ca + a + b + c + 1
// is odd (no cover exists) or a side is negative.
tiling bibone_hexagon(int a, int b, int c) {
  if (a < 0 || b < 0 || c < 0) return tiling();
  const long nv = (long)a * b + (long)b * c + (long)c * a + a + b + c + 1;
  if (nv % 2) return tiling();
  const int rows = b + c + 1;
  const int E = std::max(rows, a + b + 1) + 4;
  Edges e(E);
  auto L = [&](int i) { return -std::min(i, b); };
  auto R = [&](int i) { return a - std::max(0, i - c); };
  auto at = [&](int i, int j) { return (i + 2) * E + (j + 2 + b); };
  // Join (i, j) to (i+1, j2), j2 = j or j - 1.
  auto down = [&](int i, int j, int j2) {
    if (j2 == j) e.l[at(i, j)] = 1;
    else e.r[at(i, j2)] = 1;
  };
  int lf = 0, rf = 0;  // row ends already matched from the row above
  for (int i = 0; i < rows; i++) {
    int lo = L(i) + lf, hi = R(i) - rf;
    int nlf = 0, nrf = 0;
    if ((hi - lo + 1) % 2) {
      if (i + 1 >= rows) return tiling();  // cannot happen when nv is even
      if (!lf) {
        down(i, lo, L(i + 1));
        lo++;
        nlf = 1;
      } else {
        down(i, hi, R(i + 1));
        hi--;
        nrf = 1;
      }
    }
    for (int j = lo; j < hi; j += 2) e.h[at(i, j)] = 1;
    lf = nlf;
    rf = nrf;
  }
  return edges_to_tiling(e);
}

// TestButterfly(type) (TriangleDimerTiler.cpp:795-844): one butterfly of
// orientation 0 (H), 1 (L) or 2 (R) on a 16 x 16 tiling; any other type
// gives the all-zero tiling.
tiling bibone_test_butterfly(int type) {
  Edges e(8);
  auto set = [&](std::vector<int>& f, int i, int j) { f[i * 8 + j] = 1; };
  if (type == 0) {
    set(e.h, 2, 3), set(e.h, 4, 3), set(e.r, 2, 4), set(e.r, 3, 2);
  } else if (type == 1) {
    set(e.r, 2, 3), set(e.r, 4, 2), set(e.l, 3, 2), set(e.l, 3, 4);
  } else if (type == 2) {
    set(e.l, 2, 4), set(e.l, 4, 3), set(e.r, 3, 2), set(e.r, 3, 4);
  } else {
    return tiling(256, 0);
  }
  return edges_to_tiling(e);
}

// TilingToTriangleFlips (TriangleDimerTiler.cpp:846-873).  Output (N/2) x N;
// even columns are down triangles, odd columns up triangles; each is three
// 2-bit fields, field k = which of its 3 edges (1..3) the k-th corner's
// dimer uses, 0 = none.  The outer ring stays 0 (the tiling is padded).
std::vector<int> bibone_triangle_flips(const tiling& t, int N) {
  std::vector<int> out((size_t)(N / 2) * N, 0);
  auto T = [&](int i, int j) { return t[(size_t)i * N + j]; };
  for (int i = 1; i < N / 2 - 1; i++) {
    const int r0 = 2 * i + 1;
    for (int j = 1; j < N - 1; j++) {
      const int a = T(r0, j), b = T(r0, j + 1);
      int f0, f1, f2;
      if (j % 2 == 0) {
        const int u = T(2 * i, j + 1);
        f0 = bit(a, 2) + 2 * bit(a, 3) + 3 * bit(b, 3);
        f1 = bit(b, 2) + 2 * bit(a, 1) + 3 * bit(b, 1);
        f2 = 2 * bit(u, 0) + bit(a, 0) + 3 * bit(u, 1);
      } else {
        const int d = T(2 * i + 2, j);
        f0 = bit(a, 0) + 2 * bit(a, 3) + 3 * bit(b, 3);
        f1 = bit(b, 0) + 2 * bit(a, 1) + 3 * bit(b, 1);
        f2 = 2 * bit(d, 3) + 3 * bit(d, 2) + bit(a, 2);
      }
      out[(size_t)i * N + j] = f0 + 4 * f1 + 16 * f2;  // sums, as the reference packs
    }
  }
  return out;
}

// TilingToButterflyFlips (TriangleDimerTiler.cpp:875-921).  Output N x N,
// filled at the H, L and R cells at distance >= 2 from the border: four
// 2-bit wing fields plus bit 8.
std::vector<int> bibone_butterfly_flips(const tiling& t, int N) {
  std::vector<int> out((size_t)N * N, 0);
  auto T = [&](int i, int j) { return t[(size_t)i * N + j]; };
  // wing(x, a, b, own, k) = bit a of x + 2 * bit b of x + 3 * bit k of own
  auto wing = [](int x, int a, int b, int own, int k) {
    return bit(x, a) + 2 * bit(x, b) + 3 * bit(own, k);
  };
  for (int i = 2; i < N - 2; i++) {
    for (int j = 2; j < N - 2; j++) {
      const int o = T(i, j);
      int f0, f1, f2, f3, c;
      if (i % 2 == 0) {
        if (j % 2 == 0) continue;
        const int up = T(i - 1, j), upr = T(i - 1, j + 1), dn = T(i + 1, j), dnl = T(i + 1, j - 1);
        f0 = wing(up, 3, 0, o, 0);
        f1 = wing(upr, 0, 1, o, 1);
        f2 = wing(dn, 1, 2, o, 2);
        f3 = wing(dnl, 2, 3, o, 3);
        c = bit(upr, 2);
      } else if (j % 2 == 0) {
        const int upr = T(i - 1, j + 1), rt = T(i, j + 1), dnl = T(i + 1, j - 1), lf = T(i, j - 1);
        f0 = wing(upr, 0, 1, o, 0);
        f1 = wing(rt, 1, 2, o, 1);
        f2 = wing(dnl, 2, 3, o, 2);
        f3 = wing(lf, 3, 0, o, 3);
        c = bit(rt, 3);
      } else {
        const int up = T(i - 1, j), rt = T(i, j + 1), dn = T(i + 1, j), lf = T(i, j - 1);
        f0 = wing(up, 0, 1, o, 0);
        f1 = wing(rt, 0, 1, o, 1);
        f2 = wing(dn, 2, 3, o, 2);
        f3 = wing(lf, 2, 3, o, 3);
        c = bit(rt, 3);
      }
      out[(size_t)i * N + j] = f0 + 4 * f1 + 16 * f2 + 64 * f3 + 256 * c;
    }
  }
  return out;
}

// TilingToVertices (TriangleDimerTiler.cpp:923-949): (N/2+1)^2 flags, 1 for
// a lattice vertex touched by one of its six dimer slots.
std::vector<int> bibone_tiling_to_vertices(const tiling& t, int N) {
  const int M = N / 2 + 1;
  std::vector<int> v((size_t)M * M, 0);
  auto T = [&](int i, int j) { return t[(size_t)i * N + j]; };
  for (int i = 2; i < M - 2; i++) {
    for (int j = 2; j < M - 2; j++) {
      const int a = T(2 * (i - 1), 2 * j + 1), b = T(2 * (i + 1), 2 * j - 1);
      const int c = T(2 * i - 1, 2 * j), d = T(2 * i - 1, 2 * j + 1);
      const bool touched = (a & 12) || (b & 3) || (c & 4) || (d & 4);
      v[(size_t)i * M + j] = touched ? 1 : 0;
    }
  }
  return v;
}

// TilingToDomain (TriangleDimerTiler.cpp:951-982): (M-1) x 2(M-1) triangle
// faces, M = N/2+1; a face is in the domain when its three corners are.
domain bibone_tiling_to_domain(const tiling& t, int N) {
  const int M = N / 2 + 1;
  const std::vector<int> v = bibone_tiling_to_vertices(t, N);
  auto V = [&](int i, int j) { return v[(size_t)i * M + j] == 1; };
  const int W = 2 * (M - 1);
  domain d((size_t)(M - 1) * W, 0);
  for (int i = 0; i < M - 1; i++) {
    for (int j = 0; j < W; j++) {
      const int b = j / 2;
      const bool in = (j % 2 == 0) ? (V(i, b) && V(i, b + 1) && V(i + 1, b))
                                   : (V(i, b + 1) && V(i + 1, b) && V(i + 1, b + 1));
      d[(size_t)i * W + j] = in ? 1 : 0;
    }
  }
  return d;
}

// DimerIndicator(t, c) (TriangleDimerTiler.cpp:35-95): (N/2)^2 flags.  c = 0
// reads bit 0 of the (odd, odd) cells; c = 1 and c = 2 both read the
// (even, odd) cells, bit 3 and bit 2 respectively (as the reference does);
// any other c gives all zeros.
std::vector<int> bibone_dimer_indicator(const tiling& t, int N, int c) {
  const int M = N / 2;
  std::vector<int> out((size_t)M * M, 0);
  if (c < 0 || c > 2) return out;
  const int row_par = c == 0 ? 1 : 0;
  const int b = c == 0 ? 0 : (c == 1 ? 3 : 2);
  for (int i = row_par; i < N; i += 2)
    for (int j = 1; j < N; j += 2) out[(size_t)(i / 2) * M + j / 2] = bit(t[(size_t)i * N + j], b);
  return out;
}

// The eight planes the walk runs on (TriangleDimerTiler.cpp:122-148):
// eH, eL, eR from the tiling's (even, odd), (odd, even) and (odd, odd)
// cells, tU / tD from the triangle flips (odd / even columns), bH, bL, bR
// from the butterfly flips at the lozenge cells.  8 x M x M, M = N / 2.
std::vector<int32_t> bibone_planes(const tiling& t, int N) {
  const int M = N / 2;
  const size_t n = (size_t)M * M;
  const std::vector<int> tri = bibone_triangle_flips(t, N);
  const std::vector<int> bf = bibone_butterfly_flips(t, N);
  std::vector<int32_t> pl(NPLANE * n, 0);
  auto P = [&](int p, int i, int j) -> int32_t& { return pl[p * n + (size_t)i * M + j]; };
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int o = (i % 2 == 0) ? (j % 2 ? 0 : -1) : (j % 2 ? 2 : 1);
      if (o < 0) continue;
      P(EH + o, i / 2, j / 2) = t[(size_t)i * N + j];
      P(BH + o, i / 2, j / 2) = bf[(size_t)i * N + j];
    }
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++) P(j % 2 ? TU : TD, i, j / 2) = tri[(size_t)i * N + j];
  return pl;
}

namespace {

// Edge bit b of cell (i, j), zero outside the M x M grid.
struct EdgeBits {
  const std::vector<uint8_t>& e;
  int M;
  int operator()(int i, int j, int b) const {
    return (i >= 0 && j >= 0 && i < M && j < M) ? (e[(size_t)i * M + j] >> b) & 1 : 0;
  }
  int slot(int i, int j, const ESlot& s) const { return (*this)(i + s.di, j + s.dj, s.bit); }
  int field(int i, int j, const ESlot* f) const {
    return slot(i, j, f[0]) + 2 * slot(i, j, f[1]) + 3 * slot(i, j, f[2]);
  }
  int lozenge(int o, int i, int j) const {
    const ESlot* s = kELoz[o];
    return slot(i, j, s[0]) | slot(i, j, s[1]) << 1 | slot(i, j, s[2]) << 2 |
           slot(i, j, s[3]) << 3;
  }
};

}  // namespace

// Edge form of a start given as its eight planes (bibone_planes): one byte
// per plane cell, or an empty vector when the edge-form walk would not
// reproduce the eight-plane walk from it.  It does when
//  * the lozenge cells agree on every shared edge (they are the edges' own),
//  * the edges form a matching inside the M x M vertex grid,
//  * the triangle and butterfly start states are the edges' (zero on the
//    outer ring of cells), and
//  * every edge that a cell state of the outer ring reads touches an
//    uncovered vertex.
// The reference never refreshes the ring (its update kernels skip it), so
// its ring states keep their start values; the last condition makes the
// edges they stand for empty for good (a flip moves dimers only among
// covered vertices), so those values stay what the edges say.
std::vector<uint8_t> bibone_edge_form(const std::vector<int32_t>& pl, int M) {
  const size_t n = (size_t)M * M;
  auto P = [&](int p, int i, int j) { return pl[p * n + (size_t)i * M + j]; };
  std::vector<uint8_t> e(n);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      const int h = P(EH, i, j);
      e[(size_t)i * M + j] = (uint8_t)((P(ER, i, j) & 1) << kEdgeH | ((h >> 3) & 1) << kEdgeL |
                                       ((h >> 2) & 1) << kEdgeR);
    }
  const EdgeBits E{e, M};
  std::vector<int> deg(n, 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      // endpoints of h(i,j), l(i,j), r(i,j) must lie in the grid
      if ((E(i, j, kEdgeH) && j + 1 >= M) || (E(i, j, kEdgeL) && i + 1 >= M) ||
          (E(i, j, kEdgeR) && (i + 1 >= M || j + 1 >= M)))
        return {};
      deg[(size_t)i * M + j] = E(i, j, kEdgeH) + E(i, j - 1, kEdgeH) + E(i, j, kEdgeL) +
                               E(i - 1, j, kEdgeL) + E(i, j - 1, kEdgeR) + E(i - 1, j, kEdgeR);
      if (deg[(size_t)i * M + j] > 1) return {};
    }
  auto covered = [&](int a, int b) {
    return a >= 0 && b >= 0 && a < M && b < M && deg[(size_t)a * M + b] > 0;
  };
  // edge (a, b, bit) stays empty: one endpoint uncovered (or off the grid)
  auto empty_for_good = [&](int a, int b, int bit) {
    if (bit == kEdgeH) return !covered(a, b) || !covered(a, b + 1);
    if (bit == kEdgeL) return !covered(a, b) || !covered(a + 1, b);
    return !covered(a, b + 1) || !covered(a + 1, b);
  };
  auto ring_ok = [&](int i, int j) {
    for (int o = 0; o < 3; o++)
      for (const ESlot& s : kELoz[o])
        if (!empty_for_good(i + s.di, j + s.dj, s.bit)) return false;
    for (const auto& f : kETri)
      for (const auto& row : f)
        for (const ESlot& s : row)
          if (!empty_for_good(i + s.di, j + s.dj, s.bit)) return false;
    for (const auto& f : kEBf)
      for (const auto& row : f)
        for (const ESlot& s : row)
          if (!empty_for_good(i + s.di, j + s.dj, s.bit)) return false;
    return true;
  };
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      for (int o = 0; o < 3; o++)
        if (E.lozenge(o, i, j) != P(EH + o, i, j)) return {};
      const bool ring = i == 0 || j == 0 || i == M - 1 || j == M - 1;
      if (ring && !ring_ok(i, j)) return {};
      int want[5];
      for (int d = 0; d < 2; d++)
        want[d] = E.field(i, j, kETri[d][0]) + 4 * E.field(i, j, kETri[d][1]) +
                  16 * E.field(i, j, kETri[d][2]);
      for (int o = 0; o < 3; o++)
        want[2 + o] = E.field(i, j, kEBf[o][0]) + 4 * E.field(i, j, kEBf[o][1]) +
                      16 * E.field(i, j, kEBf[o][2]) + 64 * E.field(i, j, kEBf[o][3]) +
                      256 * E.slot(i, j, kEBf[o][4][0]);
      for (int k = 0; k < 5; k++)
        if (P(TU + k, i, j) != (ring ? 0 : want[k])) return {};
    }
  return e;
}

// Writes the lozenge cells of an N x N tiling (N = 2M) from an edge form;
// the tiling's other cells are left as they are.
void bibone_edge_merge(const std::vector<uint8_t>& e, int M, int32_t* tiling) {
  const int N = 2 * M;
  const EdgeBits E{e, M};
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++)
      for (int o = 0; o < 3; o++)
        tiling[(size_t)(2 * i + (o != 0)) * N + 2 * j + (o != 1)] = E.lozenge(o, i, j);
}

}  // namespace rt

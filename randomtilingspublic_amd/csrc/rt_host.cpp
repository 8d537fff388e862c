// rt_host.cpp -- host half of librtiling.so: colour schedules, the TinyMT
// parameter-file parser, and the per-model domain / Thurston / height
// function helpers that the reference keeps as static class methods.
#include <stdint.h>

#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "rt_internal.hpp"

namespace rt {

static const int kInf = INT_MAX;  // the reference's `infty` (common.h:47)

// ------------------------------------------------------------- schedules
void domino_schedule(uint32_t seed, long steps, bool skip_first, uint32_t* base_out,
                     std::vector<uint8_t>& colours) {
  std::mt19937 g(seed);
  if (skip_first) {
    uint32_t b = (uint32_t)g();
    if (base_out) *base_out = b;
  }
  colours.resize(steps > 0 ? (size_t)steps : 0);
  // r == 1 -> black (colour index 0); r == 0 -> white (index 1).
  for (long s = 0; s < steps; s++) colours[s] = (g() % 2 == 1) ? 0 : 1;
}

void tri_schedule(long seed, long steps, std::vector<uint8_t>& colours) {
  std::default_random_engine g;
  std::uniform_int_distribution<int> dist(0, 2);
  g.seed(seed);
  colours.resize(steps > 0 ? (size_t)steps : 0);
  for (long s = 0; s < steps; s++) colours[s] = (uint8_t)dist(g);
}

// ------------------------------------------------------ parameter file
// tinymt32dc text format as read by the reference (file_reader.cpp:51-70,
// 114-166): '#' lines are comments; mat1, mat2, tmat are the hex fields
// after the 3rd, 4th and 5th comma.  A blank line, a short line or end of
// file before `count` sets is a format error.
int parse_tinymt_file(const char* path, size_t count, uint32_t* out3, std::string* err) {
  std::ifstream in(path);
  if (!in) {
    if (err) *err = std::string("cannot open TinyMT parameter file: ") + path;
    return RT_EIO;
  }
  std::string line;
  for (size_t k = 0; k < count; k++) {
    bool got = false;
    while (std::getline(in, line)) {
      if (!line.empty() && line[0] == '#') continue;
      got = true;
      break;
    }
    if (!got || line.empty() || line[0] == '\r') {
      if (err) *err = "TinyMT parameter file: format error at set " + std::to_string(k);
      return RT_EPARAMS;
    }
    size_t pos[6] = {0};
    int commas = 0;
    for (size_t c = 0; c < line.size() && commas < 5; c++)
      if (line[c] == ',') pos[++commas] = c + 1;
    if (commas < 5) {
      if (err) *err = "TinyMT parameter file: comma not found at set " + std::to_string(k);
      return RT_EPARAMS;
    }
    // A field past unsigned long's range is a format error, as in
    // file_reader::get_params (errno checked after the three fields).
    errno = 0;
    for (int f = 0; f < 3; f++)
      out3[3 * k + f] = (uint32_t)std::strtoul(line.c_str() + pos[3 + f], nullptr, 16);
    if (errno != 0) {
      if (err) *err = "TinyMT parameter file: value out of range at set " + std::to_string(k);
      return RT_EPARAMS;
    }
  }
  return RT_OK;
}

rt_vec* make_vec(std::vector<int32_t>&& v) {
  rt_vec* r = new rt_vec();
  r->v = std::move(v);
  return r;
}

// ====================================================================
// Height-function machinery shared by the square and triangular lattices
// (Thurston's construction: boundary heights by a walk along boundary
// edges, then the extremal filling by a multi-source shortest-path pass
// over the edges on which the height moves by one in the filling
// direction).  Heights are small integers, so a bucket queue replaces the
// reference's std::priority_queue; the extremal height function is
// unique, so the filling order does not change the result.
// ====================================================================
struct Lattice {
  int N;                         // vertex grid is N x N
  int nnb;                       // 4 (square) or 6 (triangular)
  const int (*nb)[2];            // neighbour offsets
  std::vector<int> status;       // 0 outside, 1 interior, 2 boundary
  // edge (v, k): two adjacent faces, -1 if off the face grid
  std::vector<int> f1, f2;       // per (vertex * nnb + k)
  std::vector<int> dh;           // height change along a non-dimer edge
  const std::vector<int>* d;     // face indicator
};

static inline int face_at(const Lattice& L, int v, int k, int which) {
  return which ? L.f2[(size_t)v * L.nnb + k] : L.f1[(size_t)v * L.nnb + k];
}

static inline int face_val(const Lattice& L, int f) {
  return (f >= 0 && f < (int)L.d->size()) ? (*L.d)[f] : 0;
}

static inline int neighbour(const Lattice& L, int v, int k) {
  const int i = v / L.N + L.nb[k][0], j = v % L.N + L.nb[k][1];
  if (i < 0 || j < 0 || i >= L.N || j >= L.N) return -1;
  return i * L.N + j;
}

// Returns RT_OK or RT_EUNTILEABLE; hf gets the extremal height function.
static int thurston(const Lattice& L, bool max, std::vector<int>& hf) {
  const int NV = L.N * L.N;
  hf.assign(NV, kInf);
  int start = -1;
  for (int v = 0; v < NV && start < 0; v++)
    if (L.status[v] == 2) start = v;
  if (start < 0) return RT_OK;
  hf[start] = 0;
  std::deque<int> q{start};
  std::vector<int> boundary{start};
  while (!q.empty()) {
    const int v = q.front();
    q.pop_front();
    for (int k = 0; k < L.nnb; k++) {
      const int u = neighbour(L, v, k);
      if (u < 0 || L.status[u] != 2) continue;
      const int a = face_val(L, face_at(L, v, k, 0)), b = face_val(L, face_at(L, v, k, 1));
      if (!(a ^ b)) continue;  // not a boundary edge
      const int h = hf[v] + L.dh[(size_t)v * L.nnb + k];
      if (hf[u] == kInf) {
        hf[u] = h;
        q.push_back(u);
        boundary.push_back(u);
      } else if (hf[u] != h) {
        return RT_EUNTILEABLE;
      }
    }
  }
  // Bucket-queue filling: max -> lowest first, climbing +1 edges;
  // min -> highest first, descending -1 edges.
  const int sgn = max ? 1 : -1;
  int lo = INT_MAX, hi = INT_MIN;
  for (int v : boundary) {
    lo = std::min(lo, sgn * hf[v]);
    hi = std::max(hi, sgn * hf[v]);
  }
  std::vector<std::vector<int>> bucket;
  auto push = [&](int key, int v) {
    const size_t idx = (size_t)(key - lo);
    if (idx >= bucket.size()) bucket.resize(idx + 1);
    bucket[idx].push_back(v);
  };
  for (int v : boundary) push(sgn * hf[v], v);
  for (size_t b = 0; b < bucket.size(); b++) {
    for (size_t n = 0; n < bucket[b].size(); n++) {
      const int v = bucket[b][n];
      for (int k = 0; k < L.nnb; k++) {
        if (sgn * L.dh[(size_t)v * L.nnb + k] <= 0) continue;
        const int u = neighbour(L, v, k);
        if (u < 0 || L.status[u] == 0) continue;
        const int h = hf[v] + sgn;
        if (L.status[u] == 2) {
          if (L.status[v] != 2 && sgn * hf[u] > sgn * h) return RT_EUNTILEABLE;
        } else if (hf[u] == kInf) {
          hf[u] = h;
          push(sgn * h, u);
        }
      }
    }
  }
  (void)hi;
  return RT_OK;
}

// ============================================================== domino
// Square lattice, vertices (N x N), faces (N-1 x N-1).
// Neighbours: 0 down (1,0), 1 up (-1,0), 2 right (0,1), 3 left (0,-1);
// vertex bits for a dimer across that edge: 2, 1, 8, 4 (DominoTiler.cpp:335-341).
static const int kSqNb[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
static const int kSqBit[4] = {2, 1, 8, 4};
// The two faces beside edge k of vertex (i,j), as offsets of the face
// (i+di, j+dj) in the (N-1)-wide face grid (DominoTiler.cpp:336).
static const int kSqFace[4][2][2] = {
    {{0, -1}, {0, 0}}, {{-1, -1}, {-1, 0}}, {{-1, 0}, {0, 0}}, {{-1, -1}, {0, -1}}};

static Lattice square_lattice(const domain& d, int N, const std::vector<int>& status) {
  Lattice L;
  L.N = N;
  L.nnb = 4;
  L.nb = kSqNb;
  L.status = status;
  L.d = &d;
  const int F = N - 1;
  L.f1.resize((size_t)N * N * 4);
  L.f2.resize((size_t)N * N * 4);
  L.dh.resize((size_t)N * N * 4);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < 4; k++) {
        const size_t e = ((size_t)i * N + j) * 4 + k;
        for (int w = 0; w < 2; w++) {
          const int fi = i + kSqFace[k][w][0], fj = j + kSqFace[k][w][1];
          const int f = (fi < 0 || fj < 0 || fi >= F || fj >= F) ? -1 : fi * F + fj;
          (w ? L.f2 : L.f1)[e] = f;
        }
        // Crossing rule of DominoTiler.cpp:371-373: +1 leaving a black
        // square north/south..., i.e. (k > 1 ? +1 : -1) * ((i+j) even ? +1 : -1).
        L.dh[e] = (k > 1 ? 1 : -1) * ((i + j) % 2 == 0 ? 1 : -1);
      }
  return L;
}

domain domino_rectangle(int M, int N) {
  int dim = std::max(M, N) + 4;
  if (dim % 2 == 0) dim++;
  domain d((size_t)dim * dim, 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++) d[(size_t)(i + 2) * dim + j + 2] = 1;
  return d;
}

// Union of axis-aligned rectangles (i0, j0, h, w) of unit squares, shifted
// so the union's bounding box starts at (0, 0) and framed like
// domino_rectangle (two empty rows/columns on each side, odd side).  Holes
// of the union (empty cells not connected to the frame) are filled, so the
// region is simply connected whenever the union is connected.  No reference
// counterpart: a domain family for batches of uneven chains.
domain domino_rect_union(const std::vector<int>& rects) {
  const size_t n = rects.size() / 4;
  if (n == 0) return domain();
  int i_lo = INT_MAX, j_lo = INT_MAX, i_hi = INT_MIN, j_hi = INT_MIN;
  for (size_t r = 0; r < n; r++) {
    const int* q = &rects[4 * r];
    if (q[2] <= 0 || q[3] <= 0) continue;
    i_lo = std::min(i_lo, q[0]);
    j_lo = std::min(j_lo, q[1]);
    i_hi = std::max(i_hi, q[0] + q[2]);
    j_hi = std::max(j_hi, q[1] + q[3]);
  }
  if (i_lo == INT_MAX) return domain();
  int dim = std::max(i_hi - i_lo, j_hi - j_lo) + 4;
  if (dim % 2 == 0) dim++;
  domain d((size_t)dim * dim, 0);
  for (size_t r = 0; r < n; r++) {
    const int* q = &rects[4 * r];
    for (int i = 0; i < q[2]; i++)
      for (int j = 0; j < q[3]; j++)
        d[(size_t)(q[0] - i_lo + i + 2) * dim + (q[1] - j_lo + j + 2)] = 1;
  }
  // Flood the empty cells from the frame; whatever stays unreached is a hole.
  std::vector<uint8_t> seen((size_t)dim * dim, 0);
  std::vector<int> stack{0};
  seen[0] = 1;
  while (!stack.empty()) {
    const int v = stack.back();
    stack.pop_back();
    const int i = v / dim, j = v % dim;
    const int nb[4][2] = {{i + 1, j}, {i - 1, j}, {i, j + 1}, {i, j - 1}};
    for (const auto& u : nb) {
      if (u[0] < 0 || u[1] < 0 || u[0] >= dim || u[1] >= dim) continue;
      const int w = u[0] * dim + u[1];
      if (seen[w] || d[w]) continue;
      seen[w] = 1;
      stack.push_back(w);
    }
  }
  for (size_t v = 0; v < d.size(); v++)
    if (!d[v] && !seen[v]) d[v] = 1;
  return d;
}

domain domino_aztec_rectangle(int n, const std::vector<int>& remove) {
  const int k = (int)remove.size();
  const int M = 2 * n - k;
  if (M <= 0) return domain();
  domain core((size_t)M * M, 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      bool in;
      if (i == 0 || i == M - 1) {
        const int c = (n - k) + (i == M - 1 ? k : 0);
        in = (j == c - 1 || j == c);
      } else if (i < n - k) {
        in = j >= n - k - i - 1 && j <= n - k + i;
      } else if (i > n - 1) {
        in = j >= i - n + k && j <= 3 * n - k - i - 1;
      } else {
        in = j >= i - (n - k) && j <= i + (n - k);
      }
      core[(size_t)i * M + j] = in ? 1 : 0;
    }
  for (int r = 0; r < k; r++) {
    const int i0 = n - k + remove[r], j0 = remove[r];
    if (i0 >= 0 && i0 < M && j0 >= 0 && j0 < M) core[(size_t)i0 * M + j0] = 0;
  }
  // pad to an odd side >= M+4 so the vertex array is even and twice zero-padded
  const int M2 = M + 4 + (1 - M % 2);
  domain d((size_t)M2 * M2, 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) d[(size_t)(i + 2) * M2 + j + 2] = core[(size_t)i * M + j];
  return d;
}

std::vector<int> domino_domain_to_vertices(const domain& d) {
  const int M = (int)std::lround(std::sqrt((double)d.size()));
  const int N = M + 1;
  std::vector<int> v((size_t)N * N, 0);
  auto at = [&](int i, int j) { return (i >= 0 && j >= 0 && i < M && j < M) ? d[(size_t)i * M + j] : 0; };
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      if (i == 0 || j == 0 || i == M || j == M) {
        // clamp the face lookups onto the grid, as the reference does
        const int ia = i - (i == M), ib = i - 1 + (i == 0);
        const int ja = j - (j == M), jb = j - 1 + (j == 0);
        if (at(ia, ja) || at(ib, ja) || at(ia, jb) || at(ib, jb)) v[(size_t)i * N + j] = 2;
      } else {
        const int a = at(i - 1, j - 1), b = at(i - 1, j), c = at(i, j - 1), e = at(i, j);
        if (a && b && c && e) v[(size_t)i * N + j] = 1;
        else if (a || b || c || e) v[(size_t)i * N + j] = 2;
      }
    }
  return v;
}

tiling domino_heightfunc_to_tiling(const heightfunc& hf, const domain& d) {
  const int N = (int)std::lround(std::sqrt((double)hf.size()));
  tiling t((size_t)N * N, 0);
  const int F = N - 1;
  auto dv = [&](int fi, int fj) {
    return (fi >= 0 && fj >= 0 && fi < F && fj < F && (size_t)(fi * F + fj) < d.size()) ? d[(size_t)fi * F + fj] : 0;
  };
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int h0 = hf[(size_t)i * N + j];
      if (h0 == kInf) continue;
      for (int k = 0; k < 4; k++) {
        const int ii = i + kSqNb[k][0], jj = j + kSqNb[k][1];
        if (ii < 0 || jj < 0 || ii >= N || jj >= N) continue;
        const int h1 = hf[(size_t)ii * N + jj];
        if (h1 == kInf) continue;
        if (!(dv(i + kSqFace[k][0][0], j + kSqFace[k][0][1]) ||
              dv(i + kSqFace[k][1][0], j + kSqFace[k][1][1])))
          continue;
        t[(size_t)i * N + j] += kSqBit[k] * std::abs((h1 - h0) / 2);
      }
    }
  return t;
}

int domino_thurston(const domain& d, bool max, tiling& out) {
  std::vector<int> status = domino_domain_to_vertices(d);
  const int N = (int)std::lround(std::sqrt((double)status.size()));
  Lattice L = square_lattice(d, N, status);
  std::vector<int> hf;
  int rc = thurston(L, max, hf);
  if (rc) return rc;
  out = domino_heightfunc_to_tiling(hf, d);
  return RT_OK;
}

heightfunc domino_tiling_to_heightfunc(const tiling& t, const domain& d) {
  const int N = (int)std::lround(std::sqrt((double)t.size()));
  std::vector<int> status = domino_domain_to_vertices(d);
  heightfunc hf((size_t)N * N, kInf);
  if ((int)status.size() != N * N) return hf;
  Lattice L = square_lattice(d, N, status);
  int start = -1;
  for (int v = 0; v < N * N && start < 0; v++)
    if (status[v] == 2) start = v;
  if (start < 0) return hf;
  hf[start] = 0;
  std::deque<int> q{start};
  while (!q.empty()) {
    const int v = q.front();
    q.pop_front();
    for (int k = 0; k < 4; k++) {
      if (t[v] & kSqBit[k]) continue;  // a domino crosses this edge
      const int u = neighbour(L, v, k);
      if (u < 0 || status[u] == 0 || hf[u] != kInf) continue;
      if (!(face_val(L, face_at(L, v, k, 0)) || face_val(L, face_at(L, v, k, 1)))) continue;
      hf[u] = hf[v] + L.dh[(size_t)v * 4 + k];
      q.push_back(u);
    }
  }
  return hf;
}

domain domino_tiling_to_domain(const tiling& t) {
  const int N = (int)std::lround(std::sqrt((double)t.size()));
  const int M = N - 1;
  domain d((size_t)(M > 0 ? M : 0) * (M > 0 ? M : 0), 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++)
      if ((t[(size_t)i * N + j] & (2 + 8)) || (t[(size_t)(i + 1) * N + j + 1] & (1 + 4)))
        d[(size_t)i * M + j] = 1;
  return d;
}

// ============================================================= lozenge
// Triangular lattice drawn on a square array: vertex (i,j); neighbours
// (LozengeTiler.cpp:21-25) E, NE, N, W, SW, S with height changes
// -1,+1,-1,+1,-1,+1 and dimer bits 8,2,1,4,16,32.  Faces live in an
// M x 2M array (M = N-1); the two faces beside edge k of vertex (i,j) are
// (i+a, 2j+b) for the offsets below.
static const int kTriNb[6][2] = {{0, 1}, {-1, 1}, {-1, 0}, {0, -1}, {1, -1}, {1, 0}};
static const int kTriDh[6] = {-1, 1, -1, 1, -1, 1};
static const int kTriBit[6] = {8, 2, 1, 4, 16, 32};
static const int kTriFace[6][2][2] = {{{0, 0}, {-1, 1}},  {{-1, 1}, {-1, 0}}, {{-1, 0}, {-1, -1}},
                                      {{-1, -1}, {0, -2}}, {{0, -2}, {0, -1}}, {{0, -1}, {0, 0}}};

static Lattice tri_lattice(const domain& d, int N, const std::vector<int>& status) {
  Lattice L;
  L.N = N;
  L.nnb = 6;
  L.nb = kTriNb;
  L.status = status;
  L.d = &d;
  const int M = N - 1;
  L.f1.resize((size_t)N * N * 6);
  L.f2.resize((size_t)N * N * 6);
  L.dh.resize((size_t)N * N * 6);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < 6; k++) {
        const size_t e = ((size_t)i * N + j) * 6 + k;
        for (int w = 0; w < 2; w++) {
          const int fi = i + kTriFace[k][w][0], fj = 2 * j + kTriFace[k][w][1];
          const int f = (fi < 0 || fj < 0 || fi >= M || fj >= 2 * M) ? -1 : fi * 2 * M + fj;
          (w ? L.f2 : L.f1)[e] = f;
        }
        L.dh[e] = kTriDh[k];
      }
  return L;
}

domain lozenge_almost_hex_domain(int n, int m) {
  const int W0 = 4 * n;
  const int Hp = 2 * n + 6, Wp = 4 * n + 12;
  if (m > n) return domain((size_t)Hp * Wp, 0);
  domain core((size_t)2 * n * W0, 0);
  for (int i = 0; i < 2 * n; i++)
    for (int j = 0; j < W0; j++) {
      const bool in = i < n ? (2 * i + j - 1 >= 2 * n - 2) : (2 * i + j <= 6 * n - 2);
      core[(size_t)i * W0 + j] = in ? 1 : 0;
    }
  for (int i = 0; i < m; i++)
    for (int j = W0 - 2 * m; j < W0; j++) core[(size_t)i * W0 + j] = 0;
  // 3 rows / 6 face-columns of padding, then widen to a side M2+k with the
  // vertex count M2+k+1 divisible by 3 (LozengeTiler.cpp:453-465).
  const int M2 = Hp;  // sqrt(Hp * Wp / 2)
  const int k = (3 - (M2 + 1) % 3) % 3;
  const int S = M2 + k;
  domain d((size_t)S * 2 * S, 0);
  for (int i = 3; i < 2 * n + 3; i++)
    for (int j = 6; j < W0 + 6; j++) d[(size_t)i * 2 * S + j] = core[(size_t)(i - 3) * W0 + j - 6];
  return d;
}

std::vector<int> lozenge_domain_to_vertices(const domain& d) {
  const int M = (int)std::lround(std::sqrt((double)d.size() / 2));
  const int N = M + 1;
  std::vector<int> v((size_t)N * N, 0);
  auto at = [&](int i, int j) {
    return (i >= 0 && j >= 0 && i < M && j < 2 * M) ? d[(size_t)i * 2 * M + j] : 0;
  };
  for (int i = 1; i < M; i++)
    for (int j = 1; j < M; j++) {
      const int f[6] = {at(i, 2 * j), at(i, 2 * j - 1), at(i, 2 * j - 2),
                        at(i - 1, 2 * j), at(i - 1, 2 * j + 1), at(i - 1, 2 * j - 1)};
      const bool all = f[0] && f[1] && f[2] && f[3] && f[4] && f[5];
      const bool any = f[0] || f[1] || f[2] || f[3] || f[4] || f[5];
      v[(size_t)i * N + j] = all ? 1 : (any ? 2 : 0);
    }
  return v;
}

tiling lozenge_heightfunc_to_tiling(const heightfunc& hf, const domain& d) {
  const int N = (int)std::lround(std::sqrt((double)hf.size()));
  const int M = N - 1;
  tiling t((size_t)N * N, 0);
  auto at = [&](int fi, int fj) {
    return (fi >= 0 && fj >= 0 && fi < M && fj < 2 * M && (size_t)(fi * 2 * M + fj) < d.size())
               ? d[(size_t)fi * 2 * M + fj]
               : 0;
  };
  for (int i = 1; i < N - 1; i++)
    for (int j = 1; j < N - 1; j++)
      for (int k = 0; k < 6; k++) {
        if (!(at(i + kTriFace[k][0][0], 2 * j + kTriFace[k][0][1]) ||
              at(i + kTriFace[k][1][0], 2 * j + kTriFace[k][1][1])))
          continue;
        const int ii = i + kTriNb[k][0], jj = j + kTriNb[k][1];
        const long long dh = (long long)hf[(size_t)ii * N + jj] - hf[(size_t)i * N + j];
        // the reference subtracts in int; a wrapped difference is never +-2
        const int wrapped = (int)(uint32_t)(uint64_t)dh;
        if (std::llabs((long long)wrapped) == 2) t[(size_t)i * N + j] += kTriBit[k];
      }
  return t;
}

int lozenge_thurston(const domain& d, bool max, tiling& out) {
  std::vector<int> status = lozenge_domain_to_vertices(d);
  const int N = (int)std::lround(std::sqrt((double)status.size()));
  Lattice L = tri_lattice(d, N, status);
  std::vector<int> hf;
  int rc = thurston(L, max, hf);
  if (rc) return rc;
  out = lozenge_heightfunc_to_tiling(hf, d);
  return RT_OK;
}

heightfunc lozenge_tiling_to_heightfunc(const tiling& t, const domain& d) {
  const int N = (int)std::lround(std::sqrt((double)t.size()));
  std::vector<int> status = lozenge_domain_to_vertices(d);
  heightfunc hf(status.size(), kInf);
  if ((int)status.size() != N * N) return hf;
  Lattice L = tri_lattice(d, N, status);
  int start = -1;
  for (int v = 0; v < N * N && start < 0; v++)
    if (status[v] == 2) start = v;
  if (start < 0) return hf;
  hf[start] = 0;
  std::deque<int> q{start};
  while (!q.empty()) {
    const int v = q.front();
    q.pop_front();
    for (int k = 0; k < 6; k++) {
      if (t[v] & kTriBit[k]) continue;
      const int u = neighbour(L, v, k);
      if (u < 0 || status[u] == 0 || hf[u] != kInf) continue;
      if (!(face_val(L, face_at(L, v, k, 0)) || face_val(L, face_at(L, v, k, 1)))) continue;
      hf[u] = hf[v] + kTriDh[k];
      q.push_back(u);
    }
  }
  return hf;
}

domain lozenge_tiling_to_domain(const tiling& t) {
  const int N = (int)std::lround(std::sqrt((double)t.size()));
  const int M = N - 1;
  domain d((size_t)(M > 0 ? M : 0) * (M > 0 ? M : 0) * 2, 0);
  auto set = [&](long idx) {
    if (idx >= 0 && (size_t)idx < d.size()) d[(size_t)idx] = 1;
  };
  for (int i = 0; i < N - 1; i++)
    for (int j = 0; j < N - 1; j++) {
      const int v = t[(size_t)i * N + j];
      if (v & 1) {
        set((long)(i - 1) * 2 * M + 2 * j);
        set((long)(i - 1) * 2 * M + 2 * j - 1);
      }
      if (v & 2) {
        set((long)(i - 1) * 2 * M + 2 * j);
        set((long)(i - 1) * 2 * M + 2 * j + 1);
      }
      if (v & 8) {
        set((long)(i - 1) * 2 * M + 2 * j + 1);
        set((long)i * 2 * M + 2 * j);
      }
    }
  return d;
}

// =================================================== rectangle-triangle
// maxHex / minHex (RectTriangleTiler.cpp:480-600): the extremal tilings of
// a side-n hexagon, built on a (2n+1)^2 array and sheared into an array of
// side M2 = M+6+(3-M%3)%3 with three rows/columns of zero padding.
static tiling recttri_shear(const std::vector<int>& a, int n) {
  const int M = 2 * n + 1;
  const int M2 = M + 6 + (3 - M % 3) % 3;
  tiling out((size_t)M2 * M2, 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      const int i2 = i + 3, j2 = j + n + 3 - i;
      if (j2 < M2) out[(size_t)i2 * M2 + j2] = a[(size_t)i * M + j];
    }
  return out;
}

tiling recttri_max_hex(int n) {
  const int M = 2 * n + 1;
  std::vector<int> a((size_t)M * M, 0);
  for (int i = 1; i < M - 1; i++)
    for (int j = 0; j < M; j++) {
      int v = 0;
      if (j - i > n || i - j > n) v = 0;
      else if (i < n && j < n) v = 0x333333;
      else if (i - j < 0 && j > n) v = 0x111111;
      else if (i > n && i - j > 0) v = 0x222222;
      else if (i == n && j < n) v = 0x333222;
      else if (i < n && j == n) v = 0x331311;
      else if (i - j == 0 && i > n) v = 0x211221;
      else if (i == n && j == n) v = 0x331221;
      a[(size_t)i * M + j] = v;
    }
  for (int i = 1; i < n; i++) {
    a[(size_t)i * M] = 0x003033;
    a[i] = 0x000333;
  }
  for (int i = n + 1; i < M - 1; i++) {
    a[(size_t)i * M + (M - 1)] = 0x110100;
    a[(size_t)(M - 1) * M + i] = 0x222000;
  }
  for (int i = 1; i < n; i++) {
    a[(size_t)i * M + i + n] = 0x100110;
    a[(size_t)(i + n) * M + i] = 0x022002;
  }
  a[0] = 0x000033;
  a[(size_t)n * M] = 0x003002;
  a[n] = 0x000310;
  a[(size_t)(M - 1) * M + (M - 1)] = 0x210000;
  a[(size_t)n * M + (M - 1)] = 0x100100;
  a[(size_t)(M - 1) * M + n] = 0x022000;
  return recttri_shear(a, n);
}

// slopeHex(N) (RectTriangleTiler.cpp:173-478): the "sloped" start tiling.
// The hexagon's three axes through the centre and its main diagonal split
// it into six slices; each slice has a bulk code, codes for its two edge
// lines (next to an axis or the diagonal) and for their corner, chosen by a
// parity; the codes are the reference's table.
namespace {
struct SliceCodes {
  int edge1[2], edge2[2], corner[2], bulk[2];  // [0] = parity even, [1] = odd; corner by N
};
}  // namespace

tiling recttri_slope_hex(int n) {
  const int M = 2 * n + 1, N = n;
  std::vector<int> a((size_t)M * M, 0);
  const int Nodd = N & 1;
  // Slices: A i<N, j<i; B i>N, j>i; C j<N, i<j; D j>N, i>j; E i<N, j>N, i>j-N;
  // F i>N, j<N, i<j+N.  edge1 / edge2 as in the comments per slice.
  static const SliceCodes A = {{0x58881b, 0x855be8}, {0x58388e, 0x8eb558}, {0x583812, 0x8ebbe8},
                               {0x588885, 0x855558}};
  static const SliceCodes B = {{0xb18885, 0x8eb558}, {0xe88385, 0x855be8}, {0x218385, 0x8ebbe8},
                               {0x588885, 0x855558}};
  static const SliceCodes Cs = {{0x7477d2, 0x47d47a}, {0x7473a7, 0xa74d74}, {0x747312, 0xa7dd7a},
                                {0x747747, 0x474474}};
  static const SliceCodes D = {{0x2d7747, 0xa74d74}, {0x7a3747, 0x47d47a}, {0x213747, 0xa7dd7a},
                               {0x747747, 0x474474}};
  static const SliceCodes E = {{0x996f19, 0x6699cf}, {0x296c99, 0xfc9966}, {0x296319, 0xfc99cf},
                               {0x996699, 0x669966}};
  static const SliceCodes F = {{0x91f699, 0xfc9966}, {0x99c692, 0x6699cf}, {0x913692, 0xfc99cf},
                               {0x996699, 0x669966}};
  auto slice = [&](const SliceCodes& c, bool e1, bool e2, int par) {
    if (e1 && e2) return c.corner[Nodd];
    if (e1) return c.edge1[par];
    if (e2) return c.edge2[par];
    return c.bulk[par];
  };
  // Codes on the axes i == N, j == N and the diagonal i == j:
  // {far before odd, far before even, just before (N even, N odd),
  //  far after odd, far after even, just after (N even, N odd)}.
  struct Axis {
    int before[2], near_before[2], after[2], near_after[2];
  };
  static const Axis row = {{0xe81c91, 0x1be1fc}, {0x122133, 0xe81c91}, {0x19c18e, 0xcf1eb1},
                           {0x331221, 0x29c38e}};
  static const Axis col = {{0x7af229, 0xd22afc}, {0x122133, 0x7af229}, {0x922fa7, 0xcfa22d},
                           {0x331221, 0x913fa7}};
  static const Axis dia = {{0xbd7833, 0x33aebd}, {0x331221, 0xbd7812}, {0x3387db, 0xdbea33},
                           {0x122133, 0x3387db}};
  auto axis = [&](const Axis& x, int k) {
    if (k < N - 1) return x.before[k & 1];
    if (k == N - 1) return x.near_before[Nodd];
    if (k > N + 1) return x.after[k & 1];
    return x.near_after[Nodd];  // k == N + 1
  };
  for (int i = 1; i < M - 1; i++)
    for (int j = 1; j < M - 1; j++) {
      int v = 0;
      const int pe = ((i + j) % 2 != Nodd) ? 1 : 0;
      if (i == N && j == N) v = Nodd ? 0x122133 : 0x213312;
      else if (i == N) v = axis(row, j);
      else if (j == N) v = axis(col, i);
      else if (i == j) v = axis(dia, i);
      else if (i < N && j < i) v = slice(A, i == N - 1, j == i - 1, j & 1);
      else if (i > N && j > i) v = slice(B, i == N + 1, j == i + 1, j & 1);
      else if (j < N && i < j) v = slice(Cs, j == N - 1, i == j - 1, i & 1);
      else if (j > N && i > j) v = slice(D, j == N + 1, i == j + 1, i & 1);
      else if (i < N && j > N && i > j - N) v = slice(E, i == N - 1, j == N + 1, pe);
      else if (i > N && j < N && i < j + N) v = slice(F, i == N + 1, j == N - 1, pe);
      a[(size_t)i * M + j] = v;
    }
  auto at = [&](int i, int j) -> int& { return a[(size_t)i * M + j]; };
  for (int i = 2; i < N - 1; i++) at(i, 0) = 0x008085, at(0, i) = 0x000747;
  at(1, 0) = 0x00308e;
  at(0, 1) = 0x0003a7;
  at(N - 1, 0) = 0x00801b;
  at(0, N - 1) = 0x0007d2;
  for (int i = N + 2; i < M - 2; i++) at(i, M - 1) = 0x580800, at(M - 1, i) = 0x747000;
  at(N + 1, M - 1) = 0xb10800;
  at(M - 1, N + 1) = 0x2d7000;
  at(M - 2, M - 1) = 0xe80300;
  at(M - 1, M - 2) = 0x7a3000;
  for (int i = 2; i < N - 1; i++) at(i, i + N) = 0x900690, at(i + N, i) = 0x096009;
  at(1, N + 1) = 0x200c90;
  at(N + 1, 1) = 0x01f009;
  at(N - 1, M - 2) = 0x900f10;
  at(M - 2, N - 1) = 0x09c002;
  at(0, 0) = N == 1 ? 0x000012 : 0x000033;
  at(M - 1, N) = N == 1 ? 0x013000 : 0x022000;
  at(N, M - 1) = N == 1 ? 0x200300 : 0x100100;
  at(0, N) = 0x000220;
  at(N, 0) = 0x001001;
  at(M - 1, M - 1) = 0x330000;
  return recttri_shear(a, n);
}

tiling recttri_min_hex(int n) {
  const int M = 2 * n + 1;
  std::vector<int> a((size_t)M * M, 0);
  for (int i = 1; i < M - 1; i++)
    for (int j = 1; j < M - 1; j++) {
      int v = 0;
      if (j - i > n || i - j > n) v = 0;
      else if (i > n && j > n) v = 0x333333;
      else if (i - j > 0 && j < n) v = 0x111111;
      else if (i < n && i - j < 0) v = 0x222222;
      else if (i == n && j > n) v = 0x222333;
      else if (i > n && j == n) v = 0x113133;
      else if (i - j == 0 && i < n) v = 0x122112;
      else if (i == n && j == n) v = 0x122133;
      a[(size_t)i * M + j] = v;
    }
  for (int i = 1; i < n; i++) {
    a[(size_t)i * M] = 0x001011;
    a[i] = 0x000222;
  }
  for (int i = n + 1; i < M - 1; i++) {
    a[(size_t)i * M + (M - 1)] = 0x330300;
    a[(size_t)(M - 1) * M + i] = 0x333000;
  }
  for (int i = 1; i < n; i++) {
    a[(size_t)i * M + i + n] = 0x200220;
    a[(size_t)(i + n) * M + i] = 0x011001;
  }
  a[0] = 0x000012;
  a[(size_t)(M - 1) * M + n] = 0x013000;
  a[(size_t)n * M + (M - 1)] = 0x200300;
  a[n] = 0x000220;
  a[(size_t)n * M] = 0x001001;
  a[(size_t)(M - 1) * M + (M - 1)] = 0x330000;
  return recttri_shear(a, n);
}

}  // namespace rt

// ================================================================ C ABI
using namespace rt;

extern "C" {

int rt_parse_tinymt_file(const char* path, size_t count, uint32_t* out3) {
  if (!path || (!out3 && count)) return RT_EINVAL;
  return parse_tinymt_file(path, count, out3, nullptr);
}

const int32_t* rt_vec_data(const rt_vec* v) { return v ? v->v.data() : nullptr; }
size_t rt_vec_size(const rt_vec* v) { return v ? v->v.size() : 0; }
void rt_vec_free(rt_vec* v) { delete v; }

static std::vector<int> from_c(const int32_t* p, size_t n) {
  return std::vector<int>(p, p + (p ? n : 0));
}

static bool is_square(size_t n) {
  const size_t r = (size_t)std::llround(std::sqrt((double)n));
  return r * r == n && n > 0;
}

rt_vec* rt_domino_rectangle(int M, int N) {
  if (M < 0 || N < 0) return nullptr;
  return make_vec(domino_rectangle(M, N));
}

rt_vec* rt_domino_rect_union(const int* rects, int n) {
  if (n < 1 || !rects) return nullptr;
  for (int r = 0; r < n; r++)
    if (rects[4 * r + 2] < 1 || rects[4 * r + 3] < 1 || rects[4 * r + 2] > 8192 ||
        rects[4 * r + 3] > 8192 || std::abs(rects[4 * r]) > 8192 || std::abs(rects[4 * r + 1]) > 8192)
      return nullptr;
  return make_vec(domino_rect_union(std::vector<int>(rects, rects + 4 * (size_t)n)));
}

rt_vec* rt_domino_aztec_diamond(int n) {
  if (n < 1) return nullptr;
  return make_vec(domino_aztec_rectangle(n, {}));
}

rt_vec* rt_domino_aztec_rectangle(int n, const int* remove, int k) {
  if (n < 1 || k < 0 || k >= 2 * n || (k && !remove)) return nullptr;
  return make_vec(domino_aztec_rectangle(n, std::vector<int>(remove, remove + k)));
}

static int thurston_c(bool domino, bool max, const int32_t* d, size_t dsize, rt_vec** out) {
  if (!out) return RT_EINVAL;
  *out = nullptr;
  if (!d || (domino ? !is_square(dsize) : (dsize % 2 || !is_square(dsize / 2)))) return RT_EINVAL;
  tiling t;
  domain dd = from_c(d, dsize);
  int rc = domino ? domino_thurston(dd, max, t) : lozenge_thurston(dd, max, t);
  if (rc) return rc;
  *out = make_vec(std::move(t));
  return RT_OK;
}

int rt_domino_max_tiling(const int32_t* d, size_t n, rt_vec** out) {
  return thurston_c(true, true, d, n, out);
}
int rt_domino_min_tiling(const int32_t* d, size_t n, rt_vec** out) {
  return thurston_c(true, false, d, n, out);
}
int rt_lozenge_max_tiling(const int32_t* d, size_t n, rt_vec** out) {
  return thurston_c(false, true, d, n, out);
}
int rt_lozenge_min_tiling(const int32_t* d, size_t n, rt_vec** out) {
  return thurston_c(false, false, d, n, out);
}

rt_vec* rt_domino_heightfunc_to_tiling(const int32_t* hf, size_t n, const int32_t* d,
                                       size_t dsize) {
  if (!hf || !is_square(n) || !d) return nullptr;
  return make_vec(domino_heightfunc_to_tiling(from_c(hf, n), from_c(d, dsize)));
}

rt_vec* rt_domino_tiling_to_heightfunc(const int32_t* t, size_t n, const int32_t* d,
                                       size_t dsize) {
  if (!t || !is_square(n) || !d || !is_square(dsize)) return nullptr;
  return make_vec(domino_tiling_to_heightfunc(from_c(t, n), from_c(d, dsize)));
}

rt_vec* rt_domino_tiling_to_domain(const int32_t* t, size_t n) {
  if (!t || !is_square(n)) return nullptr;
  return make_vec(domino_tiling_to_domain(from_c(t, n)));
}

rt_vec* rt_domino_domain_to_vertices(const int32_t* d, size_t dsize) {
  if (!d || !is_square(dsize)) return nullptr;
  return make_vec(domino_domain_to_vertices(from_c(d, dsize)));
}

rt_vec* rt_lozenge_almost_hex_domain(int n, int m) {
  if (n < 1 || m < 0) return nullptr;
  return make_vec(lozenge_almost_hex_domain(n, m));
}

rt_vec* rt_lozenge_heightfunc_to_tiling(const int32_t* hf, size_t n, const int32_t* d,
                                        size_t dsize) {
  if (!hf || !is_square(n) || !d) return nullptr;
  return make_vec(lozenge_heightfunc_to_tiling(from_c(hf, n), from_c(d, dsize)));
}

rt_vec* rt_lozenge_tiling_to_heightfunc(const int32_t* t, size_t n, const int32_t* d,
                                        size_t dsize) {
  if (!t || !is_square(n) || !d || dsize % 2 || !is_square(dsize / 2)) return nullptr;
  return make_vec(lozenge_tiling_to_heightfunc(from_c(t, n), from_c(d, dsize)));
}

rt_vec* rt_lozenge_tiling_to_domain(const int32_t* t, size_t n) {
  if (!t || !is_square(n)) return nullptr;
  return make_vec(lozenge_tiling_to_domain(from_c(t, n)));
}

rt_vec* rt_lozenge_domain_to_vertices(const int32_t* d, size_t dsize) {
  if (!d || dsize % 2 || !is_square(dsize / 2)) return nullptr;
  return make_vec(lozenge_domain_to_vertices(from_c(d, dsize)));
}

rt_vec* rt_recttri_max_hex(int n) {
  if (n < 1) return nullptr;
  return make_vec(recttri_max_hex(n));
}

rt_vec* rt_recttri_min_hex(int n) {
  if (n < 1) return nullptr;
  return make_vec(recttri_min_hex(n));
}

rt_vec* rt_recttri_slope_hex(int n) {
  if (n < 1) return nullptr;
  return make_vec(recttri_slope_hex(n));
}

// Bibone statics.  Tilings are N x N with N even.
static bool bibone_side(size_t n, int* N) {
  if (!is_square(n)) return false;
  *N = (int)std::lround(std::sqrt((double)n));
  return *N >= 6 && *N % 2 == 0;
}

rt_vec* rt_bibone_ice_cream_cone(int N, int M) {
  if (N < 1 || M < 0) return nullptr;
  return make_vec(bibone_ice_cream_cone(N, M));
}

rt_vec* rt_bibone_test_butterfly(int type) { return make_vec(bibone_test_butterfly(type)); }

rt_vec* rt_bibone_hexagon(int a, int b, int c) {
  if (a < 0 || b < 0 || c < 0 || (long)a + b + c > 20000) return nullptr;
  tiling t = bibone_hexagon(a, b, c);
  if (t.empty()) return nullptr;
  return make_vec(std::move(t));
}

#define RT_BIBONE_STATIC(name, fn)                          \
  rt_vec* name(const int32_t* t, size_t n) {                \
    int N;                                                  \
    if (!t || !bibone_side(n, &N)) return nullptr;          \
    return make_vec(fn(from_c(t, n), N));                   \
  }
RT_BIBONE_STATIC(rt_bibone_triangle_flips, bibone_triangle_flips)
RT_BIBONE_STATIC(rt_bibone_butterfly_flips, bibone_butterfly_flips)
RT_BIBONE_STATIC(rt_bibone_tiling_to_vertices, bibone_tiling_to_vertices)
RT_BIBONE_STATIC(rt_bibone_tiling_to_domain, bibone_tiling_to_domain)
#undef RT_BIBONE_STATIC

int rt_bibone_edge_form(const int32_t* t, size_t n, uint8_t* out) {
  int N;
  if (!t || !out || !bibone_side(n, &N)) return RT_EINVAL;
  const std::vector<uint8_t> e = bibone_edge_form(bibone_planes(from_c(t, n), N), N / 2);
  if (e.empty()) return RT_EINVAL;
  std::copy(e.begin(), e.end(), out);
  return RT_OK;
}

rt_vec* rt_bibone_dimer_indicator(const int32_t* t, size_t n, int c) {
  int N;
  if (!t || !bibone_side(n, &N)) return nullptr;
  return make_vec(bibone_dimer_indicator(from_c(t, n), N, c));
}

}  // extern "C"

// rtiling_tilers.hpp -- C++ drop-in for the reference's per-model tiler
// classes (src/Domino/DominoTiler.h, src/Lozenge/LozengeTiler.h,
// src/RectTriangle/RectTriangleTiler.h), header-only over the C ABI in
// rtiling.h.  Link with librtiling.so.
//
// What changes for a caller of the reference:
//  * the constructor takes a HIP device ordinal and the reference's kernel
//    file name (which selects the flip rule) instead of
//    (cl::Context, cl::CommandQueue, devices, source path, err);
//  * LoadTinyMT(path, size) reads the same tinymt32dc text format;
//    LoadTinyMT(params) takes the table in memory;
//  * errors are thrown as rt::Error (std::runtime_error) with the engine's
//    message, where the reference threw cl::Error / runtime_error, printed,
//    or called exit();
//  * MinTiling/MaxTiling of a region that is not tileable throw instead of
//    returning the partial height function.
// Everything else -- names, argument meaning, in-place synchronous Walk,
// `tiling`/`domain`/`heightfunc` as std::vector<int> -- is the reference's.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtiling.h"

typedef std::vector<int> domain;
typedef std::vector<int> heightfunc;
typedef std::vector<int> tiling;

namespace rt {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline std::vector<int> take(rt_vec* v) {
  if (!v) throw Error(RT_EINVAL, "helper rejected its arguments");
  std::vector<int> out(rt_vec_data(v), rt_vec_data(v) + rt_vec_size(v));
  rt_vec_free(v);
  return out;
}

inline std::vector<int> thurston(int (*f)(const int32_t*, size_t, rt_vec**), const domain& d) {
  rt_vec* v = nullptr;
  int rc = f(d.data(), d.size(), &v);
  if (rc == RT_EUNTILEABLE) throw Error(rc, "!!This region is NOT tileable!!");
  if (rc) throw Error(rc, "Thurston failed");
  return take(v);
}

inline void svg(int rc) {
  if (rc) throw Error(rc, "cannot write the SVG file");
}

inline int side_of(const std::vector<int>& t) {
  int n = 0;
  while ((size_t)(n + 1) * (n + 1) <= t.size()) n++;
  return n;
}

// Owns an rt_ctx; shared plumbing of the tiler classes.
class Device {
 public:
  explicit Device(int ordinal) {
    int rc = rt_create(&ctx_, ordinal);
    if (rc) throw Error(rc, "no HIP device " + std::to_string(ordinal));
  }
  ~Device() { rt_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  rt_ctx* ctx() const { return ctx_; }
  void check(int rc) const {
    if (rc) throw Error(rc, rt_last_error(ctx_));
  }
  void LoadTinyMT(const std::string& path, int size) {
    check(rt_load_tinymt_file(ctx_, path.c_str(), (size_t)size));
  }
  void LoadTinyMT(const std::vector<uint32_t>& mat1_mat2_tmat) {
    check(rt_set_tinymt_params(ctx_, mat1_mat2_tmat.data(), mat1_mat2_tmat.size() / 3));
  }

 private:
  rt_ctx* ctx_ = nullptr;
};

}  // namespace rt

// ------------------------------------------------------- matrix text I/O
// The reference's only on-disk tiling format (src/common/common.cpp:115-179):
// rows of space-separated integers, each row ending in " \n"; the loader
// reads every maximal run of decimal digits as one non-negative integer.
inline void SaveMatrix(const std::vector<int>& m, int M, int N, const std::string& filename) {
  FILE* f = std::fopen(filename.c_str(), "w");
  if (!f) throw rt::Error(RT_EIO, "cannot open " + filename);
  for (int i = 0; i < M; i++) {
    for (int j = 0; j < N; j++) std::fprintf(f, "%d ", m[(size_t)i * N + j]);
    std::fputc('\n', f);
  }
  std::fclose(f);
}

inline void SaveMatrix(const std::vector<int>& m, const std::string& filename) {
  const int n = rt::side_of(m);
  SaveMatrix(m, n, n, filename);
}

inline std::vector<int> LoadMatrix(const std::string& filename) {
  FILE* f = std::fopen(filename.c_str(), "r");
  if (!f) throw rt::Error(RT_EIO, "Cannot open file: " + filename);
  std::vector<int> out;
  int c, v = 0;
  bool in = false;
  while ((c = std::fgetc(f)) != EOF) {
    if (c >= '0' && c <= '9') {
      v = in ? 10 * v + (c - '0') : (c - '0');
      in = true;
    } else if (in) {
      out.push_back(v);
      in = false;
    }
  }
  if (in) out.push_back(v);
  std::fclose(f);
  return out;
}

// Zero padding around an M x N matrix (common.cpp:4-38).
inline void PadMatrix(std::vector<int>& v, int M, int N) {
  std::vector<int> p((size_t)(M + 2) * (N + 2), 0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++) p[(size_t)(i + 1) * (N + 2) + j + 1] = v[(size_t)i * N + j];
  v.swap(p);
}

inline void UnPadMatrix(std::vector<int>& v, int M, int N) {
  std::vector<int> u((size_t)(M - 2) * (N - 2));
  for (int i = 0; i < M - 2; i++)
    for (int j = 0; j < N - 2; j++) u[(size_t)i * (N - 2) + j] = v[(size_t)(i + 1) * N + j + 1];
  v.swap(u);
}

// ------------------------------------------------------------ DominoTiler
class DominoTiler : public rt::Device {
 public:
  // `kernel` names the reference kernel file whose flip rule to use:
  // dominokernel.cl (toggle 0.8), dominokernelCFTP.cl (heat-bath 1/2),
  // dominokerneltwoperiodic.cl, dominokernelCFTPtwoperiodic.cl; a, b are
  // the two-periodic weights that file #defines (0.7 and 1.0).
  explicit DominoTiler(int device = 0, const std::string& kernel = "dominokernel.cl",
                       double a = 0.7, double b = 1.0)
      : rt::Device(device) {
    rule_.a = a;
    rule_.b = b;
    rule_.kind = RT_DOMINO_TOGGLE08;
    if (kernel.find("CFTPtwoperiodic") != std::string::npos) rule_.kind = RT_DOMINO_HEATBATH2P;
    else if (kernel.find("twoperiodic") != std::string::npos) rule_.kind = RT_DOMINO_METRO2P;
    else if (kernel.find("CFTP") != std::string::npos) rule_.kind = RT_DOMINO_HEATBATH;
  }

  void Walk(tiling& t, long steps, long seed) {
    check(rt_domino_walk(ctx(), t.data(), rt::side_of(t), steps, seed, &rule_));
  }
  void Walk(tiling& t, std::vector<long> steps, std::vector<long> seeds) {
    if (steps.size() != seeds.size()) throw rt::Error(RT_EINVAL, "steps/seeds length differ");
    check(rt_domino_walk_segments(ctx(), t.data(), rt::side_of(t), steps.data(), seeds.data(),
                                  (int)steps.size(), &rule_));
  }
  // Exact sample by coupling from the past (Examples/AztecDiamondCftpEx.cpp:51-77).
  tiling CFTP(const tiling& tMax, const tiling& tMin, uint32_t master_seed,
              int sched = RT_SCHED_REFERENCE, long t0 = 128, int max_rounds = 64,
              rt_cftp_stats* stats = nullptr) {
    tiling out(tMax.size());
    check(rt_domino_cftp(ctx(), tMax.data(), tMin.data(), rt::side_of(tMax), master_seed, sched,
                         t0, max_rounds, &rule_, out.data(), stats));
    return out;
  }

  static domain Rectangle(int M, int N) { return rt::take(rt_domino_rectangle(M, N)); }
  static domain AztecRectangle(int N, const std::vector<int>& remove) {
    return rt::take(rt_domino_aztec_rectangle(N, remove.data(), (int)remove.size()));
  }
  static domain AztecDiamond(int N) { return rt::take(rt_domino_aztec_diamond(N)); }
  static tiling MaxTiling(const domain& d) { return rt::thurston(rt_domino_max_tiling, d); }
  static tiling MinTiling(const domain& d) { return rt::thurston(rt_domino_min_tiling, d); }
  static std::vector<int> DomainToVertices(const domain& d) {
    return rt::take(rt_domino_domain_to_vertices(d.data(), d.size()));
  }
  static domain TilingToDomain(const tiling& t) {
    return rt::take(rt_domino_tiling_to_domain(t.data(), t.size()));
  }
  static void DomainToSVG(const domain& d, std::string f) {
    rt::svg(rt_domino_domain_svg(d.data(), d.size(), f.c_str()));
  }
  static void TilingToSVG(const tiling& t, std::string f) {
    rt::svg(rt_domino_tiling_svg(t.data(), t.size(), f.c_str()));
  }
  static void MayaToSVG(const tiling& t, std::string f) {
    rt::svg(rt_domino_maya_svg(t.data(), t.size(), f.c_str()));
  }
  static void DimerToSVG(const tiling& t, std::string f) {
    rt::svg(rt_domino_dimer_svg(t.data(), t.size(), nullptr, 0, f.c_str()));
  }
  static void DimerToSVG(const tiling& t, const domain& d, std::string f) {
    rt::svg(rt_domino_dimer_svg(t.data(), t.size(), d.data(), d.size(), f.c_str()));
  }
  static tiling HeightfuncToTiling(const heightfunc& hf, const domain& d) {
    return rt::take(rt_domino_heightfunc_to_tiling(hf.data(), hf.size(), d.data(), d.size()));
  }
  static heightfunc TilingToHeightfunc(const tiling& t, const domain& d) {
    return rt::take(rt_domino_tiling_to_heightfunc(t.data(), t.size(), d.data(), d.size()));
  }

 private:
  rt_domino_rule rule_{};
};

// ----------------------------------------------------------- LozengeTiler
class LozengeTiler : public rt::Device {
 public:
  explicit LozengeTiler(int device = 0) : rt::Device(device) {}

  void Walk(tiling& t, int steps, long seed) {
    check(rt_lozenge_walk(ctx(), t.data(), rt::side_of(t), steps, seed));
  }
  void Walk(tiling& t, std::vector<long> steps, std::vector<long> seeds) {
    if (steps.size() != seeds.size()) throw rt::Error(RT_EINVAL, "steps/seeds length differ");
    check(rt_lozenge_walk_segments(ctx(), t.data(), rt::side_of(t), steps.data(), seeds.data(),
                                   (int)steps.size()));
  }
  tiling CFTP(const tiling& tMax, const tiling& tMin, uint32_t master_seed,
              int sched = RT_SCHED_DOUBLING, long t0 = 1024, int max_rounds = 64,
              rt_cftp_stats* stats = nullptr) {
    tiling out(tMax.size());
    check(rt_lozenge_cftp(ctx(), tMax.data(), tMin.data(), rt::side_of(tMax), master_seed, sched,
                          t0, max_rounds, out.data(), stats));
    return out;
  }

  static domain AlmostHexDomain(int N, int M) { return rt::take(rt_lozenge_almost_hex_domain(N, M)); }
  static domain Hexagon(int N) { return AlmostHexDomain(N, 0); }
  static tiling MaxTiling(const domain& d) { return rt::thurston(rt_lozenge_max_tiling, d); }
  static tiling MinTiling(const domain& d) { return rt::thurston(rt_lozenge_min_tiling, d); }
  static std::vector<int> DomainToVertices(const domain& d) {
    return rt::take(rt_lozenge_domain_to_vertices(d.data(), d.size()));
  }
  static domain TilingToDomain(const tiling& t) {
    return rt::take(rt_lozenge_tiling_to_domain(t.data(), t.size()));
  }
  static void DomainToSVG(const domain& d, std::string f) {
    rt::svg(rt_lozenge_domain_svg(d.data(), d.size(), f.c_str()));
  }
  static void TilingToSVG(const tiling& t, std::string f) {
    rt::svg(rt_lozenge_tiling_svg(t.data(), t.size(), f.c_str()));
  }
  static void DimerToSVG(const tiling& t, std::string f) {
    rt::svg(rt_lozenge_dimer_svg(t.data(), t.size(), nullptr, 0, f.c_str()));
  }
  static void DimerToSVG(const tiling& t, const domain& d, std::string f) {
    rt::svg(rt_lozenge_dimer_svg(t.data(), t.size(), d.data(), d.size(), f.c_str()));
  }
  static tiling HeightfuncToTiling(const heightfunc& hf, const domain& d) {
    return rt::take(rt_lozenge_heightfunc_to_tiling(hf.data(), hf.size(), d.data(), d.size()));
  }
  static heightfunc TilingToHeightfunc(const tiling& t, const domain& d) {
    return rt::take(rt_lozenge_tiling_to_heightfunc(t.data(), t.size(), d.data(), d.size()));
  }
};

// ------------------------------------------------------ RectTriangleTiler
class RectTriangleTiler : public rt::Device {
 public:
  // weights: the kernel's 16-entry table (recttrianglekernel.cl:13); empty
  // means the reference's all-1.0 weights.
  explicit RectTriangleTiler(int device = 0, std::vector<float> weights = {})
      : rt::Device(device), w_(std::move(weights)) {}

  void Walk(tiling& t, int steps, long seed) {
    check(rt_recttri_walk(ctx(), t.data(), rt::side_of(t), steps, seed,
                          w_.size() == 16 ? w_.data() : nullptr));
  }

  static tiling maxHex(int N) { return rt::take(rt_recttri_max_hex(N)); }
  static tiling minHex(int N) { return rt::take(rt_recttri_min_hex(N)); }
  static tiling slopeHex(int N) { return rt::take(rt_recttri_slope_hex(N)); }
  static void TilingToSVG(tiling& t, std::string f) { rt::svg(rt_recttri_tiling_svg(t.data(), t.size(), f.c_str())); }

 private:
  std::vector<float> w_;
};

// ----------------------------------------------------- TriangleDimerTiler
// Bibone (triangular-lattice dimer) model (TriangleDimerTiler.h:124-209).
class TriangleDimerTiler : public rt::Device {
 public:
  explicit TriangleDimerTiler(int device = 0) : rt::Device(device) {}

  // TriangleDimerTiler.cpp:101-695.  LoadTinyMT's size must be >= (N/2)^2.
  void Walk(tiling& t, int steps, long seed) {
    check(rt_bibone_walk(ctx(), t.data(), rt::side_of(t), steps, seed));
  }

  static tiling IceCreamCone(int N, int M) { return rt::take(rt_bibone_ice_cream_cone(N, M)); }
  static tiling TestButterfly(int t) { return rt::take(rt_bibone_test_butterfly(t)); }
  // Hexagon with sides a, b, c, a, b, c (no reference counterpart).
  static tiling Hexagon(int a, int b, int c) { return rt::take(rt_bibone_hexagon(a, b, c)); }
  static std::vector<int> TilingToTriangleFlips(tiling& t) {
    return rt::take(rt_bibone_triangle_flips(t.data(), t.size()));
  }
  static std::vector<int> TilingToButterflyFlips(tiling& t) {
    return rt::take(rt_bibone_butterfly_flips(t.data(), t.size()));
  }
  static std::vector<int> TilingToVertices(tiling& t) {
    return rt::take(rt_bibone_tiling_to_vertices(t.data(), t.size()));
  }
  static std::vector<int> DimerIndicator(tiling& t, int c) {
    return rt::take(rt_bibone_dimer_indicator(t.data(), t.size(), c));
  }
  static domain TilingToDomain(tiling& t) {
    return rt::take(rt_bibone_tiling_to_domain(t.data(), t.size()));
  }
  static void TilingToSVG(tiling& t, std::string f) { rt::svg(rt_bibone_tiling_svg(t.data(), t.size(), f.c_str())); }
  static void DimerToSVG(tiling& t, std::string f) {
    rt::svg(rt_bibone_dimer_svg(t.data(), t.size(), nullptr, 0, f.c_str()));
  }
  static void DimerToSVG(tiling& t, domain& d, std::string f) {
    rt::svg(rt_bibone_dimer_svg(t.data(), t.size(), d.data(), d.size(), f.c_str()));
  }
  static void DomainToSVG(domain& d, std::string f) { rt::svg(rt_bibone_domain_svg(d.data(), d.size(), f.c_str())); }
};

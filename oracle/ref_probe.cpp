// ref_probe.cpp -- extern "C" entry points over the reference's own host
// statics, its TinyMT host generator and its parameter-file reader,
// compiled together with the reference's sources into
// oracle/_ref/libref_host.so (see oracle/Makefile).  Test infrastructure:
// only tests/golden scripts and CPU tests load it, to pin the engine's
// host helpers against the reference implementation itself.
#include <cstdint>
#include <cstring>
#include <exception>
#include <fstream>
#include <string>
#include <vector>

#include "Domino/DominoTiler.h"
#include "TinyMT/file_reader.h"
#include "TinyMT/tinymt32.h"
#include "Lozenge/LozengeTiler.h"
#include "RectTriangle/RectTriangleTiler.h"
#include "TriangleDimer/TriangleDimerTiler.h"

namespace {
// Results are returned through a caller buffer; the call returns the
// element count (or, when cap is too small, the negated required count).
long give(const std::vector<int>& v, int32_t* out, long cap) {
  if ((long)v.size() > cap) return -(long)v.size();
  for (size_t k = 0; k < v.size(); k++) out[k] = v[k];
  return (long)v.size();
}
std::vector<int> vec(const int32_t* p, long n) { return std::vector<int>(p, p + n); }
}  // namespace

extern "C" {
long ref_domino_aztec_diamond(int n, int32_t* out, long cap) {
  return give(DominoTiler::AztecDiamond(n), out, cap);
}
long ref_domino_aztec_rectangle(int n, const int* rem, int k, int32_t* out, long cap) {
  return give(DominoTiler::AztecRectangle(n, std::vector<int>(rem, rem + k)), out, cap);
}
long ref_domino_rectangle(int m, int n, int32_t* out, long cap) {
  return give(DominoTiler::Rectangle(m, n), out, cap);
}
long ref_domino_max_tiling(const int32_t* d, long nd, int32_t* out, long cap) {
  return give(DominoTiler::MaxTiling(vec(d, nd)), out, cap);
}
long ref_domino_min_tiling(const int32_t* d, long nd, int32_t* out, long cap) {
  return give(DominoTiler::MinTiling(vec(d, nd)), out, cap);
}
long ref_domino_tiling_to_heightfunc(const int32_t* t, long nt, const int32_t* d, long nd,
                                     int32_t* out, long cap) {
  return give(DominoTiler::TilingToHeightfunc(vec(t, nt), vec(d, nd)), out, cap);
}
long ref_domino_tiling_to_domain(const int32_t* t, long nt, int32_t* out, long cap) {
  return give(DominoTiler::TilingToDomain(vec(t, nt)), out, cap);
}
long ref_domino_domain_to_vertices(const int32_t* d, long nd, int32_t* out, long cap) {
  return give(DominoTiler::DomainToVertices(vec(d, nd)), out, cap);
}
long ref_lozenge_almost_hex_domain(int n, int m, int32_t* out, long cap) {
  return give(LozengeTiler::AlmostHexDomain(n, m), out, cap);
}
long ref_lozenge_max_tiling(const int32_t* d, long nd, int32_t* out, long cap) {
  return give(LozengeTiler::MaxTiling(vec(d, nd)), out, cap);
}
long ref_lozenge_min_tiling(const int32_t* d, long nd, int32_t* out, long cap) {
  return give(LozengeTiler::MinTiling(vec(d, nd)), out, cap);
}
long ref_lozenge_tiling_to_heightfunc(const int32_t* t, long nt, const int32_t* d, long nd,
                                      int32_t* out, long cap) {
  return give(LozengeTiler::TilingToHeightfunc(vec(t, nt), vec(d, nd)), out, cap);
}
long ref_lozenge_tiling_to_domain(const int32_t* t, long nt, int32_t* out, long cap) {
  return give(LozengeTiler::TilingToDomain(vec(t, nt)), out, cap);
}
long ref_recttri_max_hex(int n, int32_t* out, long cap) {
  return give(RectTriangleTiler::maxHex(n), out, cap);
}
long ref_recttri_min_hex(int n, int32_t* out, long cap) {
  return give(RectTriangleTiler::minHex(n), out, cap);
}
long ref_recttri_slope_hex(int n, int32_t* out, long cap) {
  return give(RectTriangleTiler::slopeHex(n), out, cap);
}
long ref_bibone_ice_cream_cone(int n, int m, int32_t* out, long cap) {
  return give(TriangleDimerTiler::IceCreamCone(n, m), out, cap);
}
long ref_bibone_test_butterfly(int type, int32_t* out, long cap) {
  return give(TriangleDimerTiler::TestButterfly(type), out, cap);
}
long ref_bibone_triangle_flips(const int32_t* t, long nt, int32_t* out, long cap) {
  std::vector<int> v = vec(t, nt);
  return give(TriangleDimerTiler::TilingToTriangleFlips(v), out, cap);
}
long ref_bibone_butterfly_flips(const int32_t* t, long nt, int32_t* out, long cap) {
  std::vector<int> v = vec(t, nt);
  return give(TriangleDimerTiler::TilingToButterflyFlips(v), out, cap);
}
long ref_bibone_tiling_to_vertices(const int32_t* t, long nt, int32_t* out, long cap) {
  std::vector<int> v = vec(t, nt);
  return give(TriangleDimerTiler::TilingToVertices(v), out, cap);
}
long ref_bibone_tiling_to_domain(const int32_t* t, long nt, int32_t* out, long cap) {
  std::vector<int> v = vec(t, nt);
  return give(TriangleDimerTiler::TilingToDomain(v), out, cap);
}
// SVG writers: write `path`, return 0.
int ref_svg(const char* model, const char* what, const int32_t* t, long nt, const int32_t* d,
            long nd, const char* path) {
  std::vector<int> a = vec(t, nt);
  std::vector<int> b = d ? vec(d, nd) : std::vector<int>();
  const std::string m(model), w(what), f(path);
  if (m == "domino") {
    if (w == "domain") DominoTiler::DomainToSVG(a, f);
    else if (w == "tiling") DominoTiler::TilingToSVG(a, f);
    else if (w == "maya") DominoTiler::MayaToSVG(a, f);
    else if (w == "dimer") d ? DominoTiler::DimerToSVG(a, b, f) : DominoTiler::DimerToSVG(a, f);
    else return -1;
  } else if (m == "lozenge") {
    if (w == "domain") LozengeTiler::DomainToSVG(a, f);
    else if (w == "tiling") LozengeTiler::TilingToSVG(a, f);
    else if (w == "dimer") d ? LozengeTiler::DimerToSVG(a, b, f) : LozengeTiler::DimerToSVG(a, f);
    else return -1;
  } else if (m == "recttri") {
    if (w == "tiling") RectTriangleTiler::TilingToSVG(a, f);
    else return -1;
  } else if (m == "bibone") {
    if (w == "domain") TriangleDimerTiler::DomainToSVG(a, f);
    else if (w == "tiling") TriangleDimerTiler::TilingToSVG(a, f);
    else if (w == "dimer") d ? TriangleDimerTiler::DimerToSVG(a, b, f) : TriangleDimerTiler::DimerToSVG(a, f);
    else return -1;
  } else {
    return -1;
  }
  return 0;
}
long ref_bibone_dimer_indicator(const int32_t* t, long nt, int c, int32_t* out, long cap) {
  std::vector<int> v = vec(t, nt);
  return give(TriangleDimerTiler::DimerIndicator(v, c), out, cap);
}
// The reference's host-side generator (src/TinyMT/tinymt32.h:72-136,
// 205-208): n draws of tinymt32_generate_float01 / _uint32 from the state
// s7 = {s0, s1, s2, s3, mat1, mat2, tmat}; the state advances.
void ref_h_tinymt_float01(uint32_t* s7, float* out, int n) {
  tinymt32_t t;
  for (int k = 0; k < 4; k++) t.status[k] = s7[k];
  t.mat1 = s7[4];
  t.mat2 = s7[5];
  t.tmat = s7[6];
  for (int k = 0; k < n; k++) out[k] = tinymt32_generate_float01(&t);
  for (int k = 0; k < 4; k++) s7[k] = t.status[k];
}
void ref_h_tinymt_uint32(uint32_t* s7, uint32_t* out, int n) {
  tinymt32_t t;
  for (int k = 0; k < 4; k++) t.status[k] = s7[k];
  t.mat1 = s7[4];
  t.mat2 = s7[5];
  t.tmat = s7[6];
  for (int k = 0; k < n; k++) out[k] = tinymt32_generate_uint32(&t);
  for (int k = 0; k < 4; k++) s7[k] = t.status[k];
}
// The reference's parameter-file reader (src/TinyMT/file_reader.cpp:51-70,
// 103-166): `count` successive file_reader::get calls, as get_params_buffer
// makes them (22-38).  Returns 0, -2 when the reader throws (format error,
// comma not found; the reference then ends the program) or -3 when the
// file does not open.
int ref_parse_params(const char* path, long count, uint32_t* out3) {
  std::ifstream probe(path);
  if (!probe) return -3;
  probe.close();
  try {
    tinymt::file_reader fr(path);
    for (long k = 0; k < count; k++) fr.get(&out3[3 * k], &out3[3 * k + 1], &out3[3 * k + 2]);
  } catch (const std::exception&) {
    return -2;
  }
  return 0;
}
}

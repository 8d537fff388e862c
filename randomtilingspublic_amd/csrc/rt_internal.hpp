// rt_internal.hpp -- host-side declarations shared by the translation units
// of librtiling.so (not part of the public C ABI).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/rtiling.h"

struct rt_vec {
  std::vector<int32_t> v;
};

namespace rt {

using tiling = std::vector<int>;
using domain = std::vector<int>;
using heightfunc = std::vector<int>;

// Colour schedules of the reference's host loops.
// Domino: std::mt19937(seed); single-seed Walk spends the first output on
// the TinyMT base seed (DominoTiler.cpp:51-53); colour r = out % 2, r == 1
// is black (70-79).  Colour index 0 = black, 1 = white.
void domino_schedule(uint32_t seed, long steps, bool skip_first,
                     uint32_t* base_out, std::vector<uint8_t>& colours);
// Lozenge and rectangle-triangle: std::default_random_engine(seed) with
// uniform_int_distribution<int>(0,2) (LozengeTiler.cpp:86-91,
// RectTriangleTiler.cpp:69-71,92).
void tri_schedule(long seed, long steps, std::vector<uint8_t>& colours);

int parse_tinymt_file(const char* path, size_t count, uint32_t* out3,
                      std::string* err);

rt_vec* make_vec(std::vector<int32_t>&& v);

// Host statics (rt_host.cpp)
domain domino_aztec_rectangle(int n, const std::vector<int>& remove);
domain domino_rectangle(int M, int N);
domain domino_rect_union(const std::vector<int>& rects);
std::vector<int> domino_domain_to_vertices(const domain& d);
int domino_thurston(const domain& d, bool max, tiling& out);
tiling domino_heightfunc_to_tiling(const heightfunc& hf, const domain& d);
heightfunc domino_tiling_to_heightfunc(const tiling& t, const domain& d);
domain domino_tiling_to_domain(const tiling& t);

domain lozenge_almost_hex_domain(int n, int m);
std::vector<int> lozenge_domain_to_vertices(const domain& d);
int lozenge_thurston(const domain& d, bool max, tiling& out);
tiling lozenge_heightfunc_to_tiling(const heightfunc& hf, const domain& d);
heightfunc lozenge_tiling_to_heightfunc(const tiling& t, const domain& d);
domain lozenge_tiling_to_domain(const tiling& t);

tiling recttri_max_hex(int n);
tiling recttri_min_hex(int n);
tiling recttri_slope_hex(int n);

// Bibone statics (rt_bibone_host.cpp); N is the tiling side.
tiling bibone_ice_cream_cone(int N, int M);
tiling bibone_test_butterfly(int type);
tiling bibone_hexagon(int a, int b, int c);
std::vector<int> bibone_triangle_flips(const tiling& t, int N);
std::vector<int> bibone_butterfly_flips(const tiling& t, int N);
std::vector<int> bibone_tiling_to_vertices(const tiling& t, int N);
domain bibone_tiling_to_domain(const tiling& t, int N);
std::vector<int> bibone_dimer_indicator(const tiling& t, int N, int c);
// The eight planes of the walk from a tiling, the edge form of such a start
// (empty when the edge-form walk does not apply to it) and the lozenge
// cells written back from an edge form (rt_bibone_host.cpp).
std::vector<int32_t> bibone_planes(const tiling& t, int N);
std::vector<uint8_t> bibone_edge_form(const std::vector<int32_t>& planes, int M);
void bibone_edge_merge(const std::vector<uint8_t>& e, int M, int32_t* tiling);

}  // namespace rt

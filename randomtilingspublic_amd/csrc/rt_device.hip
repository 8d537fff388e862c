// rt_device.hip -- context, device arenas, the one-launch-per-step kernels
// and the walk / CFTP entry points of the C ABI (include/rtiling.h).
//
// Formulation (identical results to the reference's Rotate-then-Update
// launch pair, DominoTiler.cpp:73-79, LozengeTiler.cpp:95-104,
// RectTriangleTiler.cpp:95-107):
//   The reference rebuilds the non-rotated colour arrays after every step.
//   Those rebuilt cells are only read when that colour is rotated next (or
//   when the tiling is copied back), so each step here is ONE launch that
//   first pulls the rotated colour's cells from the neighbours rotated since
//   its own last rotation, then applies the local move.  Colour arrays that
//   were not rotated last are brought up to date once, at the end.
//   Per step that is one TinyMT draw (28 B read, 16 B written) and ~1-5 B of
//   tiling traffic per work item instead of two launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <random>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "models_dev.hpp"
#include "rt_internal.hpp"
#include "rt_fused.hpp"
#include "rt_kernels.hpp"
#include "tinymt_dev.hpp"

using namespace rt;

// ------------------------------------------------------------------ kernels
namespace rt {

__global__ void k_tiny_init(StreamTable S, long n, uint32_t seed) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  Tiny t;
  t.mat1 = S.param[id];
  t.mat2 = S.param[S.pstride + id];
  t.tmat = S.param[2 * S.pstride + id];
  tiny_init(t, seed + (uint32_t)id);
  tiny_store(S, id, t);
}

__global__ void k_fill_u32(uint32_t* p, long n, uint32_t v) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) p[id] = v;
}

// Sets *flag to 1 if a and b differ anywhere (wavefront vote, one atomic
// per wave that saw a difference).
__global__ void k_differ(const uint8_t* a, const uint8_t* b, long n, uint32_t* flag) {
  bool diff = false;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    diff |= a[k] != b[k];
  if (__any(diff) && (threadIdx.x & (warpSize - 1)) == 0) atomicOr(flag, 1u);
}

// Domino: colour 0 = black ((i+j) even), 1 = white; the reference's kernel
// argument t is 1 for black, 0 for white (DominoTiler.cpp:73-79).
template <int NC>
__global__ void k_domino_step(DominoArrays c0, DominoArrays c1, StreamTable S, int N,
                              int colour, int pull, int rule, DominoWeights w) {
  const int H = N / 2, Wd = H - 2;
  const long P = (long)(N - 2) * Wd;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P) return;
  const int i = (int)((unsigned)gid / (unsigned)Wd) + 1, j = (int)((unsigned)gid % (unsigned)Wd) + 1;
  Tiny t = tiny_load(S, gid);
  const uint32_t rd = tiny_bits(t);
  tiny_store(S, gid, t);
  const int tpar = 1 - colour;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const DominoArrays& A = c == 0 ? c0 : c1;
    uint8_t* X = A.col[colour];
    uint32_t v = pull ? domino_rebuild(A.col[1 - colour] + i * H + j, H, i, 1 - tpar)
                      : (uint32_t)X[i * H + j];
    X[i * H + j] = (uint8_t)domino_move(v, rd, rule, i, tpar, w);
  }
}

template <int NC>
__global__ void k_domino_sync(DominoArrays c0, DominoArrays c1, int N, int last) {
  const int H = N / 2, Wd = H - 2;
  const long P = (long)(N - 2) * Wd;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P) return;
  const int i = (int)((unsigned)gid / (unsigned)Wd) + 1, j = (int)((unsigned)gid % (unsigned)Wd) + 1;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const DominoArrays& A = c == 0 ? c0 : c1;
    A.col[1 - last][i * H + j] = (uint8_t)domino_rebuild(A.col[last] + i * H + j, H, i, 1 - last);
  }
}

// Lozenge: colours 0,1,2 = R,B,G by ((j-i)%3+3)%3 (LozengeTiler.cpp:40-46).
// pull_a: colour-1 was rotated after this colour -> refresh bits {2,4,32};
// pull_b: colour-2 was rotated after this colour -> refresh bits {1,8,16}.
template <int NC>
__global__ void k_lozenge_step(TriArrays8 c0, TriArrays8 c1, StreamTable S, int N, int colour,
                               int pull_a, int pull_b) {
  const int M = N / 3;
  const long P = (long)N * M;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P) return;
  const int i = (int)((unsigned)gid / (unsigned)M), j = (int)((unsigned)gid % (unsigned)M);
  Tiny t = tiny_load(S, gid);
  const uint32_t rd = tiny_bits(t);
  tiny_store(S, gid, t);
  const bool interior = i >= 1 && i <= N - 2 && j >= 1 && j <= M - 2;
  const int ca = (colour + 2) % 3, cb = (colour + 1) % 3;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const TriArrays8& A = c == 0 ? c0 : c1;
    uint8_t* X = A.col[colour];
    uint32_t v = X[gid];
    if (interior) {
      if (pull_a) v = lozenge_pull_prev(v, A.col[ca] + gid, M, i, ca);
      if (pull_b) v = lozenge_pull_prev2(v, A.col[cb] + gid, M, i, cb);
    }
    X[gid] = (uint8_t)lozenge_move(v, rd);
  }
}

template <int NC>
__global__ void k_lozenge_sync(TriArrays8 c0, TriArrays8 c1, int N, int colour, int pull_a,
                               int pull_b) {
  const int M = N / 3;
  const long P = (long)N * M;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P) return;
  const int i = (int)((unsigned)gid / (unsigned)M), j = (int)((unsigned)gid % (unsigned)M);
  if (!(i >= 1 && i <= N - 2 && j >= 1 && j <= M - 2)) return;
  const int ca = (colour + 2) % 3, cb = (colour + 1) % 3;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const TriArrays8& A = c == 0 ? c0 : c1;
    uint32_t v = A.col[colour][gid];
    if (pull_a) v = lozenge_pull_prev(v, A.col[ca] + gid, M, i, ca);
    if (pull_b) v = lozenge_pull_prev2(v, A.col[cb] + gid, M, i, cb);
    A.col[colour][gid] = (uint8_t)v;
  }
}

// Rectangle-triangle: colours 0,1,2 = R,B,G by ((i-j)%3+3)%3
// (RectTriangleTiler.cpp:42-49).  A rotation of colour y rewrites the whole
// interior of colours y+1 (updateTiles1) and y+2 (updateTiles2).
template <int NC>
__global__ void k_recttri_step(TriArrays32 c0, TriArrays32 c1, StreamTable S, int N, int colour,
                               int prev, RTWeights W) {
  const int M = N / 3;
  const long P = (long)N * M;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P) return;
  const int i = (int)((unsigned)gid / (unsigned)M), j = (int)((unsigned)gid % (unsigned)M);
  Tiny t = tiny_load(S, gid);
  const uint32_t rd = tiny_bits(t);
  tiny_store(S, gid, t);
  const bool interior = i >= 1 && i <= N - 2 && j >= 1 && j <= M - 2;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const TriArrays32& A = c == 0 ? c0 : c1;
    int32_t* X = A.col[colour];
    uint32_t v = (uint32_t)X[gid];
    if (interior && prev >= 0 && prev != colour)
      v = colour == (prev + 1) % 3 ? recttri_rebuild1(A.col[prev] + gid, M, i, prev)
                                   : recttri_rebuild2(A.col[prev] + gid, M, i, prev);
    X[gid] = (int32_t)recttri_move(v, rd, W);
  }
}

template <int NC>
__global__ void k_recttri_sync(TriArrays32 c0, TriArrays32 c1, int N, int last) {
  const int M = N / 3;
  const long P = (long)N * M;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P) return;
  const int i = (int)((unsigned)gid / (unsigned)M), j = (int)((unsigned)gid % (unsigned)M);
  if (!(i >= 1 && i <= N - 2 && j >= 1 && j <= M - 2)) return;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const TriArrays32& A = c == 0 ? c0 : c1;
    A.col[(last + 1) % 3][gid] = (int32_t)recttri_rebuild1(A.col[last] + gid, M, i, last);
    A.col[(last + 2) % 3][gid] = (int32_t)recttri_rebuild2(A.col[last] + gid, M, i, last);
  }
}

}  // namespace rt

// ------------------------------------------------- allocation-free walk
// Device memory of rt_domino_walk_noalloc: one 16 MiB array per GPU, laid
// out per call as TinyMT states (16 B per stream) | both colour planes
// (N*N bytes) | the int32 tiling staged for the on-device split and merge
// (4*N*N bytes), each part 256-byte aligned.  Host code never uses the
// array's name: rt_create looks up its address on the context's device
// (hipGetSymbolAddress) and keeps it in the context.
constexpr size_t kNoAllocBytes = (size_t)16 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace rt {

static size_t noalloc_round(size_t b) { return (b + 255) & ~(size_t)255; }

// Offsets of the three parts for side N; returns the bytes used.
static size_t noalloc_layout(int N, size_t* planes_at, size_t* tiling_at) {
  const size_t P = (size_t)(N - 2) * (size_t)(N / 2 - 2);
  *planes_at = noalloc_round(16 * P);
  *tiling_at = *planes_at + noalloc_round((size_t)N * N);
  return *tiling_at + noalloc_round(4 * (size_t)N * N);
}

// Split / merge of DominoTiler.cpp:24-33, 109-114 on the device: cell
// (i, j) of the int32 tiling <-> byte (i, j/2) of colour (i+j) % 2.
__global__ void k_domino_split_dev(const int32_t* t, int N, uint8_t* black, uint8_t* white) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)N * N) return;
  const int i = (int)(id / N), j = (int)(id % N);
  ((i + j) % 2 == 0 ? black : white)[(long)i * (N / 2) + j / 2] = (uint8_t)t[id];
}

__global__ void k_domino_merge_dev(int32_t* t, int N, const uint8_t* black, const uint8_t* white) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)N * N) return;
  const int i = (int)(id / N), j = (int)(id % N);
  t[id] = (int32_t)(int8_t)((i + j) % 2 == 0 ? black : white)[(long)i * (N / 2) + j / 2];
}

}  // namespace rt

// ------------------------------------------------------------------ context
#define HIPCHK HIPCHK_RET

static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int rt_ctx::reserve(size_t bytes) {
  if (bytes <= arena_bytes) return RT_OK;
  if (arena) hipFree(arena);
  arena = nullptr;
  arena_bytes = 0;
  size_t want = bytes + bytes / 4 + (1u << 20);
  hipError_t e = hipMalloc(&arena, want);
  if (e != hipSuccess) {
    err = std::string("hipMalloc arena: ") + hipGetErrorString(e);
    return RT_ENOMEM;
  }
  arena_bytes = want;
  return RT_OK;
}

int rt_ctx::reserve_aux(size_t bytes) {
  if (bytes <= aux_bytes) return RT_OK;
  if (aux) hipFree(aux);
  aux = nullptr;
  aux_bytes = 0;
  size_t want = bytes + bytes / 4 + (1u << 20);
  hipError_t e = hipMalloc(&aux, want);
  if (e != hipSuccess) {
    err = std::string("hipMalloc aux: ") + hipGetErrorString(e);
    return RT_ENOMEM;
  }
  aux_bytes = want;
  return RT_OK;
}

int rt_ctx::check_params(size_t need) {
  if (!d_param || nparam < need) {
    err = "TinyMT parameter table has " + std::to_string(nparam) + " sets, the walk needs " +
          std::to_string(need) + " (LoadTinyMT size)";
    return RT_EPARAMS;
  }
  return RT_OK;
}

extern "C" int rt_create(rt_ctx** out, int device_ordinal) {
  if (!out) return RT_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RT_EDEVICE;
  if (device_ordinal < 0 || device_ordinal >= ndev) return RT_EINVAL;
  rt_ctx* c = new rt_ctx();
  c->device = device_ordinal;
  if (hipSetDevice(device_ordinal) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->evk0) != hipSuccess || hipEventCreate(&c->evk1) != hipSuccess ||
      [&] {
        for (hipEvent_t& e : c->evt)
          if (hipEventCreate(&e) != hipSuccess) return true;
        return false;
      }() ||
      hipMalloc(&c->d_flag, 64 * sizeof(uint32_t)) != hipSuccess ||
      hipHostMalloc(&c->h_flag, 64 * sizeof(uint32_t)) != hipSuccess) {
    delete c;
    return RT_EDEVICE;
  }
  // this device's copy of the allocation-free walk's scratch array
  if (hipGetSymbolAddress(&c->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess ||
      !c->noalloc_scratch) {
    delete c;
    return RT_EDEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess) {
    c->devname = prop.gcnArchName;
    c->num_cu = prop.multiProcessorCount;
    c->lds_cu = prop.maxSharedMemoryPerMultiProcessor;
    c->lds_block = prop.sharedMemPerBlock;
  }
  *out = c;
  return RT_OK;
}

int rt_ctx::ensure_streams(int n) {
  for (int k = 0; k < n && k < 4; k++) {
    if (k && !sx[k] && hipStreamCreate(&sx[k]) != hipSuccess) {
      err = "extra stream";
      return RT_EDEVICE;
    }
    if (!evj[k] && hipEventCreateWithFlags(&evj[k], hipEventDisableTiming) != hipSuccess) {
      err = "stream join event";
      return RT_EDEVICE;
    }
  }
  return RT_OK;
}

rt_ctx::~rt_ctx() {
  for (hipStream_t s : sx)
    if (s) hipStreamSynchronize(s);
  if (stream) hipStreamSynchronize(stream);
  if (arena) hipFree(arena);
  if (aux) hipFree(aux);
  if (d_param) hipFree(d_param);
  if (d_flag) hipFree(d_flag);
  if (h_flag) hipHostFree(h_flag);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  if (evk0) hipEventDestroy(evk0);
  if (evk1) hipEventDestroy(evk1);
  for (hipEvent_t e : evt)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : evj)
    if (e) hipEventDestroy(e);
  for (hipStream_t s : sx)
    if (s) hipStreamDestroy(s);
  if (stream) hipStreamDestroy(stream);
}

extern "C" void rt_destroy(rt_ctx* ctx) { delete ctx; }

extern "C" const char* rt_last_error(const rt_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" const char* rt_device_name(const rt_ctx* ctx) {
  return ctx ? ctx->devname.c_str() : "";
}

extern "C" int rt_last_timing(const rt_ctx* ctx, rt_timing* out) {
  if (!ctx || !out) return RT_EINVAL;
  *out = ctx->last;
  return RT_OK;
}

extern "C" int rt_set_engine(rt_ctx* ctx, int engine) {
  if (!ctx || engine < RT_ENGINE_AUTO || engine > RT_ENGINE_FUSED_SWAR) return RT_EINVAL;
  ctx->engine = engine;
  return RT_OK;
}

extern "C" int rt_set_tile_shape(rt_ctx* ctx, int shape) {
  if (!ctx || shape < 0 || shape > 7) return RT_EINVAL;
  ctx->tile_shape = shape;
  return RT_OK;
}

extern "C" int rt_device_limits(const rt_ctx* ctx, int* num_cu, size_t* lds_per_cu,
                                size_t* lds_per_workgroup) {
  if (!ctx) return RT_EINVAL;
  if (num_cu) *num_cu = ctx->num_cu;
  if (lds_per_cu) *lds_per_cu = ctx->lds_cu;
  if (lds_per_workgroup) *lds_per_workgroup = ctx->lds_block;
  return RT_OK;
}

extern "C" int rt_set_halo(rt_ctx* ctx, int halo) {
  if (!ctx || halo < 0 || halo > 12) return RT_EINVAL;
  const bool set = halo != 0;
  if (!set) halo = 6;
  if (halo != ctx->halo || set != ctx->halo_set) ctx->tuned.clear();
  ctx->halo = halo;
  ctx->halo_set = set;
  return RT_OK;
}

extern "C" int rt_set_streams(rt_ctx* ctx, int n) {
  if (!ctx || n < 1 || n > 4) return RT_EINVAL;
  ctx->nstreams = n;
  return RT_OK;
}

extern "C" int rt_set_round_callback(rt_ctx* ctx, rt_round_fn fn, void* user) {
  if (!ctx) return RT_EINVAL;
  ctx->round_fn = fn;
  ctx->round_user = user;
  return RT_OK;
}

extern "C" int rt_set_xcd_order(rt_ctx* ctx, int on) {
  if (!ctx || on < 0 || on > 2) return RT_EINVAL;
  ctx->xcd_order = on;
  return RT_OK;
}

extern "C" int rt_set_autotune(rt_ctx* ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return RT_EINVAL;
  if (mode == 2) {
    ctx->tuned.clear();
    mode = 1;
  }
  ctx->autotune = mode;
  return RT_OK;
}

extern "C" int rt_set_tinymt_params(rt_ctx* ctx, const uint32_t* p3, size_t count) {
  if (!ctx || (!p3 && count)) return RT_EINVAL;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (ctx->d_param) hipFree(ctx->d_param);
  ctx->d_param = nullptr;
  ctx->nparam = 0;
  if (count == 0) return RT_OK;
  std::vector<uint32_t> soa(3 * count);
  for (size_t k = 0; k < count; k++) {
    soa[k] = p3[3 * k];
    soa[count + k] = p3[3 * k + 1];
    soa[2 * count + k] = p3[3 * k + 2];
  }
  hipError_t e = hipMalloc(&ctx->d_param, soa.size() * sizeof(uint32_t));
  if (e != hipSuccess) {
    ctx->err = "hipMalloc params failed";
    return RT_ENOMEM;
  }
  HIPCHK(ctx, hipMemcpy(ctx->d_param, soa.data(), soa.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice));
  ctx->nparam = count;
  return RT_OK;
}

extern "C" int rt_load_tinymt_file(rt_ctx* ctx, const char* path, size_t count) {
  if (!ctx || !path) return RT_EINVAL;
  std::vector<uint32_t> p3(3 * count);
  int rc = parse_tinymt_file(path, count, p3.data(), &ctx->err);
  if (rc) return rc;
  return rt_set_tinymt_params(ctx, p3.data(), count);
}

// --------------------------------------------------------- walk machinery
namespace {

struct Segments {
  std::vector<long> steps;
  std::vector<long> seeds;
  bool single = false;  // DominoTiler single-seed Walk: base = first mt19937 output
};

StreamTable stream_table(rt_ctx* ctx, uint32_t* state, long nstreams) {
  StreamTable S;
  S.state = state;
  S.param = ctx->d_param;
  S.stride = nstreams;
  S.pstride = (long)ctx->nparam;
  return S;
}

int launch_init(rt_ctx* ctx, const StreamTable& S, long n, uint32_t seed) {
  hipLaunchKernelGGL(k_tiny_init, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, S, n,
                     seed);
  HIPCHK(ctx, hipGetLastError());
  return RT_OK;
}

// Host side of a walk over NC chains that share the TinyMT streams.
struct DominoRun {
  rt_ctx* ctx;
  int N;
  DominoArrays a[2];
  int nc;
  int rule;
  DominoWeights w;
};

int domino_device_walk(DominoRun& R, const Segments& seg, uint32_t* state) {
  rt_ctx* ctx = R.ctx;
  const int N = R.N;
  const long P = (long)(N - 2) * (N / 2 - 2);
  StreamTable S = stream_table(ctx, state, P);
  int prev = -1;
  long total = 0;
  std::vector<uint8_t> colours;
  for (size_t k = 0; k < seg.steps.size(); k++) {
    uint32_t base = 0;
    domino_schedule((uint32_t)seg.seeds[k], seg.steps[k], seg.single, &base, colours);
    const uint32_t tseed = seg.single ? base : (uint32_t)seg.seeds[k];
    int rc = launch_init(ctx, S, P, tseed);
    if (rc) return rc;
    for (long s = 0; s < seg.steps[k]; s++) {
      const int c = colours[s];
      const int pull = prev >= 0 && prev != c;
      if (R.nc == 1)
        hipLaunchKernelGGL(k_domino_step<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                           R.a[0], R.a[0], S, N, c, pull, R.rule, R.w);
      else
        hipLaunchKernelGGL(k_domino_step<2>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                           R.a[0], R.a[1], S, N, c, pull, R.rule, R.w);
      prev = c;
    }
    total += seg.steps[k];
    ctx->last.launches += seg.steps[k] + 1;
  }
  if (prev >= 0) {
    if (R.nc == 1)
      hipLaunchKernelGGL(k_domino_sync<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                         R.a[0], R.a[0], N, prev);
    else
      hipLaunchKernelGGL(k_domino_sync<2>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                         R.a[0], R.a[1], N, prev);
    ctx->last.launches += 1;
  }
  HIPCHK(ctx, hipGetLastError());
  ctx->last.steps += total;
  ctx->last.attempts += total * P;
  return RT_OK;
}

// Split / merge of the reference (DominoTiler.cpp:24-33, 109-114): the int
// tiling is narrowed to char per colour and widened back with sign.
void domino_split(const int32_t* t, int N, uint8_t* host) {
  const int H = N / 2;
  uint8_t* B = host;
  uint8_t* W = host + (size_t)N * H;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      ((i + j) % 2 == 0 ? B : W)[i * H + j / 2] = (uint8_t)t[(size_t)i * N + j];
}

void domino_merge(int32_t* t, int N, const uint8_t* host) {
  const int H = N / 2;
  const uint8_t* B = host;
  const uint8_t* W = host + (size_t)N * H;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      t[(size_t)i * N + j] = (int32_t)(int8_t)((i + j) % 2 == 0 ? B : W)[i * H + j / 2];
}

int domino_check(rt_ctx* ctx, const int32_t* t, int N, const rt_domino_rule* rule) {
  if (!t || N < 6 || N % 2) {
    ctx->err = "domino tiling must be N*N with N even and N >= 6";
    return RT_EINVAL;
  }
  if (rule && (rule->kind < 0 || rule->kind > 3)) {
    ctx->err = "unknown domino rule";
    return RT_EINVAL;
  }
  return ctx->check_params((size_t)N * N / 2);
}

DominoWeights domino_weights(const rt_domino_rule* rule) {
  DominoWeights w;
  const double a = rule ? rule->a : 1.0, b = rule ? rule->b : 1.0;
  w.wa2 = (float)(a * a);
  w.wb2 = (float)(b * b);
  return w;
}

void begin_timing(rt_ctx* ctx) {
  ctx->last = rt_timing{};
  hipEventRecord(ctx->ev0, ctx->stream);
}

int end_timing(rt_ctx* ctx) {
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last.device_ms = ms;
  return RT_OK;
}

int domino_walk_impl(rt_ctx* ctx, int32_t* tiling, int N, const Segments& seg,
                     const rt_domino_rule* rule) {
  int rc = domino_check(ctx, tiling, N, rule);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t half = (size_t)N * (N / 2);
  const long P = (long)(N - 2) * (N / 2 - 2);
  const size_t state_bytes = 4 * (size_t)P * sizeof(uint32_t);
  rc = ctx->reserve(state_bytes + 2 * half + 256);
  if (rc) return rc;
  uint32_t* state = (uint32_t*)ctx->arena;
  uint8_t* arr = (uint8_t*)ctx->arena + state_bytes;
  std::vector<uint8_t> host(2 * half);
  domino_split(tiling, N, host.data());
  HIPCHK(ctx, hipMemcpyAsync(arr, host.data(), 2 * half, hipMemcpyHostToDevice, ctx->stream));
  const uint8_t* pl[2] = {host.data(), host.data() + half};
  const std::vector<int> spans = fused_live_spans(pl, 2, N, N / 2);
  DominoRun R;
  R.ctx = ctx;
  R.N = N;
  R.a[0].col[0] = arr;
  R.a[0].col[1] = arr + half;
  R.a[1] = R.a[0];
  R.nc = 1;
  R.rule = rule ? rule->kind : RT_DOMINO_TOGGLE08;
  R.w = domino_weights(rule);
  begin_timing(ctx);
  rc = ctx->engine == 1 ? domino_device_walk(R, seg, state)
                        : domino_fused_walk(ctx, R.a, 1, N, R.rule, R.w, seg.steps, seg.seeds,
                                            seg.single, state, &spans);
  if (rc) return rc;
  rc = end_timing(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(host.data(), arr, 2 * half, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  domino_merge(tiling, N, host.data());
  return RT_OK;
}

}  // namespace

int rt::launch_domino_sync(rt_ctx* ctx, DominoArrays* a, int nc, int N, int last) {
  const long P = (long)(N - 2) * (N / 2 - 2);
  if (nc == 1)
    hipLaunchKernelGGL(k_domino_sync<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                       a[0], a[0], N, last);
  else
    hipLaunchKernelGGL(k_domino_sync<2>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                       a[0], a[1], N, last);
  ctx->last.launches += 1;
  HIPCHK(ctx, hipGetLastError());
  return RT_OK;
}

int rt::launch_recttri_sync(rt_ctx* ctx, TriArrays32* a, int nc, int N, int last) {
  const long P = (long)N * (N / 3);
  if (nc == 1)
    hipLaunchKernelGGL(k_recttri_sync<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                       a[0], a[0], N, last);
  else
    hipLaunchKernelGGL(k_recttri_sync<2>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                       a[0], a[1], N, last);
  ctx->last.launches += 1;
  HIPCHK(ctx, hipGetLastError());
  return RT_OK;
}

int rt::domino_fused_walk(rt_ctx* ctx, DominoArrays* a, int nc, int N, int rule, DominoWeights w,
                          const std::vector<long>& steps, const std::vector<long>& seeds,
                          bool single, uint32_t* state, const std::vector<int>* spans) {
  if (ctx->engine == 1) {
    DominoRun R;
    R.ctx = ctx;
    R.N = N;
    R.a[0] = a[0];
    R.a[1] = nc > 1 ? a[1] : a[0];
    R.nc = nc;
    R.rule = rule;
    R.w = w;
    Segments s;
    s.steps = steps;
    s.seeds = seeds;
    s.single = single;
    return domino_device_walk(R, s, state);
  }
  FusedJob job;
  job.model = M_DOMINO;
  job.nc = nc;
  job.mp.rule = rule;
  job.mp.dw = w;
  UnitDesc u{};
  for (int ch = 0; ch < nc; ch++)
    for (int k = 0; k < 2; k++) u.plane[ch][k] = a[ch].col[k];
  u.state = state;
  u.param = ctx->d_param;
  u.nstream = (long)(N - 2) * (N / 2 - 2);
  u.pstride = (long)ctx->nparam;
  u.rows = N;
  u.cols = N / 2;
  job.units.push_back(u);
  if (spans) job.spans.push_back(*spans);
  job.seg_steps = steps;
  job.seg_seeds.resize(1);
  job.ops.resize(1);
  job.hist.resize(1);
  job.hist[0].model = M_DOMINO;
  std::vector<uint8_t> colours;
  for (size_t k = 0; k < steps.size(); k++) {
    uint32_t base = 0;
    domino_schedule((uint32_t)seeds[k], steps[k], single, &base, colours);
    job.seg_seeds[0].push_back(single ? base : (uint32_t)seeds[k]);
    job.hist[0].append(colours, job.ops[0]);
  }
  int rc = ctx->reserve_aux(fused_scratch_bytes(ctx, job));
  if (rc) return rc;
  rc = fused_run(ctx, job, ctx->aux);
  if (rc) return rc;
  return fused_sync(ctx, job);
}

extern "C" int rt_domino_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps, long seed,
                              const rt_domino_rule* rule) {
  if (!ctx) return RT_EINVAL;
  if (steps < 0) {
    ctx->err = "negative step count";
    return RT_EINVAL;
  }
  Segments s;
  s.steps = {steps};
  s.seeds = {seed};
  s.single = true;
  return domino_walk_impl(ctx, tiling, N, s, rule);
}

extern "C" int rt_domino_walk_segments(rt_ctx* ctx, int32_t* tiling, int N, const long* steps,
                                       const long* seeds, int nseg,
                                       const rt_domino_rule* rule) {
  if (!ctx) return RT_EINVAL;
  if (nseg < 0 || (nseg > 0 && (!steps || !seeds))) {
    ctx->err = "bad segment list";
    return RT_EINVAL;
  }
  Segments s;
  for (int k = 0; k < nseg; k++) {
    if (steps[k] < 0) {
      ctx->err = "negative step count";
      return RT_EINVAL;
    }
    s.steps.push_back(steps[k]);
    s.seeds.push_back(seeds[k]);
  }
  return domino_walk_impl(ctx, tiling, N, s, rule);
}

extern "C" int rt_domino_walk_noalloc(rt_ctx* ctx, int32_t* tiling, int N, long steps,
                                      long seed, const rt_domino_rule* rule) {
  if (!ctx) return RT_EINVAL;
  if (steps < 0) {
    ctx->err = "negative step count";
    return RT_EINVAL;
  }
  int rc = domino_check(ctx, tiling, N, rule);
  if (rc) return rc;
  size_t planes_at, tiling_at;
  const size_t need = noalloc_layout(N, &planes_at, &tiling_at);
  if (need > kNoAllocBytes) {
    ctx->err = "rt_domino_walk_noalloc: side " + std::to_string(N) + " needs " +
               std::to_string(need) + " bytes of device scratch, the fixed array holds " +
               std::to_string(kNoAllocBytes);
    return RT_EINVAL;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  uint8_t* base = (uint8_t*)ctx->noalloc_scratch;
  uint32_t* state = (uint32_t*)base;
  DominoArrays a;
  a.col[0] = base + planes_at;
  a.col[1] = a.col[0] + (size_t)N * (N / 2);
  int32_t* staged = (int32_t*)(base + tiling_at);
  const long cells = (long)N * N;
  const long P = (long)(N - 2) * (N / 2 - 2);
  const int rk = rule ? rule->kind : RT_DOMINO_TOGGLE08;
  const DominoWeights w = domino_weights(rule);
  const StreamTable S = stream_table(ctx, state, P);
  begin_timing(ctx);
  HIPCHK(ctx, hipMemcpyAsync(staged, tiling, 4 * (size_t)cells, hipMemcpyHostToDevice,
                             ctx->stream));
  hipLaunchKernelGGL(k_domino_split_dev, dim3(blocks_for(cells, 256)), dim3(256), 0, ctx->stream,
                     staged, N, a.col[0], a.col[1]);
  // DominoTiler::Walk(t, steps, seed): mt19937(seed), first output = TinyMT
  // base seed, then one colour per step (out % 2 == 1: black, index 0).
  std::mt19937 g((uint32_t)seed);
  rc = launch_init(ctx, S, P, (uint32_t)g());
  if (rc) return rc;
  int prev = -1;
  for (long s = 0; s < steps; s++) {
    const int c = (g() % 2 == 1) ? 0 : 1;
    hipLaunchKernelGGL(k_domino_step<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream, a,
                       a, S, N, c, prev >= 0 && prev != c ? 1 : 0, rk, w);
    prev = c;
  }
  if (prev >= 0)
    hipLaunchKernelGGL(k_domino_sync<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream, a,
                       a, N, prev);
  hipLaunchKernelGGL(k_domino_merge_dev, dim3(blocks_for(cells, 256)), dim3(256), 0, ctx->stream,
                     staged, N, a.col[0], a.col[1]);
  HIPCHK(ctx, hipGetLastError());
  rc = end_timing(ctx);
  if (rc) return rc;
  ctx->last.launches = steps + (prev >= 0 ? 1 : 0) + 3;
  ctx->last.steps = steps;
  ctx->last.attempts = steps * P;
  HIPCHK(ctx, hipMemcpyAsync(tiling, staged, 4 * (size_t)cells, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return RT_OK;
}

// ----------------------------------------------------------------- lozenge
namespace {

void tri_split8(const int32_t* t, int N, uint8_t* host, bool lozenge_colouring) {
  const int M = N / 3;
  const size_t sz = (size_t)N * M;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int c = lozenge_colouring ? ((j - i) % 3 + 3) % 3 : ((i - j) % 3 + 3) % 3;
      host[c * sz + (size_t)i * M + j / 3] = (uint8_t)t[(size_t)i * N + j];
    }
}

void tri_merge8(int32_t* t, int N, const uint8_t* host) {
  const int M = N / 3;
  const size_t sz = (size_t)N * M;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const int c = ((j - i) % 3 + 3) % 3;
      t[(size_t)i * N + j] = (int32_t)(int8_t)host[c * sz + (size_t)i * M + j / 3];
    }
}

}  // namespace

int rt::lozenge_device_walk(rt_ctx* ctx, TriArrays8* a, int nc, int N,
                            const std::vector<long>& steps, const std::vector<long>& seeds,
                            uint32_t* state, const std::vector<int>* spans) {
  if (ctx->engine != 1) {
    FusedJob job;
    job.model = M_LOZENGE;
    job.nc = nc;
    UnitDesc u{};
    for (int ch = 0; ch < nc; ch++)
      for (int k = 0; k < 3; k++) u.plane[ch][k] = a[ch].col[k];
    u.state = state;
    u.param = ctx->d_param;
    u.nstream = (long)N * (N / 3);
    u.pstride = (long)ctx->nparam;
    u.rows = N;
    u.cols = N / 3;
    job.units.push_back(u);
    if (spans) job.spans.push_back(*spans);
    job.seg_steps = steps;
    job.seg_seeds.resize(1);
    job.ops.resize(1);
    job.hist.resize(1);
    job.hist[0].model = M_LOZENGE;
    std::vector<uint8_t> colours;
    for (size_t k = 0; k < steps.size(); k++) {
      tri_schedule(seeds[k], steps[k], colours);
      job.seg_seeds[0].push_back((uint32_t)seeds[k]);
      job.hist[0].append(colours, job.ops[0]);
    }
    int rc = ctx->reserve_aux(fused_scratch_bytes(ctx, job));
    if (rc) return rc;
    rc = fused_run(ctx, job, ctx->aux);
    if (rc) return rc;
    return fused_sync(ctx, job);
  }
  const long P = (long)N * (N / 3);
  StreamTable S = stream_table(ctx, state, P);
  long last[3] = {-1, -1, -1};
  long g = 0;
  std::vector<uint8_t> colours;
  for (size_t k = 0; k < steps.size(); k++) {
    tri_schedule(seeds[k], steps[k], colours);
    int rc = launch_init(ctx, S, P, (uint32_t)seeds[k]);
    if (rc) return rc;
    for (long s = 0; s < steps[k]; s++, g++) {
      const int c = colours[s];
      const int ca = (c + 2) % 3, cb = (c + 1) % 3;
      const int pa = last[ca] > last[c], pb = last[cb] > last[c];
      if (nc == 1)
        hipLaunchKernelGGL(k_lozenge_step<1>, dim3(blocks_for(P, 256)), dim3(256), 0,
                           ctx->stream, a[0], a[0], S, N, c, pa, pb);
      else
        hipLaunchKernelGGL(k_lozenge_step<2>, dim3(blocks_for(P, 256)), dim3(256), 0,
                           ctx->stream, a[0], a[1], S, N, c, pa, pb);
      last[c] = g;
    }
    ctx->last.launches += steps[k] + 1;
    ctx->last.steps += steps[k];
    ctx->last.attempts += steps[k] * P;
  }
  return lozenge_device_sync(ctx, a, nc, N, last);
}

int rt::lozenge_device_sync(rt_ctx* ctx, TriArrays8* a, int nc, int N, const long* last) {
  const long P = (long)N * (N / 3);
  // Each colour refreshes the bits facing colours rotated after it; the two
  // directions of a colour pair never both apply, so the order is free.
  for (int c = 0; c < 3; c++) {
    const int ca = (c + 2) % 3, cb = (c + 1) % 3;
    const int pa = last[ca] > last[c], pb = last[cb] > last[c];
    if (!pa && !pb) continue;
    if (nc == 1)
      hipLaunchKernelGGL(k_lozenge_sync<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                         a[0], a[0], N, c, pa, pb);
    else
      hipLaunchKernelGGL(k_lozenge_sync<2>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                         a[0], a[1], N, c, pa, pb);
    ctx->last.launches += 1;
  }
  HIPCHK(ctx, hipGetLastError());
  return RT_OK;
}

static int lozenge_walk_impl(rt_ctx* ctx, int32_t* tiling, int N, const std::vector<long>& steps,
                             const std::vector<long>& seeds) {
  if (!tiling || N < 9 || N % 3) {
    ctx->err = "lozenge tiling must be N*N with N divisible by 3 and N >= 9";
    return RT_EINVAL;
  }
  const size_t sz = (size_t)N * (N / 3);
  int rc = ctx->check_params(sz);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t state_bytes = 4 * sz * sizeof(uint32_t);
  rc = ctx->reserve(state_bytes + 3 * sz + 256);
  if (rc) return rc;
  uint32_t* state = (uint32_t*)ctx->arena;
  uint8_t* arr = (uint8_t*)ctx->arena + state_bytes;
  std::vector<uint8_t> host(3 * sz);
  tri_split8(tiling, N, host.data(), true);
  HIPCHK(ctx, hipMemcpyAsync(arr, host.data(), 3 * sz, hipMemcpyHostToDevice, ctx->stream));
  const uint8_t* pl[3] = {host.data(), host.data() + sz, host.data() + 2 * sz};
  const std::vector<int> spans = fused_live_spans(pl, 3, N, N / 3);
  TriArrays8 a;
  for (int c = 0; c < 3; c++) a.col[c] = arr + c * sz;
  begin_timing(ctx);
  rc = lozenge_device_walk(ctx, &a, 1, N, steps, seeds, state, &spans);
  if (rc) return rc;
  rc = end_timing(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(host.data(), arr, 3 * sz, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  tri_merge8(tiling, N, host.data());
  return RT_OK;
}

extern "C" int rt_lozenge_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps, long seed) {
  if (!ctx) return RT_EINVAL;
  if (steps < 0) {
    ctx->err = "negative step count";
    return RT_EINVAL;
  }
  return lozenge_walk_impl(ctx, tiling, N, {steps}, {seed});
}

extern "C" int rt_lozenge_walk_segments(rt_ctx* ctx, int32_t* tiling, int N, const long* steps,
                                        const long* seeds, int nseg) {
  if (!ctx) return RT_EINVAL;
  if (nseg < 0 || (nseg > 0 && (!steps || !seeds))) {
    ctx->err = "bad segment list";
    return RT_EINVAL;
  }
  std::vector<long> st(steps, steps + nseg), se(seeds, seeds + nseg);
  for (long v : st)
    if (v < 0) {
      ctx->err = "negative step count";
      return RT_EINVAL;
    }
  return lozenge_walk_impl(ctx, tiling, N, st, se);
}

// -------------------------------------------------------- rect-triangle
int rt::recttri_device_walk(rt_ctx* ctx, TriArrays32* a, int nc, int N,
                            const std::vector<long>& steps, const std::vector<long>& seeds,
                            const RTWeights& W, uint32_t* state, int* last_out) {
  if (ctx->engine != 1) {
    FusedJob job;
    job.model = M_RECTTRI;
    job.nc = nc;
    job.mp.rw = W;
    UnitDesc u{};
    for (int ch = 0; ch < nc; ch++)
      for (int k = 0; k < 3; k++) u.plane[ch][k] = a[ch].col[k];
    u.state = state;
    u.param = ctx->d_param;
    u.nstream = (long)N * (N / 3);
    u.pstride = (long)ctx->nparam;
    u.rows = N;
    u.cols = N / 3;
    job.units.push_back(u);
    job.seg_steps = steps;
    job.seg_seeds.resize(1);
    job.ops.resize(1);
    job.hist.resize(1);
    job.hist[0].model = M_RECTTRI;
    std::vector<uint8_t> colours;
    for (size_t k = 0; k < steps.size(); k++) {
      tri_schedule(seeds[k], steps[k], colours);
      job.seg_seeds[0].push_back((uint32_t)seeds[k]);
      job.hist[0].append(colours, job.ops[0]);
    }
    int rc = ctx->reserve_aux(fused_scratch_bytes(ctx, job));
    if (rc) return rc;
    rc = fused_run(ctx, job, ctx->aux);
    if (rc) return rc;
    if (last_out) *last_out = job.hist[0].prev;
    return fused_sync(ctx, job);
  }
  const long P = (long)N * (N / 3);
  StreamTable S = stream_table(ctx, state, P);
  int prev = -1;
  std::vector<uint8_t> colours;
  for (size_t k = 0; k < steps.size(); k++) {
    tri_schedule(seeds[k], steps[k], colours);
    int rc = launch_init(ctx, S, P, (uint32_t)seeds[k]);
    if (rc) return rc;
    for (long s = 0; s < steps[k]; s++) {
      const int c = colours[s];
      if (nc == 1)
        hipLaunchKernelGGL(k_recttri_step<1>, dim3(blocks_for(P, 256)), dim3(256), 0,
                           ctx->stream, a[0], a[0], S, N, c, prev, W);
      else
        hipLaunchKernelGGL(k_recttri_step<2>, dim3(blocks_for(P, 256)), dim3(256), 0,
                           ctx->stream, a[0], a[1], S, N, c, prev, W);
      prev = c;
    }
    ctx->last.launches += steps[k] + 1;
    ctx->last.steps += steps[k];
    ctx->last.attempts += steps[k] * P;
  }
  if (prev >= 0) {
    if (nc == 1)
      hipLaunchKernelGGL(k_recttri_sync<1>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                         a[0], a[0], N, prev);
    else
      hipLaunchKernelGGL(k_recttri_sync<2>, dim3(blocks_for(P, 256)), dim3(256), 0, ctx->stream,
                         a[0], a[1], N, prev);
    ctx->last.launches += 1;
  }
  if (last_out) *last_out = prev;
  HIPCHK(ctx, hipGetLastError());
  return RT_OK;
}

RTWeights rt::recttri_weights(const float* w16) {
  RTWeights W;
  for (int k = 0; k < 16; k++) W.w[k] = w16 ? w16[k] : (k == 0 ? 0.0f : 1.0f);
  return W;
}

extern "C" int rt_recttri_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps, long seed,
                               const float* weights16) {
  if (!ctx) return RT_EINVAL;
  if (!tiling || N < 9 || N % 3 || steps < 0) {
    ctx->err = "rectangle-triangle tiling must be N*N with N divisible by 3 and N >= 9";
    return RT_EINVAL;
  }
  const int M = N / 3;
  const size_t sz = (size_t)N * M;
  int rc = ctx->check_params(sz);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t state_bytes = 4 * sz * sizeof(uint32_t);
  rc = ctx->reserve(state_bytes + 3 * sz * sizeof(int32_t) + 256);
  if (rc) return rc;
  uint32_t* state = (uint32_t*)ctx->arena;
  int32_t* arr = (int32_t*)((uint8_t*)ctx->arena + state_bytes);
  std::vector<int32_t> host(3 * sz);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      host[(((i - j) % 3 + 3) % 3) * sz + (size_t)i * M + j / 3] = tiling[(size_t)i * N + j];
  HIPCHK(ctx, hipMemcpyAsync(arr, host.data(), 3 * sz * sizeof(int32_t), hipMemcpyHostToDevice,
                             ctx->stream));
  TriArrays32 a;
  for (int c = 0; c < 3; c++) a.col[c] = arr + c * sz;
  begin_timing(ctx);
  rc = recttri_device_walk(ctx, &a, 1, N, {steps}, {seed}, recttri_weights(weights16), state,
                           nullptr);
  if (rc) return rc;
  rc = end_timing(ctx);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(host.data(), arr, 3 * sz * sizeof(int32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      tiling[(size_t)i * N + j] = host[(((i - j) % 3 + 3) % 3) * sz + (size_t)i * M + j / 3];
  return RT_OK;
}

// ------------------------------------------------------------------- CFTP
// Examples/AztecDiamondCftpEx.cpp:51-77 on the device: both chains of a
// round live in one arena and share every TinyMT draw; coalescence is a
// wavefront-vote compare of the two chains' colour arrays.
namespace {

long next_segment(const std::vector<long>& steps, int sched) {
  if (sched == RT_SCHED_REFERENCE) return steps.back() * 2;
  long tot = 0;
  for (long s : steps) tot += s;
  return tot;
}

int differ(rt_ctx* ctx, const uint8_t* a, const uint8_t* b, size_t n, bool* out) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_flag, 1L, 0u);
  const unsigned nb = std::min<unsigned>(blocks_for((long)n, 256), 4u * (unsigned)ctx->num_cu);
  hipLaunchKernelGGL(k_differ, dim3(nb ? nb : 1), dim3(256), 0, ctx->stream, a, b, (long)n,
                     ctx->d_flag);
  HIPCHK(ctx, hipMemcpyAsync(ctx->h_flag, ctx->d_flag, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *out = ctx->h_flag[0] != 0;
  return RT_OK;
}

// Generic CFTP round loop; `round` runs one round from the initial arrays
// and reports whether the chains differ.
template <class RoundFn>
int cftp_loop(rt_ctx* ctx, uint32_t master_seed, int sched, long t0, int max_rounds,
              rt_cftp_stats* stats, RoundFn round) {
  if (t0 <= 0 || max_rounds <= 0 || (sched != RT_SCHED_REFERENCE && sched != RT_SCHED_DOUBLING)) {
    ctx->err = "bad CFTP schedule";
    return RT_EINVAL;
  }
  std::mt19937 m(master_seed);
  std::vector<long> steps{t0};
  std::vector<long> seeds{(long)m()};
  long run = 0;
  double ms = 0;
  rt_timing acc{};
  for (int r = 1; r <= max_rounds; r++) {
    bool differs = true;
    int rc = round(steps, seeds, &differs);
    if (rc) return rc;
    long tot = 0;
    for (long s : steps) tot += s;
    run += tot;
    ms += ctx->last.device_ms;
    acc.launches += ctx->last.launches;
    acc.steps += ctx->last.steps;
    acc.attempts += ctx->last.attempts;
    acc.shape = ctx->last.shape;
    acc.split = ctx->last.split;
    if (ctx->last.shape_source == 1 || acc.shape_source != 1)
      acc.shape_source = ctx->last.shape_source;
    if (ctx->round_fn) ctx->round_fn(ctx->round_user, r, tot, differs ? 0 : 1, 1);
    if (!differs) {
      if (stats) {
        stats->rounds = r;
        stats->total_steps = tot;
        stats->steps_run = run;
        stats->device_ms = ms;
      }
      acc.device_ms = ms;
      ctx->last = acc;
      return RT_OK;
    }
    if (r == max_rounds) break;
    long ns = next_segment(steps, sched);
    steps.insert(steps.begin(), ns);
    seeds.insert(seeds.begin(), (long)m());
  }
  if (stats) {
    stats->rounds = max_rounds;
    stats->total_steps = 0;
    stats->steps_run = run;
    stats->device_ms = ms;
  }
  acc.device_ms = ms;
  ctx->last = acc;
  ctx->err = "CFTP did not coalesce within max_rounds";
  return RT_ENOCOUPLE;
}

}  // namespace

extern "C" int rt_domino_cftp(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                              uint32_t master_seed, int sched, long t0, int max_rounds,
                              const rt_domino_rule* rule, int32_t* out, rt_cftp_stats* stats) {
  if (!ctx) return RT_EINVAL;
  int rc = domino_check(ctx, tmax, N, rule);
  if (rc) return rc;
  if (!domino_cftp_rule(rule)) {
    ctx->err = "CFTP needs a heat-bath rule (RT_DOMINO_HEATBATH or RT_DOMINO_HEATBATH2P)";
    return RT_EINVAL;
  }
  if (!tmin || !out) {
    ctx->err = "null tiling";
    return RT_EINVAL;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t half = (size_t)N * (N / 2);
  const long P = (long)(N - 2) * (N / 2 - 2);
  const size_t state_bytes = 4 * (size_t)P * sizeof(uint32_t);
  rc = ctx->reserve(state_bytes + 8 * half + 256);
  if (rc) return rc;
  uint32_t* state = (uint32_t*)ctx->arena;
  uint8_t* base = (uint8_t*)ctx->arena + state_bytes;
  uint8_t* init_top = base;
  uint8_t* init_bot = base + 2 * half;
  uint8_t* top = base + 4 * half;
  uint8_t* bot = base + 6 * half;
  std::vector<uint8_t> host(2 * half), host2(2 * half);
  domino_split(tmax, N, host.data());
  HIPCHK(ctx, hipMemcpy(init_top, host.data(), 2 * half, hipMemcpyHostToDevice));
  domino_split(tmin, N, host2.data());
  HIPCHK(ctx, hipMemcpy(init_bot, host2.data(), 2 * half, hipMemcpyHostToDevice));
  const uint8_t* pl[4] = {host.data(), host.data() + half, host2.data(), host2.data() + half};
  const std::vector<int> spans = fused_live_spans(pl, 4, N, N / 2);
  DominoArrays a[2];
  a[0].col[0] = top;
  a[0].col[1] = top + half;
  a[1].col[0] = bot;
  a[1].col[1] = bot + half;
  const int kind = rule ? rule->kind : RT_DOMINO_HEATBATH;
  const DominoWeights w = domino_weights(rule);
  rc = cftp_loop(ctx, master_seed, sched, t0, max_rounds, stats,
                 [&](const std::vector<long>& steps, const std::vector<long>& seeds,
                     bool* differs) -> int {
                   begin_timing(ctx);
                   HIPCHK(ctx, hipMemcpyAsync(top, init_top, 2 * half, hipMemcpyDeviceToDevice,
                                              ctx->stream));
                   HIPCHK(ctx, hipMemcpyAsync(bot, init_bot, 2 * half, hipMemcpyDeviceToDevice,
                                              ctx->stream));
                   int r = domino_fused_walk(ctx, a, 2, N, kind, w, steps, seeds, false, state,
                                             &spans);
                   if (r) return r;
                   r = differ(ctx, top, bot, 2 * half, differs);
                   if (r) return r;
                   return end_timing(ctx);
                 });
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(host.data(), top, 2 * half, hipMemcpyDeviceToHost));
  domino_merge(out, N, host.data());
  return RT_OK;
}

extern "C" int rt_lozenge_cftp(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                               uint32_t master_seed, int sched, long t0, int max_rounds,
                               int32_t* out, rt_cftp_stats* stats) {
  if (!ctx) return RT_EINVAL;
  if (!tmax || !tmin || !out || N < 9 || N % 3) {
    ctx->err = "lozenge tiling must be N*N with N divisible by 3 and N >= 9";
    return RT_EINVAL;
  }
  const size_t sz = (size_t)N * (N / 3);
  int rc = ctx->check_params(sz);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t state_bytes = 4 * sz * sizeof(uint32_t);
  rc = ctx->reserve(state_bytes + 12 * sz + 256);
  if (rc) return rc;
  uint32_t* state = (uint32_t*)ctx->arena;
  uint8_t* base = (uint8_t*)ctx->arena + state_bytes;
  uint8_t* init_top = base;
  uint8_t* init_bot = base + 3 * sz;
  uint8_t* top = base + 6 * sz;
  uint8_t* bot = base + 9 * sz;
  std::vector<uint8_t> host(3 * sz), host2(3 * sz);
  tri_split8(tmax, N, host.data(), true);
  HIPCHK(ctx, hipMemcpy(init_top, host.data(), 3 * sz, hipMemcpyHostToDevice));
  tri_split8(tmin, N, host2.data(), true);
  HIPCHK(ctx, hipMemcpy(init_bot, host2.data(), 3 * sz, hipMemcpyHostToDevice));
  std::vector<const uint8_t*> pl;
  for (int c = 0; c < 3; c++) {
    pl.push_back(host.data() + c * sz);
    pl.push_back(host2.data() + c * sz);
  }
  const std::vector<int> spans = fused_live_spans(pl.data(), 6, N, N / 3);
  TriArrays8 a[2];
  for (int c = 0; c < 3; c++) {
    a[0].col[c] = top + c * sz;
    a[1].col[c] = bot + c * sz;
  }
  rc = cftp_loop(ctx, master_seed, sched, t0, max_rounds, stats,
                 [&](const std::vector<long>& steps, const std::vector<long>& seeds,
                     bool* differs) -> int {
                   begin_timing(ctx);
                   HIPCHK(ctx, hipMemcpyAsync(top, init_top, 3 * sz, hipMemcpyDeviceToDevice,
                                              ctx->stream));
                   HIPCHK(ctx, hipMemcpyAsync(bot, init_bot, 3 * sz, hipMemcpyDeviceToDevice,
                                              ctx->stream));
                   int r = lozenge_device_walk(ctx, a, 2, N, steps, seeds, state, &spans);
                   if (r) return r;
                   r = differ(ctx, top, bot, 3 * sz, differs);
                   if (r) return r;
                   return end_timing(ctx);
                 });
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(host.data(), top, 3 * sz, hipMemcpyDeviceToHost));
  tri_merge8(out, N, host.data());
  return RT_OK;
}

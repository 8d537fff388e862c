// rt_kernels.hpp -- device-side types and the host drivers shared by the
// walk, CFTP, batching and serving translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "models_dev.hpp"
#include "rt_internal.hpp"
#include "tinymt_dev.hpp"

#define HIPCHK_RET(ctx, expr)                                              \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);        \
      return RT_EDEVICE;                                                    \
    }                                                                       \
  } while (0)

struct rt_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  // second stream (default priority, created on first use) and its fork /
  // join events, for batch launches split in two halves
  hipStream_t sx[4] = {};          // sx[1..3]; part 0 runs on `stream`
  hipEvent_t evj[4] = {};
  int nstreams = 2;                // streams a batch job's launches are split over (1..4)
  int ensure_streams(int n);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evk0 = nullptr, evk1 = nullptr;  // around sweep-kernel sequences
  std::string err;
  std::string devname;
  uint32_t* d_param = nullptr;  // SoA: mat1[n], mat2[n], tmat[n]
  size_t nparam = 0;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* aux = nullptr;  // schedules, unit tables, work lists
  size_t aux_bytes = 0;
  uint32_t* d_flag = nullptr;
  uint32_t* h_flag = nullptr;
  rt_timing last{};
  int engine = 0;
  int tile_shape = 0;  // sweep-block tile shape: 0 auto, else a TileShape
  int autotune = 1;    // time candidate tile shapes on a job's first windows
  int xcd_order = 1;   // each XCD takes a contiguous share of a launch's tiles
  int halo = 6;        // sweep-block halo depth (byte-cell models)
  bool halo_set = false;  // rt_set_halo was called (else lozenge pairs take 9)
  size_t lds_cu = 0;   // LDS bytes per CU
  size_t lds_block = 0;  // LDS bytes one workgroup may hold
  std::map<uint32_t, int> tuned;  // tuning key (model, chains, tile count) -> shape
  rt_round_fn round_fn = nullptr;  // CFTP round-end callback
  void* round_user = nullptr;
  hipEvent_t evt[32] = {};        // tuner trial boundaries
  void* noalloc_scratch = nullptr;  // this device's rt_domino_walk_noalloc array

  int reserve(size_t bytes);
  int reserve_aux(size_t bytes);
  int check_params(size_t need);
  ~rt_ctx();
};

namespace rt {

struct DominoArrays {
  uint8_t* col[2];  // 0 = black ((i+j) even), 1 = white; each N x N/2
};
struct TriArrays8 {
  uint8_t* col[3];  // lozenge R, B, G; each N x N/3
};
struct TriArrays32 {
  int32_t* col[3];  // rectangle-triangle R, B, G; each N x N/3
};

// spans: fused_live_spans of the start planes (tiles that stay zero are
// left out), or null.
int domino_fused_walk(rt_ctx* ctx, DominoArrays* a, int nc, int N, int rule, DominoWeights w,
                      const std::vector<long>& steps, const std::vector<long>& seeds, bool single,
                      uint32_t* state, const std::vector<int>* spans = nullptr);
int lozenge_device_walk(rt_ctx* ctx, TriArrays8* a, int nc, int N, const std::vector<long>& steps,
                        const std::vector<long>& seeds, uint32_t* state,
                        const std::vector<int>* spans = nullptr);
int lozenge_device_sync(rt_ctx* ctx, TriArrays8* a, int nc, int N, const long* last);
int recttri_device_walk(rt_ctx* ctx, TriArrays32* a, int nc, int N,
                        const std::vector<long>& steps, const std::vector<long>& seeds,
                        const RTWeights& W, uint32_t* state, int* last_out);
RTWeights recttri_weights(const float* w16);

// CFTP runs the heat-bath rules only (dominokernelCFTP.cl,
// dominokernelCFTPtwoperiodic.cl, the kernels the reference's CFTP driver
// uses): their top/bottom coupling is monotone, so coalescence of the two
// extreme chains is an exact sample.  The 0.8 toggle and the two-periodic
// Metropolis rule are not monotone couplings.  NULL means heat bath.
inline bool domino_cftp_rule(const rt_domino_rule* rule) {
  return !rule || rule->kind == RT_DOMINO_HEATBATH || rule->kind == RT_DOMINO_HEATBATH2P;
}

// Final lazy sync of the colour planes (step-path kernels).
int launch_domino_sync(rt_ctx* ctx, DominoArrays* a, int nc, int N, int last);
int launch_recttri_sync(rt_ctx* ctx, TriArrays32* a, int nc, int N, int last);

}  // namespace rt

// rt_batch.hip -- batched CFTP (many master seeds per launch sequence) and
// batched domino walks on uneven domains, on top of the fused engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "rt_fused.hpp"

namespace rt {

// Per-unit coalescence test: flags[u] |= 1 if chain 0 and chain 1 of unit u
// differ anywhere in their `bytes`-long plane blocks.
__global__ void k_differ_units(const uint8_t* const* tops, const uint8_t* const* bots, long bytes,
                               uint32_t* flags) {
  const int u = blockIdx.y;
  const uint8_t* a = tops[u];
  const uint8_t* b = bots[u];
  bool diff = false;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < bytes; k += stride)
    diff |= a[k] != b[k];
  if (__any(diff) && (threadIdx.x & (warpSize - 1)) == 0) atomicOr(&flags[u], 1u);
}

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_differ_pair(const uint8_t* a, const uint8_t* b, long n, uint32_t* flag) {
  bool diff = false;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    diff |= a[k] != b[k];
  if (__any(diff) && (threadIdx.x & (warpSize - 1)) == 0) atomicOr(flag, 1u);
}

// Sets *flag to 0, then to 1 if the two byte blocks differ anywhere.
int launch_differ_pair(rt_ctx* ctx, const uint8_t* a, const uint8_t* b, long bytes,
                       uint32_t* flag) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, flag, 1);
  const unsigned nb = (unsigned)std::max<long>(1, std::min<long>((bytes + 4095) / 4096, 512));
  hipLaunchKernelGGL(k_differ_pair, dim3(nb), dim3(256), 0, ctx->stream, a, b, bytes, flag);
  HIPCHK_RET(ctx, hipGetLastError());
  return RT_OK;
}

namespace {

// Host split / merge between the int32 N x N tiling and the per-colour
// byte planes (domino: DominoTiler.cpp:24-33, lozenge: LozengeTiler.cpp:38-48).
void split_planes(int model, const int32_t* t, int N, uint8_t* planes) {
  if (model == M_DOMINO) {
    const int H = N / 2;
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        planes[(size_t)((i + j) % 2) * N * H + (size_t)i * H + j / 2] = (uint8_t)t[(size_t)i * N + j];
  } else {
    const int M = N / 3;
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        planes[(size_t)(((j - i) % 3 + 3) % 3) * N * M + (size_t)i * M + j / 3] =
            (uint8_t)t[(size_t)i * N + j];
  }
}

void merge_planes(int model, int32_t* t, int N, const uint8_t* planes) {
  if (model == M_DOMINO) {
    const int H = N / 2;
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        t[(size_t)i * N + j] =
            (int8_t)planes[(size_t)((i + j) % 2) * N * H + (size_t)i * H + j / 2];
  } else {
    const int M = N / 3;
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        t[(size_t)i * N + j] =
            (int8_t)planes[(size_t)(((j - i) % 3 + 3) % 3) * N * M + (size_t)i * M + j / 3];
  }
}

struct Geometry {
  int ncol;
  long cols;
  long plane;     // cells per colour plane
  long nstream;   // TinyMT streams
};

Geometry geometry_of(int model, int N) {
  Geometry g;
  if (model == M_DOMINO) {
    g.ncol = 2;
    g.cols = N / 2;
    g.nstream = (long)(N - 2) * (N / 2 - 2);
  } else {
    g.ncol = 3;
    g.cols = N / 3;
    g.nstream = (long)N * (N / 3);
  }
  g.plane = (long)N * g.cols;
  return g;
}

struct Sample {
  std::mt19937 master;
  std::vector<long> seeds;                    // newest first, like the segment list
  std::vector<std::vector<uint8_t>> colours;  // per segment, oldest first
  bool done = false;
};

}  // namespace

// fixed_rounds > 0: every sample runs exactly that many rounds, coalesced
// or not, and out / bot_out receive its top and bottom chains after the
// last round (both planes of the pair, merged back to N x N tilings).
static int cftp_batch(rt_ctx* ctx, int model, const int32_t* tmax, const int32_t* tmin, int N,
                      const uint32_t* masters, int ns, int sched, long t0, int max_rounds,
                      const ModelParams& mp, int32_t* out, rt_cftp_stats* stats,
                      int fixed_rounds = 0, int32_t* bot_out = nullptr) {
  const bool fixed = fixed_rounds > 0;
  if (fixed) max_rounds = fixed_rounds;
  if (!tmax || !tmin || !out || !masters || ns <= 0 || (fixed && !bot_out)) {
    ctx->err = "bad batch arguments";
    return RT_EINVAL;
  }
  if (model == M_DOMINO ? (N < 6 || N % 2) : (N < 9 || N % 3)) {
    ctx->err = "bad tiling side for the model";
    return RT_EINVAL;
  }
  if (t0 <= 0 || max_rounds <= 0 || (sched != RT_SCHED_REFERENCE && sched != RT_SCHED_DOUBLING)) {
    ctx->err = "bad CFTP schedule";
    return RT_EINVAL;
  }
  const Geometry g = geometry_of(model, N);
  int rc = ctx->check_params(model == M_DOMINO ? (size_t)N * N / 2 : (size_t)g.plane);
  if (rc) return rc;
  HIPCHK_RET(ctx, hipSetDevice(ctx->device));
  const size_t pbytes = (size_t)g.ncol * g.plane;          // one chain's planes
  const size_t sbytes = 4 * (size_t)g.nstream * 4;         // one sample's TinyMT state
  const size_t per = sbytes + 2 * pbytes;
  const size_t head = 2 * pbytes + 64 * 1024 + (size_t)ns * 4 * sizeof(void*) + (size_t)ns * 4;
  rc = ctx->reserve(head + (size_t)ns * ((per + 255) & ~(size_t)255) + 4096);
  if (rc) return rc;
  uint8_t* base = (uint8_t*)ctx->arena;
  uint8_t* init_top = base;
  uint8_t* init_bot = base + pbytes;
  uint8_t* cur = base + 2 * pbytes;
  cur = (uint8_t*)(((uintptr_t)cur + 255) & ~(uintptr_t)255);
  uint32_t* d_flags = (uint32_t*)cur;
  cur += ((size_t)ns * 4 + 255) & ~(size_t)255;
  const uint8_t** d_tops = (const uint8_t**)cur;
  cur += ((size_t)ns * sizeof(void*) + 255) & ~(size_t)255;
  const uint8_t** d_bots = (const uint8_t**)cur;
  cur += ((size_t)ns * sizeof(void*) + 255) & ~(size_t)255;
  std::vector<uint32_t*> state(ns);
  std::vector<uint8_t*> top(ns), bot(ns);
  for (int k = 0; k < ns; k++) {
    state[k] = (uint32_t*)cur;
    top[k] = cur + sbytes;
    bot[k] = cur + sbytes + pbytes;
    cur += (per + 255) & ~(size_t)255;
  }
  std::vector<int> spans;  // shared by every sample (FusedJob::spans)
  {
    std::vector<uint8_t> h(pbytes, 0), h2(pbytes, 0);
    split_planes(model, tmax, N, h.data());
    HIPCHK_RET(ctx, hipMemcpy(init_top, h.data(), pbytes, hipMemcpyHostToDevice));
    split_planes(model, tmin, N, h2.data());
    HIPCHK_RET(ctx, hipMemcpy(init_bot, h2.data(), pbytes, hipMemcpyHostToDevice));
    std::vector<const uint8_t*> pl;
    for (int c = 0; c < g.ncol; c++) {
      pl.push_back(h.data() + (size_t)c * g.plane);
      pl.push_back(h2.data() + (size_t)c * g.plane);
    }
    spans = fused_live_spans(pl.data(), (int)pl.size(), N, (int)g.cols);
  }
  std::vector<Sample> S(ns);
  for (int k = 0; k < ns; k++) {
    S[k].master.seed(masters[k]);
    S[k].seeds.push_back((long)S[k].master());
  }
  std::vector<long> steps{t0};
  auto gen = [&](int k, long seed, long n) {
    std::vector<uint8_t> c;
    if (model == M_DOMINO) domino_schedule((uint32_t)seed, n, false, nullptr, c);
    else tri_schedule(seed, n, c);
    S[k].colours.push_back(std::move(c));
  };
  parallel_for(ns, [&](int k) { gen(k, S[k].seeds[0], t0); });
  // A sample's op stream over all its segments, newest first (colours are
  // stored oldest first).
  auto build_ops = [&](int k, std::vector<uint8_t>& ops, OpBuilder& h) {
    h = OpBuilder{};
    h.model = model;
    size_t tot = 0;
    for (const std::vector<uint8_t>& c : S[k].colours) tot += c.size();
    ops.clear();
    ops.reserve(tot);
    const int nseg = (int)S[k].seeds.size();
    for (int s = 0; s < nseg; s++) h.append(S[k].colours[nseg - 1 - s], ops);
  };
  // The next round's seeds, colours and op streams are made on a host
  // thread while the device runs the current round (every sample still
  // running draws its next seed then; one that coalesces in this round
  // never uses it, and its master stream serves nobody else).
  std::vector<std::vector<uint8_t>> next_ops(ns);
  std::vector<OpBuilder> next_hist(ns);
  std::vector<char> has_next(ns, 0);
  struct Joiner {
    std::thread t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  };

  rt_timing acc{};
  double ms_total = 0;
  long run = 0;
  int remaining = ns;
  for (int r = 1; r <= max_rounds && remaining > 0; r++) {
    long tot = 0;
    for (long s : steps) tot += s;
    std::vector<int> act;
    for (int k = 0; k < ns; k++)
      if (!S[k].done) act.push_back(k);
    FusedJob job;
    job.model = model;
    job.nc = 2;
    job.mp = mp;
    job.seg_steps = steps;
    job.units.resize(act.size());
    job.seg_seeds.resize(act.size());
    job.ops.resize(act.size());
    job.hist.resize(act.size());
    job.spans.assign(act.size(), spans);
    parallel_for((int)act.size(), [&](int u) {
      const int k = act[u];
      UnitDesc d{};
      for (int c = 0; c < g.ncol; c++) {
        d.plane[0][c] = top[k] + (size_t)c * g.plane;
        d.plane[1][c] = bot[k] + (size_t)c * g.plane;
      }
      d.state = state[k];
      d.param = ctx->d_param;
      d.nstream = g.nstream;
      d.pstride = (long)ctx->nparam;
      d.rows = N;
      d.cols = (int)g.cols;
      job.units[u] = d;
      for (long s : S[k].seeds) job.seg_seeds[u].push_back((uint32_t)s);
      if (has_next[k]) {
        job.ops[u] = std::move(next_ops[k]);
        job.hist[u] = next_hist[k];
        has_next[k] = 0;
      } else {
        build_ops(k, job.ops[u], job.hist[u]);
      }
    });
    rc = ctx->reserve_aux(fused_scratch_bytes(ctx, job));
    if (rc) return rc;
    ctx->last = rt_timing{};
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k : act) {
      HIPCHK_RET(ctx, hipMemcpyAsync(top[k], init_top, pbytes, hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK_RET(ctx, hipMemcpyAsync(bot[k], init_bot, pbytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    FusedPlan plan;
    rc = fused_upload(ctx, job, ctx->aux, &plan);
    if (!rc) rc = fused_tune(ctx, job, &plan);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->evk0, ctx->stream));
    const long ns_len = sched == RT_SCHED_REFERENCE ? steps.back() * 2 : tot;
    Joiner prep;
    if (r < max_rounds)
      prep.t = std::thread([&, ns_len] {
        for (int k : act) S[k].seeds.insert(S[k].seeds.begin(), (long)S[k].master());
        // a few threads: the launch loop shares the host
        parallel_for((int)act.size(), [&](int i) {
          const int k = act[i];
          gen(k, S[k].seeds[0], ns_len);
          build_ops(k, next_ops[k], next_hist[k]);
          has_next[k] = 1;
        }, 6);
      });
    rc = fused_launch(ctx, job, plan, true);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    rc = fused_sync(ctx, job);
    if (rc) return rc;
    std::vector<const uint8_t*> ht(act.size()), hb(act.size());
    for (size_t u = 0; u < act.size(); u++) {
      ht[u] = top[act[u]];
      hb[u] = bot[act[u]];
    }
    HIPCHK_RET(ctx, hipMemcpyAsync(d_tops, ht.data(), ht.size() * sizeof(void*),
                                   hipMemcpyHostToDevice, ctx->stream));
    HIPCHK_RET(ctx, hipMemcpyAsync(d_bots, hb.data(), hb.size() * sizeof(void*),
                                   hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((act.size() + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_flags, (int)act.size());
    const unsigned bx = (unsigned)std::min<long>((long)(pbytes + 1023) / 1024, 64);
    hipLaunchKernelGGL(k_differ_units, dim3(bx, (unsigned)act.size()), dim3(256), 0, ctx->stream,
                       d_tops, d_bots, (long)pbytes, d_flags);
    HIPCHK_RET(ctx, hipGetLastError());
    std::vector<uint32_t> flags(act.size());
    HIPCHK_RET(ctx, hipMemcpyAsync(flags.data(), d_flags, flags.size() * 4, hipMemcpyDeviceToHost,
                                   ctx->stream));
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK_RET(ctx, hipEventSynchronize(ctx->ev1));
    if (prep.t.joinable()) prep.t.join();
    float ms = 0;
    HIPCHK_RET(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ms_total += ms;
    float kms = 0;
    if (plan.nlaunch) HIPCHK_RET(ctx, hipEventElapsedTime(&kms, ctx->evk0, ctx->evk1));
    const long nlaunch = (long)(plan.nlaunch - std::min(plan.first, plan.nlaunch));
    const long nsites = plan.sites(plan.first, plan.nlaunch);
    acc.kernel_ms += kms;
    acc.kernel_launches += nlaunch;
    acc.kernel_sites += nsites;
    if ((job.shape >= 1 && job.shape <= 4) || job.shape == 7) {
      const int sl = shape_slot(job.shape, job.split);
      acc.shape_ms[sl] += kms;
      acc.shape_launches[sl] += nlaunch;
      acc.shape_sites[sl] += nsites;
    }
    fused_add_trials(plan, &acc);
    acc.shape = job.shape;
    acc.split = job.split ? 1 : 0;
    if (job.shape_source == 1 || acc.shape_source != 1) acc.shape_source = job.shape_source;
    acc.launches += ctx->last.launches;
    acc.steps += ctx->last.steps;
    acc.attempts += ctx->last.attempts;
    run += tot;
    std::vector<uint8_t> h(pbytes);
    for (size_t u = 0; u < act.size(); u++) {
      const int k = act[u];
      if (fixed) {
        if (r < fixed_rounds) continue;
        HIPCHK_RET(ctx, hipMemcpy(h.data(), bot[k], pbytes, hipMemcpyDeviceToHost));
        merge_planes(model, bot_out + (size_t)k * N * N, N, h.data());
      } else if (flags[u]) {
        continue;
      }
      HIPCHK_RET(ctx, hipMemcpy(h.data(), top[k], pbytes, hipMemcpyDeviceToHost));
      merge_planes(model, out + (size_t)k * N * N, N, h.data());
      S[k].done = true;
      remaining--;
      std::vector<uint8_t>().swap(next_ops[k]);  // its speculative next round
      has_next[k] = 0;
      if (stats) {
        stats[k].rounds = r;
        stats[k].total_steps = tot;
        stats[k].steps_run = run;
        stats[k].device_ms = ms_total;
      }
    }
    if (ctx->round_fn) ctx->round_fn(ctx->round_user, r, tot, ns - remaining, ns);
    if (remaining == 0 || r == max_rounds) break;
    steps.insert(steps.begin(), ns_len);
  }
  acc.device_ms = ms_total;
  ctx->last = acc;
  if (fixed) return RT_OK;
  if (remaining > 0) {
    for (int k = 0; k < ns; k++)
      if (!S[k].done && stats) {
        stats[k].rounds = max_rounds;
        stats[k].total_steps = 0;
        stats[k].steps_run = run;
        stats[k].device_ms = ms_total;
      }
    ctx->err = "CFTP did not coalesce within max_rounds for every sample";
    return RT_ENOCOUPLE;
  }
  return RT_OK;
}

}  // namespace rt

using namespace rt;

extern "C" int rt_lozenge_cftp_batch(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                                     const uint32_t* masters, int nsamples, int sched, long t0,
                                     int max_rounds, int32_t* out, rt_cftp_stats* stats) {
  if (!ctx) return RT_EINVAL;
  ModelParams mp{};
  return cftp_batch(ctx, M_LOZENGE, tmax, tmin, N, masters, nsamples, sched, t0, max_rounds, mp,
                    out, stats);
}

extern "C" int rt_lozenge_cftp_batch_rounds(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin,
                                            int N, const uint32_t* masters, int nsamples,
                                            int sched, long t0, int rounds, int32_t* top_out,
                                            int32_t* bot_out, rt_cftp_stats* stats) {
  if (!ctx) return RT_EINVAL;
  if (rounds <= 0) {
    ctx->err = "rounds must be positive";
    return RT_EINVAL;
  }
  ModelParams mp{};
  return cftp_batch(ctx, M_LOZENGE, tmax, tmin, N, masters, nsamples, sched, t0, rounds, mp,
                    top_out, stats, rounds, bot_out);
}

extern "C" int rt_domino_cftp_batch(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                                    const uint32_t* masters, int nsamples, int sched, long t0,
                                    int max_rounds, const rt_domino_rule* rule, int32_t* out,
                                    rt_cftp_stats* stats) {
  if (!ctx) return RT_EINVAL;
  if (!domino_cftp_rule(rule)) {
    ctx->err = "CFTP needs a heat-bath rule (RT_DOMINO_HEATBATH or RT_DOMINO_HEATBATH2P)";
    return RT_EINVAL;
  }
  ModelParams mp{};
  mp.rule = rule ? rule->kind : RT_DOMINO_HEATBATH;
  const double a = rule ? rule->a : 1.0, b = rule ? rule->b : 1.0;
  mp.dw.wa2 = (float)(a * a);
  mp.dw.wb2 = (float)(b * b);
  return cftp_batch(ctx, M_DOMINO, tmax, tmin, N, masters, nsamples, sched, t0, max_rounds, mp, out,
                    stats);
}

extern "C" int rt_domino_cftp_batch_rounds(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin,
                                           int N, const uint32_t* masters, int nsamples, int sched,
                                           long t0, int rounds, const rt_domino_rule* rule,
                                           int32_t* top_out, int32_t* bot_out,
                                           rt_cftp_stats* stats) {
  if (!ctx) return RT_EINVAL;
  if (rounds <= 0) {
    ctx->err = "rounds must be positive";
    return RT_EINVAL;
  }
  if (!domino_cftp_rule(rule)) {
    ctx->err = "CFTP needs a heat-bath rule (RT_DOMINO_HEATBATH or RT_DOMINO_HEATBATH2P)";
    return RT_EINVAL;
  }
  ModelParams mp{};
  mp.rule = rule ? rule->kind : RT_DOMINO_HEATBATH;
  const double a = rule ? rule->a : 1.0, b = rule ? rule->b : 1.0;
  mp.dw.wa2 = (float)(a * a);
  mp.dw.wb2 = (float)(b * b);
  return cftp_batch(ctx, M_DOMINO, tmax, tmin, N, masters, nsamples, sched, t0, rounds, mp,
                    top_out, stats, rounds, bot_out);
}

// Observables of a domino tiling read straight off its two device colour
// planes (N x N/2 bytes each, vertex states of DominoTiler.h:28-32): the
// number of vertices with each edge bit set (1 up, 2 down, 4 right, 8 left)
// and of rotatable faces (vertex state 3 or 12: two parallel dominoes).
// One block row of the grid per chain; 64-bit counters, one atomic per
// wave and counter.
__global__ void k_domino_counts(const uint8_t* const* planes, const long* cells, int nchains,
                                unsigned long long* counts) {
  for (int c = blockIdx.y; c < nchains; c += gridDim.y) {
    const uint8_t* p = planes[c];
    const long n = cells[c];  // both planes, contiguous
    uint32_t k[5] = {0, 0, 0, 0, 0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
      const uint32_t v = p[i];
      k[0] += v & 1u;
      k[1] += (v >> 1) & 1u;
      k[2] += (v >> 2) & 1u;
      k[3] += (v >> 3) & 1u;
      k[4] += (v == 3u || v == 12u) ? 1u : 0u;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) {
      uint32_t s = k[q];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if ((threadIdx.x & 63) == 0 && s)
        atomicAdd(&counts[(size_t)c * 5 + q], (unsigned long long)s);
    }
  }
}

// ---------------------------------------------------------------------
// Height function of a domino tiling on the device, equal to
// DominoTiler::TilingToHeightfunc(t, TilingToDomain(t)) (DominoTiler.cpp:
// 555-593, 595-631): heights 0 at the first boundary vertex in row-major
// order, spread over every edge no domino crosses whose two ends are
// vertices of the region and one of whose two faces is in the domain,
// changing by (east/west ? +1 : -1) * ((i+j) even ? +1 : -1) from (i, j)
// (DominoTiler.cpp:335-341, 578-586); vertices not reached keep INT_MAX.
// For a tiling, heights do not depend on the path (every face loop sums to
// zero), so the breadth-first search becomes: (1) per row, prefix sums
// along runs of traversable east edges; (2) per chain, offsets of those
// runs linked through traversable north-south edges, by sweeps over the
// rows (down, then up, repeated until a sweep adds nothing); (3) height =
// run offset + prefix.  The domain and vertex classes come from the tiling
// itself (TilingToDomain / DomainToVertices).
struct HfChain {
  const uint8_t* planes;  // black N x N/2, then white N x N/2
  int N;
  int first;              // row-major index of the first boundary vertex
  int32_t* rr;            // N x N: prefix along the run; heights at the end
  int32_t* run;           // N x N: column where the vertex's run starts, -1 outside
  int32_t* off;           // N x N: offset of the run starting at (i, j)
};

constexpr int32_t kHfUnknown = INT32_MIN;

__device__ __forceinline__ uint32_t hf_cell(const HfChain& c, int i, int j) {
  const int H = c.N / 2;
  return c.planes[(size_t)((i + j) & 1) * c.N * H + (size_t)i * H + j / 2];
}

// Face (fi, fj) of the (N-1) x (N-1) face grid is in the domain when a
// domino covers it (TilingToDomain, DominoTiler.cpp:595-609); faces off
// the grid are not.
__device__ __forceinline__ bool hf_face(const HfChain& c, int fi, int fj) {
  const int M = c.N - 1;
  if (fi < 0 || fj < 0 || fi >= M || fj >= M) return false;
  return (hf_cell(c, fi, fj) & (2u + 8u)) || (hf_cell(c, fi + 1, fj + 1) & (1u + 4u));
}

// DomainToVertices (DominoTiler.cpp:611-631): 2 boundary, 1 interior, 0
// outside; grid-edge vertices look at their faces clamped into the grid.
__device__ __forceinline__ int hf_status(const HfChain& c, int i, int j) {
  const int M = c.N - 1;
  if (i == 0 || j == 0 || i == M || j == M) {
    const int a = i - (i == M), b = i - 1 + (i == 0), x = j - (j == M), y = j - 1 + (j == 0);
    return (hf_face(c, a, x) || hf_face(c, b, x) || hf_face(c, a, y) || hf_face(c, b, y)) ? 2 : 0;
  }
  const bool f0 = hf_face(c, i - 1, j - 1), f1 = hf_face(c, i - 1, j), f2 = hf_face(c, i, j - 1),
             f3 = hf_face(c, i, j);
  if (f0 && f1 && f2 && f3) return 1;
  return (f0 || f1 || f2 || f3) ? 2 : 0;
}

__device__ __forceinline__ int hf_sign(int i, int j) { return ((i + j) & 1) ? -1 : 1; }

// (1) one thread per row: runs, prefixes, first boundary vertex.
__global__ void k_hf_rows(HfChain* chains, int nchains) {
  for (int k = blockIdx.y; k < nchains; k += gridDim.y) {
    HfChain& c = chains[k];
    const int N = c.N, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) continue;
    int start = 0, r = 0, first = INT32_MAX;
    bool cont = false;
    int st = hf_status(c, i, 0);
    for (int j = 0; j < N; j++) {
      const size_t v = (size_t)i * N + j;
      const int st_next = j + 1 < N ? hf_status(c, i, j + 1) : 0;
      c.off[v] = kHfUnknown;
      if (st == 0) {
        c.run[v] = -1;
        cont = false;
      } else {
        if (st == 2 && first == INT32_MAX) first = (int)v;
        if (!cont) {
          start = j;
          r = 0;
        }
        c.run[v] = start;
        c.rr[v] = r;
        // east edge: no domino across (bit 8 of (i, j)), a face above or below
        cont = st_next > 0 && !(hf_cell(c, i, j) & 8u) && (hf_face(c, i - 1, j) || hf_face(c, i, j));
        r += hf_sign(i, j);
      }
      st = st_next;
    }
    if (first != INT32_MAX) atomicMin(&c.first, first);
  }
}

// (2) one workgroup per chain: run offsets through north-south edges.
__global__ void k_hf_link(HfChain* chains, int nchains, int max_sweeps) {
  for (int k = blockIdx.x; k < nchains; k += gridDim.x) {
    HfChain& c = chains[k];
    const int N = c.N;
    const int f = c.first;
    if (f == INT32_MAX) continue;  // no boundary vertex: nothing is reached
    if (threadIdx.x == 0) {
      const int fi = f / N;
      c.off[(size_t)fi * N + c.run[f]] = -c.rr[f];  // height 0 at the first boundary vertex
    }
    __syncthreads();
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
      int changed = 0;
      for (int pass = 0; pass < 2; pass++) {  // down, then up
        for (int q = 0; q + 1 < N; q++) {
          const int i = pass == 0 ? q : N - 2 - q;  // edge between rows i and i+1
          for (int j = threadIdx.x; j < N; j += blockDim.x) {
            const size_t a = (size_t)i * N + j, b = a + N;
            const int ra = c.run[a], rb = c.run[b];
            if (ra < 0 || rb < 0) continue;
            if (!(hf_face(c, i, j - 1) || hf_face(c, i, j))) continue;
            int32_t* oa = &c.off[(size_t)i * N + ra];
            int32_t* ob = &c.off[(size_t)(i + 1) * N + rb];
            if (pass == 0) {  // down from (i, j): bit 2, step -sign(i, j)
              if (*oa == kHfUnknown || *ob != kHfUnknown || (hf_cell(c, i, j) & 2u)) continue;
              *ob = *oa + c.rr[a] - hf_sign(i, j) - c.rr[b];
            } else {  // up from (i+1, j): bit 1, step -sign(i+1, j)
              if (*ob == kHfUnknown || *oa != kHfUnknown || (hf_cell(c, i + 1, j) & 1u)) continue;
              *oa = *ob + c.rr[b] - hf_sign(i + 1, j) - c.rr[a];
            }
            changed = 1;
          }
          __syncthreads();
        }
      }
      if (!__syncthreads_or(changed)) break;
    }
  }
}

// (3) heights.
__global__ void k_hf_final(HfChain* chains, int nchains) {
  for (int k = blockIdx.y; k < nchains; k += gridDim.y) {
    HfChain& c = chains[k];
    const long n = (long)c.N * c.N;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
      const int r = c.run[v];
      const int32_t o = r >= 0 ? c.off[(v / c.N) * c.N + r] : kHfUnknown;
      c.rr[v] = o == kHfUnknown ? INT32_MAX : o + c.rr[v];
    }
  }
}

__global__ void k_zero_u64(unsigned long long* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0;
}

static int walk_batch(rt_ctx* ctx, int32_t* const* tilings, const int* Ns, int nchains, long steps,
                      const long* seeds, const rt_domino_rule* rule, int use_graph, int download,
                      long* counts, int32_t* const* heights = nullptr);

extern "C" int rt_domino_walk_batch(rt_ctx* ctx, int32_t* const* tilings, const int* Ns,
                                    int nchains, long steps, const long* seeds,
                                    const rt_domino_rule* rule, int use_graph) {
  return walk_batch(ctx, tilings, Ns, nchains, steps, seeds, rule, use_graph, 1, nullptr);
}

extern "C" int rt_domino_walk_batch_observe(rt_ctx* ctx, int32_t* const* tilings, const int* Ns,
                                            int nchains, long steps, const long* seeds,
                                            const rt_domino_rule* rule, int use_graph,
                                            int download, long* counts) {
  if (!counts) {
    if (ctx) ctx->err = "counts output missing";
    return RT_EINVAL;
  }
  return walk_batch(ctx, tilings, Ns, nchains, steps, seeds, rule, use_graph, download, counts);
}

extern "C" int rt_domino_walk_batch_heights(rt_ctx* ctx, int32_t* const* tilings, const int* Ns,
                                            int nchains, long steps, const long* seeds,
                                            const rt_domino_rule* rule, int use_graph,
                                            int download, long* counts,
                                            int32_t* const* heights) {
  if (!heights) {
    if (ctx) ctx->err = "heights output missing";
    return RT_EINVAL;
  }
  for (int k = 0; k < nchains && heights; k++)
    if (!heights[k]) {
      if (ctx) ctx->err = "heights output of a chain missing";
      return RT_EINVAL;
    }
  return walk_batch(ctx, tilings, Ns, nchains, steps, seeds, rule, use_graph, download, counts,
                    heights);
}

static int walk_batch(rt_ctx* ctx, int32_t* const* tilings, const int* Ns, int nchains, long steps,
                      const long* seeds, const rt_domino_rule* rule, int use_graph, int download,
                      long* counts, int32_t* const* heights) {
  if (!ctx) return RT_EINVAL;
  if (!tilings || !Ns || !seeds || nchains <= 0 || steps < 0) {
    ctx->err = "bad batch arguments";
    return RT_EINVAL;
  }
  if (rule && (rule->kind < 0 || rule->kind > 3)) {
    ctx->err = "unknown domino rule";
    return RT_EINVAL;
  }
  size_t need = 0, bytes = 0;
  for (int k = 0; k < nchains; k++) {
    const int N = Ns[k];
    if (!tilings[k] || N < 6 || N % 2) {
      ctx->err = "chain " + std::to_string(k) + ": side must be even and >= 6";
      return RT_EINVAL;
    }
    need = std::max(need, (size_t)N * N / 2);
    const Geometry g = geometry_of(M_DOMINO, N);
    bytes += ((4 * (size_t)g.nstream * 4 + 2 * (size_t)g.plane) + 255) & ~(size_t)255;
  }
  int rc = ctx->check_params(need);
  if (rc) return rc;
  HIPCHK_RET(ctx, hipSetDevice(ctx->device));
  // tail of the arena: per-chain counters, plane pointers, cell counts
  const size_t tail = ((size_t)nchains * (5 * 8 + 8 + 8) + 255) & ~(size_t)255;
  rc = ctx->reserve(bytes + tail + 4096);
  if (rc) return rc;
  FusedJob job;
  job.model = M_DOMINO;
  job.nc = 1;
  job.mp.rule = rule ? rule->kind : RT_DOMINO_TOGGLE08;
  const double a = rule ? rule->a : 1.0, b = rule ? rule->b : 1.0;
  job.mp.dw.wa2 = (float)(a * a);
  job.mp.dw.wb2 = (float)(b * b);
  job.seg_steps = {steps};
  job.units.resize(nchains);
  job.seg_seeds.resize(nchains);
  job.ops.resize(nchains);
  job.hist.resize(nchains);
  std::vector<uint8_t*> planes(nchains);
  uint8_t* cur = (uint8_t*)ctx->arena;
  for (int k = 0; k < nchains; k++) {
    const Geometry g = geometry_of(M_DOMINO, Ns[k]);
    UnitDesc d{};
    d.state = (uint32_t*)cur;
    planes[k] = cur + 4 * (size_t)g.nstream * 4;
    d.plane[0][0] = planes[k];
    d.plane[0][1] = planes[k] + g.plane;
    d.param = ctx->d_param;
    d.nstream = g.nstream;
    d.pstride = (long)ctx->nparam;
    d.rows = Ns[k];
    d.cols = (int)g.cols;
    job.units[k] = d;
    cur += ((4 * (size_t)g.nstream * 4 + 2 * (size_t)g.plane) + 255) & ~(size_t)255;
  }
  parallel_for(nchains, [&](int k) {
    std::vector<uint8_t> colours;
    uint32_t base = 0;
    domino_schedule((uint32_t)seeds[k], steps, true, &base, colours);
    job.seg_seeds[k] = {base};
    job.hist[k].model = M_DOMINO;
    job.hist[k].append(colours, job.ops[k]);
  });
  job.spans.resize(nchains);
  for (int k = 0; k < nchains; k++) {
    const Geometry g = geometry_of(M_DOMINO, Ns[k]);
    std::vector<uint8_t> h(2 * (size_t)g.plane);
    split_planes(M_DOMINO, tilings[k], Ns[k], h.data());
    HIPCHK_RET(ctx, hipMemcpy(planes[k], h.data(), h.size(), hipMemcpyHostToDevice));
    const uint8_t* pl[2] = {h.data(), h.data() + g.plane};
    job.spans[k] = fused_live_spans(pl, 2, Ns[k], (int)g.cols);
  }
  rc = ctx->reserve_aux(fused_scratch_bytes(ctx, job));
  if (rc) return rc;
  FusedPlan plan;
  ctx->last = rt_timing{};
  rc = fused_upload(ctx, job, ctx->aux, &plan);
  if (!rc) rc = fused_tune(ctx, job, &plan);  // eager: its timing cannot be captured
  if (rc) return rc;
  if (use_graph) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK_RET(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    // the parts' streams join the capture through the fork event and leave
    // it through the join: the graph holds one chain of launches per part
    int lrc = fused_launch(ctx, job, plan, true);
    if (!lrc) lrc = fused_sync(ctx, job);
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (lrc) {
      if (graph) hipGraphDestroy(graph);
      return lrc;
    }
    HIPCHK_RET(ctx, e);
    HIPCHK_RET(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK_RET(ctx, hipGraphLaunch(exec, ctx->stream));
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK_RET(ctx, hipEventSynchronize(ctx->ev1));
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    ctx->last.kernel_launches = (long)(plan.nlaunch - std::min(plan.first, plan.nlaunch));
    ctx->last.kernel_sites = plan.sites(plan.first, plan.nlaunch);
  } else {
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = fused_launch(ctx, job, plan, true);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    rc = fused_sync(ctx, job);
    if (rc) return rc;
    HIPCHK_RET(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK_RET(ctx, hipEventSynchronize(ctx->ev1));
    float kms = 0;
    HIPCHK_RET(ctx, hipEventElapsedTime(&kms, ctx->ev0, ctx->evk1));
    const long nlaunch = (long)(plan.nlaunch - std::min(plan.first, plan.nlaunch));
    const long nsites = plan.sites(plan.first, plan.nlaunch);
    ctx->last.kernel_ms = kms;
    ctx->last.kernel_launches = nlaunch;
    ctx->last.kernel_sites = nsites;
    if ((job.shape >= 1 && job.shape <= 4) || job.shape == 7) {
      const int sl = shape_slot(job.shape, job.split);
      ctx->last.shape_ms[sl] = kms;
      ctx->last.shape_launches[sl] = nlaunch;
      ctx->last.shape_sites[sl] = nsites;
    }
  }
  fused_add_trials(plan, &ctx->last);
  float ms = 0;
  HIPCHK_RET(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last.device_ms = ms;
  for (double t : plan.trial_ms) ctx->last.device_ms += t;  // tuning windows ran before ev0
  if (counts) {
    unsigned long long* d_counts = (unsigned long long*)cur;
    const uint8_t** d_planes = (const uint8_t**)(cur + (size_t)nchains * 5 * 8);
    long* d_cells = (long*)(cur + (size_t)nchains * (5 * 8 + 8));
    std::vector<const uint8_t*> hp(nchains);
    std::vector<long> hc(nchains);
    for (int k = 0; k < nchains; k++) {
      hp[k] = planes[k];
      hc[k] = 2 * (long)geometry_of(M_DOMINO, Ns[k]).plane;
    }
    HIPCHK_RET(ctx, hipMemcpy(d_planes, hp.data(), nchains * sizeof(void*), hipMemcpyHostToDevice));
    HIPCHK_RET(ctx, hipMemcpy(d_cells, hc.data(), nchains * sizeof(long), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, ctx->stream, d_counts, 5L * nchains);
    hipLaunchKernelGGL(k_domino_counts, dim3(16, (unsigned)std::min(nchains, 4096)), dim3(256),
                       0, ctx->stream, d_planes, d_cells, nchains, d_counts);
    HIPCHK_RET(ctx, hipGetLastError());
    std::vector<unsigned long long> h((size_t)nchains * 5);
    HIPCHK_RET(ctx, hipMemcpyAsync(h.data(), d_counts, h.size() * 8, hipMemcpyDeviceToHost,
                                   ctx->stream));
    HIPCHK_RET(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t q = 0; q < h.size(); q++) counts[q] = (long)h[q];
  }
  if (heights) {
    // workspace in the aux block (the walk's launch tables are done with):
    // chain table, then rr | run | off per chain
    size_t bytes = ((size_t)nchains * sizeof(HfChain) + 255) & ~(size_t)255;
    std::vector<size_t> at(nchains);
    int maxN = 0;
    for (int k = 0; k < nchains; k++) {
      at[k] = bytes;
      bytes += 3 * ((4 * (size_t)Ns[k] * Ns[k] + 255) & ~(size_t)255);
      maxN = std::max(maxN, Ns[k]);
    }
    HIPCHK_RET(ctx, hipStreamSynchronize(ctx->stream));
    rc = ctx->reserve_aux(bytes);
    if (rc) return rc;
    uint8_t* base = (uint8_t*)ctx->aux;
    std::vector<HfChain> hc(nchains);
    for (int k = 0; k < nchains; k++) {
      const size_t nn = (4 * (size_t)Ns[k] * Ns[k] + 255) & ~(size_t)255;
      hc[k] = HfChain{planes[k], Ns[k], INT32_MAX, (int32_t*)(base + at[k]),
                      (int32_t*)(base + at[k] + nn), (int32_t*)(base + at[k] + 2 * nn)};
    }
    HfChain* d_chains = (HfChain*)base;
    HIPCHK_RET(ctx, hipMemcpy(d_chains, hc.data(), nchains * sizeof(HfChain),
                              hipMemcpyHostToDevice));
    const unsigned gy = (unsigned)std::min(nchains, 4096);
    hipLaunchKernelGGL(k_hf_rows, dim3((unsigned)((maxN + 63) / 64), gy), dim3(64), 0,
                       ctx->stream, d_chains, nchains);
    hipLaunchKernelGGL(k_hf_link, dim3((unsigned)std::min(nchains, 4096)), dim3(1024), 0,
                       ctx->stream, d_chains, nchains, 64);
    hipLaunchKernelGGL(k_hf_final, dim3(64, gy), dim3(256), 0, ctx->stream, d_chains, nchains);
    HIPCHK_RET(ctx, hipGetLastError());
    for (int k = 0; k < nchains; k++)
      HIPCHK_RET(ctx, hipMemcpy(heights[k], hc[k].rr, 4 * (size_t)Ns[k] * Ns[k],
                                hipMemcpyDeviceToHost));
  }
  if (download)
    for (int k = 0; k < nchains; k++) {
      const Geometry g = geometry_of(M_DOMINO, Ns[k]);
      std::vector<uint8_t> h(2 * (size_t)g.plane);
      HIPCHK_RET(ctx, hipMemcpy(h.data(), planes[k], h.size(), hipMemcpyDeviceToHost));
      merge_planes(M_DOMINO, tilings[k], Ns[k], h.data());
    }
  return RT_OK;
}

// rt_serve.hip -- continuous-batching CFTP (and walk) server.
//
// Requests (domino or lozenge CFTP samples of any size, and fixed-length
// walks of any model, rectangle-triangle included) arrive at any time.
// A device loop thread admits them by HBM footprint into slots of one
// pre-allocated arena, keeps every slot's chain state resident, and advances
// ALL active slots together: one fused launch per group per tick, each slot
// at its own position in its own schedule (per-unit windows).  A slot whose
// round ends is synced and tested for coalescence two ticks later (the
// test overlaps the next tick's sweep); it then either returns its sample
// or starts its next round while the other slots keep running.  A result
// depends only on its request (RNG streams are keyed by the request's seeds
// and site indices), so it is the same alone and in any mix.  A walk
// request is one round of one chain whose op stream is the model's plain
// Walk schedule; it ends after that round, with no coalescence test.
// Launches are grouped by (model, chains): CFTP pairs of domino and
// lozenge, one-chain walks of domino, lozenge and rectangle-triangle.
//
// Synchronisation: per-tick host tables are double-buffered in pinned
// memory; tick t waits for the event of tick t-2 before reusing its buffer.
// Any change of the active set (admission, completion, cancellation)
// drains the stream first, so no kernel can touch freed slot memory.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "rt_fused.hpp"

namespace rt {

int launch_differ_pair(rt_ctx* ctx, const uint8_t* a, const uint8_t* b, long bytes,
                       uint32_t* flag);

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Req {
  uint64_t id = 0;
  int model = M_LOZENGE;
  bool walk = false;     // a fixed-length walk of tmax, not a CFTP sample
  long walk_steps = 0;
  int N = 0;
  std::vector<int32_t> tmax, tmin;
  uint32_t master = 0;
  int sched = RT_SCHED_DOUBLING;
  long t0 = 1024;
  int max_rounds = 32;
  ModelParams mp{};
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = RT_EAGAIN;
  // set by rt_server_cancel (any thread), read by the loop thread
  std::atomic<bool> cancel{false};
  bool active = false;
  std::vector<int32_t> out;
  rt_cftp_stats st{};
  size_t bytes = 0;
};

struct Slot {
  Req* req = nullptr;
  int id = -1;  // flag index
  size_t off = 0, bytes = 0;
  uint8_t *init_top = nullptr, *init_bot = nullptr, *top = nullptr, *bot = nullptr;
  uint8_t *top_b = nullptr, *bot_b = nullptr;  // second buffer set of the sweep
  uint32_t* state = nullptr;
  uint32_t* state_b = nullptr;
  int par = 0;         // buffer set holding the current cells and states
  int ncol = 0;
  int cb = 1;          // bytes per cell (int32 for rectangle-triangle)
  int group = 0;       // launch group (kGroups)
  long plane = 0, cols = 0, nstream = 0;
  std::mt19937 master;
  std::vector<long> steps, seeds;             // newest first
  std::vector<std::vector<uint8_t>> colours;  // oldest first
  std::vector<uint8_t> ops;                   // this round's op stream
  std::vector<long> seg_start;                // first step of each segment
  OpBuilder hist;
  long cursor = 0, total = 0, run = 0;
  int round = 0;
  int phase = 0;       // 0 sweeping, 1 coalescence flag pending
  bool fresh = true;   // planes must be reset from the initial ones
  long flag_tick = -1;
  std::vector<int> spans;  // fused_live_spans of the start planes (tiles left out)
};

// First-fit allocator over [0, cap).
struct Arena {
  std::map<size_t, size_t> free_;  // offset -> size
  void init(size_t c) {
    free_.clear();
    free_[0] = c;
  }
  bool alloc(size_t n, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
      }
    return false;
  }
  void release(size_t off, size_t n) {
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

constexpr int kMaxSlots = 256;
// Launch groups: CFTP pairs (domino, lozenge), one-chain walks (domino,
// lozenge, rectangle-triangle).
constexpr int kGroups = 5;
int group_of(int model, bool walk) {
  if (!walk) return model == M_DOMINO ? 0 : 1;
  return model == M_DOMINO ? 2 : (model == M_LOZENGE ? 3 : 4);
}
int group_model(int g) { return g == 0 || g == 2 ? M_DOMINO : (g == 4 ? M_RECTTRI : M_LOZENGE); }
int group_nc(int g) { return g < 2 ? 2 : 1; }
// stream of a launch group within a tick: domino pairs, lozenge pairs, walks
int group_stream(int g) { return g < 2 ? g : 2; }
constexpr int SH_SERVE_COUNT = 1;  // tile shape whose tile count picks the launch shape (64 rows)
constexpr int kOpsPerUnit = 64;  // = the fused kernel's window limit
// Windows per tick: each tick launches kSub sweep kernels back to back from
// one upload of window tables, so the host work per tick (tables, event,
// flag reads) is spread over kSub windows of every live request.  Burst
// capacity of the default served stream (domino 64-256 and lozenge 64-160
// CFTP, rectangle-triangle walks), requests/s at kSub 4 / 8 / 16 / 32:
// 4.06 / 4.54 / 4.83 / 5.00, p99 latency 11.5 / 10.0 / 9.3 / 9.0 s.
constexpr int kSub = 32;

// A tick's tables travel to the device in one copy from one pinned block
// per tick buffer: the window records of the live units, [sub][group][unit]
// packed with no gaps, then (from the next 256-byte boundary) the op bytes
// of those windows, K bytes each back to back.  Only the bytes used are
// copied, so the upload shrinks with the number of live requests.
constexpr size_t kTickWinBytes = (size_t)kSub * kGroups * kMaxSlots * sizeof(WindowRec);
constexpr size_t kTickBytes = kTickWinBytes + (size_t)kSub * kGroups * kMaxSlots * kOpsPerUnit;
static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes of a slot: two TinyMT state sets and, per chain (two for
// CFTP, one for a walk), the start planes and two working sets.
size_t footprint(int model, int N, int chains, int* ncol, long* plane, long* nstream,
                 long* cols) {
  if (model == M_DOMINO) {
    *ncol = 2;
    *cols = N / 2;
    *nstream = (long)(N - 2) * (N / 2 - 2);
  } else {
    *ncol = 3;
    *cols = N / 3;
    *nstream = (long)N * (N / 3);
  }
  *plane = (long)N * *cols;
  const size_t cb = model == M_RECTTRI ? 4 : 1;
  const size_t pb = round_up((size_t)*ncol * (size_t)*plane * cb, 256);
  return 2 * round_up(16 * (size_t)*nstream, 256) + 3 * (size_t)chains * pb;
}

// Index of tiling cell (i, j) in the model's colour planes: domino black /
// white by (i+j) % 2 (DominoTiler.cpp:24-33), lozenge by (j-i) mod 3
// (LozengeTiler.cpp:38-48), rectangle-triangle by (i-j) mod 3
// (RectTriangleTiler.cpp:40-50).
size_t plane_index(int model, int i, int j, long plane, long cols) {
  if (model == M_DOMINO) return (size_t)((i + j) % 2) * plane + (size_t)i * cols + j / 2;
  const int c = model == M_LOZENGE ? ((j - i) % 3 + 3) % 3 : ((i - j) % 3 + 3) % 3;
  return (size_t)c * plane + (size_t)i * cols + j / 3;
}

}  // namespace

struct Server {
  rt_ctx* ctx = nullptr;
  size_t budget = 0;
  int max_active = 0;
  uint8_t* d_arena = nullptr;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv;
  std::thread loop;
  bool stop = false;
  std::string err;
  int loop_rc = RT_OK;
  uint64_t next_id = 1;
  // Waiting requests in arrival order.  A list: the admission loop keeps an
  // iterator across the unlocked admit() call, and a concurrent
  // rt_server_submit's push_back must not invalidate it.
  std::list<Req*> queue;
  std::map<uint64_t, Req*> reqs;
  // loop-thread state
  std::vector<Slot*> active;
  std::vector<int> free_ids;
  bool dirty = true;
  UnitDesc* d_units[kGroups] = {};  // [group]
  WorkItem* d_work[kGroups] = {};
  size_t work_cap[kGroups] = {};
  unsigned nwork[kGroups] = {};
  int shape[kGroups] = {};                // tile shape of each group's launches
  // flip rule shared by every live unit of a group, -1 when they differ
  // (the kernels read each unit's own; this picks the group's kernel)
  int group_rule[kGroups] = {};
  std::vector<Slot*> unit_slot[kGroups];  // unit index -> slot, per group
  uint8_t* d_tick = nullptr;              // [buf][kTickBytes]: window records, op bytes
  uint32_t* d_flags = nullptr;      // [buf][kMaxSlots]
  uint8_t* h_tick = nullptr;              // pinned twin of d_tick
  uint32_t* h_flags = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ev_used[2] = {false, false};
  long tick = 0;
  // statistics
  long completed = 0;
  std::vector<double> latencies;
  double t_first = 0, t_last = 0;
  long ticks = 0;
};

static int fail(Server* S, const std::string& m, int rc) {
  std::lock_guard<std::mutex> g(S->mu);
  S->err = m;
  S->loop_rc = rc;
  return rc;
}

static void gen_colours(Slot& s, long seed, long n) {
  std::vector<uint8_t> c;
  if (s.req->model == M_DOMINO) domino_schedule((uint32_t)seed, n, false, nullptr, c);
  else tri_schedule(seed, n, c);
  s.colours.push_back(std::move(c));
}

// Op stream of the slot's current segment list (segments newest first).
static void build_round(Slot& s) {
  s.ops.clear();
  s.seg_start.clear();
  s.hist = OpBuilder();
  s.hist.model = s.req->model;
  s.total = 0;
  for (long v : s.steps) s.total += v;
  s.ops.reserve((size_t)s.total);
  const int nseg = (int)s.seeds.size();
  for (int k = 0; k < nseg; k++) {
    s.seg_start.push_back((long)s.ops.size());
    s.hist.append(s.colours[nseg - 1 - k], s.ops);
  }
  s.cursor = 0;
  s.fresh = true;
  s.phase = 0;
}

static int admit(Server* S, Req* r) {
  int ncol;
  long plane, nstream, cols;
  const int chains = r->walk ? 1 : 2;
  const size_t bytes = footprint(r->model, r->N, chains, &ncol, &plane, &nstream, &cols);
  size_t off;
  if (S->free_ids.empty() || !S->arena.alloc(bytes, &off)) return RT_EAGAIN;
  Slot* s = new Slot();
  s->req = r;
  s->id = S->free_ids.back();
  S->free_ids.pop_back();
  s->off = off;
  s->bytes = bytes;
  s->ncol = ncol;
  s->cb = r->model == M_RECTTRI ? 4 : 1;
  s->group = group_of(r->model, r->walk);
  s->plane = plane;
  s->cols = cols;
  s->nstream = nstream;
  uint8_t* base = S->d_arena + off;
  s->state = (uint32_t*)base;
  base += round_up(16 * (size_t)nstream, 256);
  s->state_b = (uint32_t*)base;
  base += round_up(16 * (size_t)nstream, 256);
  const size_t pb = round_up((size_t)ncol * (size_t)plane * s->cb, 256);
  auto take = [&]() {
    uint8_t* q = base;
    base += pb;
    return q;
  };
  s->init_top = take();
  if (!r->walk) s->init_bot = take();
  s->top = take();
  if (!r->walk) s->bot = take();
  s->top_b = take();
  if (!r->walk) s->bot_b = take();
  std::vector<uint8_t> h((size_t)ncol * (size_t)plane * s->cb);
  const int N = r->N;
  for (int which = 0; which < chains; which++) {
    const std::vector<int32_t>& t = which ? r->tmin : r->tmax;
    std::fill(h.begin(), h.end(), 0);
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) {
        const size_t idx = plane_index(r->model, i, j, plane, cols);
        if (s->cb == 4) std::memcpy(&h[idx * 4], &t[(size_t)i * N + j], 4);
        else h[idx] = (uint8_t)t[(size_t)i * N + j];
      }
    if (s->cb == 1) {  // the union of both chains' extents
      std::vector<const uint8_t*> pl;
      for (int c = 0; c < ncol; c++) pl.push_back(h.data() + (size_t)c * plane);
      std::vector<int> sp = fused_live_spans(pl.data(), ncol, N, (int)cols);
      if (s->spans.empty()) {
        s->spans = std::move(sp);
      } else {
        for (int i = 0; i < N; i++) {
          s->spans[2 * i] = std::min(s->spans[2 * i], sp[2 * i]);
          s->spans[2 * i + 1] = std::max(s->spans[2 * i + 1], sp[2 * i + 1]);
        }
      }
    }
    if (hipMemcpy(which ? s->init_bot : s->init_top, h.data(), h.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
      S->arena.release(off, bytes);
      S->free_ids.push_back(s->id);
      delete s;
      return fail(S, "upload of start tilings failed", RT_EDEVICE);
    }
  }
  s->round = 1;
  if (r->walk) {
    // the model's plain Walk: domino seeds TinyMT with the first mt19937
    // output and draws its colours after it (DominoTiler.cpp:51-70);
    // lozenge and rectangle-triangle seed TinyMT with the seed itself
    // (LozengeTiler.cpp:86-91, RectTriangleTiler.cpp:69-71)
    std::vector<uint8_t> c;
    uint32_t tseed = r->master;
    if (r->model == M_DOMINO) domino_schedule(r->master, r->walk_steps, true, &tseed, c);
    else tri_schedule((long)r->master, r->walk_steps, c);
    s->steps = {r->walk_steps};
    s->seeds = {(long)tseed};
    s->colours.push_back(std::move(c));
  } else {
    s->master.seed(r->master);
    s->steps = {r->t0};
    s->seeds = {(long)s->master()};
    gen_colours(*s, s->seeds[0], r->t0);
  }
  build_round(*s);
  r->t_start = now_s();
  r->active = true;
  S->active.push_back(s);
  S->dirty = true;
  return RT_OK;
}

// Slot leaves the device: the stream is drained first.
static void retire(Server* S, Slot* s, int status) {
  hipStreamSynchronize(S->ctx->stream);
  Req* r = s->req;
  {
    std::lock_guard<std::mutex> g(S->mu);
    r->status = status;
    r->t_done = now_s();
    r->active = false;
    if (status == RT_OK) {
      S->completed++;
      S->latencies.push_back(r->t_done - r->t_submit);
      if (S->t_first == 0 || r->t_submit < S->t_first) S->t_first = r->t_submit;
      S->t_last = std::max(S->t_last, r->t_done);
    }
  }
  S->arena.release(s->off, s->bytes);
  S->free_ids.push_back(s->id);
  S->active.erase(std::find(S->active.begin(), S->active.end(), s));
  delete s;
  S->dirty = true;
  S->cv.notify_all();
}

static int upload_tables(Server* S) {
  hipStreamSynchronize(S->ctx->stream);
  for (int m = 0; m < kGroups; m++) {
    const int model = group_model(m), nc = group_nc(m);
    std::vector<UnitDesc> units;
    std::vector<WorkItem> work;
    S->unit_slot[m].clear();
    // tile shape from the live tile count (64-row tiles)
    long t64 = 0;
    int rule = -2;
    for (Slot* s : S->active) {
      if (s->group != m) continue;
      int tr, tc;
      fused_tiles(S->ctx, model, nc, -1, SH_SERVE_COUNT, s->req->N, (int)s->cols, &tr, &tc);
      t64 += (long)tr * tc;
      rule = rule == -2 || rule == s->req->mp.rule ? s->req->mp.rule : -1;
    }
    S->group_rule[m] = rule < 0 ? -1 : rule;
    S->shape[m] = fused_serve_shape(S->ctx, model, nc, S->group_rule[m], t64);
    for (Slot* s : S->active) {
      if (s->group != m) continue;
      UnitDesc d{};
      const size_t pc = (size_t)s->plane * s->cb;  // bytes of one colour plane
      for (int c = 0; c < s->ncol; c++) {
        d.plane[0][c] = s->top + c * pc;
        d.plane_b[0][c] = s->top_b + c * pc;
        if (nc == 2) {
          d.plane[1][c] = s->bot + c * pc;
          d.plane_b[1][c] = s->bot_b + c * pc;
        }
      }
      d.state = s->state;
      d.state_b = s->state_b;
      d.param = S->ctx->d_param;
      d.nstream = s->nstream;
      d.pstride = (long)S->ctx->nparam;
      d.rows = s->req->N;
      d.cols = (int)s->cols;
      d.rule = s->req->mp.rule;
      d.dw = s->req->mp.dw;
      const int u = (int)units.size();
      d.ops = S->d_tick;  // a window's s0 is its op bytes' offset in d_tick
      int tr, tc;
      fused_tiles(S->ctx, model, nc, S->group_rule[m], S->shape[m], d.rows, d.cols, &tr, &tc);
      d.tiles_c = tc;
      fused_unit_work(S->ctx, model, nc, S->group_rule[m], S->shape[m], d, s->spans, u, work);
      units.push_back(d);
      S->unit_slot[m].push_back(s);
    }
    if (work.size() > S->work_cap[m]) {
      if (S->d_work[m]) hipFree(S->d_work[m]);
      S->work_cap[m] = work.size() * 2;
      if (hipMalloc(&S->d_work[m], S->work_cap[m] * sizeof(WorkItem)) != hipSuccess)
        return fail(S, "hipMalloc work list", RT_ENOMEM);
    }
    if (!units.empty()) {
      if (hipMemcpy(S->d_units[m], units.data(), units.size() * sizeof(UnitDesc),
                    hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(S->d_work[m], work.data(), work.size() * sizeof(WorkItem),
                    hipMemcpyHostToDevice) != hipSuccess)
        return fail(S, "upload of unit tables failed", RT_EDEVICE);
    }
    S->nwork[m] = (unsigned)work.size();
  }
  S->dirty = false;
  return RT_OK;
}

// Host merge of a finished slot's top planes into the request's tiling.
static int collect(Server* S, Slot* s) {
  Req* r = s->req;
  std::vector<uint8_t> h((size_t)s->ncol * (size_t)s->plane * s->cb);
  if (hipMemcpy(h.data(), s->top, h.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(S, "download of a sample failed", RT_EDEVICE);
  const int N = r->N;
  r->out.assign((size_t)N * N, 0);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const size_t idx = plane_index(r->model, i, j, s->plane, s->cols);
      int32_t v;
      if (s->cb == 4) std::memcpy(&v, &h[idx * 4], 4);
      else v = (int8_t)h[idx];
      r->out[(size_t)i * N + j] = v;
    }
  r->st.rounds = s->round;
  r->st.total_steps = s->total;
  r->st.steps_run = s->run;
  return RT_OK;
}

// Flags of the tick that last used buffer b are ready: finish or re-round.
static int process_flags(Server* S, int b) {
  std::vector<Slot*> done;
  for (Slot* s : std::vector<Slot*>(S->active)) {
    if (s->phase != 1 || (s->flag_tick & 1) != b || s->flag_tick > S->tick - 2) continue;
    // a walk has no coalescence test: its one round is its result
    const uint32_t differs = s->req->walk ? 0 : S->h_flags[(size_t)b * kMaxSlots + s->id];
    s->run += s->total;
    if (!differs) {
      int rc = collect(S, s);
      if (rc) return rc;
      retire(S, s, RT_OK);
      continue;
    }
    if (s->round >= s->req->max_rounds) {
      s->req->st.rounds = s->round;
      s->req->st.steps_run = s->run;
      retire(S, s, RT_ENOCOUPLE);
      continue;
    }
    const long ns = s->req->sched == RT_SCHED_REFERENCE ? s->steps.back() * 2 : s->total;
    s->steps.insert(s->steps.begin(), ns);
    s->seeds.insert(s->seeds.begin(), (long)s->master());
    gen_colours(*s, s->seeds[0], ns);
    s->round++;
    build_round(*s);
  }
  return RT_OK;
}

static int run_tick(Server* S) {
  rt_ctx* ctx = S->ctx;
  const int b = (int)(S->tick & 1);
  if (S->ev_used[b]) {
    if (hipEventSynchronize(S->ev[b]) != hipSuccess) return fail(S, "event wait", RT_EDEVICE);
    int rc = process_flags(S, b);
    if (rc) return rc;
  }
  // cancellations
  for (Slot* s : std::vector<Slot*>(S->active))
    if (s->req->cancel) retire(S, s, RT_ECANCELLED);
  if (S->dirty) {
    int rc = upload_tables(S);
    if (rc) return rc;
  }
  if (S->active.empty()) return RT_OK;
  // resets of slots starting a round (both buffer sets: left-out tiles'
  // cells are never written)
  for (Slot* s : S->active)
    if (s->phase == 0 && s->fresh) {
      const size_t pb = (size_t)s->ncol * (size_t)s->plane * s->cb;
      const bool pair = !s->req->walk;
      if (hipMemcpyAsync(s->top, s->init_top, pb, hipMemcpyDeviceToDevice, ctx->stream) !=
              hipSuccess ||
          hipMemcpyAsync(s->top_b, s->init_top, pb, hipMemcpyDeviceToDevice, ctx->stream) !=
              hipSuccess ||
          (pair && hipMemcpyAsync(s->bot, s->init_bot, pb, hipMemcpyDeviceToDevice,
                                  ctx->stream) != hipSuccess) ||
          (pair && hipMemcpyAsync(s->bot_b, s->init_bot, pb, hipMemcpyDeviceToDevice,
                                  ctx->stream) != hipSuccess))
        return fail(S, "plane reset", RT_EDEVICE);
      s->fresh = false;
      s->par = 0;
    }
  // windows: kSub per tick for every live unit
  std::vector<Slot*> ended;
  const size_t base = (size_t)b * kTickBytes;
  WindowRec* hw = reinterpret_cast<WindowRec*>(S->h_tick + base);
  size_t wbase[kSub][kGroups];
  size_t nw = 0;
  for (int k = 0; k < kSub; k++)
    for (int m = 0; m < kGroups; m++) {
      wbase[k][m] = nw;
      nw += S->unit_slot[m].size();
    }
  size_t op = base + align256(nw * sizeof(WindowRec));  // next op byte, offset in the block
  for (int k = 0; k < kSub; k++) {
    for (int m = 0; m < kGroups; m++) {
      const int H = fused_shape_halo(S->ctx, group_model(m), group_nc(m), S->group_rule[m],
                                     S->shape[m]);
      for (size_t u = 0; u < S->unit_slot[m].size(); u++) {
        Slot* s = S->unit_slot[m][u];
        WindowRec& w = hw[wbase[k][m] + u];
        w = WindowRec{0, 0, 0, 0, 0};
        if (s->phase != 0 || s->cursor >= s->total) continue;
        // segment containing the cursor
        int seg = 0;
        while (seg + 1 < (int)s->seg_start.size() && s->seg_start[seg + 1] <= s->cursor) seg++;
        const long seg_end = seg + 1 < (int)s->seg_start.size() ? s->seg_start[seg + 1] : s->total;
        int K = 0, pulls = 0;
        while (s->cursor + K < seg_end && K < kOpsPerUnit) {
          const int p = (s->ops[s->cursor + K] & (OP_PULL_A | OP_PULL_B)) ? 1 : 0;
          if (pulls + p > H) break;
          pulls += p;
          K++;
        }
        std::memcpy(S->h_tick + op, s->ops.data() + s->cursor, (size_t)K);
        w.s0 = (long)op;
        op += (size_t)K;
        w.K = K;
        w.init = s->cursor == s->seg_start[seg] ? 1 : 0;
        w.seed = (uint32_t)s->seeds[seg];
        w.flags = s->par;
        s->par ^= 1;
        s->cursor += K;
        if (s->cursor >= s->total) ended.push_back(s);
      }
    }
  }
  if (hipMemcpyAsync(S->d_tick + base, S->h_tick + base, op - base, hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess)
    return fail(S, "tick upload", RT_EDEVICE);
  const WindowRec* dw = reinterpret_cast<const WindowRec*>(S->d_tick + base);
  // Launch groups are independent: domino pairs, lozenge pairs and the
  // one-chain walks each get a stream (forked after the uploads above and
  // joined before the ended slots below), so one group's tile loads
  // overlap another's steps.
  int live[3] = {0, 0, 0};
  for (int m = 0; m < kGroups; m++)
    if (!S->unit_slot[m].empty()) live[group_stream(m)] = 1;
  const bool fork = live[0] + live[1] + live[2] > 1;
  if (fork) {
    int rc = ctx->ensure_streams(3);
    if (rc) return fail(S, ctx->err, rc);
    if (hipEventRecord(ctx->evj[0], ctx->stream) != hipSuccess ||
        hipStreamWaitEvent(ctx->sx[1], ctx->evj[0], 0) != hipSuccess ||
        hipStreamWaitEvent(ctx->sx[2], ctx->evj[0], 0) != hipSuccess)
      return fail(S, "stream fork", RT_EDEVICE);
  }
  for (int k = 0; k < kSub; k++)
    for (int m = 0; m < kGroups; m++) {
      if (S->unit_slot[m].empty()) continue;
      // per-unit rule and weights live in the unit table; the launch
      // parameters carry the rectangle-triangle weights (the reference
      // kernel's table for every served walk)
      ModelParams mp = S->unit_slot[m][0]->req->mp;
      mp.rule = S->group_rule[m];  // the group's kernel choice (units keep their own rules)
      const int gs = fork ? group_stream(m) : 0;
      int rc = fused_launch_windows(ctx, gs ? ctx->sx[gs] : ctx->stream, group_model(m),
                                    group_nc(m), S->shape[m], mp, S->d_units[m], S->d_work[m],
                                    S->nwork[m], dw + wbase[k][m]);
      if (rc) return fail(S, ctx->err, rc);
    }
  if (fork)
    for (int g = 1; g < 3; g++)
      if (hipEventRecord(ctx->evj[g], ctx->sx[g]) != hipSuccess ||
          hipStreamWaitEvent(ctx->stream, ctx->evj[g], 0) != hipSuccess)
        return fail(S, "stream join", RT_EDEVICE);
  // slots whose round ended: cells back into the first buffer set, lazy
  // sync of their planes, then (CFTP) coalescence flag
  for (Slot* s : ended) {
    const bool pair = !s->req->walk;
    const size_t pb = (size_t)s->ncol * (size_t)s->plane * s->cb;
    if (s->par) {
      if (hipMemcpyAsync(s->top, s->top_b, pb, hipMemcpyDeviceToDevice, ctx->stream) !=
              hipSuccess ||
          (pair && hipMemcpyAsync(s->bot, s->bot_b, pb, hipMemcpyDeviceToDevice, ctx->stream) !=
                       hipSuccess))
        return fail(S, "plane copy-back", RT_EDEVICE);
      s->par = 0;
    }
    FusedJob one;
    one.model = s->req->model;
    one.nc = pair ? 2 : 1;
    UnitDesc d{};
    for (int c = 0; c < s->ncol; c++) {
      d.plane[0][c] = s->top + (size_t)c * s->plane * s->cb;
      if (pair) d.plane[1][c] = s->bot + (size_t)c * s->plane * s->cb;
    }
    d.rows = s->req->N;
    one.units.push_back(d);
    one.hist.push_back(s->hist);
    int rc = fused_sync(ctx, one);
    if (rc) return fail(S, ctx->err, rc);
    if (pair) {
      rc = launch_differ_pair(ctx, s->top, s->bot, (long)pb,
                              S->d_flags + (size_t)b * kMaxSlots + s->id);
      if (rc) return fail(S, ctx->err, rc);
    }
    s->phase = 1;
    s->flag_tick = S->tick;
  }
  if (!ended.empty() &&
      hipMemcpyAsync(S->h_flags + (size_t)b * kMaxSlots, S->d_flags + (size_t)b * kMaxSlots,
                     kMaxSlots * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    return fail(S, "flag download", RT_EDEVICE);
  if (hipEventRecord(S->ev[b], ctx->stream) != hipSuccess) return fail(S, "event", RT_EDEVICE);
  S->ev_used[b] = true;
  S->tick++;
  S->ticks++;
  return RT_OK;
}

static void loop_main(Server* S) {
  hipSetDevice(S->ctx->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> g(S->mu);
      if (S->stop) break;
      // admit in arrival order, letting smaller requests pass a blocked one
      for (auto it = S->queue.begin(); it != S->queue.end();) {
        Req* r = *it;
        if (r->cancel) {
          r->status = RT_ECANCELLED;
          it = S->queue.erase(it);
          S->cv.notify_all();
          continue;
        }
        if ((int)S->active.size() >= S->max_active) break;
        g.unlock();
        int rc = admit(S, r);
        g.lock();
        if (rc == RT_OK) {
          it = S->queue.erase(it);
        } else if (rc == RT_EAGAIN) {
          ++it;
        } else {
          for (auto& kv : S->reqs)
            if (kv.second->status == RT_EAGAIN) kv.second->status = rc;
          S->cv.notify_all();
          return;
        }
      }
      const bool pending_flags = S->ev_used[0] || S->ev_used[1];
      if (S->active.empty() && !pending_flags) {
        S->cv.wait(g, [&] { return S->stop || !S->queue.empty(); });
        continue;
      }
    }
    if (S->active.empty()) {
      // drain the last ticks' events so their flags are processed
      for (int b = 0; b < 2; b++) S->ev_used[b] = false;
      continue;
    }
    if (run_tick(S) != RT_OK) {
      std::lock_guard<std::mutex> g(S->mu);
      for (auto& kv : S->reqs)
        if (kv.second->status == RT_EAGAIN) kv.second->status = S->loop_rc;
      S->cv.notify_all();
      return;
    }
  }
}

}  // namespace rt

using namespace rt;

struct rt_server {
  Server s;
};

extern "C" int rt_server_create(rt_ctx* ctx, size_t hbm_budget, int max_active, rt_server** out) {
  if (!ctx || !out || hbm_budget == 0 || max_active <= 0) return RT_EINVAL;
  rt_server* sv = new rt_server();
  Server* S = &sv->s;
  S->ctx = ctx;
  S->budget = hbm_budget;
  S->max_active = std::min(max_active, kMaxSlots);
  hipSetDevice(ctx->device);
  bool ok = hipMalloc(&S->d_arena, hbm_budget) == hipSuccess;
  for (int m = 0; m < kGroups && ok; m++)
    ok = hipMalloc(&S->d_units[m], kMaxSlots * sizeof(UnitDesc)) == hipSuccess;
  ok = ok && hipMalloc(&S->d_tick, 2 * kTickBytes) == hipSuccess;
  ok = ok && hipMalloc(&S->d_flags, 2 * kMaxSlots * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipHostMalloc(&S->h_tick, 2 * kTickBytes) == hipSuccess;
  ok = ok && hipHostMalloc(&S->h_flags, 2 * kMaxSlots * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipEventCreate(&S->ev[0]) == hipSuccess && hipEventCreate(&S->ev[1]) == hipSuccess;
  if (!ok) {
    ctx->err = "server: device allocation failed (budget too large?)";
    rt_server_destroy(sv);
    return RT_ENOMEM;
  }
  std::memset(S->h_flags, 0, 2 * kMaxSlots * sizeof(uint32_t));
  S->arena.init(hbm_budget);
  for (int k = kMaxSlots - 1; k >= 0; k--) S->free_ids.push_back(k);
  S->loop = std::thread(loop_main, S);
  *out = sv;
  return RT_OK;
}

extern "C" void rt_server_destroy(rt_server* sv) {
  if (!sv) return;
  Server* S = &sv->s;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->loop.joinable()) S->loop.join();
  if (S->ctx && S->ctx->stream) hipStreamSynchronize(S->ctx->stream);
  for (Slot* s : S->active) delete s;
  for (auto& kv : S->reqs) delete kv.second;
  if (S->d_arena) hipFree(S->d_arena);
  for (int m = 0; m < kGroups; m++) {
    if (S->d_units[m]) hipFree(S->d_units[m]);
    if (S->d_work[m]) hipFree(S->d_work[m]);
  }
  if (S->d_tick) hipFree(S->d_tick);
  if (S->d_flags) hipFree(S->d_flags);
  if (S->h_tick) hipHostFree(S->h_tick);
  if (S->h_flags) hipHostFree(S->h_flags);
  for (int b = 0; b < 2; b++)
    if (S->ev[b]) hipEventDestroy(S->ev[b]);
  delete sv;
}

extern "C" int rt_server_submit(rt_server* sv, const rt_request* q, uint64_t* id) {
  if (!sv || !q || !id || !q->tmax) return RT_EINVAL;
  Server* S = &sv->s;
  const bool walk = q->kind == RT_REQ_WALK;
  if (!walk && q->kind != RT_REQ_CFTP) return RT_EINVAL;
  const int model = q->model == RT_MODEL_DOMINO    ? M_DOMINO
                    : q->model == RT_MODEL_LOZENGE ? M_LOZENGE
                    : q->model == RT_MODEL_RECTTRI ? M_RECTTRI
                                                   : -1;
  // CFTP is served for the monotone models only (no rectangle-triangle)
  if (model < 0 || (!walk && model == M_RECTTRI)) return RT_EINVAL;
  const int N = q->N;
  if (model == M_DOMINO ? (N < 6 || N % 2) : (N < 9 || N % 3)) return RT_EINVAL;
  if (walk) {
    if (q->steps <= 0) return RT_EINVAL;
    if (model == M_DOMINO && (q->rule.kind < RT_DOMINO_TOGGLE08 ||
                              q->rule.kind > RT_DOMINO_HEATBATH2P))
      return RT_EINVAL;
  } else {
    if (!q->tmin || q->t0 <= 0 || q->max_rounds <= 0 ||
        (q->sched != RT_SCHED_REFERENCE && q->sched != RT_SCHED_DOUBLING))
      return RT_EINVAL;
    if (model == M_DOMINO && !domino_cftp_rule(&q->rule)) return RT_EINVAL;
  }
  int ncol;
  long plane, nstream, cols;
  const size_t bytes = footprint(model, N, walk ? 1 : 2, &ncol, &plane, &nstream, &cols);
  if (bytes > S->budget) return RT_EADMIT;
  const size_t need_params = model == M_DOMINO ? (size_t)N * N / 2 : (size_t)N * (N / 3);
  if (S->ctx->nparam < need_params) return RT_EPARAMS;
  Req* r = new Req();
  r->model = model;
  r->walk = walk;
  r->walk_steps = walk ? q->steps : 0;
  r->N = N;
  r->tmax.assign(q->tmax, q->tmax + (size_t)N * N);
  if (!walk) r->tmin.assign(q->tmin, q->tmin + (size_t)N * N);
  r->master = q->master_seed;
  r->sched = q->sched;
  r->t0 = q->t0;
  r->max_rounds = walk ? 1 : q->max_rounds;
  r->bytes = bytes;
  if (model == M_DOMINO) {
    r->mp.rule = q->rule.kind;
    r->mp.dw.wa2 = (float)(q->rule.a * q->rule.a);
    r->mp.dw.wb2 = (float)(q->rule.b * q->rule.b);
  }
  if (model == M_RECTTRI) r->mp.rw = recttri_weights(nullptr);
  {
    std::lock_guard<std::mutex> g(S->mu);
    if (S->loop_rc != RT_OK) {
      delete r;
      return S->loop_rc;
    }
    r->id = S->next_id++;
    r->t_submit = now_s();
    S->reqs[r->id] = r;
    S->queue.push_back(r);
    *id = r->id;
  }
  S->cv.notify_all();
  return RT_OK;
}

extern "C" int rt_server_poll(rt_server* sv, uint64_t id, int32_t* out, rt_cftp_stats* st,
                              double* latency_s) {
  if (!sv) return RT_EINVAL;
  Server* S = &sv->s;
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return RT_EINVAL;
  Req* r = it->second;
  if (r->status == RT_EAGAIN) return RT_EAGAIN;
  const int status = r->status;
  if (status == RT_OK && out) std::memcpy(out, r->out.data(), r->out.size() * sizeof(int32_t));
  if (st) *st = r->st;
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  S->reqs.erase(it);
  delete r;
  return status;
}

extern "C" int rt_server_wait(rt_server* sv, uint64_t id, double timeout_s) {
  if (!sv) return RT_EINVAL;
  Server* S = &sv->s;
  std::unique_lock<std::mutex> g(S->mu);
  auto done = [&] {
    auto it = S->reqs.find(id);
    return it == S->reqs.end() || it->second->status != RT_EAGAIN;
  };
  if (timeout_s < 0) S->cv.wait(g, done);
  else S->cv.wait_for(g, std::chrono::duration<double>(timeout_s), done);
  return done() ? RT_OK : RT_EAGAIN;
}

extern "C" int rt_server_cancel(rt_server* sv, uint64_t id) {
  if (!sv) return RT_EINVAL;
  Server* S = &sv->s;
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return RT_EINVAL;
    it->second->cancel = true;
  }
  S->cv.notify_all();
  return RT_OK;
}

extern "C" int rt_server_get_stats(rt_server* sv, rt_server_stats* out) {
  if (!sv || !out) return RT_EINVAL;
  Server* S = &sv->s;
  std::lock_guard<std::mutex> g(S->mu);
  std::vector<double> l = S->latencies;
  std::sort(l.begin(), l.end());
  auto q = [&](double p) {
    if (l.empty()) return 0.0;
    const size_t k = std::min(l.size() - 1, (size_t)(p * (double)(l.size() - 1) + 0.5));
    return l[k];
  };
  out->completed = S->completed;
  out->queued = (long)S->queue.size();
  out->ticks = S->ticks;
  out->p50_latency_s = q(0.50);
  out->p99_latency_s = q(0.99);
  out->span_s = S->t_last > S->t_first ? S->t_last - S->t_first : 0.0;
  out->samples_per_s = out->span_s > 0 ? (double)S->completed / out->span_s : 0.0;
  out->status = S->loop_rc;
  return RT_OK;
}

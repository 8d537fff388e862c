// rt_fused.hpp -- the fused sweep-block engine shared by plain walks, CFTP
// and batched chains.
//
// A "unit" is one TinyMT stream table driving NC chains (NC = 1 for a walk,
// 2 for a CFTP top/bottom pair that share every draw) on one domain.  A
// launch advances every unit of a job by one window of K colour steps.
// Each workgroup owns a tile of the unit's compressed cell grid, loads the
// tile plus an H-cell halo into LDS (all colour planes of all NC chains)
// and the TinyMT states of those cells into VGPRs, runs the K steps with one
// workgroup barrier per step, and writes back the tile interior.
//
// Exactness of the halo: a step only reads neighbours when it refreshes the
// rotated colour from a colour rotated since (a "pull"); the local move is
// pointwise.  Cells whose neighbourhood reaches past the loaded region go
// stale, one cell deeper per pull step, so a window with at most H pull
// steps leaves the tile interior exact.  Halo cells also advance their
// (owners') TinyMT streams, redundantly, so every draw is the reference's.
#pragma once
#include <algorithm>
#include <thread>
#include <vector>

#include "rt_kernels.hpp"

namespace rt {

enum ModelKind { M_DOMINO = 0, M_LOZENGE = 1, M_RECTTRI = 2 };

// Columns of halo a tile needs for windows of at most H pulling steps.  A
// pull reads rows i-1 and i+1, so rows go stale one deeper per pull; its
// column offsets are 1 only in some rows: domino (dominokernel*.cl
// UpdateTiles, j -/+ (i+t) parity) in one row parity per side, so at most
// ceil(H/2) stale columns per side; lozenge (lozengekernel.cl:38-53, p2 p4
// p8 p16) in one row residue mod 3 each, at most ceil(H/3).  Both bounds
// hold over every colour sequence (host tests, exhaustively).
// Rectangle-triangle keeps H.
__host__ __device__ constexpr int col_halo(int model, int H) {
  return model == M_DOMINO ? (H + 1) / 2 : model == M_LOZENGE ? (H + 2) / 3 : H;
}

// One op byte per colour step:
//   bits 0-1 colour rotated;
//   bit 2    domino: pull the rotated colour from the other one;
//            lozenge: refresh bits facing colour-1 ("pull A");
//            rect-tri: rebuild from the previous step's colour;
//   bit 3    lozenge: refresh bits facing colour-2 ("pull B");
//   bits 4-5 rect-tri: previous step's colour.
enum : uint8_t { OP_COLOUR = 3u, OP_PULL_A = 4u, OP_PULL_B = 8u };

// Cells and TinyMT states are double-buffered: a window reads set `plane`
// / `state` and writes set `plane_b` / `state_b`, or the reverse when its
// parity bit is set.  A workgroup reads its halo from neighbouring tiles,
// and with more tiles than fit on the chip at once some workgroups start
// after a neighbour has finished its window, so one buffer updated in place
// would hand them that neighbour's new cells.
struct UnitDesc {
  void* plane[2][3];       // [chain][colour] -> rows x cols cells (uint8 or int32)
  void* plane_b[2][3];     // second buffer set
  uint32_t* state;         // SoA s0..s3, stride = nstream
  uint32_t* state_b;       // second buffer set
  int no_state_write;      // the unit's streams are written by another unit
  const uint32_t* param;   // SoA mat1, mat2, tmat, stride = pstride
  long nstream;
  long pstride;
  const uint8_t* ops;      // op byte per global step of this unit
  int rows, cols;          // compressed grid
  int tiles_c;             // tile columns
  int rule;                // domino flip rule of this unit
  DominoWeights dw;        // domino two-periodic weights of this unit
};

// Per-unit window of one launch: ops are read from unit.ops[s0 .. s0+K);
// init re-seeds the unit's TinyMT streams with `seed` before the first
// step; flags bit 0 is the window's buffer parity (set: read the _b
// buffers, write the others).  K = 0 without init leaves the unit idle.
struct WindowRec {
  long s0;
  int K;
  int init;
  uint32_t seed;
  int flags;
};

// A window's source buffers in plane / state, its destination in the _b
// fields (device side, on a private copy of the unit).
__device__ __forceinline__ void window_buffers(UnitDesc& U, int parity) {
  if (parity & 1) {
#pragma unroll
    for (int ch = 0; ch < 2; ch++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        void* t = U.plane[ch][k];
        U.plane[ch][k] = U.plane_b[ch][k];
        U.plane_b[ch][k] = t;
      }
    uint32_t* t = U.state;
    U.state = U.state_b;
    U.state_b = t;
  }
}

struct WorkItem {
  int unit;
  int tile;  // tile index within the unit (row-major over tiles_c)
  int cs;    // column shift of the tile: its interior starts at column tc * TW + cs
};

struct ModelParams {
  int rule;
  DominoWeights dw;
  RTWeights rw;
};

// Unit colour-step history used to build op bytes and the final sync.
struct OpBuilder {
  int model;
  long last[3] = {-1, -1, -1};
  int prev = -1;
  long g = 0;
  // Append the op bytes of `colours` (colour | pull flags) to ops.
  void append(const std::vector<uint8_t>& colours, std::vector<uint8_t>& ops);
};

// One launch's share of a job: K steps from global step s0; init re-seeds
// the TinyMT streams with segment seg's seed before the first step.
struct FusedWindow {
  long s0;
  int K;
  int init;
  int seg;
};

// f(0) .. f(n-1) on up to one host thread per hardware thread (at most
// max_threads when that is positive).
template <class F>
void parallel_for(int n, F f, int max_threads = 0) {
  int nt = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
  if (max_threads > 0) nt = std::min(nt, max_threads);
  if (nt <= 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (int i = t; i < n; i += nt) f(i);
    });
  for (auto& x : th) x.join();
}

// Host job: units on one device advancing through a shared segment list.
struct FusedJob {
  int model = M_DOMINO;
  int nc = 1;
  ModelParams mp{};
  std::vector<UnitDesc> units;        // host copy (ops/seeds filled by the driver)
  std::vector<long> seg_steps;        // shared segment lengths
  std::vector<std::vector<uint32_t>> seg_seeds;   // [unit][segment]
  std::vector<std::vector<uint8_t>> ops;          // [unit] concatenated op bytes
  std::vector<OpBuilder> hist;        // [unit]
  int shape = 0;                      // tile shape, resolved by fused_scratch_bytes
  bool split = false;                 // pairs run as two one-chain units (see below)
  int shape_source = 0;               // 0 rule of thumb / forced, 1 tuned now, 2 tuned before
  bool tune = false;                  // time the candidate shapes on the first launches
  int halo = 6;                       // tile halo H: windows hold at most H pulling steps
  // Launch windows per unit, cut by fused_scratch_bytes from the unit's own
  // pull steps: launch l advances unit u by windows[u][l] (nothing once
  // the unit's windows run out), so units with independent schedules do
  // not shorten each other's windows.
  std::vector<std::vector<FusedWindow>> windows;
  size_t nlaunch = 0;                 // longest window list
  // Per unit (or empty): fused_live_spans of the unit's start planes.
  // Domino and byte-cell lozenge work lists then hold, per row of tiles,
  // only the tiles whose interiors cover the cells within one row and one
  // column of those spans, shifted to start at the first such column
  // (WorkItem::cs).  The cells no tile covers are zero in every tiling the
  // chain can reach, so leaving them at zero in both buffer sets (their
  // TinyMT streams never advanced) gives the same planes as running them.
  std::vector<std::vector<int>> spans;
};

// Row extents of the nonzero cells of a unit's start planes: `nplanes`
// compressed planes of rows x cols bytes (every colour plane of every
// chain's start tiling); entry 2r / 2r+1 is the first / last column of row
// r holding a nonzero cell in any of them (first > last: none).  A cell
// value's bits are lattice edges between two faces of the domain (a
// domino's or lozenge's inner edge), and every face of the domain is
// covered in every tiling, so a cell that is nonzero at any time has a
// lattice neighbour (one row and one compressed column away at most) that
// is nonzero in any valid tiling of the domain.
std::vector<int> fused_live_spans(const uint8_t* const* planes, int nplanes, int rows, int cols);
// Work items of unit index `unit` (rows, cols and tiles_c set) under the
// launch geometry of (model, nc, rule, shape): every tile, or with the
// unit's fused_live_spans (domino and byte-cell lozenge units) per row of
// tiles only those covering the cells within one row and one column of the
// spans, shifted to start at the first such column (FusedJob::spans).
void fused_unit_work(const rt_ctx* ctx, int model, int nc, int rule, int shape,
                     const UnitDesc& u, const std::vector<int>& spans, int unit,
                     std::vector<WorkItem>& out);

// Run all windows of the job on ctx->stream (no final sync).  `scratch`
// must hold fused_scratch_bytes(ctx, job) bytes of device memory; that call
// also cuts the launch windows and fixes the job's tile shape (forced by
// ctx->tile_shape, remembered from an earlier tuning of jobs like it, or a
// rule of thumb on the tile count, with job.tune set when the first windows
// should time the candidates).
size_t fused_scratch_bytes(rt_ctx* ctx, FusedJob& job);
int fused_run(rt_ctx* ctx, FusedJob& job, void* scratch);

// The same in phases, so the launch sequence can be captured in a
// hipGraph: fused_upload copies unit tables, work lists, ops and seeds to
// the device; fused_tune (a no-op unless job.tune) runs the first launches
// split among the candidate tile shapes, timed, keeps the fastest for the
// rest of the job and for later jobs like it; fused_launch enqueues only
// kernels, from launch plan.first on.  Results never depend on the shape.
struct FusedPlan {
  // One way to launch the job: a tile shape, and for chain pairs whether
  // the two chains run in one workgroup per tile (sharing each draw) or in
  // two ("split": two one-chain units per pair reading the same TinyMT
  // states, the first of them writing them back), which doubles the
  // workgroups when the chip has idle CUs.
  struct Tiles {
    const UnitDesc* d_units;
    const WorkItem* d_work;
    unsigned nwork;
    int shape;
    bool split;
    const WindowRec* d_wins;  // [launch][unit of this launch shape]
    size_t nunits;
    std::vector<unsigned> ends;  // work-list size after each unit (stream split points)
    // per job unit: the unit's sites scaled by the share of its cells that
    // the tiles of this shape compute (FusedJob::spans)
    std::vector<long> unit_sites;
  };
  const UnitDesc* d_units = nullptr;  // the tiles of job.shape
  const WorkItem* d_work = nullptr;
  unsigned nwork = 0;
  std::vector<Tiles> cand;            // every uploaded shape, job.shape first
  const WindowRec* d_wins = nullptr;  // [launch][unit]
  const struct CopyDesc* d_copies = nullptr;  // second set -> first, after the last launch
  unsigned ncopies = 0;
  size_t nlaunch = 0;
  size_t first = 0;                   // launches before this one ran in fused_tune
  long total_steps = 0;               // steps per chain
  long streams = 0;                   // TinyMT sites over all units
  std::vector<long> launch_sites;     // per launch: sites it computes (Tiles::unit_sites)
  std::vector<long> launch_work;      // per launch: sum over units of K x sites
  long sites(size_t l0, size_t l1) const {
    long n = 0;
    for (size_t l = l0; l < l1 && l < launch_sites.size(); l++) n += launch_sites[l];
    return n;
  }
  // launches run by fused_tune, by tile shape (index shape - 1)
  double trial_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long trial_launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long trial_sites[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
int fused_upload(rt_ctx* ctx, FusedJob& job, void* scratch, FusedPlan* plan);
int fused_tune(rt_ctx* ctx, FusedJob& job, FusedPlan* plan);
// The plan's launches after the tuner's; with split_streams the units are
// split into rt_set_streams parts run on that many streams, forked from and
// joined into the context's stream with events (under graph capture the
// parts become parallel chains of the graph).
int fused_launch(rt_ctx* ctx, const FusedJob& job, const FusedPlan& plan, bool split_streams);
// Slot of a launch shape in rt_timing.shape_*: shape - 1, +4 when split;
// 7 for 16x4 pairs at eight waves per SIMD (shape 7).
inline int shape_slot(int shape, bool split) {
  return shape == 7 ? 7 : shape - 1 + (split ? 4 : 0);
}
// Add the plan's tuning launches to per-shape kernel timing.
void fused_add_trials(const FusedPlan& plan, rt_timing* acc);

// Tile shape for launches over `tiles64` 64-row tiles of (model, nc)
// units whose windows are cut by the caller (the serving loop, with the
// halo of fused_shape_halo).
int fused_serve_shape(const rt_ctx* ctx, int model, int nc, int rule, long tiles64);
// One launch over `nwork` work items of tile shape `shape` where each unit
// follows its own window (wins[unit]); model / chains / params as in a job.
int fused_launch_windows(rt_ctx* ctx, hipStream_t st, int model, int nc, int shape,
                         const ModelParams& mp, const UnitDesc* d_units, const WorkItem* d_work,
                         unsigned nwork, const WindowRec* d_wins);
// Tile counts of a unit under a tile shape's geometry.
void fused_tiles(const rt_ctx* ctx, int model, int nc, int rule, int shape, int rows, int cols,
                 int* tiles_r, int* tiles_c);
// Halo depth of a serving launch group on a tile shape (the most pulling
// steps one window may take): 9 for lozenge and 8 for domino chain pairs on
// shape 7 (the packed-pair kernels), else the shape's own (6;
// rect-triangle pairs 4).
int fused_shape_halo(const rt_ctx* ctx, int model, int nc, int rule, int shape);
// Bring every colour plane of every unit up to date (end of a walk).
int fused_sync(rt_ctx* ctx, const FusedJob& job);


}  // namespace rt

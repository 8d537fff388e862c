/*
 * rtiling.h -- C ABI of the MI355X random-tiling engine (librtiling.so).
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (RT_OK = 0, negative on error) and never throws; rt_last_error() gives
 * the message of the last failing call on a context.  Calls on one context
 * are synchronous and must come from one host thread at a time.  The
 * batched calls use helper host threads of their own while they run (op
 * schedules are built in parallel, and the next CFTP round's while the
 * device runs the current one); none outlives the call.
 *
 * Tilings are int32 N*N row-major arrays, exactly the reference's `tiling`
 * (std::vector<int>, src/common/common.h:50-52).  Walk calls mutate the
 * tiling in place and return after the device work has finished, like the
 * reference's blocking cl::copy (src/Domino/DominoTiler.cpp:100-114).
 */
#ifndef RTILING_H
#define RTILING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define RT_OK 0
#define RT_EINVAL (-1)      /* bad argument: size, shape, null pointer */
#define RT_EPARAMS (-2)     /* TinyMT parameter table missing, short, malformed */
#define RT_EIO (-3)         /* file cannot be opened */
#define RT_ENOCOUPLE (-4)   /* CFTP did not coalesce within max_rounds */
#define RT_EDEVICE (-5)     /* HIP runtime error */
#define RT_ENOMEM (-6)      /* device or host allocation failed */
#define RT_EUNTILEABLE (-7) /* Thurston found the region not tileable */
#define RT_EADMIT (-8)      /* serving: request exceeds the HBM budget */
#define RT_ECANCELLED (-9)  /* serving: request cancelled */
#define RT_EAGAIN (-10)     /* serving: result not ready yet */

typedef struct rt_ctx rt_ctx;

/* Context: one HIP device, one stream, the TinyMT parameter table and the
 * device arenas.  Replaces the reference's (cl::Context, cl::CommandQueue,
 * devices, kernel source) constructor arguments (DominoTiler.h:70-76,
 * LozengeTiler.h:55-61, RectTriangleTiler.h:54-61). */
int rt_create(rt_ctx** out, int device_ordinal);
void rt_destroy(rt_ctx* ctx);
const char* rt_last_error(const rt_ctx* ctx);
/* Name of the device the context runs on (e.g. "gfx950"). */
const char* rt_device_name(const rt_ctx* ctx);

/* TinyMT parameter table: `count` (mat1, mat2, tmat) triples, interleaved.
 * Replaces LoadTinyMT -> get_params_buffer (src/TinyMT/file_reader.cpp:22-38). */
int rt_set_tinymt_params(rt_ctx* ctx, const uint32_t* mat1_mat2_tmat,
                         size_t count);
/* Parse `count` parameter lines in the tinymt32dc text format (hex fields
 * after the 3rd/4th/5th comma, '#' lines skipped) and load them.
 * Mirrors DominoTiler::LoadTinyMT(path, size) (DominoTiler.cpp:235-237,
 * file_reader.cpp:51-70,139-166). */
int rt_load_tinymt_file(rt_ctx* ctx, const char* path, size_t count);
/* Parse only (host), into out3[3*count]. */
int rt_parse_tinymt_file(const char* path, size_t count, uint32_t* out3);

/* Device time (ms, HIP events on the context's stream) and kernel launch
 * count of the last walk / CFTP / batch call on this context. */
typedef struct {
  double device_ms;
  long launches;
  long steps;         /* colour steps run, per chain */
  long attempts;      /* flip attempts (TinyMT draws) */
  double kernel_ms;   /* time inside the sweep kernels alone (batched calls) */
  long kernel_launches;
  long kernel_sites;  /* sum over those launches of the TinyMT sites they compute
                       * (the tiles left out below excluded) */
  /* The same three split by sweep-block launch shape: index 0, 1, 2 =
   * 64 rows / 8 waves, 64 rows / 16 waves, 32 rows / 16 waves, 3 = 32 rows /
   * 8 waves for rect-triangle pairs, 4, 5, 6 = shapes 0, 1, 2 with CFTP
   * pairs split over two workgroups per tile, 7 = tile shape 7 below; no
   * reference counterpart. */
  double shape_ms[8];
  long shape_launches[8];
  long shape_sites[8];
  /* Tile shape of the bulk of the last sweep-block job and how it was
   * picked: 0 forced or rule of thumb, 1 tuned in this call, 2 tuned by an
   * earlier call on a job like it. */
  int shape;
  int shape_source;
  /* 1 when the chain pairs of that job ran one chain per workgroup (two
   * workgroups per tile, each drawing the pair's shared numbers). */
  int split;
} rt_timing;
int rt_last_timing(const rt_ctx* ctx, rt_timing* out);

/* Kernel selection for walks: 0 = auto (fused LDS sweep blocks where the
 * shape allows), 1 = one launch per colour step, 2 = fused only, 3 = fused,
 * with the lozenge kernel that packs four cells per lane-word (fewer VALU
 * operations per cell, slower on the side-300 hexagon because its per-cell
 * stream loads are strided); results never depend on it. */
enum { RT_ENGINE_AUTO = 0, RT_ENGINE_STEP = 1, RT_ENGINE_FUSED = 2, RT_ENGINE_FUSED_SWAR = 3 };
int rt_set_engine(rt_ctx* ctx, int engine);

/* Tile shape of the sweep-block kernels (no reference counterpart; a tuning
 * knob for walks and batched CFTP, results do not depend on it).  Tiles are
 * 64 columns wide: 1 = 64 rows worked by 8 waves x 8 rows per lane, 2 = 64
 * rows by 16 waves x 4 rows, 3 = 32 rows by 16 waves x 2 rows; 4, 5, 6 =
 * shapes 1, 2, 3 with CFTP chain pairs split over two workgroups per tile
 * (each runs one chain, both draw the pair's shared numbers); 7 = 48 rows
 * by 16 waves x 3 rows at eight waves per SIMD (two workgroups per CU,
 * VGPRs capped at 64; lozenge and domino chain pairs there hold both
 * chains in one LDS word per cell, on 96-row x 32-column tiles, when the
 * halo is a multiple of 3 (lozenge) or even (domino), and so do one-chain
 * domino walks with a heat-bath rule, in the low halves); 0 = chosen per
 * job (autotuner, else a rule of thumb on the tile count; jobs that the
 * packed kernels take, at their halo, go to shape 7 directly).  The
 * serving loop puts domino and lozenge units on the packed kernels at
 * every size, and other groups on 64-row tiles while their live tiles fit
 * the CUs, shape 7 beyond. */
int rt_set_tile_shape(rt_ctx* ctx, int shape);

/* Tile-shape autotuner (no reference counterpart; on by default).  With
 * shape 0 above, a sweep-block job of at least 194 launch windows (98
 * for one-chain walks) whose kind (model, chains, tile count to a factor
 * of two) has not been tuned on this context runs its first windows split
 * among the candidate shapes, timed with events, and keeps the fastest for
 * the rest of the job and for later jobs of that kind.  Results never depend on the shape.
 * mode 0 = off (rule of thumb on the tile count), 1 = on, 2 = on and forget
 * earlier choices. */
int rt_set_autotune(rt_ctx* ctx, int mode);

/* Sweep-block launches with no more workgroups than CUs (and lozenge-pair
 * launches on tile shape 7 with more than four per CU) hand each of the
 * chip's 8 XCDs (which receive workgroups round robin) a contiguous share of
 * the tile list, so tiles that read each other's halo rows share an L2
 * (on = 1, the default; 2: every launch; 0: none; no reference counterpart;
 * results never depend on it). */
int rt_set_xcd_order(rt_ctx* ctx, int on);

/* Halo depth H of the sweep-block tiles (no reference counterpart; default
 * 6, 1..12): a launch runs a window of at most H pulling colour steps of
 * each unit, tiles overlap by 2H rows and by 2H/3 (lozenge) or H (domino)
 * columns.  Larger H means fewer launches and more redundant halo work.
 * Results never depend on it.  Until it is called, lozenge chain-pair
 * jobs whose 64-row tiles would fill over a third of the CUs take H = 9,
 * and domino units on the packed kernel (chain pairs, heat-bath walks)
 * H = 8 at every size; halo 0 returns to that default.  Setting it
 * forgets the autotuner's earlier choices; the serving loop keeps 6 (9
 * and 8 for lozenge and domino units on the packed kernels) and
 * rectangle-triangle pairs 4. */
int rt_set_halo(rt_ctx* ctx, int halo);

/* Streams of a batched CFTP or (eager) batched walk job (no reference
 * counterpart; default 2, 1..4): a job of at least two workgroups per CU
 * splits its units (samples, chains) into that many parts of about equal
 * tile count, each launched on its own stream, so one part's tile loads
 * and stores overlap another's colour steps; the parts join before the
 * call returns.  Results never depend on it.  Streams run concurrently only
 * on distinct hardware queues, of which a process has GPU_MAX_HW_QUEUES
 * (4 by default): every context holds one stream, up to three more once it
 * splits a job or serves, so keep few contexts open at once (a server's
 * stream groups measured 27 % slower with another context's two streams
 * open in the process). */
int rt_set_streams(rt_ctx* ctx, int n);

/* Round callback of the CFTP entry points (no reference counterpart: the
 * reference's driver loop, AztecDiamondCftpEx.cpp:57-77, reports nothing
 * until it has coupled).  After each round's coalescence test, in the
 * calling thread: fn(user, round, steps of the round, samples finished so
 * far, samples in the call).  fn must not call into the library with the
 * same context.  fn = NULL turns it off (the default). */
typedef void (*rt_round_fn)(void* user, int round, long round_steps, int finished, int total);
int rt_set_round_callback(rt_ctx* ctx, rt_round_fn fn, void* user);

/* The device's CU count and LDS bytes per CU and per workgroup. */
int rt_device_limits(const rt_ctx* ctx, int* num_cu, size_t* lds_per_cu,
                     size_t* lds_per_workgroup);

/* ---------------------------------------------------------------- domino */
/* Flip rules: the four reference kernel files. */
#define RT_DOMINO_TOGGLE08 0    /* dominokernel.cl:4-18 */
#define RT_DOMINO_HEATBATH 1    /* dominokernelCFTP.cl:11-31 */
#define RT_DOMINO_METRO2P 2     /* dominokerneltwoperiodic.cl:7-35 */
#define RT_DOMINO_HEATBATH2P 3  /* dominokernelCFTPtwoperiodic.cl:15-44 */

typedef struct {
  int kind;   /* RT_DOMINO_* */
  double a;   /* two-periodic weights (reference: #define a 0.7, b 1.) */
  double b;
} rt_domino_rule;

/* DominoTiler::Walk(tiling&, long steps, long seed) (DominoTiler.cpp:18-122,
 * DominoTiler.h:89). Needs >= N*N/2 TinyMT parameter sets. */
int rt_domino_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps,
                   long seed, const rt_domino_rule* rule);
/* rt_domino_walk with no device or host allocation in the call: no
 * reference counterpart; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as rt_domino_walk, in the reference's launch
 * shape (one launch per colour step).  All device memory of the call -- the
 * TinyMT states (16 B per stream), the two colour planes (N*N bytes) and
 * the int32 tiling staged for the on-device split and merge (4*N*N bytes),
 * each rounded up to 256 bytes -- lives in one fixed 16 MiB array per GPU
 * (N up to 1136, i.e. Aztec diamonds up to order 565); larger N returns
 * RT_EINVAL.  The array keeps the last call's contents and is shared by
 * every context on that GPU: calls that overlap on one GPU race.  Needs
 * the TinyMT parameter table set beforehand (rt_set_tinymt_params). */
int rt_domino_walk_noalloc(rt_ctx* ctx, int32_t* tiling, int N, long steps,
                           long seed, const rt_domino_rule* rule);

/* DominoTiler::Walk(tiling&, vector<long> steps, vector<long> seeds)
 * (DominoTiler.cpp:125-231, DominoTiler.h:92). */
int rt_domino_walk_segments(rt_ctx* ctx, int32_t* tiling, int N,
                            const long* steps, const long* seeds, int nseg,
                            const rt_domino_rule* rule);

/* CFTP (Examples/AztecDiamondCftpEx.cpp:51-77): master mt19937 -> segment
 * seeds; top from tmax, bottom from tmin; both chains share every TinyMT
 * draw on the device. */
#define RT_SCHED_REFERENCE 0  /* prepend 2*t0 per round (AztecDiamondCftpEx.cpp:76) */
#define RT_SCHED_DOUBLING 1   /* prepend the current total: look-back doubles */

typedef struct {
  int rounds;        /* rounds run until coalescence */
  long total_steps;  /* look-back of the coalesced round (sum of segments) */
  long steps_run;    /* colour steps run over all rounds, per chain */
  double device_ms;  /* device time over all rounds */
} rt_cftp_stats;

/* rule: RT_DOMINO_HEATBATH or RT_DOMINO_HEATBATH2P only (the reference's
 * CFTP kernels, dominokernelCFTP*.cl; NULL = heat bath); the 0.8 toggle and
 * the two-periodic Metropolis rule are not monotone couplings, so they
 * return RT_EINVAL here, in rt_domino_cftp_batch and in rt_server_submit. */
int rt_domino_cftp(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin,
                   int N, uint32_t master_seed, int sched, long t0,
                   int max_rounds, const rt_domino_rule* rule, int32_t* out,
                   rt_cftp_stats* stats);

/* --------------------------------------------------------------- lozenge */
/* LozengeTiler::Walk(tiling&, int steps, long seed) (LozengeTiler.cpp:27-159).
 * Needs >= N*(N/3) parameter sets. */
int rt_lozenge_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps,
                    long seed);
/* Segmented lozenge walk: per segment, TinyMT and the colour engine are
 * both seeded with seeds[k] (LozengeTiler.cpp:69,86-88).  The reference
 * has it for domino only (DominoTiler.cpp:125-231). */
int rt_lozenge_walk_segments(rt_ctx* ctx, int32_t* tiling, int N,
                             const long* steps, const long* seeds, int nseg);
int rt_lozenge_cftp(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin,
                    int N, uint32_t master_seed, int sched, long t0,
                    int max_rounds, int32_t* out, rt_cftp_stats* stats);

/* ------------------------------------------------------------- batching */
/* Domino and lozenge walks and CFTP on the sweep-block kernels (the batch
 * entry points below, rt_domino_walk, rt_lozenge_walk and the one-sample
 * CFTP, the server) leave out the tiles whose cells all lie more than one
 * row and one column away from every nonzero cell of the start tilings:
 * in a tiling of a domain a cell's bits are lattice edges between two of
 * the domain's faces, so such cells stay zero in every tiling the chain
 * reaches, and the results equal the reference's, which visits them.  The
 * start tilings must be tilings of their domain, which the reference's
 * kernels assume as well. */
/* Batched CFTP: one exact sample per master seed, all samples of the batch
 * advanced by the same fused launches (one launch covers every tile of
 * every unfinished sample).  Sample k is bit-identical to
 * rt_*_cftp(master_seeds[k]) run alone.  out: nsamples*N*N, stats:
 * nsamples entries.  No reference counterpart: the reference draws one
 * sample per driver run (AztecDiamondCftpEx.cpp:58-77). */
int rt_lozenge_cftp_batch(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                          const uint32_t* master_seeds, int nsamples, int sched, long t0,
                          int max_rounds, int32_t* out, rt_cftp_stats* stats);
int rt_domino_cftp_batch(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                         const uint32_t* master_seeds, int nsamples, int sched, long t0,
                         int max_rounds, const rt_domino_rule* rule, int32_t* out,
                         rt_cftp_stats* stats);

/* Both coupled chains of each sample after exactly `rounds` rounds of the
 * CFTP schedule, coalesced or not: top_out / bot_out (nsamples*N*N each)
 * equal LozengeTiler::Walk over the round's segment list (seeds newest
 * first, AztecDiamondCftpEx.cpp:57-76) from tmax / tmin.  For coupling
 * diagnostics and for checking the batched kernels against plain walks;
 * no reference counterpart.  stats[k].rounds = rounds. */
int rt_lozenge_cftp_batch_rounds(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                                 const uint32_t* master_seeds, int nsamples, int sched, long t0,
                                 int rounds, int32_t* top_out, int32_t* bot_out,
                                 rt_cftp_stats* stats);
/* The domino counterpart: top_out / bot_out equal DominoTiler::Walk over
 * the round's segment list (DominoTiler.cpp:125-231, seeds newest first)
 * with the heat-bath rule `rule` (RT_DOMINO_HEATBATH or
 * RT_DOMINO_HEATBATH2P, dominokernelCFTP*.cl).  Lets a caller run a bounded
 * number of rounds of a CFTP whose coupling time is out of reach and watch
 * the chains approach each other; no reference counterpart. */
int rt_domino_cftp_batch_rounds(rt_ctx* ctx, const int32_t* tmax, const int32_t* tmin, int N,
                                const uint32_t* master_seeds, int nsamples, int sched, long t0,
                                int rounds, const rt_domino_rule* rule, int32_t* top_out,
                                int32_t* bot_out, rt_cftp_stats* stats);

/* Batched domino walks on domains of uneven size: chain k is
 * DominoTiler::Walk(tilings[k], steps, seeds[k]) with side Ns[k]; all
 * chains advance together.  use_graph != 0 captures the launch sequence in
 * a hipGraph and replays it. */
int rt_domino_walk_batch(rt_ctx* ctx, int32_t* const* tilings, const int* Ns, int nchains,
                         long steps, const long* seeds, const rt_domino_rule* rule,
                         int use_graph);
/* The same, plus observables computed on the device from each chain's final
 * colour planes: counts[5*k .. 5*k+4] = vertices of chain k with edge bit
 * 1, 2, 4, 8 set (DominoTiler.h:28-32; each domino sets two) and rotatable
 * faces (vertex state 3 or 12).  download = 0 leaves tilings[] as they were
 * (observables only, no device-to-host copy of the tilings). */
int rt_domino_walk_batch_observe(rt_ctx* ctx, int32_t* const* tilings, const int* Ns,
                                 int nchains, long steps, const long* seeds,
                                 const rt_domino_rule* rule, int use_graph, int download,
                                 long* counts);

/* The same, plus each chain's height function computed on the device from
 * its final colour planes: heights[k] (Ns[k]^2 int32) receives
 * DominoTiler::TilingToHeightfunc(t_k, TilingToDomain(t_k))
 * (DominoTiler.cpp:555-593, 595-609): 0 at the first boundary vertex in
 * row-major order, INT_MAX at vertices the reference's search does not
 * reach.  counts may be NULL; download = 0 leaves tilings[] untouched. */
int rt_domino_walk_batch_heights(rt_ctx* ctx, int32_t* const* tilings, const int* Ns,
                                 int nchains, long steps, const long* seeds,
                                 const rt_domino_rule* rule, int use_graph, int download,
                                 long* counts, int32_t* const* heights);

/* -------------------------------------------------------------- serving */
/* Continuous-batching CFTP server on one context.  A device loop thread
 * admits requests by HBM footprint into slots of an arena of
 * `hbm_budget` bytes (at most `max_active` at once), keeps their chains
 * resident and advances all of them in shared launches, each at its own
 * point of its own schedule.  Request k's sample equals rt_*_cftp on the
 * same inputs, alone or mixed with any others.  Walk requests (kind
 * RT_REQ_WALK) run `steps` steps of the model's plain Walk from tmax with
 * seed master_seed and return rt_domino_walk / rt_lozenge_walk /
 * rt_recttri_walk (the reference kernel's weights) of the same inputs; they
 * are how rectangle-triangle tilings are served, since that model has no
 * monotone coupling for CFTP.  While a server exists, its context must not
 * be used by other calls.  No reference counterpart. */
#define RT_MODEL_DOMINO 0
#define RT_MODEL_LOZENGE 1
#define RT_MODEL_RECTTRI 2
#define RT_REQ_CFTP 0
#define RT_REQ_WALK 1

typedef struct {
  int model;               /* RT_MODEL_*; RT_MODEL_RECTTRI for walks only */
  int N;                   /* tiling side */
  const int32_t* tmax;     /* N*N, copied at submit; a walk's start */
  const int32_t* tmin;     /* CFTP only (may be NULL for walks) */
  uint32_t master_seed;    /* CFTP master seed; a walk's seed */
  int sched;               /* RT_SCHED_* (CFTP) */
  long t0;                 /* CFTP */
  int max_rounds;          /* CFTP */
  rt_domino_rule rule;     /* domino only: heat-bath rules for CFTP, any for walks */
  int kind;                /* RT_REQ_CFTP (0) or RT_REQ_WALK */
  long steps;              /* walk length (>= 1) */
} rt_request;

typedef struct {
  long completed;
  long queued;
  long ticks;              /* fused launches issued per model */
  double p50_latency_s;    /* submit -> result, completed requests */
  double p99_latency_s;
  double span_s;           /* first submit -> last completion */
  double samples_per_s;    /* completed / span_s */
  int status;              /* RT_OK unless the device loop failed */
} rt_server_stats;

typedef struct rt_server rt_server;
int rt_server_create(rt_ctx* ctx, size_t hbm_budget, int max_active, rt_server** out);
void rt_server_destroy(rt_server* srv);
/* RT_EADMIT: larger than the whole budget; RT_EPARAMS: parameter table too
 * small for the request; RT_EINVAL: bad shape or model. */
int rt_server_submit(rt_server* srv, const rt_request* req, uint64_t* id);
/* RT_EAGAIN until done; then the request's status (RT_OK, RT_ENOCOUPLE,
 * RT_ECANCELLED, ...), its sample in out[N*N] and stats; forgets the id. */
int rt_server_poll(rt_server* srv, uint64_t id, int32_t* out, rt_cftp_stats* stats,
                   double* latency_s);
/* Block until request `id` is done (timeout_s < 0: no limit). */
int rt_server_wait(rt_server* srv, uint64_t id, double timeout_s);
int rt_server_cancel(rt_server* srv, uint64_t id);
int rt_server_get_stats(rt_server* srv, rt_server_stats* out);

/* ---------------------------------------------------- rectangle-triangle */
/* RectTriangleTiler::Walk (RectTriangleTiler.cpp:33-171); weights16 is the
 * kernel's weight table (recttrianglekernel.cl:3-13; NULL = all 1.0 with
 * weights[0] = 0). Needs >= N*N/3 parameter sets. */
int rt_recttri_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps,
                    long seed, const float* weights16);

/* ------------------------------------------------ bibone (triangle dimer) */
/* TriangleDimerTiler::Walk (TriangleDimerTiler.cpp:101-695): in place on
 * the N*N lozenge-type tiling (N even); the triangle- and butterfly-type
 * states are rebuilt from it on entry, as the reference does.  Same
 * schedule (default_random_engine(seed): move type, r1, r2 and, for
 * butterfly steps, p1, p2) and TinyMT streams (id j + (N/2)*i, seeded
 * seed + id).  Needs >= (N/2)^2 parameter sets.  Engines: RT_ENGINE_AUTO
 * and RT_ENGINE_FUSED_SWAR run starts that rt_bibone_edge_form accepts on
 * one bit per lattice edge (each move a cycle flip, H/2 steps per launch;
 * rt_set_tile_shape 1-7 picks its tile, 0 the default, reported in
 * rt_timing.shape) and other starts on the eight planes in sweep blocks
 * (rt_timing.shape 0); RT_ENGINE_FUSED forces the eight-plane sweep
 * blocks, RT_ENGINE_STEP one launch per reference kernel.  Results never
 * depend on the engine. */
int rt_bibone_walk(rt_ctx* ctx, int32_t* tiling, int N, long steps, long seed);

/* ---------------------------------------------- host-side domain helpers */
/* An owned int32 vector returned by the helpers below. */
typedef struct rt_vec rt_vec;
const int32_t* rt_vec_data(const rt_vec* v);
size_t rt_vec_size(const rt_vec* v);
void rt_vec_free(rt_vec* v);

/* DominoTiler statics (DominoTiler.h:99-123). */
rt_vec* rt_domino_rectangle(int M, int N);               /* DominoTiler.cpp:249-260 */
rt_vec* rt_domino_aztec_diamond(int n);                  /* DominoTiler.cpp:262-265 */
rt_vec* rt_domino_aztec_rectangle(int n, const int* remove, int k); /* 267-324 */
/* Union of n rectangles {i0, j0, h, w} of unit squares, framed like
 * Rectangle (DominoTiler.cpp:249-260), holes filled.  No reference
 * counterpart: a domain family for batches of uneven chains; tileability is
 * what rt_domino_min_tiling reports. */
rt_vec* rt_domino_rect_union(const int* rects, int n);
int rt_domino_max_tiling(const int32_t* d, size_t dsize, rt_vec** out);   /* 345-438 */
int rt_domino_min_tiling(const int32_t* d, size_t dsize, rt_vec** out);   /* 440-531 */
rt_vec* rt_domino_heightfunc_to_tiling(const int32_t* hf, size_t n,
                                       const int32_t* d, size_t dsize);   /* 534-553 */
rt_vec* rt_domino_tiling_to_heightfunc(const int32_t* t, size_t n,
                                       const int32_t* d, size_t dsize);   /* 555-593 */
rt_vec* rt_domino_tiling_to_domain(const int32_t* t, size_t n);           /* 595-609 */
rt_vec* rt_domino_domain_to_vertices(const int32_t* d, size_t dsize);     /* 611-631 */

/* LozengeTiler statics (LozengeTiler.h:77-92). */
rt_vec* rt_lozenge_almost_hex_domain(int n, int m);      /* LozengeTiler.cpp:422-467 */
int rt_lozenge_max_tiling(const int32_t* d, size_t dsize, rt_vec** out);  /* 308-396 */
int rt_lozenge_min_tiling(const int32_t* d, size_t dsize, rt_vec** out);  /* 218-306 */
rt_vec* rt_lozenge_heightfunc_to_tiling(const int32_t* hf, size_t n,
                                        const int32_t* d, size_t dsize);  /* 399-420 */
rt_vec* rt_lozenge_tiling_to_heightfunc(const int32_t* t, size_t n,
                                        const int32_t* d, size_t dsize);  /* 173-216 */
rt_vec* rt_lozenge_tiling_to_domain(const int32_t* t, size_t n);          /* 514-540 */
rt_vec* rt_lozenge_domain_to_vertices(const int32_t* d, size_t dsize);    /* 483-512 */

/* RectTriangleTiler statics (RectTriangleTiler.h:77-82). */
rt_vec* rt_recttri_max_hex(int n);                       /* RectTriangleTiler.cpp:480-538 */
rt_vec* rt_recttri_min_hex(int n);                       /* 540-600 */
rt_vec* rt_recttri_slope_hex(int n);                     /* 173-478 */

/* TriangleDimerTiler statics (TriangleDimerTiler.h:194-209). */
rt_vec* rt_bibone_ice_cream_cone(int N, int M);          /* TriangleDimerTiler.cpp:697-793 */
rt_vec* rt_bibone_test_butterfly(int type);              /* 795-844 */
/* Hexagon with sides a, b, c, a, b, c and a dimer cover, in the IceCreamCone
 * tiling format (TriangleDimerTiler.cpp:697-793).  No reference counterpart.
 * NULL when the hexagon has an odd number of vertices (ab+bc+ca+a+b+c+1). */
rt_vec* rt_bibone_hexagon(int a, int b, int c);
rt_vec* rt_bibone_triangle_flips(const int32_t* t, size_t n);    /* 846-873 */
rt_vec* rt_bibone_butterfly_flips(const int32_t* t, size_t n);   /* 875-921 */
rt_vec* rt_bibone_tiling_to_vertices(const int32_t* t, size_t n); /* 923-949 */
rt_vec* rt_bibone_tiling_to_domain(const int32_t* t, size_t n);  /* 951-982 */
rt_vec* rt_bibone_dimer_indicator(const int32_t* t, size_t n, int c); /* 35-95 */
/* Edge form of a bibone tiling (N x N, n = N*N): out[(N/2)*(N/2)] gets one
 * byte per plane cell (i, j), bit 0 = dimer on (i,j)-(i,j+1), bit 1 on
 * (i,j)-(i+1,j), bit 2 on (i,j+1)-(i+1,j) (vertex coordinates).  Returns
 * RT_EINVAL when rt_bibone_walk would not run this start on its edge form
 * (inconsistent cells, not a matching, a covered vertex within two cells
 * of the border); such starts run on the eight planes.  No reference
 * counterpart: the reference keeps only the eight planes. */
int rt_bibone_edge_form(const int32_t* t, size_t n, uint8_t* out);

/* SVG writers (drop-in drawing helpers; files are the reference's byte for
 * byte).  Return RT_OK or RT_EIO.  A NULL domain in a *_dimer_svg call means
 * TilingToDomain(t), as in the reference's one-argument overloads. */
int rt_domino_domain_svg(const int32_t* d, size_t nd, const char* path);   /* DominoTiler.cpp:640-668 */
int rt_domino_tiling_svg(const int32_t* t, size_t nt, const char* path);   /* 745-774 */
int rt_domino_maya_svg(const int32_t* t, size_t nt, const char* path);     /* 776-816 */
int rt_domino_dimer_svg(const int32_t* t, size_t nt, const int32_t* d, size_t nd,
                        const char* path);                                 /* 670-743 */
int rt_lozenge_domain_svg(const int32_t* d, size_t nd, const char* path);  /* LozengeTiler.cpp:669-693 */
int rt_lozenge_tiling_svg(const int32_t* t, size_t nt, const char* path);  /* 643-667 */
int rt_lozenge_dimer_svg(const int32_t* t, size_t nt, const int32_t* d, size_t nd,
                         const char* path);                                /* 558-641 */
int rt_recttri_tiling_svg(const int32_t* t, size_t nt, const char* path);  /* RectTriangleTiler.cpp:646-776 */
int rt_bibone_domain_svg(const int32_t* d, size_t nd, const char* path);   /* TriangleDimerTiler.cpp:984-1006 */
int rt_bibone_tiling_svg(const int32_t* t, size_t nt, const char* path);   /* 1013-1044 */
int rt_bibone_dimer_svg(const int32_t* t, size_t nt, const int32_t* d, size_t nd,
                        const char* path);                                 /* 1008-1011, 1046-1117 */

#ifdef __cplusplus
}
#endif
#endif

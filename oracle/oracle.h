/*
 * oracle.h -- CPU restatement of the reference sampler, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library. The product
 * (randomtilingspublic_amd/) never links, calls or falls back to it.
 *
 * Every function restates the reference literally, in the reference's own
 * order: one Rotate pass then one Update pass per step, both colour arrays
 * materialised after every step, exactly as the OpenCL NDRanges run
 * (file:line citations on each function in oracle.c).  The device path uses
 * a different (lazy-update, fused) formulation, so agreement between the two
 * is a real check and not a restatement of the same code.
 *
 * Pinning (see tests/): TinyMT is cross-checked against the independent
 * host formulation of the generator (reference src/TinyMT/tinymt32.h) and
 * its invariants; mt19937 / minstd_rand0 against the C++ standard's
 * 10000th-output known answers and against libstdc++ itself; the step
 * kernels against invariants (valid tiling, domain preserved, monotone
 * coupling) and exact enumeration on tiny domains.
 *
 * Status codes: 0 ok, negative on error (see ORC_E*).
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_EINVAL = -1,   /* bad argument (size, N not divisible, ...) */
  ORC_EPARAMS = -2,  /* parameter table too small / malformed */
  ORC_EIO = -3,      /* file could not be opened */
  ORC_ENOCOUPLE = -4 /* CFTP did not coalesce within max_rounds */
};

/* ---- TinyMT32 (reference src/TinyMT/tinymt32.clh) ---- */
typedef struct {
  uint32_t s0, s1, s2, s3, mat1, mat2, tmat;
} orc_tinymt;

void orc_tinymt_init(orc_tinymt* t, uint32_t seed);
uint32_t orc_tinymt_uint32(orc_tinymt* t);
float orc_tinymt_single01(orc_tinymt* t);
/* Second formulation, following the host header tinymt32.h
 * (branch-free masks instead of branches). */
float orc_tinymt_h_float01(orc_tinymt* t);

/* ---- std::mt19937 and libstdc++ default_random_engine ---- */
typedef struct {
  uint32_t mt[624];
  int idx;
} orc_mt19937;
void orc_mt_seed(orc_mt19937* g, uint32_t s);
uint32_t orc_mt_next(orc_mt19937* g);

typedef struct {
  uint64_t x;
} orc_minstd;
void orc_minstd_seed(orc_minstd* g, unsigned long s);
uint32_t orc_minstd_next(orc_minstd* g);
/* std::uniform_int_distribution<int>(lo, hi) over minstd_rand0, libstdc++ */
int orc_uniform_int(orc_minstd* g, int lo, int hi);

/* ---- parameter table (reference src/TinyMT/file_reader.cpp) ----
 * out3 receives count triples (mat1, mat2, tmat). */
int orc_parse_params(const char* path, size_t count, uint32_t* out3);

/* ---- domino ---- */
enum {
  ORC_DOMINO_TOGGLE08 = 0,     /* dominokernel.cl */
  ORC_DOMINO_HEATBATH = 1,     /* dominokernelCFTP.cl */
  ORC_DOMINO_METRO2P = 2,      /* dominokerneltwoperiodic.cl */
  ORC_DOMINO_HEATBATH2P = 3    /* dominokernelCFTPtwoperiodic.cl */
};

/* t: N*N int32, in place.  params3: nparams triples. a, b: weights. */
int orc_domino_walk(int32_t* t, int N, long steps, long seed,
                    const uint32_t* params3, size_t nparams, int rule,
                    double a, double b);
int orc_domino_walk_segments(int32_t* t, int N, const long* steps,
                             const long* seeds, int nseg,
                             const uint32_t* params3, size_t nparams,
                             int rule, double a, double b);

/* CFTP schedules */
enum { ORC_SCHED_REFERENCE = 0, ORC_SCHED_DOUBLING = 1 };

typedef struct {
  int rounds;          /* rounds run (1 = the first segment coalesced) */
  long total_steps;    /* sum of the final segment list */
  long steps_run;      /* steps run over all rounds, per chain */
} orc_cftp_stats;

int orc_domino_cftp(const int32_t* tmax, const int32_t* tmin, int N,
                    uint32_t master_seed, int sched, long t0, int max_rounds,
                    const uint32_t* params3, size_t nparams, int rule,
                    double a, double b, int32_t* out, orc_cftp_stats* st);

/* ---- lozenge (lozengekernel.cl, LozengeTiler.cpp) ---- */
int orc_lozenge_walk(int32_t* t, int N, long steps, long seed,
                     const uint32_t* params3, size_t nparams);
int orc_lozenge_walk_segments(int32_t* t, int N, const long* steps,
                              const long* seeds, int nseg,
                              const uint32_t* params3, size_t nparams);
int orc_lozenge_cftp(const int32_t* tmax, const int32_t* tmin, int N,
                     uint32_t master_seed, int sched, long t0,
                     int max_rounds, const uint32_t* params3,
                     size_t nparams, int32_t* out, orc_cftp_stats* st);

/* ---- rectangle-triangle (recttrianglekernel.cl, RectTriangleTiler.cpp) */
int orc_recttri_walk(int32_t* t, int N, long steps, long seed,
                     const uint32_t* params3, size_t nparams,
                     const float* weights16);

/* ---- bibone (triangledimerkernel.cl, TriangleDimerTiler.cpp) ---- */
/* Derived flip-state arrays of an N x N tiling (TriangleDimerTiler.cpp:
 * 846-921): tri is (N/2) x N, bf is N x N. */
void orc_bibone_triangle_flips(const int32_t* t, int N, int32_t* tri);
void orc_bibone_butterfly_flips(const int32_t* t, int N, int32_t* bf);
/* Move schedule: five ints per step (type, r1, r2, p1, p2; -1 = none). */
void orc_bibone_schedule(long seed, long steps, int32_t* out5);
/* Walk (TriangleDimerTiler.cpp:101-695), in place.  With `planes` non-NULL
 * the eight final M x M planes (eH, eL, eR, tU, tD, bH, bL, bR; M = N/2)
 * are copied out as well. */
int orc_bibone_walk(int32_t* t, int N, long steps, long seed,
                    const uint32_t* params3, size_t nparams, int32_t* planes);

int orc_recttri_walk_segments(int32_t* t, int N, const long* steps, const long* seeds, int nseg,
                              const uint32_t* params3, size_t nparams, const float* weights16);
int orc_recttri_cftp(const int32_t* tmax, const int32_t* tmin, int N, uint32_t master_seed,
                     int sched, long t0, int max_rounds, const uint32_t* params3,
                     size_t nparams, const float* weights16, int32_t* out, orc_cftp_stats* st);

/* Number of OpenMP threads the oracle uses (1 = serial). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif

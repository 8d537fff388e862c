/*
 * oracle.c -- CPU restatement of the reference sampler (TEST INFRASTRUCTURE).
 *
 * This file is the checker, never the product: only tests/, the smoke check
 * and the cpu_baseline leg of bench.py load it.  Each function follows the
 * reference file:line given above it, literally and in the reference's own
 * order (Rotate pass, then Update pass, every step; both colour arrays
 * materialised).  OpenMP parallelises the work items of one NDRange, the way
 * a CPU OpenCL runtime would run it; results do not depend on the thread
 * count because every work item writes only its own cell.
 */
#include "oracle.h"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ */
/* TinyMT32 -- reference src/TinyMT/tinymt32.clh                       */
/* ------------------------------------------------------------------ */

/* tinymt32.clh:39-55 */
static inline void tm_next_state(orc_tinymt* t) {
  uint32_t x = (t->s0 & 0x7fffffffu) ^ t->s1 ^ t->s2;
  uint32_t y = t->s3;
  x ^= x << 1;
  y ^= (y >> 1) ^ x;
  t->s0 = t->s1;
  t->s1 = t->s2;
  t->s2 = x ^ (y << 10);
  t->s3 = y;
  if (y & 1u) {
    t->s1 ^= t->mat1;
    t->s2 ^= t->mat2;
  }
}

/* tinymt32.clh:62-73 */
static inline uint32_t tm_temper(const orc_tinymt* t) {
  uint32_t t0 = t->s3;
  uint32_t t1 = t->s0 + (t->s2 >> 8);
  t0 ^= t1;
  if (t1 & 1u) t0 ^= t->tmat;
  return t0;
}

/* tinymt32.clh:81-93 */
static inline float tm_temper12(const orc_tinymt* t) {
  uint32_t u = (tm_temper(t) >> 9) | 0x3f800000u;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* tinymt32.clh:100-105 */
uint32_t orc_tinymt_uint32(orc_tinymt* t) {
  tm_next_state(t);
  return tm_temper(t);
}

/* tinymt32.clh:112-129 */
float orc_tinymt_single01(orc_tinymt* t) {
  tm_next_state(t);
  return tm_temper12(t) - 1.0f;
}

/* tinymt32.clh:163-174 and 182-203 */
void orc_tinymt_init(orc_tinymt* t, uint32_t seed) {
  uint32_t st[4];
  st[0] = seed;
  st[1] = t->mat1;
  st[2] = t->mat2;
  st[3] = t->tmat;
  for (int i = 1; i < 8; i++)
    st[i & 3] ^= (uint32_t)i +
                 1812433253u * (st[(i - 1) & 3] ^ (st[(i - 1) & 3] >> 30));
  t->s0 = st[0];
  t->s1 = st[1];
  t->s2 = st[2];
  t->s3 = st[3];
  if ((t->s0 & 0x7fffffffu) == 0 && t->s1 == 0 && t->s2 == 0 && t->s3 == 0) {
    t->s0 = 'T';
    t->s1 = 'I';
    t->s2 = 'N';
    t->s3 = 'Y';
  }
  for (int i = 0; i < 8; i++) orc_tinymt_uint32(t);
}

/* Second formulation after the host header src/TinyMT/tinymt32.h:72-136,
 * 205-208: the conditional xors are written as sign-extended masks. */
float orc_tinymt_h_float01(orc_tinymt* t) {
  uint32_t st[4] = {t->s0, t->s1, t->s2, t->s3};
  uint32_t y = st[3];
  uint32_t x = (st[0] & 0x7fffffffu) ^ st[1] ^ st[2];
  x ^= (x << 1);
  y ^= (y >> 1) ^ x;
  st[0] = st[1];
  st[1] = st[2];
  st[2] = x ^ (y << 10);
  st[3] = y;
  st[1] ^= (uint32_t)(-(int32_t)(y & 1)) & t->mat1;
  st[2] ^= (uint32_t)(-(int32_t)(y & 1)) & t->mat2;
  t->s0 = st[0];
  t->s1 = st[1];
  t->s2 = st[2];
  t->s3 = st[3];
  uint32_t t0 = st[3];
  uint32_t t1 = st[0] + (st[2] >> 8);
  t0 ^= t1;
  union {
    uint32_t u;
    float f;
  } conv;
  conv.u = ((t0 ^ ((uint32_t)(-(int32_t)(t1 & 1)) & t->tmat)) >> 9) |
           0x3f800000u;
  return conv.f - 1.0f;
}

/* ------------------------------------------------------------------ */
/* std::mt19937 (the algorithm the C++ standard specifies)             */
/* ------------------------------------------------------------------ */
void orc_mt_seed(orc_mt19937* g, uint32_t s) {
  g->mt[0] = s;
  for (int i = 1; i < 624; i++)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) +
               (uint32_t)i;
  g->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937* g) {
  if (g->idx >= 624) {
    for (int k = 0; k < 624; k++) {
      uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->mt[(k + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->mt[k] = v;
    }
    g->idx = 0;
  }
  uint32_t z = g->mt[g->idx++];
  z ^= (z >> 11);
  z ^= (z << 7) & 0x9d2c5680u;
  z ^= (z << 15) & 0xefc60000u;
  z ^= (z >> 18);
  return z;
}

/* ------------------------------------------------------------------ */
/* libstdc++ default_random_engine = minstd_rand0, and                 */
/* uniform_int_distribution<int> (bits/uniform_int_dist.h, GCC 11):    */
/* the generator's range 2^31-3 is neither 2^32-1 nor 2^64-1, so the   */
/* "scaling" rejection path is taken.                                  */
/* ------------------------------------------------------------------ */
void orc_minstd_seed(orc_minstd* g, unsigned long s) {
  uint64_t x = (uint64_t)(s % 2147483647ul);
  g->x = x == 0 ? 1 : x;
}

uint32_t orc_minstd_next(orc_minstd* g) {
  g->x = (g->x * 16807u) % 2147483647u;
  return (uint32_t)g->x;
}

int orc_uniform_int(orc_minstd* g, int lo, int hi) {
  const uint64_t urngmin = 1, urngmax = 2147483646;
  const uint64_t urngrange = urngmax - urngmin;
  const uint64_t urange = (uint64_t)(uint32_t)hi - (uint64_t)(uint32_t)lo;
  /* only the downscaling branch is reachable for the ranges used here */
  const uint64_t uerange = urange + 1;
  const uint64_t scaling = urngrange / uerange;
  const uint64_t past = uerange * scaling;
  uint64_t ret;
  do {
    ret = (uint64_t)orc_minstd_next(g) - urngmin;
  } while (ret >= past);
  ret /= scaling;
  return (int)(ret + (uint64_t)lo);
}

/* ------------------------------------------------------------------ */
/* Parameter file -- reference src/TinyMT/file_reader.cpp:51-70,       */
/* 114-166: lines starting with '#' are skipped; mat1/mat2/tmat are    */
/* the hex fields after the 3rd/4th/5th comma; a blank line, a missing */
/* comma or end of file is a format error.                             */
/* ------------------------------------------------------------------ */
static const char* after_comma(const char* s, int count) {
  for (int i = 0; i < 500; i++) {
    if (s[i] == '\0') return NULL;
    if (s[i] == ',') count--;
    if (count == 0) return s + i + 1;
  }
  return NULL;
}

int orc_parse_params(const char* path, size_t count, uint32_t* out3) {
  FILE* f = fopen(path, "r");
  if (!f) return ORC_EIO;
  char buf[512];
  for (size_t k = 0; k < count; k++) {
    for (;;) {
      if (!fgets(buf, 500, f)) {
        fclose(f);
        return ORC_EPARAMS;
      }
      if (buf[0] != '#') break;
    }
    if (buf[0] == '\n' || buf[0] == '\0' || buf[0] == '\r') {
      fclose(f);
      return ORC_EPARAMS;
    }
    const char* p3 = after_comma(buf, 3);
    const char* p4 = after_comma(buf, 4);
    const char* p5 = after_comma(buf, 5);
    if (!p3 || !p4 || !p5) {
      fclose(f);
      return ORC_EPARAMS;
    }
    errno = 0;
    unsigned long m1 = strtoul(p3, NULL, 16);
    unsigned long m2 = strtoul(p4, NULL, 16);
    unsigned long tm = strtoul(p5, NULL, 16);
    if (errno != 0) {
      fclose(f);
      return ORC_EPARAMS;
    }
    out3[3 * k + 0] = (uint32_t)m1;
    out3[3 * k + 1] = (uint32_t)m2;
    out3[3 * k + 2] = (uint32_t)tm;
  }
  fclose(f);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Shared: InitTinyMT kernel (every .cl file, e.g. dominokernel.cl:33-40)*/
/* state[id] = tinymt32_init(params[id], seed + id), id < nstreams     */
/* ------------------------------------------------------------------ */
static orc_tinymt* init_streams(const uint32_t* params3, size_t nstreams,
                                uint32_t seed) {
  orc_tinymt* st = (orc_tinymt*)malloc(sizeof(orc_tinymt) * (nstreams ? nstreams : 1));
  if (!st) return NULL;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (long id = 0; id < (long)nstreams; id++) {
    st[id].mat1 = params3[3 * id + 0];
    st[id].mat2 = params3[3 * id + 1];
    st[id].tmat = params3[3 * id + 2];
    orc_tinymt_init(&st[id], seed + (uint32_t)id);
  }
  return st;
}

static inline int32_t sx8(uint8_t v) { return (int32_t)(int8_t)v; }

/* ------------------------------------------------------------------ */
/* Domino                                                              */
/* ------------------------------------------------------------------ */

/* RotateTiles: dominokernel.cl:4-18 (toggle08), dominokernelCFTP.cl:51-71
 * (heat-bath), dominokerneltwoperiodic.cl:7-35 (Metropolis),
 * dominokernelCFTPtwoperiodic.cl:15-44 (heat-bath Metropolis).
 * NDRange (N-2, N/2-2); i = gid0+1, j = gid1+1;
 * TinyMT id = gid1 + (N/2-2)*gid0 (tinymt.clh:23-30). */
static void domino_rotate(uint8_t* tl, orc_tinymt* st, int N, int t,
                          int rule, double a, double b) {
  const int H = N / 2;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 1; i <= N - 2; i++) {
    for (int j = 1; j <= H - 2; j++) {
      orc_tinymt* s = &st[(size_t)(j - 1) + (size_t)(H - 2) * (size_t)(i - 1)];
      float rd = orc_tinymt_single01(s);
      uint8_t* c = &tl[i * H + j];
      if (rule == ORC_DOMINO_TOGGLE08) {
        if ((double)rd < .8) {
          if (*c == 3) *c = 12;
          else if (*c == 12) *c = 3;
        }
      } else if (rule == ORC_DOMINO_HEATBATH) {
        if ((double)rd < 0.5) {
          if (*c == 3) *c = 12;
        } else {
          if (*c == 12) *c = 3;
        }
      } else if (rule == ORC_DOMINO_METRO2P) {
        float prob_propose = 0.8f;
        if (rd < prob_propose) {
          float threshold = rd / prob_propose;
          if (*c == 3) {
            float wInit = (float)(b * b * ((i + t + 1) % 2) + a * a * ((i + t) % 2));
            float wFin = (float)(a * a * (i % 2) + b * b * ((i + 1) % 2));
            if (wFin / wInit > threshold) *c = 12;
          } else if (*c == 12) {
            float wInit = (float)(a * a * (i % 2) + b * b * ((i + 1) % 2));
            float wFin = (float)(b * b * ((i + t + 1) % 2) + a * a * ((i + t) % 2));
            if (wFin / wInit > threshold) *c = 3;
          }
        }
      } else { /* ORC_DOMINO_HEATBATH2P */
        if ((double)rd < 0.5) {
          float threshold = (float)((double)rd / 0.5);
          if (*c == 3) {
            float wInit = (float)(b * b * ((i + t + 1) % 2) + a * a * ((i + t) % 2));
            float wFin = (float)(a * a * (i % 2) + b * b * ((i + 1) % 2));
            if (wFin / wInit > threshold) *c = 12;
          }
        } else {
          float threshold = (float)(((double)rd - 0.5) / 0.5);
          if (*c == 12) {
            float wInit = (float)(a * a * (i % 2) + b * b * ((i + 1) % 2));
            float wFin = (float)(b * b * ((i + t + 1) % 2) + a * a * ((i + t) % 2));
            if (wFin / wInit > threshold) *c = 3;
          }
        }
      }
    }
  }
}

/* UpdateTiles: dominokernel.cl:22-31 (identical in all four .cl files). */
static void domino_update(uint8_t* tl, const uint8_t* ref, int N, int t) {
  const int H = N / 2;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 1; i <= N - 2; i++) {
    for (int j = 1; j <= H - 2; j++) {
      int v = (sx8(ref[(i - 1) * H + j]) & 2) / 2 +
              2 * (sx8(ref[(i + 1) * H + j]) & 1) +
              (sx8(ref[i * H + j - (i + t + 1) % 2]) & 8) / 2 +
              2 * (sx8(ref[i * H + j + (i + t) % 2]) & 4);
      tl[i * H + j] = (uint8_t)v;
    }
  }
}

/* DominoTiler.cpp:24-33 (split) and 109-114 (merge) */
static void domino_split(const int32_t* t, int N, uint8_t* B, uint8_t* W) {
  const int H = N / 2;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      if ((i + j) % 2 == 0) B[i * H + j / 2] = (uint8_t)t[i * N + j];
      else W[i * H + j / 2] = (uint8_t)t[i * N + j];
    }
}

static void domino_merge(int32_t* t, int N, const uint8_t* B, const uint8_t* W) {
  const int H = N / 2;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      t[i * N + j] = (i + j) % 2 == 0 ? sx8(B[i * H + j / 2]) : sx8(W[i * H + j / 2]);
}

static int domino_check(int N, size_t nparams, int rule) {
  if (N < 6 || N % 2 != 0) return ORC_EINVAL;
  if (rule < 0 || rule > 3) return ORC_EINVAL;
  if (nparams < (size_t)N * (size_t)N / 2) return ORC_EPARAMS;
  return ORC_OK;
}

/* One step of DominoTiler::Walk (DominoTiler.cpp:70-79):
 * r == 1 -> black (t = 1), else white (t = 0). */
static void domino_step(uint8_t* B, uint8_t* W, orc_tinymt* st, int N, int r,
                        int rule, double a, double b) {
  if (r == 1) {
    domino_rotate(B, st, N, 1, rule, a, b);
    domino_update(W, B, N, 1);
  } else {
    domino_rotate(W, st, N, 0, rule, a, b);
    domino_update(B, W, N, 0);
  }
}

/* DominoTiler::Walk(tiling&, long steps, long seed): DominoTiler.cpp:18-122.
 * TinyMT is seeded with the FIRST mt19937(seed) output (51-53); each step's
 * colour is the next output % 2 (70). */
int orc_domino_walk(int32_t* t, int N, long steps, long seed,
                    const uint32_t* params3, size_t nparams, int rule,
                    double a, double b) {
  int rc = domino_check(N, nparams, rule);
  if (rc) return rc;
  const size_t half = (size_t)N * (size_t)(N / 2);
  uint8_t* B = (uint8_t*)calloc(half, 1);
  uint8_t* W = (uint8_t*)calloc(half, 1);
  orc_mt19937 g;
  orc_mt_seed(&g, (uint32_t)seed);
  uint32_t base = orc_mt_next(&g);
  orc_tinymt* st = init_streams(params3, (size_t)N * N / 2, base);
  if (!B || !W || !st) {
    free(B); free(W); free(st);
    return ORC_EINVAL;
  }
  domino_split(t, N, B, W);
  for (long s = 0; s < steps; s++) {
    int r = (int)(orc_mt_next(&g) % 2);
    domino_step(B, W, st, N, r, rule, a, b);
  }
  domino_merge(t, N, B, W);
  free(B); free(W); free(st);
  return ORC_OK;
}

/* DominoTiler::Walk(tiling&, vector<long> steps, vector<long> seeds):
 * DominoTiler.cpp:125-231.  Per segment k: InitTinyMT(seeds[k]) and
 * mt19937(seeds[k]) for the colours -- no extra first draw (157-159). */
int orc_domino_walk_segments(int32_t* t, int N, const long* steps,
                             const long* seeds, int nseg,
                             const uint32_t* params3, size_t nparams,
                             int rule, double a, double b) {
  int rc = domino_check(N, nparams, rule);
  if (rc) return rc;
  const size_t half = (size_t)N * (size_t)(N / 2);
  uint8_t* B = (uint8_t*)calloc(half, 1);
  uint8_t* W = (uint8_t*)calloc(half, 1);
  if (!B || !W) {
    free(B); free(W);
    return ORC_EINVAL;
  }
  domino_split(t, N, B, W);
  for (int k = 0; k < nseg; k++) {
    orc_mt19937 g;
    orc_mt_seed(&g, (uint32_t)seeds[k]);
    orc_tinymt* st = init_streams(params3, (size_t)N * N / 2, (uint32_t)seeds[k]);
    for (long s = 0; s < steps[k]; s++) {
      int r = (int)(orc_mt_next(&g) % 2);
      domino_step(B, W, st, N, r, rule, a, b);
    }
    free(st);
  }
  domino_merge(t, N, B, W);
  free(B); free(W);
  return ORC_OK;
}

/* Segment-list growth between CFTP rounds.  Reference schedule:
 * Examples/AztecDiamondCftpEx.cpp:76 prepends steps.back()*2 (always 2*t0
 * because back() is the first segment).  Doubling: prepend the current
 * total, so the look-back doubles each round. */
static long next_segment(const long* steps, int n, int sched) {
  if (sched == ORC_SCHED_REFERENCE) return steps[n - 1] * 2;
  long tot = 0;
  for (int k = 0; k < n; k++) tot += steps[k];
  return tot;
}

typedef int (*walkseg_fn)(int32_t*, int, const long*, const long*, int,
                          const void*);

/* CFTP driver: Examples/AztecDiamondCftpEx.cpp:51-77.  Master mt19937(S)
 * supplies the segment seeds; each round restarts top = max, bottom = min,
 * walks both through the whole segment list and compares the tilings. */
static int cftp_driver(const int32_t* tmax, const int32_t* tmin, int N,
                       uint32_t master_seed, int sched, long t0,
                       int max_rounds, walkseg_fn walk, const void* ctx,
                       int32_t* out, orc_cftp_stats* stats) {
  if (t0 <= 0 || max_rounds <= 0) return ORC_EINVAL;
  const size_t n2 = (size_t)N * N;
  long* steps = (long*)malloc(sizeof(long) * (size_t)max_rounds);
  long* seeds = (long*)malloc(sizeof(long) * (size_t)max_rounds);
  int32_t* top = (int32_t*)malloc(sizeof(int32_t) * n2);
  int32_t* bot = (int32_t*)malloc(sizeof(int32_t) * n2);
  orc_mt19937 m;
  orc_mt_seed(&m, master_seed);
  int n = 1;
  steps[0] = t0;
  seeds[0] = (long)orc_mt_next(&m);
  long run = 0;
  int rc = ORC_ENOCOUPLE;
  for (int round = 1; round <= max_rounds; round++) {
    memcpy(top, tmax, sizeof(int32_t) * n2);
    memcpy(bot, tmin, sizeof(int32_t) * n2);
    rc = walk(bot, N, steps, seeds, n, ctx);
    if (rc) break;
    rc = walk(top, N, steps, seeds, n, ctx);
    if (rc) break;
    long tot = 0;
    for (int k = 0; k < n; k++) tot += steps[k];
    run += tot;
    int coupled = memcmp(top, bot, sizeof(int32_t) * n2) == 0;
    if (coupled) {
      memcpy(out, top, sizeof(int32_t) * n2);
      if (stats) {
        stats->rounds = round;
        stats->total_steps = tot;
        stats->steps_run = run;
      }
      rc = ORC_OK;
      break;
    }
    rc = ORC_ENOCOUPLE;
    if (round == max_rounds) break;
    long ns = next_segment(steps, n, sched);
    memmove(steps + 1, steps, sizeof(long) * (size_t)n);
    memmove(seeds + 1, seeds, sizeof(long) * (size_t)n);
    steps[0] = ns;
    seeds[0] = (long)orc_mt_next(&m);
    n++;
  }
  if (rc == ORC_ENOCOUPLE && stats) {
    stats->rounds = max_rounds;
    stats->total_steps = 0;
    stats->steps_run = run;
  }
  free(steps); free(seeds); free(top); free(bot);
  return rc;
}

typedef struct {
  const uint32_t* params3;
  size_t nparams;
  int rule;
  double a, b;
} domino_ctx;

static int domino_walkseg_cb(int32_t* t, int N, const long* steps,
                             const long* seeds, int n, const void* c) {
  const domino_ctx* d = (const domino_ctx*)c;
  return orc_domino_walk_segments(t, N, steps, seeds, n, d->params3,
                                  d->nparams, d->rule, d->a, d->b);
}

int orc_domino_cftp(const int32_t* tmax, const int32_t* tmin, int N,
                    uint32_t master_seed, int sched, long t0, int max_rounds,
                    const uint32_t* params3, size_t nparams, int rule,
                    double a, double b, int32_t* out, orc_cftp_stats* st) {
  int rc = domino_check(N, nparams, rule);
  if (rc) return rc;
  domino_ctx c = {params3, nparams, rule, a, b};
  return cftp_driver(tmax, tmin, N, master_seed, sched, t0, max_rounds,
                     domino_walkseg_cb, &c, out, st);
}

/* ------------------------------------------------------------------ */
/* Lozenge                                                             */
/* ------------------------------------------------------------------ */

/* RotateTiles: lozengekernel.cl:3-34.  NDRange (N, N/3), id = j + (N/3)*i. */
static void lozenge_rotate(uint8_t* tl, orc_tinymt* st, int N) {
  const int M = N / 3;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < M; j++) {
      float rd = orc_tinymt_single01(&st[(size_t)j + (size_t)M * i]);
      uint8_t* c = &tl[i * M + j];
      if ((double)rd < .4) {
        if (*c == 25) *c = 38;
      } else if ((double)rd < .8) {
        if (*c == 38) *c = 25;
      }
    }
  }
}

/* UpdateTiles: lozengekernel.cl:36-54. */
static void lozenge_update(const uint8_t* t1, uint8_t* t2, uint8_t* t3, int N,
                           int t) {
  const int M = N / 3;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 1; i < N - 1; i++) {
    for (int j = 1; j < M - 1; j++) {
      int p2 = i % 3 == (2 * t + 1) % 3;
      int p4 = i % 3 == (2 * t + 2) % 3;
      int p8 = i % 3 == (2 * t) % 3;
      int p16 = i % 3 == (2 * t + 1) % 3;
      int v2 = sx8(t2[i * M + j]);
      v2 &= ~((v2 & 2) + (v2 & 4) + (v2 & 32));
      v2 += (sx8(t1[(i - 1) * M + j + p2]) & 16) / 8 +
            (sx8(t1[i * M + j - p4]) & 8) / 2 +
            (sx8(t1[(i + 1) * M + j]) & 1) * 32;
      t2[i * M + j] = (uint8_t)v2;
      int v3 = sx8(t3[i * M + j]);
      v3 &= ~((v3 & 1) + (v3 & 8) + (v3 & 16));
      v3 += (sx8(t1[(i - 1) * M + j]) & 32) / 32 +
            (sx8(t1[i * M + j + p8]) & 4) * 2 +
            (sx8(t1[(i + 1) * M + j - p16]) & 2) * 8;
      t3[i * M + j] = (uint8_t)v3;
    }
  }
}

/* LozengeTiler.cpp:38-48 / 140-150: colour ((j-i)%3+3)%3 -> R, B, G. */
static void lozenge_split(const int32_t* t, int N, uint8_t* C[3]) {
  const int M = N / 3;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      C[((j - i) % 3 + 3) % 3][i * M + j / 3] = (uint8_t)t[i * N + j];
}

static void lozenge_merge(int32_t* t, int N, uint8_t* C[3]) {
  const int M = N / 3;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      t[i * N + j] = sx8(C[((j - i) % 3 + 3) % 3][i * M + j / 3]);
}

/* LozengeTiler.cpp:95-104: r picks R/B/G; Update(t1 = rotated, next,
 * next-next, t = r). */
static void lozenge_step(uint8_t* C[3], orc_tinymt* st, int N, int r) {
  lozenge_rotate(C[r], st, N);
  lozenge_update(C[r], C[(r + 1) % 3], C[(r + 2) % 3], N, r);
}

static int lozenge_check(int N, size_t nparams) {
  if (N < 9 || N % 3 != 0) return ORC_EINVAL;
  if (nparams < (size_t)N * (size_t)(N / 3)) return ORC_EPARAMS;
  return ORC_OK;
}

/* LozengeTiler::Walk: LozengeTiler.cpp:27-159.  TinyMT seeded with `seed`
 * itself (69); colours from default_random_engine(seed) +
 * uniform_int_distribution(0,2) (86-91). */
int orc_lozenge_walk(int32_t* t, int N, long steps, long seed,
                     const uint32_t* params3, size_t nparams) {
  return orc_lozenge_walk_segments(t, N, &steps, &seed, 1, params3, nparams);
}

/* Segmented lozenge walk: the reference has none; this follows the domino
 * segmented Walk (DominoTiler.cpp:155-209) with the lozenge per-segment
 * seeding of LozengeTiler.cpp:69,86-88 (TinyMT and the colour engine both
 * seeded with seeds[k]).  A single segment is exactly LozengeTiler::Walk. */
int orc_lozenge_walk_segments(int32_t* t, int N, const long* steps,
                              const long* seeds, int nseg,
                              const uint32_t* params3, size_t nparams) {
  int rc = lozenge_check(N, nparams);
  if (rc) return rc;
  const size_t sz = (size_t)N * (size_t)(N / 3);
  uint8_t* C[3];
  for (int c = 0; c < 3; c++) C[c] = (uint8_t*)calloc(sz, 1);
  lozenge_split(t, N, C);
  for (int k = 0; k < nseg; k++) {
    orc_tinymt* st = init_streams(params3, sz, (uint32_t)seeds[k]);
    orc_minstd g;
    orc_minstd_seed(&g, (unsigned long)seeds[k]);
    for (long s = 0; s < steps[k]; s++) {
      int r = orc_uniform_int(&g, 0, 2);
      lozenge_step(C, st, N, r);
    }
    free(st);
  }
  lozenge_merge(t, N, C);
  for (int c = 0; c < 3; c++) free(C[c]);
  return ORC_OK;
}

typedef struct {
  const uint32_t* params3;
  size_t nparams;
} lozenge_ctx;

static int lozenge_walkseg_cb(int32_t* t, int N, const long* steps,
                              const long* seeds, int n, const void* c) {
  const lozenge_ctx* d = (const lozenge_ctx*)c;
  return orc_lozenge_walk_segments(t, N, steps, seeds, n, d->params3,
                                   d->nparams);
}

int orc_lozenge_cftp(const int32_t* tmax, const int32_t* tmin, int N,
                     uint32_t master_seed, int sched, long t0,
                     int max_rounds, const uint32_t* params3,
                     size_t nparams, int32_t* out, orc_cftp_stats* st) {
  int rc = lozenge_check(N, nparams);
  if (rc) return rc;
  lozenge_ctx c = {params3, nparams};
  return cftp_driver(tmax, tmin, N, master_seed, sched, t0, max_rounds,
                     lozenge_walkseg_cb, &c, out, st);
}

/* ------------------------------------------------------------------ */
/* Rectangle-triangle                                                  */
/* ------------------------------------------------------------------ */

/* getWeightRatio: recttrianglekernel.cl:157-160 (float products, left to
 * right, then one division). */
static float rt_weight_ratio(const float* w, int a, int b) {
  float nb = w[(b & 0xf00000) / 0x100000] * w[(b & 0x0f0000) / 0x010000] *
             w[(b & 0x00f000) / 0x001000] * w[(b & 0x000f00) / 0x000100] *
             w[(b & 0x0000f0) / 0x000010] * w[(b & 0x00000f) / 0x000001];
  float na = w[(a & 0xf00000) / 0x100000] * w[(a & 0x0f0000) / 0x010000] *
             w[(a & 0x00f000) / 0x001000] * w[(a & 0x000f00) / 0x000100] *
             w[(a & 0x0000f0) / 0x000010] * w[(a & 0x00000f) / 0x000001];
  return nb / na;
}

#define SEL(a, b, c) ((c) ? (b) : (a)) /* OpenCL scalar select(a,b,c) */

/* flipTiles: recttrianglekernel.cl:25-118.  NDRange (N, N/3). */
static void recttri_flip(int32_t* tl, orc_tinymt* st, int N, const float* w) {
  const int M = N / 3;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < M; j++) {
      float rd = orc_tinymt_single01(&st[(size_t)j + (size_t)M * i]);
      float bs = 0.9f;
      if (rd < bs) {
        rd /= bs;
        int hexType = tl[i * M + j];
        int hex1, hex2, newType = hexType;
        float sc1 = SEL(0.75f, 0.5f, rd < 0.75f);
        float sc2 = SEL(sc1, 0.25f, rd < 0.5f);
        float sc = SEL(sc2, 0.0f, rd < 0.25f);
        if (hexType == 0x122133) {
          hex1 = SEL(0x331221, 0x8ebbe8, (double)rd < 0.75);
          hex2 = SEL(hex1, 0xfc99cf, (double)rd < 0.5);
          newType = SEL(hex2, 0xa7dd7a, (double)rd < 0.25);
        } else if (hexType == 0x331221) {
          hex1 = SEL(0x122133, 0x8ebbe8, (double)rd < 0.75);
          hex2 = SEL(hex1, 0xfc99cf, (double)rd < 0.5);
          newType = SEL(hex2, 0xa7dd7a, (double)rd < 0.25);
        } else if (hexType == 0x8ebbe8) {
          hex1 = SEL(0x331221, 0x122133, (double)rd < 0.75);
          hex2 = SEL(hex1, 0xfc99cf, (double)rd < 0.5);
          newType = SEL(hex2, 0xa7dd7a, (double)rd < 0.25);
        } else if (hexType == 0xfc99cf) {
          hex1 = SEL(0x331221, 0x8ebbe8, (double)rd < 0.75);
          hex2 = SEL(hex1, 0x122133, (double)rd < 0.5);
          newType = SEL(hex2, 0xa7dd7a, (double)rd < 0.25);
        } else if (hexType == 0xa7dd7a) {
          hex1 = SEL(0x331221, 0x8ebbe8, (double)rd < 0.75);
          hex2 = SEL(hex1, 0xfc99cf, (double)rd < 0.5);
          newType = SEL(hex2, 0x122133, (double)rd < 0.25);
        } else if (hexType == 0x47d47a) { newType = 0xd22a33; sc = 1.0f; }
        else if (hexType == 0xd22a33) { newType = 0x47d47a; sc = 1.0f; }
        else if (hexType == 0xa74d74) { newType = 0x33a22d; sc = 1.0f; }
        else if (hexType == 0x33a22d) { newType = 0xa74d74; sc = 1.0f; }
        else if (hexType == 0x8eb558) { newType = 0x331eb1; sc = 1.0f; }
        else if (hexType == 0x331eb1) { newType = 0x8eb558; sc = 1.0f; }
        else if (hexType == 0x855be8) { newType = 0x1be133; sc = 1.0f; }
        else if (hexType == 0x1be133) { newType = 0x855be8; sc = 1.0f; }
        else if (hexType == 0x6699cf) { newType = 0xcf1221; sc = 1.0f; }
        else if (hexType == 0xcf1221) { newType = 0x6699cf; sc = 1.0f; }
        else if (hexType == 0xfc9966) { newType = 0x1221fc; sc = 1.0f; }
        else if (hexType == 0x1221fc) { newType = 0xfc9966; sc = 1.0f; }
        float weightratio = rt_weight_ratio(w, hexType, newType);
        rd = SEL((rd - sc) * 4, rd, (double)sc == 1.0);
        tl[i * M + j] = SEL(hexType, newType, weightratio >= rd);
      }
    }
  }
}

/* updateTiles1: recttrianglekernel.cl:120-132 */
static void recttri_update1(const int32_t* t1, int32_t* t2, int N, int t) {
  const int M = N / 3;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 1; i < N - 1; i++) {
    for (int j = 1; j < M - 1; j++) {
      int p1 = i % 3 == t;
      int p2 = i % 3 == (t + 1) % 3;
      t2[i * M + j] = (t1[(i - 1) * M + j] & 0x0000f0) * 0x010000 +
                      (t1[(i - 1) * M + j] & 0x00000f) * 0x010000 +
                      (t1[i * M + j + p1] & 0xf00000) / 0x000100 +
                      (t1[(i + 1) * M + j - p2] & 0x0f0000) / 0x000100 +
                      (t1[(i + 1) * M + j - p2] & 0x00f000) / 0x000100 +
                      (t1[i * M + j + p1] & 0x000f00) / 0x000100;
    }
  }
}

/* updateTiles2: recttrianglekernel.cl:134-146 */
static void recttri_update2(const int32_t* t1, int32_t* t2, int N, int t) {
  const int M = N / 3;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 1; i < N - 1; i++) {
    for (int j = 1; j < M - 1; j++) {
      int p1 = i % 3 == (t + 2) % 3;
      int p2 = i % 3 == (t + 1) % 3;
      t2[i * M + j] = (t1[i * M + j - p1] & 0x00f000) * 0x000100 +
                      (t1[(i - 1) * M + j + p2] & 0x000f00) * 0x000100 +
                      (t1[(i - 1) * M + j + p2] & 0x0000f0) * 0x000100 +
                      (t1[i * M + j - p1] & 0x00000f) * 0x000100 +
                      (t1[(i + 1) * M + j] & 0xf00000) / 0x010000 +
                      (t1[(i + 1) * M + j] & 0x0f0000) / 0x010000;
    }
  }
}

/* RectTriangleTiler::Walk: RectTriangleTiler.cpp:33-171.  Colour
 * ((i-j)%3+3)%3 -> R, B, G (42-49); TinyMT seeded with `seed` (74); colours
 * from default_random_engine(seed) + uniform_int_distribution(0,2) (69-71,
 * 92); per step flip + updateTiles1 + updateTiles2 (95-107). */
int orc_recttri_walk(int32_t* t, int N, long steps, long seed,
                     const uint32_t* params3, size_t nparams,
                     const float* weights16) {
  if (N < 9 || N % 3 != 0) return ORC_EINVAL;
  const int M = N / 3;
  const size_t sz = (size_t)N * (size_t)M;
  if (nparams < sz) return ORC_EPARAMS;
  int32_t* C[3];
  for (int c = 0; c < 3; c++) C[c] = (int32_t*)calloc(sz, sizeof(int32_t));
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      C[((i - j) % 3 + 3) % 3][i * M + j / 3] = t[i * N + j];
  orc_minstd g;
  orc_minstd_seed(&g, (unsigned long)seed);
  orc_tinymt* st = init_streams(params3, sz, (uint32_t)seed);
  for (long s = 0; s < steps; s++) {
    int r = orc_uniform_int(&g, 0, 2);
    int c0 = r % 3;
    recttri_flip(C[c0], st, N, weights16);
    recttri_update1(C[c0], C[(c0 + 1) % 3], N, c0);
    recttri_update2(C[c0], C[(c0 + 2) % 3], N, c0);
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      t[i * N + j] = C[((i - j) % 3 + 3) % 3][i * M + j / 3];
  free(st);
  for (int c = 0; c < 3; c++) free(C[c]);
  return ORC_OK;
}

/* Segmented walk: one RectTriangleTiler::Walk per segment, seeded with
 * seeds[k] (TinyMT and the colour engine), as the lozenge segments are. */
int orc_recttri_walk_segments(int32_t* t, int N, const long* steps, const long* seeds, int nseg,
                              const uint32_t* params3, size_t nparams, const float* weights16) {
  for (int k = 0; k < nseg; k++) {
    int rc = orc_recttri_walk(t, N, steps[k], seeds[k], params3, nparams, weights16);
    if (rc) return rc;
  }
  return ORC_OK;
}

typedef struct {
  const uint32_t* params3;
  size_t nparams;
  const float* weights16;
} recttri_ctx;

static int recttri_walkseg_cb(int32_t* t, int N, const long* steps, const long* seeds, int n,
                              const void* c) {
  const recttri_ctx* d = (const recttri_ctx*)c;
  return orc_recttri_walk_segments(t, N, steps, seeds, n, d->params3, d->nparams, d->weights16);
}

/* CFTP for the rectangle-triangle model from maxHex / minHex (no reference
 * counterpart: the same driver as the other models). */
int orc_recttri_cftp(const int32_t* tmax, const int32_t* tmin, int N, uint32_t master_seed,
                     int sched, long t0, int max_rounds, const uint32_t* params3,
                     size_t nparams, const float* weights16, int32_t* out, orc_cftp_stats* st) {
  if (N < 9 || N % 3 != 0) return ORC_EINVAL;
  if (nparams < (size_t)N * (size_t)(N / 3)) return ORC_EPARAMS;
  recttri_ctx c = {params3, nparams, weights16};
  return cftp_driver(tmax, tmin, N, master_seed, sched, t0, max_rounds, recttri_walkseg_cb, &c,
                     out, st);
}

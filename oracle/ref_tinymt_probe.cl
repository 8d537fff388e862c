/*
 * ref_tinymt_probe.cl -- host entry points over the reference's own OpenCL
 * TinyMT code (src/TinyMT/tinymt32.clh, included where it lies), compiled
 * for the host CPU by clang's OpenCL C front end into
 * oracle/_ref/libref_host.so (see oracle/Makefile).  Test infrastructure:
 * only CPU tests load it, to pin the oracle's TinyMT restatement against
 * the reference generator itself.
 *
 * s7 = {s0, s1, s2, s3, mat1, mat2, tmat}, the field order of
 * tinymt32wp_t (tinymt32def.clh:28-36).  Plain (private) pointers are
 * ordinary host pointers on the x86-64 target.
 */
#include "TinyMT/tinymt32.clh"

static void load7(tinymt32wp_t* t, const uint* s7) {
  t->s0 = s7[0];
  t->s1 = s7[1];
  t->s2 = s7[2];
  t->s3 = s7[3];
  t->mat1 = s7[4];
  t->mat2 = s7[5];
  t->tmat = s7[6];
}

static void store7(uint* s7, const tinymt32wp_t* t) {
  s7[0] = t->s0;
  s7[1] = t->s1;
  s7[2] = t->s2;
  s7[3] = t->s3;
}

/* tinymt32_init(params, seed) (tinymt32.clh:182-203): s7[4..6] in, s7[0..3] out. */
void ref_cl_tinymt_init(uint* s7, uint seed) {
  tinymt32wp_t t;
  load7(&t, s7);
  tinymt32_init(&t, seed);
  store7(s7, &t);
}

/* n draws of tinymt32_single01 (tinymt32.clh:125-129); the state advances. */
void ref_cl_tinymt_single01(uint* s7, float* out, int n) {
  tinymt32wp_t t;
  load7(&t, s7);
  for (int k = 0; k < n; k++) out[k] = tinymt32_single01(&t);
  store7(s7, &t);
}

/* n draws of tinymt32_uint32 (tinymt32.clh:100-105). */
void ref_cl_tinymt_uint32(uint* s7, uint* out, int n) {
  tinymt32wp_t t;
  load7(&t, s7);
  for (int k = 0; k < n; k++) out[k] = tinymt32_uint32(&t);
  store7(s7, &t);
}

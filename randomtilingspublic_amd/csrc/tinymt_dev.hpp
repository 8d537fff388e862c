// tinymt_dev.hpp -- TinyMT32 on gfx950, state held in VGPRs.
//
// Same generator as the reference's per-work-item TinyMT
// (src/TinyMT/tinymt32.clh: next_state 39-55, temper12 81-93, single01
// 125-129, init 182-203), written for registers: the conditional xors are
// lane-wise selects, so a wavefront never diverges on the low state bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {

struct Tiny {
  uint32_t s0, s1, s2, s3;
  uint32_t mat1, mat2, tmat;
};

// a ^ b ^ c in one instruction (gfx950's three-input bit operation with
// the parity truth table 0x96); the compiler splits some of these chains.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

__device__ __forceinline__ void tiny_next(Tiny& t) {
  uint32_t x = (t.s0 & 0x7fffffffu) ^ t.s1 ^ t.s2;
  uint32_t y = t.s3;
  x ^= x << 1;
  y = xor3(y, y >> 1, x);
  const uint32_t m = 0u - (y & 1u);
  t.s0 = t.s1;
  t.s1 = t.s2 ^ (m & t.mat1);
  t.s2 = xor3(x, y << 10, m & t.mat2);
  t.s3 = y;
}

// Tempered output of the next state (tinymt32_temper).
__device__ __forceinline__ uint32_t tiny_bits(Tiny& t) {
  tiny_next(t);
  const uint32_t t1 = t.s0 + (t.s2 >> 8);
  return xor3(t.s3, t1, (0u - (t1 & 1u)) & t.tmat);
}

// The float the reference draws from tempered bits: ((b >> 9) | 1.0f) - 1.0f,
// i.e. exactly (b >> 9) * 2^-23.
__device__ __forceinline__ float tiny_to_float01(uint32_t b) {
  return __uint_as_float((b >> 9) | 0x3f800000u) - 1.0f;
}

__device__ __forceinline__ float tiny_float01(Tiny& t) { return tiny_to_float01(tiny_bits(t)); }

// Integer forms of the reference's threshold tests on rd = (b>>9)*2^-23:
// rd < 0.8f <=> (b>>9) <= 6710886; rd < 0.4f <=> (b>>9) <= 3355443;
// rd < 0.5f <=> b < 2^31.
constexpr uint32_t kBelow08 = 6710887u << 9;
constexpr uint32_t kBelow04 = 3355444u << 9;
constexpr uint32_t kBelow05 = 1u << 31;

// tinymt32_init(params, seed): 7 mixing rounds, period certification,
// 8 discarded outputs.
__device__ __forceinline__ void tiny_init(Tiny& t, uint32_t seed) {
  uint32_t st[4] = {seed, t.mat1, t.mat2, t.tmat};
#pragma unroll
  for (int i = 1; i < 8; i++)
    st[i & 3] ^= (uint32_t)i + 1812433253u * (st[(i - 1) & 3] ^ (st[(i - 1) & 3] >> 30));
  t.s0 = st[0];
  t.s1 = st[1];
  t.s2 = st[2];
  t.s3 = st[3];
  if ((t.s0 & 0x7fffffffu) == 0 && t.s1 == 0 && t.s2 == 0 && t.s3 == 0) {
    t.s0 = 'T';
    t.s1 = 'I';
    t.s2 = 'N';
    t.s3 = 'Y';
  }
#pragma unroll
  for (int i = 0; i < 8; i++) tiny_next(t);
}

// Stream table in HBM, structure-of-arrays so that consecutive work items
// read consecutive words: state[k*stride + id] for k = 0..3 (s0..s3),
// param[k*pstride + id] for k = 0..2 (mat1, mat2, tmat).
struct StreamTable {
  uint32_t* state;
  const uint32_t* param;
  long stride;
  long pstride;
};

__device__ __forceinline__ Tiny tiny_load(const StreamTable& S, long id) {
  Tiny t;
  t.s0 = S.state[id];
  t.s1 = S.state[S.stride + id];
  t.s2 = S.state[2 * S.stride + id];
  t.s3 = S.state[3 * S.stride + id];
  t.mat1 = S.param[id];
  t.mat2 = S.param[S.pstride + id];
  t.tmat = S.param[2 * S.pstride + id];
  return t;
}

__device__ __forceinline__ void tiny_store(const StreamTable& S, long id, const Tiny& t) {
  S.state[id] = t.s0;
  S.state[S.stride + id] = t.s1;
  S.state[2 * S.stride + id] = t.s2;
  S.state[3 * S.stride + id] = t.s3;
}

}  // namespace rt

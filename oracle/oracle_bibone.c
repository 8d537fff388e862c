/*
 * oracle_bibone.c -- CPU restatement of the bibone (TriangleDimer) sampler.
 * TEST INFRASTRUCTURE (see oracle.h): only tests and the CPU baseline load
 * this.  Each function follows the reference kernel or host function named
 * above it, in the reference's launch order.
 *
 * Arrays: eH, eL, eR (lozenge-type states), tU, tD (triangle-type), bH, bL,
 * bR (butterfly-type), each M x M int32 with M = N/2; every kernel's NDRange
 * is (M, M) with i = gid0, j = gid1 and TinyMT id j + M*i.  Every update
 * kernel writes only cells that it does not read (the flip classes and the
 * slices are chosen that way), so a serial loop is the parallel result.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define A(t, i, j) (t)[(size_t)(i) * M + (j)]
#define SEL(a, b, c) ((c) ? (b) : (a))
#define INTERIOR(i, j) ((i) < M - 1 && (j) < M - 1 && (i) > 0 && (j) > 0)

static inline int mod3(int x) { return ((x % 3) + 3) % 3; }

/* RotateLozenges: triangledimerkernel.cl:4-22 */
static void rot_loz(int* tl, orc_tinymt* st, int M, int t, int c) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      float rd = orc_tinymt_single01(&st[(size_t)j + (size_t)M * i]);
      if ((double)rd < .5 && ((1 - (t & 1)) * i + ((1 + t) / 2) * j) % 2 == c) {
        if (A(tl, i, j) == 5) A(tl, i, j) = 10;
        else if (A(tl, i, j) == 10) A(tl, i, j) = 5;
      }
    }
}

/* UpdateLozengesFlipped: 24-34 */
static void upd_loz_flipped(int* tl, int M, int t, int c) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++)
      if (INTERIOR(i, j) && ((1 - (t & 1)) * i + ((1 + t) / 2) * j) % 2 != c)
        A(tl, i, j) = (A(tl, i - 1, j + (t & 1)) & 4) / 4 + (A(tl, i - (2 - t) / 2, j + 1) & 8) / 4 +
                      (A(tl, i + 1, j - (t & 1)) & 1) * 4 + (A(tl, i + (2 - t) / 2, j - 1) & 2) * 4;
}

/* UpdateLozenges0/1/2: 36-79 */
static void upd_loz(int* t1, int* t2, int* t3, int M, int which) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      if (which == 0) {
        A(t2, i, j) &= ~10;
        A(t2, i, j) |= (A(t1, i + 1, j) & 1) * 2 + (A(t1, i + 1, j - 1) & 1) * 8;
        A(t3, i, j) &= ~10;
        A(t3, i, j) |= (A(t1, i + 1, j) & 2) + (A(t1, i, j) & 8);
      } else if (which == 1) {
        A(t2, i, j) &= ~5;
        A(t2, i, j) |= (A(t1, i, j) & 1) + (A(t1, i + 1, j) & 1) * 4;
        A(t3, i, j) &= ~5;
        A(t3, i, j) |= (A(t1, i - 1, j) & 2) / 2 + (A(t1, i, j) & 2) * 2;
      } else {
        A(t2, i, j) &= ~10;
        A(t2, i, j) |= (A(t1, i - 1, j) & 2) + (A(t1, i, j - 1) & 2) * 4;
        A(t3, i, j) &= ~5;
        A(t3, i, j) |= (A(t1, i, j) & 1) + (A(t1, i, j - 1) & 4);
      }
    }
}

/* UpdateTriangleUFromLozenges: 97-119 */
static void upd_triU_from_loz(const int* H, const int* L, int* U, int M) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int L1 = SEL(0, 1, (A(L, i, j) & 1) == 1);
      int L2 = SEL(0, 2, (A(H, i, j) & 8) == 8);
      int L3 = SEL(0, 3, (A(H, i + 1, j) & 1) == 1);
      int R1 = SEL(0, 1, (A(L, i, j + 1) & 1) == 1);
      int R2 = SEL(0, 2, (A(H, i + 1, j) & 2) == 2);
      int R3 = SEL(0, 3, (A(L, i, j + 1) & 2) == 2);
      int V1 = SEL(0, 1, (A(L, i, j + 1) & 4) == 4);
      int V2 = SEL(0, 2, (A(H, i + 1, j) & 8) == 8);
      int V3 = SEL(0, 3, (A(H, i + 1, j) & 4) == 4);
      A(U, i, j) = L1 + L2 + L3 + 4 * (R1 + R2 + R3) + 16 * (V1 + V2 + V3);
    }
}

/* UpdateButterflys{H,L,R}FromLozenge: 121-188 */
static void upd_bf_from_loz(int* B, const int* LH, const int* LL, const int* LR, int M, int o) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int c, t1, t4, t16, t64;
      if (o == 0) {
        int lh = A(LH, i, j), lr1 = A(LR, i - 1, j), ll1 = A(LL, i - 1, j + 1), lr2 = A(LR, i, j),
            ll2 = A(LL, i, j);
        c = (ll1 & 4) / 4;
        t1 = (lr1 & 8) / 8 + 2 * (lr1 & 1) + 3 * (lh & 1);
        t4 = (ll1 & 1) + (ll1 & 2) + ((lh & 2) / 2) * 3;
        t16 = (lr2 & 2) / 2 + (lr2 & 4) / 2 + ((lh & 4) / 4) * 3;
        t64 = (ll2 & 4) / 4 + (ll2 & 8) / 4 + ((lh & 8) / 8) * 3;
      } else if (o == 1) {
        int ll = A(LL, i, j), lh1 = A(LH, i, j), lr1 = A(LR, i, j), lh2 = A(LH, i + 1, j - 1),
            lr2 = A(LR, i, j - 1);
        c = (lr1 & 8) / 8;
        t1 = (lh1 & 1) + (lh1 & 2) + 3 * (ll & 1);
        t4 = (lr1 & 2) / 2 + (lr1 & 4) / 2 + ((ll & 2) / 2) * 3;
        t16 = (lh2 & 4) / 4 + (lh2 & 8) / 4 + ((ll & 4) / 4) * 3;
        t64 = (lr2 & 8) / 8 + 2 * (lr2 & 1) + ((ll & 8) / 8) * 3;
      } else {
        int lr = A(LR, i, j), lh1 = A(LH, i, j), ll1 = A(LL, i, j + 1), lh2 = A(LH, i + 1, j),
            ll2 = A(LL, i, j);
        c = (ll1 & 8) / 8;
        t1 = (lh1 & 1) + (lh1 & 2) + 3 * (lr & 1);
        t4 = (ll1 & 1) + (ll1 & 2) + ((lr & 2) / 2) * 3;
        t16 = (lh2 & 4) / 4 + (lh2 & 8) / 4 + ((lr & 4) / 4) * 3;
        t64 = (ll2 & 4) / 4 + (ll2 & 8) / 4 + ((lr & 8) / 8) * 3;
      }
      A(B, i, j) = t1 + 4 * t4 + 16 * t16 + 64 * t64 + 256 * c;
    }
}

/* RotateTriangles: 190-211 */
static void rot_tri(int* tl, orc_tinymt* st, int M, int c) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      float rd = orc_tinymt_single01(&st[(size_t)j + (size_t)M * i]);
      if ((double)rd < 0.5 && mod3(j - i) == c) {
        if (A(tl, i, j) == 45) A(tl, i, j) = 54;
        else if (A(tl, i, j) == 54) A(tl, i, j) = 45;
      }
    }
}

/* UpdateTrianglesFlipped0 (213-233) and UpdateTrianglesFlipped1 (235-255) */
static void upd_tri_flipped(int* tl, int M, int c, int down) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j) || mod3(j - i) == c) continue;
      int p = mod3(j - i) == (c + 1) % 3;
      int t1 = (4 - 3 * p) * 12, m1 = (4 - 3 * p) * 4;
      int t4 = (15 * p + 1) * 3, m4 = 15 * p + 1;
      int t16 = (4 - 3 * p) * 3, m16 = 4 - 3 * p;
      if (!down)
        A(tl, i, j) = (A(tl, i - (1 - p), j - p) & t1) / m1 +
                      ((A(tl, i - p, j + 1) & t4) / m4) * 4 +
                      ((A(tl, i + 1, j - 1 + p) & t16) / m16) * 16;
      else
        A(tl, i, j) = (A(tl, i + 1 - p, j - 1) & t1) / m1 +
                      ((A(tl, i + p, j + 1 - p) & t4) / m4) * 4 +
                      ((A(tl, i - 1, j + p) & t16) / m16) * 16;
    }
}

/* UpdateTriangles: 257-278 */
static void upd_tri(const int* t1, int* t2, int M, int t) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int p = SEL(1, -1, t == 0);
      int a = A(t1, i, j - 1 + t), b = A(t1, i, j + t), v = A(t1, i + p, j);
      int L1 = SEL(0, 1, (a & 48) == 16), L2 = SEL(0, 2, (a & 3) == 3), L3 = SEL(0, 3, (a & 12) == 8);
      int R1 = SEL(0, 1, (b & 48) == 16), R2 = SEL(0, 2, (b & 3) == 3), R3 = SEL(0, 3, (b & 12) == 8);
      int V1 = SEL(0, 1, (v & 48) == 16), V2 = SEL(0, 2, (v & 3) == 3), V3 = SEL(0, 3, (v & 12) == 8);
      A(t2, i, j) = L1 + L2 + L3 + 4 * (R1 + R2 + R3) + 16 * (V1 + V2 + V3);
    }
}

/* UpdateLozenge{H,L,R}FromTriangles: 280-328 */
static void upd_loz_from_tri(const int* U, const int* D, int* X, int M, int o) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int tD = A(D, i, j);
      if (o == 0) {
        int tU = A(U, i - 1, j);
        A(X, i, j) = ((tD & 48) == 32) + 2 * ((tD & 48) == 48) + 4 * ((tU & 48) == 48) +
                     8 * ((tU & 48) == 32);
      } else if (o == 1) {
        int tU = A(U, i, j - 1);
        A(X, i, j) = ((tU & 12) == 4) + 2 * ((tU & 12) == 12) + 4 * ((tD & 3) == 1) +
                     8 * ((tD & 3) == 2);
      } else {
        int tU = A(U, i, j);
        A(X, i, j) = ((tU & 3) == 1) + 2 * ((tD & 12) == 12) + 4 * ((tD & 12) == 4) +
                     8 * ((tU & 3) == 2);
      }
    }
}

/* RotateButterflys: 330-352 */
static void rot_bf(int* tl, orc_tinymt* st, int M, int t, int p1, int p2) {
  int b1 = 1 - (t & 1), b2 = (1 + t) / 2, b3 = t & 1;
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      float rd = orc_tinymt_single01(&st[(size_t)j + (size_t)M * i]);
      if ((double)rd < 0.5 && (b1 * i + b2 * j) % 3 == p1) {
        if (A(tl, i, j) == 170) A(tl, i, j) = SEL(170, 85, ((b1 * j + b3 * i) & 1) == p2);
        else if (A(tl, i, j) == 85) A(tl, i, j) = SEL(85, 170, ((b1 * j + b3 * i) & 1) == p2);
      }
    }
}

static inline int enc(int t1, int t4, int t16, int t64, int c) {
  return t1 + 4 * t4 + 16 * t16 + 64 * t64 + 256 * c;
}

/* Butterfly "same slice" refresh after a flip: UpdateButterflysFlipped{H,L,R}1
 * (355-376, 459-478, 552-573).  o = 0 H (slices i%3), 1 L (j%3), 2 R ((i+j)%3). */
static void upd_bf_flipped1(int* tl, int M, int o, int p1, int p2) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int c = (A(tl, i, j) & 256) / 256;
      if (o == 0 && i % 3 == p1 && (j & 1) != p2) {
        int bf1 = A(tl, i, j - 1), bf2 = A(tl, i, j + 1);
        int t1 = SEL(((bf1 & 12) / 4) % 3 + 1, 0, (bf1 & 12) / 4 == 0);
        int t4 = SEL(((bf2 & 3) + 1) % 3 + 1, 0, (bf2 & 3) == 0);
        int t16 = SEL(((bf2 & 192) / 64) % 3 + 1, 0, (bf2 & 192) / 64 == 0);
        int t64 = SEL(((bf1 & 48) / 16 + 1) % 3 + 1, 0, (bf1 & 48) / 16 == 0);
        A(tl, i, j) = enc(t1, t4, t16, t64, c);
      } else if (o == 1 && j % 3 == p1 && (i & 1) != p2) {
        int bf1 = A(tl, i - 1, j), bf2 = A(tl, i + 1, j);
        int t1 = SEL(((bf1 & 12) / 4) % 3 + 1, 0, (bf1 & 12) / 4 == 0);
        int t4 = SEL(((bf2 & 3) + 1) % 3 + 1, 0, (bf2 & 3) == 0);
        int t16 = SEL(((bf2 & 192) / 64) % 3 + 1, 0, (bf2 & 192) / 64 == 0);
        int t64 = SEL(((bf1 & 48) / 16 + 1) % 3 + 1, 0, (bf1 & 48) / 16 == 0);
        A(tl, i, j) = enc(t1, t4, t16, t64, c);
      } else if (o == 2 && (i + j) % 3 == p1 && (j & 1) != p2) {
        int bf1 = A(tl, i + 1, j - 1), bf2 = A(tl, i - 1, j + 1);
        int t1 = SEL(((bf2 & 192) / 64 + 1) % 3 + 1, 0, (bf2 & 192) / 64 == 0);
        int t4 = SEL(((bf2 & 48) / 16) % 3 + 1, 0, (bf2 & 48) / 16 == 0);
        int t16 = SEL(((bf1 & 12) / 4 + 1) % 3 + 1, 0, (bf1 & 12) / 4 == 0);
        int t64 = SEL(((bf1 & 3)) % 3 + 1, 0, (bf1 & 3) == 0);
        A(tl, i, j) = enc(t1, t4, t16, t64, c);
      }
    }
}

/* Adjacent-slice updates: H21/H22/H23 (378-457), L21/L22/L23 (480-550),
 * R21/R22/R23 (575-649).  stage = 1, 2, 3. */
static void upd_bf_flipped2(int* tl, int M, int o, int p1, int stage) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int bf = A(tl, i, j);
      if (o == 0) {
        if (stage == 1 && i % 3 == (p1 + 1) % 3) {
          int bf1 = A(tl, i - 1, j), bf2 = A(tl, i - 1, j + 1);
          int c = SEL(0, 1, (bf1 & 48) / 16 == 2);
          int t1 = SEL((bf1 & 192) / 192 + ((bf1 & 48) / 48) * 3, (bf & 3), (bf & 3) == 2);
          int t4 = SEL(((bf2 & 192) / 192) * 3 + ((bf2 & 48) / 48) * 2, (bf & 12) / 4,
                       (bf & 12) / 4 == 1);
          A(tl, i, j) = enc(t1, t4, (bf & 48) / 16, (bf & 192) / 64, c);
        } else if (stage == 2 && i % 3 == (p1 + 2) % 3) {
          int s1 = SEL(0, 1, (bf & 3) != 2), s4 = SEL(0, 1, (bf & 12) / 4 != 1);
          int bf1 = A(tl, i - 1, j), bf2 = A(tl, i - 1, j + 1), bf3 = A(tl, i + 1, j),
              bf4 = A(tl, i + 1, j - 1);
          int c = SEL(0, 1, (bf3 & 3) == 2);
          int t1 = SEL(s1 * (bf & 3), 2, (bf1 & 256) == 256);
          int t4 = SEL(s4 * ((bf & 12) / 4), 1, (bf2 & 256) == 256);
          int t16 = SEL(((bf3 & 3) / 3) * 3 + (bf3 & 12) / 12, (bf & 48) / 16, (bf & 48) / 16 == 2);
          int t64 = SEL(((bf4 & 3) / 3) * 2 + ((bf4 & 12) / 12) * 3, (bf & 192) / 64,
                        (bf & 192) / 64 == 1);
          A(tl, i, j) = enc(t1, t4, t16, t64, c);
        } else if (stage == 3 && i % 3 == (p1 + 1) % 3) {
          int s16 = SEL(0, 1, (bf & 48) / 16 != 2), s64 = SEL(0, 1, (bf & 192) / 64 != 1);
          int bf3 = A(tl, i + 1, j), bf4 = A(tl, i + 1, j - 1);
          int c = (bf & 256) / 256;
          int t16 = SEL(s16 * ((bf & 48) / 16), 2, (bf3 & 256) == 256);
          int t64 = SEL(s64 * ((bf & 192) / 64), 1, (bf4 & 256) == 256);
          A(tl, i, j) = enc(bf & 3, (bf & 12) / 4, t16, t64, c);
        }
      } else if (o == 1) {
        if (stage == 1 && j % 3 == (p1 + 2) % 3) {
          int bf1 = A(tl, i - 1, j + 1), bf2 = A(tl, i, j + 1);
          int c = SEL(0, 1, (bf1 & 48) / 16 == 2);
          int t1 = SEL((bf1 & 192) / 192 + ((bf1 & 48) / 48) * 3, (bf & 3), (bf & 3) == 2);
          int t4 = SEL(((bf2 & 192) / 192) * 3 + ((bf2 & 48) / 48) * 2, (bf & 12) / 4,
                       (bf & 12) / 4 == 1);
          A(tl, i, j) = enc(t1, t4, (bf & 48) / 16, (bf & 192) / 64, c);
        } else if (stage == 2 && j % 3 == (p1 + 1) % 3) {
          int s1 = SEL(0, 1, (bf & 3) != 2), s4 = SEL(0, 1, (bf & 12) / 4 != 1);
          int bf1 = A(tl, i - 1, j + 1), bf2 = A(tl, i, j + 1), bf3 = A(tl, i + 1, j - 1),
              bf4 = A(tl, i, j - 1);
          int c = SEL(0, 1, (bf3 & 3) == 2);
          int t1 = SEL(s1 * (bf & 3), 2, (bf1 & 256) / 256 == 1);
          int t4 = SEL(s4 * (bf & 12) / 4, 1, (bf2 & 256) / 256 == 1);
          int t16 = SEL(((bf3 & 3) / 3) * 3 + (bf3 & 12) / 12, (bf & 48) / 16, (bf & 48) / 16 == 2);
          int t64 = SEL(((bf4 & 3) / 3) * 2 + ((bf4 & 12) / 12) * 3, (bf & 192) / 64,
                        (bf & 192) / 64 == 1);
          A(tl, i, j) = enc(t1, t4, t16, t64, c);
        } else if (stage == 3 && j % 3 == (p1 + 2) % 3) {
          int s16 = SEL(0, 1, (bf & 48) / 16 != 2), s64 = SEL(0, 1, (bf & 192) / 64 != 1);
          int bf3 = A(tl, i + 1, j - 1), bf4 = A(tl, i, j - 1);
          int c = (bf & 256) / 256;
          int t16 = SEL(s16 * (bf & 48) / 16, 2, (bf3 & 256) / 256 == 1);
          int t64 = SEL(s64 * (bf & 192) / 64, 1, (bf4 & 256) / 256 == 1);
          A(tl, i, j) = enc(bf & 3, (bf & 12) / 4, t16, t64, c);
        }
      } else {
        if (stage == 1 && (i + j) % 3 == (p1 + 2) % 3) {
          int bf2 = A(tl, i, j + 1), bf3 = A(tl, i + 1, j);
          int c = SEL(0, 1, (bf3 & 3) == 1);
          int t4 = SEL(((bf2 & 3) / 3) + ((bf2 & 192) / 192) * 3, (bf & 12) / 4, (bf & 12) / 4 == 2);
          int t16 = SEL(((bf3 & 3) / 3) * 3 + ((bf3 & 192) / 192) * 2, (bf & 48) / 16,
                        (bf & 48) / 16 == 1);
          A(tl, i, j) = enc(bf & 3, t4, t16, (bf & 192) / 64, c);
        } else if (stage == 2 && (i + j) % 3 == (p1 + 1) % 3) {
          int s4 = SEL(0, 1, (bf & 12) / 4 != 2), s16 = SEL(0, 1, (bf & 48) / 16 != 1);
          int bf1 = A(tl, i - 1, j), bf2 = A(tl, i, j + 1), bf3 = A(tl, i + 1, j),
              bf4 = A(tl, i, j - 1);
          int c = SEL(0, 1, (bf1 & 48) / 16 == 1);
          int t1 = SEL(((bf1 & 48) / 48) * 3 + ((bf1 & 12) / 12) * 2, (bf & 3), (bf & 3) == 1);
          int t4 = SEL(s4 * (bf & 12) / 4, 2, (bf2 & 256) / 256 == 1);
          int t16 = SEL(s16 * (bf & 48) / 16, 1, (bf3 & 256) / 256 == 1);
          int t64 = SEL(((bf4 & 48) / 48) + ((bf4 & 12) / 12) * 3, (bf & 192) / 64,
                        (bf & 192) / 64 == 2);
          A(tl, i, j) = enc(t1, t4, t16, t64, c);
        } else if (stage == 3 && (i + j) % 3 == (p1 + 2) % 3) {
          int s1 = SEL(0, 1, (bf & 3) != 1), s64 = SEL(0, 1, (bf & 192) / 64 != 2);
          int bf1 = A(tl, i - 1, j), bf4 = A(tl, i, j - 1);
          int c = (bf & 256) / 256;
          int t1 = SEL(s1 * (bf & 3), 1, (bf1 & 256) / 256 == 1);
          int t64 = SEL(s64 * (bf & 192) / 64, 2, (bf4 & 256) / 256 == 1);
          A(tl, i, j) = enc(t1, (bf & 12) / 4, (bf & 48) / 16, t64, c);
        }
      }
    }
}

/* UpdateLozengeFromButterflys{H,L,R}: 652-749 */
static void upd_loz_from_bf(const int* B, int* LH, int* LL, int* LR, int M, int o) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      if (!INTERIOR(i, j)) continue;
      int bf = A(B, i, j);
      if (o == 0) {
        A(LH, i, j) = ((bf & 3) == 3) + 2 * ((bf & 12) == 12) + 4 * ((bf & 48) == 48) +
                      8 * ((bf & 192) == 192);
        A(LL, i, j) = ((bf & 256) == 256) + 2 * ((bf & 48) == 48) + 4 * ((bf & 192) == 64) +
                      8 * ((bf & 192) == 128);
        A(LR, i, j) = ((bf & 256) == 256) + 2 * ((bf & 48) == 16) + 4 * ((bf & 48) == 32) +
                      8 * ((bf & 192) == 192);
      } else if (o == 1) {
        A(LL, i, j) = ((bf & 3) == 3) + 2 * ((bf & 12) == 12) + 4 * ((bf & 48) == 48) +
                      8 * ((bf & 192) == 192);
        A(LR, i, j) = ((bf & 3) == 3) + 2 * ((bf & 12) == 4) + 4 * ((bf & 12) == 8) +
                      8 * ((bf & 256) == 256);
        A(LH, i, j) = ((bf & 3) == 1) + 2 * ((bf & 3) == 2) + 4 * ((bf & 12) == 12) +
                      8 * ((bf & 256) == 256);
      } else {
        A(LR, i, j) = ((bf & 3) == 3) + 2 * ((bf & 12) == 12) + 4 * ((bf & 48) == 48) +
                      8 * ((bf & 192) == 192);
        A(LH, i, j) = ((bf & 3) == 1) + 2 * ((bf & 3) == 2) + 4 * ((bf & 256) == 256) +
                      8 * ((bf & 192) == 192);
        A(LL, i, j) = ((bf & 3) == 3) + 2 * ((bf & 256) == 256) + 4 * ((bf & 192) == 64) +
                      8 * ((bf & 192) == 128);
      }
    }
}

#undef A
#define T(i, j) t[(size_t)(i) * N + (j)]

/* TriangleDimerTiler::TilingToTriangleFlips: TriangleDimerTiler.cpp:846-873.
 * out: (N/2) x N. */
void orc_bibone_triangle_flips(const int32_t* t, int N, int32_t* out) {
  memset(out, 0, sizeof(int32_t) * (size_t)(N / 2) * N);
  for (int i = 1; i < N / 2 - 1; i++)
    for (int j = 1; j < N - 1; j++) {
      int v;
      if (j % 2 == 0)
        v = ((T(2 * i + 1, j) & 4) / 4 + (T(2 * i + 1, j) & 8) / 4 + ((T(2 * i + 1, j + 1) & 8) / 8) * 3) +
            4 * ((T(2 * i + 1, j + 1) & 4) / 4 + (T(2 * i + 1, j) & 2) + ((T(2 * i + 1, j + 1) & 2) / 2) * 3) +
            16 * ((T(2 * i, j + 1) & 1) * 2 + (T(2 * i + 1, j) & 1) + ((T(2 * i, j + 1) & 2) / 2) * 3);
      else
        v = ((T(2 * i + 1, j) & 1) + (T(2 * i + 1, j) & 8) / 4 + ((T(2 * i + 1, j + 1) & 8) / 8) * 3) +
            4 * ((T(2 * i + 1, j + 1) & 1) + (T(2 * i + 1, j) & 2) + ((T(2 * i + 1, j + 1) & 2) / 2) * 3) +
            16 * ((T(2 * i + 2, j) & 8) / 4 + ((T(2 * i + 2, j) & 4) / 4) * 3 + (T(2 * i + 1, j) & 4) / 4);
      out[(size_t)i * N + j] = v;
    }
}

/* TriangleDimerTiler::TilingToButterflyFlips: TriangleDimerTiler.cpp:875-921. */
void orc_bibone_butterfly_flips(const int32_t* t, int N, int32_t* tB) {
  memset(tB, 0, sizeof(int32_t) * (size_t)N * N);
  for (int i = 2; i < N - 2; i++)
    for (int j = 2; j < N - 2; j++) {
      int f1, f4, f16, f64, c;
      if (i % 2 == 0) {
        if (j % 2 != 1) continue;
        f1 = (T(i - 1, j) & 8) / 8 + (T(i - 1, j) & 1) * 2 + (T(i, j) & 1) * 3;
        f4 = (T(i - 1, j + 1) & 1) + (T(i - 1, j + 1) & 2) + ((T(i, j) & 2) / 2) * 3;
        f16 = (T(i + 1, j) & 2) / 2 + (T(i + 1, j) & 4) / 2 + ((T(i, j) & 4) / 4) * 3;
        f64 = (T(i + 1, j - 1) & 4) / 4 + (T(i + 1, j - 1) & 8) / 4 + ((T(i, j) & 8) / 8) * 3;
        c = (T(i - 1, j + 1) & 4) / 4;
      } else if (j % 2 == 0) {
        f1 = (T(i - 1, j + 1) & 1) + (T(i - 1, j + 1) & 2) + (T(i, j) & 1) * 3;
        f4 = (T(i, j + 1) & 2) / 2 + (T(i, j + 1) & 4) / 2 + ((T(i, j) & 2) / 2) * 3;
        f16 = (T(i + 1, j - 1) & 4) / 4 + (T(i + 1, j - 1) & 8) / 4 + ((T(i, j) & 4) / 4) * 3;
        f64 = (T(i, j - 1) & 8) / 8 + (T(i, j - 1) & 1) * 2 + ((T(i, j) & 8) / 8) * 3;
        c = (T(i, j + 1) & 8) / 8;
      } else {
        f1 = (T(i - 1, j) & 1) + (T(i - 1, j) & 2) + (T(i, j) & 1) * 3;
        f4 = (T(i, j + 1) & 1) + (T(i, j + 1) & 2) + ((T(i, j) & 2) / 2) * 3;
        f16 = (T(i + 1, j) & 4) / 4 + (T(i + 1, j) & 8) / 4 + ((T(i, j) & 4) / 4) * 3;
        f64 = (T(i, j - 1) & 4) / 4 + (T(i, j - 1) & 8) / 4 + ((T(i, j) & 8) / 8) * 3;
        c = (T(i, j + 1) & 8) / 8;
      }
      tB[(size_t)i * N + j] = f1 + 4 * f4 + 16 * f16 + 64 * f64 + 256 * c;
    }
}

/* Move schedule of TriangleDimerTiler::Walk (TriangleDimerTiler.cpp:
 * 272-298, 447-448): per step five ints (type, r1, r2, p1, p2), p1 = p2 =
 * -1 unless type == 2. */
void orc_bibone_schedule(long seed, long steps, int32_t* out5) {
  orc_minstd g;
  orc_minstd_seed(&g, (unsigned long)seed);
  for (long s = 0; s < steps; s++) {
    int32_t* m = out5 + 5 * s;
    m[0] = orc_uniform_int(&g, 0, 2);
    m[1] = orc_uniform_int(&g, 0, 2);
    m[2] = orc_uniform_int(&g, 0, 1);
    m[3] = m[4] = -1;
    if (m[0] == 2) {
      m[3] = orc_uniform_int(&g, 0, 2);
      m[4] = orc_uniform_int(&g, 0, 1);
    }
  }
}

/* TriangleDimerTiler::Walk: TriangleDimerTiler.cpp:101-695.  The host draws
 * rFlipType, r1 (both uniform 0..2) and r2 (uniform 0..1) per step, plus p1
 * (0..2), p2 (0..1) for butterfly steps, from one default_random_engine
 * seeded with `seed`; TinyMT is seeded with `seed` (M*M streams). */
int orc_bibone_walk(int32_t* t, int N, long steps, long seed, const uint32_t* params3,
                    size_t nparams, int32_t* planes) {
  if (N < 6 || N % 2) return ORC_EINVAL;
  const int M = N / 2;
  const size_t sz = (size_t)M * M;
  if (nparams < sz) return ORC_EPARAMS;
  int32_t* tTri = (int32_t*)calloc((size_t)(N / 2) * N, sizeof(int32_t));
  int32_t* tB = (int32_t*)calloc((size_t)N * N, sizeof(int32_t));
  int* a[8];
  for (int k = 0; k < 8; k++) a[k] = (int*)calloc(sz, sizeof(int));
  int *eR = a[0], *eL = a[1], *eH = a[2], *tU = a[3], *tD = a[4], *bR = a[5], *bL = a[6], *bH = a[7];
  orc_bibone_triangle_flips(t, N, tTri);
  orc_bibone_butterfly_flips(t, N, tB);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const size_t k = (size_t)(i / 2) * M + j / 2;
      if (i % 2 == 0) {
        if (j % 2 == 1) {
          eH[k] = T(i, j);
          bH[k] = tB[(size_t)i * N + j];
        }
      } else if (j % 2 == 0) {
        eL[k] = T(i, j);
        bL[k] = tB[(size_t)i * N + j];
      } else {
        eR[k] = T(i, j);
        bR[k] = tB[(size_t)i * N + j];
      }
    }
  for (int i = 0; i < N / 2; i++)
    for (int j = 0; j < N; j++) {
      if (j % 2 == 0) tD[(size_t)i * M + j / 2] = tTri[(size_t)i * N + j];
      else tU[(size_t)i * M + j / 2] = tTri[(size_t)i * N + j];
    }
  int32_t* sched = (int32_t*)malloc(sizeof(int32_t) * 5 * (size_t)(steps > 0 ? steps : 1));
  orc_bibone_schedule(seed, steps, sched);
  orc_tinymt* st = (orc_tinymt*)malloc(sizeof(orc_tinymt) * sz);
  for (size_t id = 0; id < sz; id++) {
    st[id].mat1 = params3[3 * id];
    st[id].mat2 = params3[3 * id + 1];
    st[id].tmat = params3[3 * id + 2];
    orc_tinymt_init(&st[id], (uint32_t)seed + (uint32_t)id);
  }
  for (long s = 0; s < steps; s++) {
    const int type = sched[5 * s], r1 = sched[5 * s + 1], r2 = sched[5 * s + 2];
    if (type == 0) { /* lozenge flips, TriangleDimerTiler.cpp:300-336 */
      int* e = r1 == 0 ? eH : (r1 == 1 ? eL : eR);
      rot_loz(e, st, M, r1, r2);
      upd_loz_flipped(e, M, r1, r2);
      if (r1 == 0) upd_loz(eH, eL, eR, M, 0);
      else if (r1 == 1) upd_loz(eL, eR, eH, M, 1);
      else upd_loz(eR, eH, eL, M, 2);
      upd_triU_from_loz(eH, eL, tU, M);
      upd_tri(tU, tD, M, 0);
      upd_bf_from_loz(bH, eH, eL, eR, M, 0);
      upd_bf_from_loz(bL, eH, eL, eR, M, 1);
      upd_bf_from_loz(bR, eH, eL, eR, M, 2);
    } else if (type == 1) { /* triangle flips, 356-379 */
      if (r2 == 1) {
        rot_tri(tD, st, M, r1);
        upd_tri_flipped(tD, M, r1, 1);
        upd_tri(tD, tU, M, r2);
      } else {
        rot_tri(tU, st, M, r1);
        upd_tri_flipped(tU, M, r1, 0);
        upd_tri(tU, tD, M, r2);
      }
      upd_loz_from_tri(tU, tD, eH, M, 0);
      upd_loz_from_tri(tU, tD, eL, M, 1);
      upd_loz_from_tri(tU, tD, eR, M, 2);
      upd_bf_from_loz(bH, eH, eL, eR, M, 0);
      upd_bf_from_loz(bL, eH, eL, eR, M, 1);
      upd_bf_from_loz(bR, eH, eL, eR, M, 2);
    } else { /* butterfly flips, 447-500 */
      const int p1 = sched[5 * s + 3], p2 = sched[5 * s + 4];
      int* b = r1 == 0 ? bH : (r1 == 1 ? bL : bR);
      rot_bf(b, st, M, r1, p1, p2);
      upd_bf_flipped1(b, M, r1, p1, p2);
      upd_bf_flipped2(b, M, r1, p1, 1);
      upd_bf_flipped2(b, M, r1, p1, 2);
      upd_bf_flipped2(b, M, r1, p1, 3);
      upd_loz_from_bf(b, eH, eL, eR, M, r1);
      if (r1 == 0) {
        upd_bf_from_loz(bL, eH, eL, eR, M, 1);
        upd_bf_from_loz(bR, eH, eL, eR, M, 2);
      } else if (r1 == 1) {
        upd_bf_from_loz(bH, eH, eL, eR, M, 0);
        upd_bf_from_loz(bR, eH, eL, eR, M, 2);
      } else {
        upd_bf_from_loz(bL, eH, eL, eR, M, 1);
        upd_bf_from_loz(bH, eH, eL, eR, M, 0);
      }
      upd_triU_from_loz(eH, eL, tU, M);
      upd_tri(tU, tD, M, 0);
    }
  }
  /* merge the lozenge-type states back (TriangleDimerTiler.cpp:563-577) */
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      const size_t k = (size_t)(i / 2) * M + j / 2;
      if (i % 2 == 0) {
        if (j % 2 == 1) T(i, j) = eH[k];
      } else {
        T(i, j) = j % 2 == 0 ? eL[k] : eR[k];
      }
    }
  if (planes) {
    int* order[8] = {eH, eL, eR, tU, tD, bH, bL, bR};
    for (int k = 0; k < 8; k++) memcpy(planes + (size_t)k * sz, order[k], sz * sizeof(int));
  }
  free(sched);
  free(st);
  for (int k = 0; k < 8; k++) free(a[k]);
  free(tTri);
  free(tB);
  return ORC_OK;
}

#undef T

/* CPU oracle (test infrastructure only; never linked into the product).
 *
 * SIMD-oriented Fast Mersenne Twister, MEXP=19937, restated from the published
 * algorithm (Saito & Matsumoto, "SIMD-oriented Fast Mersenne Twister", 2006)
 * as vendored by the reference, source/libSteve/SFMT/SFMT.c (v1.3.3) with the
 * 19937 parameter set of SFMT-params19937.h:4-16.  The reference keeps one
 * global state (random2.h:21-58); the oracle keeps one state per simulation,
 * which is identical when one simulation runs per process.
 *
 * Parity: the u32 stream is checked against the compiled reference
 * (oracle/_ref/libsteve_ref.so) by tests/test_oracle_ref.py.
 */
#include <stdint.h>
#include <string.h>

#include "oracle_internal.h"

#define SF_N 156
#define SF_N32 (SF_N * 4)
#define SF_POS1 122
#define SF_SL1 18
#define SF_SL2 1
#define SF_SR1 11
#define SF_SR2 1

static const uint32_t sf_msk[4] = {0xdfffffefU, 0xddfecb7fU, 0xbffaffffU, 0xbffffff6U};
static const uint32_t sf_parity[4] = {0x00000001U, 0x00000000U, 0x00000000U, 0x13c9e684U};

/* 128-bit little-endian word shifted by 8*nbytes bits. */
static void sf_rshift(const uint32_t* in, uint32_t* out, int nbytes) {
  uint64_t th = ((uint64_t)in[3] << 32) | in[2];
  uint64_t tl = ((uint64_t)in[1] << 32) | in[0];
  uint64_t oh = th >> (nbytes * 8);
  uint64_t ol = (tl >> (nbytes * 8)) | (th << (64 - nbytes * 8));
  out[0] = (uint32_t)ol;
  out[1] = (uint32_t)(ol >> 32);
  out[2] = (uint32_t)oh;
  out[3] = (uint32_t)(oh >> 32);
}

static void sf_lshift(const uint32_t* in, uint32_t* out, int nbytes) {
  uint64_t th = ((uint64_t)in[3] << 32) | in[2];
  uint64_t tl = ((uint64_t)in[1] << 32) | in[0];
  uint64_t oh = (th << (nbytes * 8)) | (tl >> (64 - nbytes * 8));
  uint64_t ol = tl << (nbytes * 8);
  out[0] = (uint32_t)ol;
  out[1] = (uint32_t)(ol >> 32);
  out[2] = (uint32_t)oh;
  out[3] = (uint32_t)(oh >> 32);
}

/* r = a ^ (a << SL2 bytes) ^ ((b >> SR1) & MSK) ^ (c >> SR2 bytes) ^ (d << SL1) */
static void sf_recursion(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* c,
                         const uint32_t* d) {
  uint32_t x[4], y[4];
  int k;
  sf_lshift(a, x, SF_SL2);
  sf_rshift(c, y, SF_SR2);
  for (k = 0; k < 4; k++)
    r[k] = a[k] ^ x[k] ^ ((b[k] >> SF_SR1) & sf_msk[k]) ^ y[k] ^ (d[k] << SF_SL1);
}

static void sf_refill(osfmt* s) {
  uint32_t* st = s->state;
  uint32_t* r1 = st + 4 * (SF_N - 2);
  uint32_t* r2 = st + 4 * (SF_N - 1);
  int i;
  for (i = 0; i < SF_N - SF_POS1; i++) {
    sf_recursion(st + 4 * i, st + 4 * i, st + 4 * (i + SF_POS1), r1, r2);
    r1 = r2;
    r2 = st + 4 * i;
  }
  for (; i < SF_N; i++) {
    sf_recursion(st + 4 * i, st + 4 * i, st + 4 * (i + SF_POS1 - SF_N), r1, r2);
    r1 = r2;
    r2 = st + 4 * i;
  }
}

void osfmt_seed(osfmt* s, uint32_t seed) {
  uint32_t* st = s->state;
  uint32_t inner = 0;
  int i, j;
  st[0] = seed;
  for (i = 1; i < SF_N32; i++) st[i] = 1812433253UL * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
  s->idx = SF_N32;
  /* period certification */
  for (i = 0; i < 4; i++) inner ^= st[i] & sf_parity[i];
  for (i = 16; i > 0; i >>= 1) inner ^= inner >> i;
  inner &= 1;
  if (inner == 1) return;
  for (i = 0; i < 4; i++) {
    uint32_t work = 1;
    for (j = 0; j < 32; j++) {
      if (work & sf_parity[i]) {
        st[i] ^= work;
        return;
      }
      work <<= 1;
    }
  }
}

uint32_t osfmt_u32(osfmt* s) {
  if (s->idx >= SF_N32) {
    sf_refill(s);
    s->idx = 0;
  }
  return s->state[s->idx++];
}

/*
 * philox_ref.h -- TEST INFRASTRUCTURE (oracle). Host-only Philox4x32-10 and the
 * random-stream layout shared by the oracle and the device library.
 *
 * Philox4x32-10 follows Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11 (multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds).  Pinned by the published
 * known-answer vectors in tests/test_philox.py.
 *
 * Stream layout (the reference draws from Numba's hidden per-thread MT19937,
 * IMC/IMC2D.py:756,761-765,697-707; a counter-based stream replaces it so the
 * device and this oracle consume identical numbers):
 *   key = { lo32(seed), hi32(seed) }
 *   ctr = { lo32(idx), (hi32(idx) & 0x00FFFFFF) | tag << 24, step, block }
 * Each block gives 4 words = 2 uniform "slots" (a,b):
 *   double  u = ((a >> 5) * 2^26 + (b >> 6)) * 2^-53      in [0,1)
 *   float   u = (a >> 8) * 2^-24                          in [0,1)
 * Slot q lives in block q/2, half q%2.
 */
#ifndef IMC_ORACLE_PHILOX_REF_H
#define IMC_ORACLE_PHILOX_REF_H
#include <stdint.h>

enum {
  ORC_TAG_TRANSPORT = 1,
  ORC_TAG_EMISSION = 2,
  ORC_TAG_SOURCE = 3,
  ORC_TAG_COMB = 4,
  ORC_TAG_IC = 5,
  ORC_TAG_BOUNDARY0 = 8 /* + side (0..3) */
};

static inline void orc_philox_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

static inline void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    orc_philox_round(c, k);
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = c[3];
}

typedef struct {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  int32_t cached_block; /* block held in buf, -1 none */
} orc_stream;

static inline void orc_stream_init(orc_stream* s, uint64_t seed, int tag, uint64_t idx, uint32_t step) {
  s->key[0] = (uint32_t)seed;
  s->key[1] = (uint32_t)(seed >> 32);
  s->ctr[0] = (uint32_t)idx;
  s->ctr[1] = ((uint32_t)(idx >> 32) & 0x00FFFFFFu) | ((uint32_t)tag << 24);
  s->ctr[2] = step;
  s->ctr[3] = 0;
  s->cached_block = -1;
  s->buf[0] = s->buf[1] = s->buf[2] = s->buf[3] = 0;
}

static inline void orc_stream_block(orc_stream* s, uint32_t block) {
  if (s->cached_block == (int32_t)block) return;
  s->ctr[3] = block;
  orc_philox4x32_10(s->ctr, s->key, s->buf);
  s->cached_block = (int32_t)block;
}

/* 53-bit uniform in [0,1) from slot q */
static inline double orc_u01(orc_stream* s, uint32_t q) {
  orc_stream_block(s, q >> 1);
  uint32_t a = s->buf[(q & 1u) * 2u], b = s->buf[(q & 1u) * 2u + 1u];
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

#endif

/*
 * lm_philox.h -- counter-based Philox-4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the uniform draws
 * of the all-or-nothing collision step.  Shared by the gfx950 kernels and the
 * CPU oracle so both consume the identical stream; pinned by the published
 * known-answer vectors of Random123 (tests/test_oracle_cpu.py).
 *
 * Consumption structure follows the reference: one uniform per unordered SIP
 * pair (i<j) of a cell (collision/all_or_nothing.py:545-633), none for cells
 * with fewer than two SIPs.  The uniform of pair (i, j), d = i + j, has 53 bits:
 *   u = (G * 2^45 + A * 2^13 + (B >> 19)) * 2^-53,
 * its leading byte G = byte (d & 3) of the guard word
 *   g = fmix32(K + (i<<16 | d>>2) * 0x9E3779B9),  K = guard key of (seed, run, step, cell),
 * then A = word (d & 3) of Philox(ctr = (i<<16 | d>>2, cell, step, run)) and the
 * top 13 bits of B = word (d & 3) of Philox(ctr = (i<<16 | d>>2, cell, step, run | 2^31)),
 * key = (seed lo, seed hi).  fmix32 is the 32-bit finaliser of Austin Appleby's
 * MurmurHash3 (public domain): a bijection of 32-bit words with full avalanche,
 * so the guard words of the quads of one cell are distinct and their bytes
 * behave as independent uniform bytes (checked in tests/test_oracle_cpu.py),
 * independent of the Philox words.
 * Decisions from the leading bits: G/256 <= u < (G+1)/256 and
 * L = (G*2^32 + A) * 2^-40 <= u < L + 2^-40, so p > u is decided by G alone when
 * p <= G/256 (no Philox block at all: the collision kernel rejects a quad of
 * four pairs, one guard word, when its pairs' bound is below its smallest byte),
 * by (G, A) unless p falls inside [L, L + 2^-40), and only then is the B block
 * needed.  Blocks follow the anti-diagonals d, so the gfx950 collision kernel
 * (one row per lane, diagonals in order) refreshes its blocks on the same
 * diagonal in every lane, once per four diagonals.  Valid for cells of up to
 * 65536 SIPs.
 */
#ifndef LM_PHILOX_H
#define LM_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LMP_FN __host__ __device__ static inline
#else
#define LMP_FN static inline
#endif

typedef struct {
  uint32_t v[4];
} lm_u32x4;

LMP_FN lm_u32x4 lm_philox4x32_10(lm_u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIPCC__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.v[0];
    const uint64_t p1 = (uint64_t)M1 * c.v[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    lm_u32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

LMP_FN double lm_res53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Philox block of the pair words of row i, diagonals 4*(d>>2) .. +3: half 0 = A words, 1 = B words. */
LMP_FN lm_u32x4 lm_pair_block(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell, uint32_t i, uint32_t d,
                              int half) {
  lm_u32x4 c;
  c.v[0] = (i << 16) | (d >> 2);
  c.v[1] = cell;
  c.v[2] = step;
  c.v[3] = half ? (run | 0x80000000u) : run;
  return lm_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

/*
 * Uniform [0, 1) with 53 bits for the initial SIP ensemble, keyed by the run's own seed: counter (a, b, level,
 * 0x40000000 | stream << 24), so these draws never meet the collision draws (whose last counter word is below
 * 2^31 with bit 30 clear for runs below 2^30) even under the same key.
 */
LMP_FN double lm_init_uniform(uint64_t seed, uint32_t level, uint32_t a, uint32_t b, uint32_t stream) {
  lm_u32x4 c;
  c.v[0] = a;
  c.v[1] = b;
  c.v[2] = level;
  c.v[3] = 0x40000000u | ((stream & 0x3Fu) << 24);
  const lm_u32x4 o = lm_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return lm_res53(o.v[0], o.v[1]);
}

LMP_FN uint32_t lm_word(lm_u32x4 o, uint32_t k) {
  return k == 0 ? o.v[0] : (k == 1 ? o.v[1] : (k == 2 ? o.v[2] : o.v[3]));
}

/* MurmurHash3's 32-bit finaliser (fmix32) */
LMP_FN uint32_t lm_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

/* guard key of one cell's collision step: the counter words absorbed one after another */
LMP_FN uint32_t lm_guard_key(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell) {
  uint32_t h = lm_fmix32((uint32_t)seed ^ 0x6A09E667u);
  h = lm_fmix32(h ^ (uint32_t)(seed >> 32));
  h = lm_fmix32(h ^ run);
  h = lm_fmix32(h ^ step);
  return lm_fmix32(h ^ cell);
}

/* guard word of the block (i, d>>2): byte (d & 3) is the leading byte of the uniform of pair (i, d - i) */
LMP_FN uint32_t lm_guard_word(uint32_t key, uint32_t i, uint32_t d) {
  return lm_fmix32(key + ((i << 16) | (d >> 2)) * 0x9E3779B9u);
}

LMP_FN uint32_t lm_guard_byte(uint32_t g, uint32_t d) { return (g >> (8u * (d & 3u))) & 0xFFu; }

/* (G * 2^32 + A) * 2^-40: the 40-bit lower bound of the pair's uniform */
LMP_FN double lm_pair_lo40(uint32_t G, uint32_t a) {
  return ((double)G * 4294967296.0 + (double)a) * (1.0 / 1099511627776.0);
}

/* the pair's uniform from its guard byte and its A and B words */
LMP_FN double lm_pair_u53(uint32_t G, uint32_t a, uint32_t b) {
  return ((double)G * 35184372088832.0 + (double)a * 8192.0 + (double)(b >> 19)) * (1.0 / 9007199254740992.0);
}

/* Uniform in [0,1) of pair (i, j), i < j, of cell `cell` in collision step `step` of run `run`. */
LMP_FN double lm_pair_uniform(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell, uint32_t i, uint32_t j) {
  const uint32_t d = i + j;
  const uint32_t G = lm_guard_byte(lm_guard_word(lm_guard_key(seed, run, step, cell), i, d), d);
  const lm_u32x4 a = lm_pair_block(seed, run, step, cell, i, d, 0);
  const lm_u32x4 b = lm_pair_block(seed, run, step, cell, i, d, 1);
  return lm_pair_u53(G, lm_word(a, d & 3), lm_word(b, d & 3));
}

#endif /* LM_PHILOX_H */

/*
 * lm_philox.h -- counter-based Philox-4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the uniform draws
 * of the all-or-nothing collision step.  Shared by the gfx950 kernels and the
 * CPU oracle so both consume the identical stream; pinned by the published
 * known-answer vectors of Random123 (tests/test_oracle_cpu.py).
 *
 * Consumption structure follows the reference: one uniform per unordered SIP
 * pair (i<j) of a cell (collision/all_or_nothing.py:545-633), none for cells
 * with fewer than two SIPs.  The uniform of pair (i, j), d = i + j, is
 * res53(A, B) -- NumPy's random_sample construction from two 32-bit words,
 * ((A>>5)*2^26 + (B>>6)) * 2^-53 -- with
 *   A = word (d & 3) of Philox(ctr = (i<<16 | d>>2, cell, step, run)),
 *   B = word (d & 3) of Philox(ctr = (i<<16 | d>>2, cell, step, run | 2^31)),
 * key = (seed lo, seed hi).  Blocks follow the anti-diagonals d, so the
 * gfx950 collision kernel (one row per lane, diagonals in order) refreshes
 * its A block on the same diagonal in every lane, once per four diagonals.
 * Since (A>>5)*2^-27 <= u < (A>>5)*2^-27 + 2^-27, the comparison p > u is
 * decided by A alone unless p falls inside that interval; only then is the
 * B block needed.  Valid for cells of up to 65536 SIPs.
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

/* Uniform in [0,1) of pair (i, j), i < j, of cell `cell` in collision step `step` of run `run`. */
LMP_FN double lm_pair_uniform(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell, uint32_t i, uint32_t j) {
  const uint32_t d = i + j;
  const lm_u32x4 a = lm_pair_block(seed, run, step, cell, i, d, 0);
  const lm_u32x4 b = lm_pair_block(seed, run, step, cell, i, d, 1);
  return lm_res53(lm_word(a, d & 3), lm_word(b, d & 3));
}

#endif /* LM_PHILOX_H */

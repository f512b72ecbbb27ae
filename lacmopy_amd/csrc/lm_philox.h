/*
 * lm_philox.h -- counter-based Philox-4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the uniform draws
 * of the all-or-nothing collision step.  Shared by the gfx950 kernels and the
 * CPU oracle so both consume the identical stream; pinned by the published
 * known-answer vectors of Random123 (tests/test_oracle_cpu.py).
 *
 * Consumption structure follows the reference: one uniform per unordered SIP
 * pair (i<j) of a cell, in row-major order, pair number
 * cnt = i*N - i*(i+1)/2 + (j-i-1) (collision/all_or_nothing.py:545-633).
 * The uniform of pair cnt is res53(A, B) (NumPy's random_sample construction
 * from two 32-bit words: (A>>5)*2^26 + (B>>6), scaled by 2^-53) with
 *   A = word (cnt & 3) of Philox(ctr = (cnt >> 2, cell, step, run)),
 *   B = word (cnt & 3) of Philox(ctr = (cnt >> 2, cell, step, run | 2^31)),
 * key = (seed lo, seed hi).  One block of A words serves four consecutive
 * pairs, and since (A>>5)*2^-27 <= u < (A>>5)*2^-27 + 2^-27, the comparison
 * p > u is decided by A alone unless p falls in that 2^-27-wide interval;
 * only then is the B block needed.
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

/* Philox block of pair words: half = 0 gives the A words, 1 the B words. */
LMP_FN lm_u32x4 lm_pair_block(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell, uint64_t blk, int half) {
  lm_u32x4 c;
  c.v[0] = (uint32_t)blk;
  c.v[1] = cell;
  c.v[2] = step;
  c.v[3] = half ? (run | 0x80000000u) : run;
  return lm_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

LMP_FN uint32_t lm_word(lm_u32x4 o, int k) { return k == 0 ? o.v[0] : (k == 1 ? o.v[1] : (k == 2 ? o.v[2] : o.v[3])); }

/* Uniform in [0,1) for pair number cnt of cell `cell` in collision step `step` of run `run`. */
LMP_FN double lm_pair_uniform(uint64_t seed, uint32_t run, uint32_t step, uint32_t cell, uint64_t cnt) {
  const lm_u32x4 a = lm_pair_block(seed, run, step, cell, cnt >> 2, 0);
  const lm_u32x4 b = lm_pair_block(seed, run, step, cell, cnt >> 2, 1);
  return lm_res53(lm_word(a, (int)(cnt & 3)), lm_word(b, (int)(cnt & 3)));
}

#endif /* LM_PHILOX_H */

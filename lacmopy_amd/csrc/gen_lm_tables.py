#!/usr/bin/env python3
"""Generate lm_tables.h: the constant tables of the lm_math FP64 library.

The tables are computed with mpmath at 200 bits and rounded to double-double
(hi + lo).  They are compiled into both the host (oracle) build and the gfx950
device build, so exp/log/pow give bit-identical results on CPU and GPU.

  exp: T_j = 2^(j/128), j = 0..127, as (hi, lo)
  log: for i = LOG_IMIN..LOG_IMAX, c_i = fl(1 / (1 + i/128)) and
       -log(c_i) as (hi, lo)

Run:  python3 gen_lm_tables.py > lm_tables.h
"""
import mpmath as mp

mp.mp.prec = 200

EXP_N = 128
LOG_IMIN = -40
LOG_IMAX = 56


def dd(x):
    hi = float(x)
    lo = float(x - mp.mpf(hi))
    return hi, lo


def main():
    out = []
    out.append("/* Generated by gen_lm_tables.py -- do not edit. */")
    out.append("#ifndef LM_TABLES_H")
    out.append("#define LM_TABLES_H")
    out.append(f"#define LM_EXP_N {EXP_N}")
    out.append(f"#define LM_LOG_IMIN ({LOG_IMIN})")
    out.append(f"#define LM_LOG_IMAX ({LOG_IMAX})")
    out.append("LM_TABLE double lm_exp_tab[2 * LM_EXP_N] = {")
    for j in range(EXP_N):
        hi, lo = dd(mp.power(2, mp.mpf(j) / EXP_N))
        out.append(f"  {hi.hex()}, {lo.hex()},")
    out.append("};")
    out.append("LM_TABLE double lm_log_tab[3 * (LM_LOG_IMAX - LM_LOG_IMIN + 1)] = {")
    for i in range(LOG_IMIN, LOG_IMAX + 1):
        c = float(1 / (1 + mp.mpf(i) / 128))
        hi, lo = dd(-mp.log(mp.mpf(c)))
        out.append(f"  {c.hex()}, {hi.hex()}, {lo.hex()},")
    out.append("};")
    out.append("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()

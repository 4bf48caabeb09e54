#!/usr/bin/env python3
"""Writes nf_math_tables.h: the constant tables of the table-driven log and exp in nf_math.h.

Every entry is the correctly rounded double of a value computed with 300-bit mpmath, so the
tables (and therefore the functions) are the same bit patterns on every compiler and target.
  LOG_TBL[j - 96] = {1/F, hi(log F), lo(log F)}   for F = j/128, j = 96 .. 192
  EXP_TBL[j]      = {hi(2^(j/64)), lo(2^(j/64))}  for j = 0 .. 63
plus ln2/64 split so that n * hi is exact for |n| < 2^18.
Run: python gen_math_tables.py > nf_math_tables.h
"""
import mpmath

mpmath.mp.prec = 300


def d(x):
    return float(x)


def hexd(x):
    return float.hex(float(x))


def split(x):
    hi = d(x)
    lo = d(x - mpmath.mpf(hi))
    return hi, lo


def main():
    out = ["// Generated by gen_math_tables.py; do not edit.", "#pragma once", ""]
    out.append("// {1/F, log(F) hi, log(F) lo} for F = j/128, j = 96..192")
    out.append("static constexpr double NFM_LOG_TBL[97][3] = {")
    for j in range(96, 193):
        F = mpmath.mpf(j) / 128
        inv = d(1 / F)
        hi, lo = split(mpmath.log(F))
        out.append(f"    {{{hexd(inv)}, {hexd(hi)}, {hexd(lo)}}},  // j = {j}")
    out.append("};")
    out.append("")
    out.append("// {2^(j/64) hi, lo}, j = 0..63")
    out.append("static constexpr double NFM_EXP_TBL[64][2] = {")
    for j in range(64):
        hi, lo = split(mpmath.power(2, mpmath.mpf(j) / 64))
        out.append(f"    {{{hexd(hi)}, {hexd(lo)}}},  // j = {j}")
    out.append("};")
    ln2_64 = mpmath.log(2) / 64
    # hi with 35 significant bits (18 trailing zero bits): n * hi exact for |n| < 2^18
    m, e = mpmath.frexp(ln2_64)
    hi = mpmath.ldexp(mpmath.floor(m * 2 ** 35), e - 35)
    lo = d(ln2_64 - hi)
    out.append("")
    out.append(f"#define NFM_L64_HI {hexd(hi)}")
    out.append(f"#define NFM_L64_LO {hexd(lo)}")
    out.append(f"#define NFM_INV_L64 {hexd(1 / ln2_64)}  // 64 / ln 2, rounded")
    ln2 = mpmath.log(2)
    h2 = float.fromhex("0x1.62e42fefa0000p-1")
    out.append(f"#define NFM_LN2_HI {h2.hex()}  // 37 significant bits")
    out.append(f"#define NFM_LN2_LO {hexd(ln2 - h2)}")
    print("\n".join(out))


if __name__ == "__main__":
    main()

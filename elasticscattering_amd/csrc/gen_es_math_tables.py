"""Writes es_math_tables.h: the constant tables of the correctly rounded
elementary functions in es_math.h.

    python elasticscattering_amd/csrc/gen_es_math_tables.py

Every entry is a value of a mathematical constant or function rounded from a
300-bit mpmath evaluation, split into a double-double (hi = RN(v),
lo = RN(v - hi)) where the header says so:

* ESM_ASIN_TAB[128][16]: Taylor coefficients of asin about the centre of each
  interval [i/256, (i+1)/256) of [0, 1/2] (centre 0 for i = 0, else
  (i + 1/2)/256): a0 and a1 as double-doubles, a2..a9 as doubles, padded to
  16 doubles (one 128-byte line per interval).  The script checks the
  truncation error of the degree-9 polynomial on every interval.
* ESM_SC_TAB[202][4]: sin(j/256) and cos(j/256) as double-doubles, j = 0..201
  (201/256 > pi/4).
* ESM_EXP_TAB[128][2]: 2^(i/128) as double-doubles.
* Splits of pi/2 and ln2/128 used by the argument reductions, and the Taylor
  coefficients 1/k! of the accurate phases as double-doubles.
"""
import os

import mpmath as mp

mp.mp.prec = 300
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "es_math_tables.h")

ASIN_N = 128   # intervals over [0, 1/2]
ASIN_K = 9     # polynomial degree
SC_N = 256     # sin/cos table step 1/256


def rn(v):
    """Round an mpf to the nearest double (ties to even)."""
    return float(mp.mpf(v))


def dd(v):
    h = rn(v)
    return h, rn(mp.mpf(v) - h)


def hexd(x):
    return float.hex(x)


def asin_table():
    rows = []
    worst = mp.mpf(0)
    for i in range(ASIN_N):
        c = mp.mpf(0) if i == 0 else (mp.mpf(i) + mp.mpf("0.5")) / (2 * ASIN_N)
        co = mp.taylor(mp.asin, c, ASIN_K + 2)
        a0 = dd(co[0])
        a1 = dd(co[1])
        rest = [rn(co[k]) for k in range(2, ASIN_K + 1)]
        row = [a0[0], a0[1], a1[0], a1[1]] + rest
        row += [0.0] * (16 - len(row))
        rows.append(row)
        lo, hi = mp.mpf(i) / (2 * ASIN_N), mp.mpf(i + 1) / (2 * ASIN_N)
        for t in range(41):
            u = lo + (hi - lo) * t / 40
            if u == 0:
                continue
            h = u - c
            approx = sum(co[k] * h ** k for k in range(ASIN_K + 1))
            worst = max(worst, abs(approx - mp.asin(u)) / mp.asin(u))
    return rows, worst


def sc_table():
    rows = []
    for j in range(202):
        r = mp.mpf(j) / SC_N
        s, c = dd(mp.sin(r)), dd(mp.cos(r))
        rows.append([s[0], s[1], c[0], c[1]])
    return rows


def exp_table():
    return [list(dd(mp.power(2, mp.mpf(i) / 128))) for i in range(128)]


def split_bits(v, parts, first_bits=53):
    """v as a sum of `parts` doubles; the first holds `first_bits` significant bits."""
    out = []
    rem = mp.mpf(v)
    for n in range(parts):
        if n == 0 and first_bits < 53:
            e = int(mp.floor(mp.log(abs(rem), 2)))
            q = mp.mpf(2) ** (e - first_bits + 1)
            h = float(mp.nint(rem / q) * q)
        else:
            h = rn(rem)
        out.append(h)
        rem -= h
    return out


def main():
    asin_rows, worst = asin_table()
    sc_rows = sc_table()
    exp_rows = exp_table()
    pio2 = split_bits(mp.pi / 2, 4)
    pio2_dd = dd(mp.pi / 2)
    ln2_128 = split_bits(mp.log(2) / 128, 3, first_bits=35)
    inv_ln2_128 = rn(128 / mp.log(2))
    inv_pio2 = rn(2 / mp.pi)
    fact = [dd(1 / mp.factorial(k)) for k in range(0, 16)]

    L = []
    w = L.append
    w("/* es_math_tables.h — generated by gen_es_math_tables.py (mpmath, 300 bits); do not edit. */")
    w("#ifndef ES_MATH_TABLES_H")
    w("#define ES_MATH_TABLES_H")
    w("")
    w(f"/* asin: {ASIN_N} intervals of [0, 1/2], degree {ASIN_K}; max relative truncation error 2^{float(mp.log(worst, 2)):.1f}. */")
    w(f"#define ESM_ASIN_N {ASIN_N}")
    w("#define ESM_ASIN_TAB_INIT { \\")
    for row in asin_rows:
        w("    " + ", ".join(hexd(v) for v in row) + ", \\")
    w("}")
    w("")
    w("/* sin(j/256) hi, lo, cos(j/256) hi, lo for j = 0..201. */")
    w("#define ESM_SC_TAB_INIT { \\")
    for row in sc_rows:
        w("    " + ", ".join(hexd(v) for v in row) + ", \\")
    w("}")
    w("")
    w("/* 2^(i/128) hi, lo for i = 0..127. */")
    w("#define ESM_EXP_TAB_INIT { \\")
    for row in exp_rows:
        w("    " + ", ".join(hexd(v) for v in row) + ", \\")
    w("}")
    w("")
    w("/* pi/2 = P1 + P2 + P3 + P4 (each the RN of what is left), and RN-split double-double. */")
    for n, v in enumerate(pio2, 1):
        w(f"#define ESM_PIO2_{n} {hexd(v)}")
    w(f"#define ESM_PIO2_H {hexd(pio2_dd[0])}")
    w(f"#define ESM_PIO2_L {hexd(pio2_dd[1])}")
    w(f"#define ESM_INVPIO2 {hexd(inv_pio2)}")
    w("/* ln2/128 = L1 + L2 + L3; L1 has 35 significant bits, so k * L1 is exact for |k| < 2^18. */")
    for n, v in enumerate(ln2_128, 1):
        w(f"#define ESM_LN2_128_{n} {hexd(v)}")
    w(f"#define ESM_INV_LN2_128 {hexd(inv_ln2_128)}")
    w("/* 1/k! as double-doubles, k = 0..15. */")
    w("#define ESM_INVFACT_INIT { \\")
    for h, l in fact:
        w(f"    {hexd(h)}, {hexd(l)}, \\")
    w("}")
    w("")
    w("#endif /* ES_MATH_TABLES_H */")
    with open(OUT, "w") as f:
        f.write("\n".join(L) + "\n")
    print(f"wrote {OUT}; asin truncation 2^{float(mp.log(worst, 2)):.2f}")


if __name__ == "__main__":
    main()

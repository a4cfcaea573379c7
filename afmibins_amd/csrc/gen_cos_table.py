"""Writes afm_cos_table.h: cos(j/64) and sin(j/64), j = -51..51, each as a
double-double (hi = the double nearest the value, lo = the double nearest
the remainder), for the table-driven fast path of afm_delta.h's cos_cr.

Row j + 51 of parity 0 (a cosine of the reduced argument) holds
(U, V) = (cos(j/64), -sin(j/64)); of parity 1 (a sine) (sin(j/64),
cos(j/64)); so that for y = j/64 + d either function is U cos d + V sin d.

The values come from mpmath at 300 bits.  Run from this directory:
    python gen_cos_table.py
"""
import os

import mpmath

mpmath.mp.prec = 300
JMAX = 51


def dd(v):
    hi = float(v)
    lo = float(v - mpmath.mpf(hi))
    return hi, lo


def main():
    rows = {0: [], 1: []}
    for j in range(-JMAX, JMAX + 1):
        c, s = mpmath.cos(mpmath.mpf(j) / 64), mpmath.sin(mpmath.mpf(j) / 64)
        rows[0].append(dd(c) + dd(-s))
        rows[1].append(dd(s) + dd(c))
    out = ["// Generated by gen_cos_table.py: cos(j/64), sin(j/64) for j = -%d..%d as" % (JMAX, JMAX),
           "// double-double pairs (U_hi, U_lo, V_hi, V_lo); parity 0: (cos, -sin), parity 1:",
           "// (sin, cos).  Do not edit.",
           "#pragma once",
           "#define AFM_COS_TAB_JMAX %d" % JMAX,
           "#define AFM_COS_TAB_INIT { \\"]
    for p in (0, 1):
        out.append("\t{ \\")
        for r in rows[p]:
            out.append("\t\t{" + ", ".join(x.hex() for x in r) + "}, \\")
        out.append("\t}, \\")
    out.append("}")
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "afm_cos_table.h"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()

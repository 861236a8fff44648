#!/usr/bin/env python3
"""Writes trig_tables.h: cos/sin(2 pi m / N) for the in-register DFT codelets.

Every N the codelets take twiddles from gets a table: the radices 3..16, the
prime codelets 17, 19, 23 and
every 13-smooth composite N <= 64 (the two-factor splits of the register
line passes).  Values come from mpmath at 40 digits (cospi/sinpi, so that the
exact points 0, +-1/2, +-1 are exact) and are printed with 21 significant
digits, more than a double holds.

    python3 gen_trig_tables.py > trig_tables.h
"""
import mpmath as mp

mp.mp.dps = 40


def smooth13(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def fmt(x):
    return mp.nstr(x, 21, min_fixed=1, max_fixed=1)


def table_sizes():
    base = [3, 5, 7, 8, 9, 11, 13, 16, 17, 19, 23]
    comp = [n for n in range(6, 65) if smooth13(n) and n not in base and n not in (2, 4)]
    return sorted(set(base + comp))


def main():
    out = ["// trig_tables.h -- cos/sin(2*pi*m/N) for the in-register DFT codelets.",
           "// Generated by gen_trig_tables.py (mpmath, 40 digits; values rounded to 21",
           "// significant digits).",
           "#pragma once",
           "namespace fftahip {",
           "template <int R> struct Trig;"]
    for n in table_sizes():
        c = [fmt(mp.cospi(mp.mpf(2 * m) / n)) for m in range(n)]
        s = [fmt(mp.sinpi(mp.mpf(2 * m) / n)) for m in range(n)]
        out.append(f"template <> struct Trig<{n}> {{")
        out.append(f"  static constexpr double c[{n}] = {{{', '.join(c)}}};")
        out.append(f"  static constexpr double s[{n}] = {{{', '.join(s)}}};")
        out.append("};")
    out.append("}  // namespace fftahip")
    print("\n".join(out))


if __name__ == "__main__":
    main()

"""Writes gkm_erfcx.h: a polynomial approximation of the scaled complementary
error function erfcx(x) = exp(x^2) erfc(x) on [0, B], used by the FP32 flux
kernel as erfc(x) = erfcx(|x|) exp(-x^2) (the Gaussian factor is shared with the
half-space moment N1, which needs the same exponential), with
erfc(-|x|) = 2 - erfc(|x|).  Beyond B, erfc(|x|) < 2.2e-17 and the asymptotic
series erfcx(x) ~ (1 - 1/(2x^2) + 3/(4x^4) - 15/(8x^6)) / (x sqrt(pi)) (relative
error below 105/(16 x^8) < 4e-6) is used.

Method: Chebyshev interpolation of erfcx at the n+1 Chebyshev points of the
first kind on [0, B], computed in 50-digit arithmetic (mpmath), converted
to the monomial basis of u = 2x/B - 1 in the same precision and rounded once
to float, so the device evaluates it with Horner's rule (v_fmaak_f32 with the
coefficient as an inline literal).  The degree is the smallest whose
truncation error (measured against mpmath on a dense grid) is below 2^-24.
(The FP64 flux keeps the device library's erfc: a 35-term FP64 polynomial
needs its coefficients in registers, which that register-bound kernel cannot
spare; measured 2.2x slower.)

Run from this folder:  python3 gen_erfcx.py   (needs mpmath; not needed to
build or run the library).
"""
from __future__ import annotations

from pathlib import Path

import mpmath as mp

mp.mp.dps = 50
B = 6.0  # |x| <= B takes the polynomial; beyond it a four-term asymptotic series


def cheb_monomial(deg: int):
    n = deg + 1
    nodes = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
    vals = [mp.erfc(B * (u + 1) / 2) * mp.exp((B * (u + 1) / 2) ** 2) for u in nodes]
    # Chebyshev coefficients c_j = (2/n) sum_k f_k T_j(u_k), c_0 halved
    c = []
    for j in range(n):
        s = mp.fsum(vals[k] * mp.cos(j * mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n))
        c.append(2 * s / n)
    c[0] /= 2
    # monomial coefficients: sum_j c_j T_j(u) = sum_i a_i u^i
    a = [mp.mpf(0)] * n
    Tm, T = [mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]  # T_0, T_1 as coefficient lists
    for j in range(n):
        Tj = Tm if j == 0 else (T if j == 1 else None)
        if j >= 2:
            Tj = [mp.mpf(0)] + [2 * t for t in T]
            for i, t in enumerate(Tm):
                Tj[i] -= t
            Tm, T = T, Tj
        for i, t in enumerate(Tj):
            a[i] += c[j] * t
    return a


def max_rel_err(a, ntest=4001):
    worst = mp.mpf(0)
    for i in range(ntest):
        x = B * mp.mpf(i) / (ntest - 1)
        u = 2 * x / B - 1
        p = mp.mpf(0)
        for coef in reversed(a):
            p = p * u + coef
        ref = mp.erfc(x) * mp.exp(x * x)
        worst = max(worst, abs(p / ref - 1))
    return worst


def pick(target):
    for deg in range(6, 40):
        a = cheb_monomial(deg)
        if max_rel_err(a, 801) < target and max_rel_err(a) < target:
            return deg, a
    raise RuntimeError("no degree below 40 reaches the target")


def main():
    out = Path(__file__).resolve().parent / "gkm_erfcx.h"
    d32, a32 = pick(mp.mpf(2) ** -24)
    lines = [
        "// gkm_erfcx.h — written by gen_erfcx.py; do not edit.",
        "// erfcx(x) = exp(x^2) erfc(x) on [0, kErfcxB] as a polynomial in u = 2x/kErfcxB - 1",
        "// (Chebyshev interpolant at the first-kind points, converted to monomials in",
        f"// 50-digit arithmetic, rounded once), degree {d32}, truncation error below 2^-24.",
        "#pragma once",
        "",
        "namespace gkm {",
        "",
        f"constexpr double kErfcxB = {B!r};",
        f"constexpr int kErfcxDegF = {d32};",
        "",
        "// coefficients of u^0 .. u^deg",
        f"__device__ constexpr float kErfcxF[{d32 + 1}] = {{",
    ]
    lines += [f"    {mp.nstr(c, 12, min_fixed=-1, max_fixed=-1)}f," for c in a32]
    lines += ["};", "", "}  // namespace gkm", ""]
    out.write_text("\n".join(lines))
    print(f"wrote {out}: degree {d32}")


if __name__ == "__main__":
    main()

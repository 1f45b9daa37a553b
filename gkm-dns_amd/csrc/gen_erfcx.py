"""Writes gkm_erfcx.h: a rational approximation P(x)/Q(x) of the scaled
complementary error function erfcx(x) = exp(x^2) erfc(x) on [0, CAP], used by
the FP32 flux kernel as erfc(x) = erfcx(min(|x|, CAP)) exp(-x^2) (the Gaussian
factor is shared with the half-space moment N1, which needs the same
exponential), with erfc(-|x|) = 2 - erfc(|x|).  Beyond CAP = 12, exp(-x^2)
is 0 in FP32 and so is erfc.  deg Q = deg P + 1, so P/Q decays like 1/x as
erfcx does and one rational covers the whole range: 4 + 5 FMAs, one
reciprocal and one multiply per evaluation, no branch.

Method: linearised least squares on P(x) - f(x) Q(x) = 0 with Q(0) = 1,
relative weighting 1/f and Lawson reweighting towards the largest relative
error (60 passes), on 6000 uniform points of [0, CAP]; f = erfcx from mpmath
(30 digits).  The fitted coefficients are rounded once to float.  Relative
error of the fit in FP64 evaluation: 8e-9; in the FP32 evaluation the kernel
performs: 4e-7 (tests/test_erfcx.py checks both against scipy).
(The FP64 flux keeps the device library's erfc: a polynomial for FP64 needs
its coefficients in registers, which that register-bound kernel cannot
spare; measured 2.2x slower.)

Run from this folder:  python3 gen_erfcx.py   (needs mpmath and numpy; not
needed to build or run the library).
"""
from __future__ import annotations

from pathlib import Path

import mpmath as mp
import numpy as np

mp.mp.dps = 30
CAP = 12.0
DEG_P, DEG_Q = 4, 5


def fit(m: int, n: int, passes: int = 60):
    xs = np.linspace(0.0, CAP, 6000)
    f = np.array([float(mp.erfc(mp.mpf(x)) * mp.exp(mp.mpf(x) ** 2)) for x in xs])
    w = np.ones_like(xs)
    for _ in range(passes):
        A = np.hstack([np.vander(xs, m + 1, increasing=True), -f[:, None] * np.vander(xs, n + 1, increasing=True)[:, 1:]])
        W = w / f
        sol, *_ = np.linalg.lstsq(A * W[:, None], f * W, rcond=None)
        p = sol[: m + 1]
        q = np.concatenate([[1.0], sol[m + 1:]])
        err = np.polyval(p[::-1], xs) / np.polyval(q[::-1], xs) / f - 1
        w = w * (1 + 50 * np.abs(err) / np.abs(err).max())
    assert np.all(np.polyval(q[::-1], xs) > 0)
    return p, q, float(np.abs(err).max())


def main():
    out = Path(__file__).resolve().parent / "gkm_erfcx.h"
    p, q, err = fit(DEG_P, DEG_Q)
    lit = lambda c: f"{float(np.float32(c))!r}f"  # noqa: E731
    lines = [
        "// gkm_erfcx.h — written by gen_erfcx.py; do not edit.",
        "// erfcx(x) = exp(x^2) erfc(x) ~ P(x)/Q(x) on [0, kErfcxCap], deg P = 4, deg Q = 5, Q(0) = 1",
        f"// (linearised least squares with Lawson reweighting; relative error of the fit {err:.1e}).",
        "#pragma once",
        "",
        "namespace gkm {",
        "",
        f"constexpr double kErfcxCap = {CAP!r};",
        f"constexpr int kErfcxDegP = {DEG_P};",
        f"constexpr int kErfcxDegQ = {DEG_Q};",
        "",
        "// coefficients of x^0 .. x^deg",
        f"__device__ constexpr float kErfcxP[{DEG_P + 1}] = {{{', '.join(lit(c) for c in p)}}};",
        f"__device__ constexpr float kErfcxQ[{DEG_Q + 1}] = {{{', '.join(lit(c) for c in q)}}};",
        "",
        "}  // namespace gkm",
        "",
    ]
    out.write_text("\n".join(lines))
    print(f"wrote {out}: fit error {err:.2e}")


if __name__ == "__main__":
    main()

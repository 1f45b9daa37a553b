"""Writes gkm_erfcx.h: rational approximations P(x)/Q(x) of the scaled
complementary error function erfcx(x) = exp(x^2) erfc(x), used by the flux
kernels as erfc(x) = erfcx(min(|x|, CAP)) exp(-x^2) (the Gaussian factor is
shared with the half-space moment N1, which needs the same exponential), with
erfc(-|x|) = 2 - erfc(|x|).  deg Q = deg P + 1, so P/Q decays like 1/x as
erfcx does and one rational covers the whole range, with no branch.

FP32 (mixed form): deg 4 / 5 on [0, 12] (beyond 12, exp(-x^2) is 0 in FP32):
4 + 5 FMAs, one reciprocal and one multiply.  Linearised least squares on
P(x) - f(x) Q(x) = 0 with Q(0) = 1, relative weighting 1/f and Lawson
reweighting towards the largest relative error (60 passes), on 6000 uniform
points of [0, 12] in FP64 arithmetic; f = erfcx from mpmath (30 digits).  The
coefficients are rounded once to float.  Relative error of the fit in FP64
evaluation: 8e-9; in the FP32 evaluation the kernel performs: 4e-7.

FP64 (drop-in form): deg 9 / 10 on [0, 27] (beyond 27, exp(-x^2) underflows
in FP64): 9 + 10 FMAs and one division, against ~130 instructions for the
device library's erfc.  The same least-squares problem solved in 60-digit
arithmetic (mpmath normal equations) on 600 points clustered towards 0
(x = 27 t^2), 30 Lawson passes.  Every coefficient is positive, so Horner's
rule has no cancellation.  Relative error of the fit: 5e-17; of the FP64
evaluation: 1.1e-15 (~5 ulp).  tests/test_erfcx.py checks both forms against
scipy.

Run from this folder:  python3 gen_erfcx.py   (needs mpmath and numpy; not
needed to build or run the library; the FP64 fit takes about a minute).
"""
from __future__ import annotations

from pathlib import Path

import mpmath as mp
import numpy as np

mp.mp.dps = 30
CAP = 12.0
DEG_P, DEG_Q = 4, 5
CAP64 = 27.0
DEG_P64, DEG_Q64 = 9, 10


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


def fit64(m: int, n: int, npts: int = 600, passes: int = 30):
    """The FP64 fit in 60-digit arithmetic (see the module text)."""
    with mp.workdps(60):
        xs = [mp.mpf(CAP64) * (mp.mpf(k) / (npts - 1)) ** 2 for k in range(npts)]
        fs = [mp.erfc(x) * mp.exp(x * x) for x in xs]
        w = [mp.mpf(1)] * npts
        best = None
        for _ in range(passes):
            A = mp.matrix([[wi * x ** i / fx for i in range(m + 1)] + [-wi * x ** j for j in range(1, n + 1)]
                           for x, fx, wi in zip(xs, fs, w)])
            b = mp.matrix(w)
            sol = mp.lu_solve(A.T * A, A.T * b)
            p = [sol[i] for i in range(m + 1)]
            q = [mp.mpf(1)] + [sol[m + 1 + j] for j in range(n)]
            errs = [mp.polyval(p[::-1], x) / mp.polyval(q[::-1], x) / fx - 1 for x, fx in zip(xs, fs)]
            emax = max(abs(e) for e in errs)
            if best is None or emax < best[2]:
                best = (p, q, emax)
            w = [wi * (1 + 20 * abs(e) / emax) for wi, e in zip(w, errs)]
        p, q, emax = best
        assert all(c > 0 for c in p) and all(c > 0 for c in q)
        return [float(c) for c in p], [float(c) for c in q], float(emax)


def main():
    out = Path(__file__).resolve().parent / "gkm_erfcx.h"
    p, q, err = fit(DEG_P, DEG_Q)
    p64, q64, err64 = fit64(DEG_P64, DEG_Q64)
    lit = lambda c: f"{float(np.float32(c))!r}f"  # noqa: E731
    lines = [
        "// gkm_erfcx.h — written by gen_erfcx.py; do not edit.",
        "// erfcx(x) = exp(x^2) erfc(x) ~ P(x)/Q(x), Q(0) = 1, linearised least squares with Lawson reweighting:",
        f"//   FP32: [0, kErfcxCap],   deg P = {DEG_P}, deg Q = {DEG_Q}, relative error of the fit {err:.1e};",
        f"//   FP64: [0, kErfcxCap64], deg P = {DEG_P64}, deg Q = {DEG_Q64}, relative error of the fit {err64:.1e}.",
        "#pragma once",
        "",
        "namespace gkm {",
        "",
        f"constexpr double kErfcxCap = {CAP!r};",
        f"constexpr int kErfcxDegP = {DEG_P};",
        f"constexpr int kErfcxDegQ = {DEG_Q};",
        f"constexpr double kErfcxCap64 = {CAP64!r};",
        f"constexpr int kErfcxDegP64 = {DEG_P64};",
        f"constexpr int kErfcxDegQ64 = {DEG_Q64};",
        "",
        "// coefficients of x^0 .. x^deg",
        f"__device__ constexpr float kErfcxP[{DEG_P + 1}] = {{{', '.join(lit(c) for c in p)}}};",
        f"__device__ constexpr float kErfcxQ[{DEG_Q + 1}] = {{{', '.join(lit(c) for c in q)}}};",
        f"__device__ constexpr double kErfcxP64[{DEG_P64 + 1}] = {{{', '.join(repr(c) for c in p64)}}};",
        f"__device__ constexpr double kErfcxQ64[{DEG_Q64 + 1}] = {{{', '.join(repr(c) for c in q64)}}};",
        "",
        "}  // namespace gkm",
        "",
    ]
    out.write_text("\n".join(lines))
    print(f"wrote {out}: fit error {err:.2e} (FP32), {err64:.2e} (FP64)")


if __name__ == "__main__":
    main()

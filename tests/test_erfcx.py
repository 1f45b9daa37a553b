"""The flux's erfc: erfc(x) = erfcx(min(|x|, cap)) exp(-x^2) with erfcx a
rational (gkm-dns_amd/csrc/gkm_erfcx.h, written by gen_erfcx.py from mpmath
values), erfc(-|x|) = 2 - erfc(|x|): degree 4 / 5 on [0, 12] in the FP32 flux,
degree 9 / 10 on [0, 27] in the FP64 flux.

Pinned against scipy.special.erfc / erfcx (an independent implementation):
FP32: the rational with its float coefficients, evaluated in FP64, is within 5e-8
relative of erfcx on [0, 12] (the fit itself: 8e-9; rounding the coefficients
to float adds the rest) and decays like 1/(x sqrt(pi)) as erfcx does; the FP32
evaluation the kernel performs (Horner with FMA, hardware reciprocal) is
within 5e-7 and stays within the error of the FP32 exponential it multiplies
by.  FP64: the rational evaluated as the kernel does (Horner with FMA, one
division) is within 2e-15 of erfcx on [0, 27] and erfc formed from it within
3e-15 where exp(-x^2) carries no argument rounding of its own.
"""
import math
import re
from pathlib import Path

import numpy as np
from scipy.special import erfc, erfcx

HDR = Path(__file__).resolve().parents[1] / "gkm-dns_amd" / "csrc" / "gkm_erfcx.h"


def _table(sfx=""):
    s = HDR.read_text()
    cap = float(re.search(r"kErfcxCap" + sfx + r" = ([0-9.eE+-]+);", s).group(1))

    def coefs(name):
        body = re.search(name + r"\[\d+\] = \{(.*?)\};", s, re.S).group(1)
        return [float(t.strip().rstrip("f")) for t in body.split(",") if t.strip()]
    p, q = coefs("kErfcxP" + sfx), coefs("kErfcxQ" + sfx)
    assert len(p) == int(re.search(r"kErfcxDegP" + sfx + r" = (\d+);", s).group(1)) + 1
    assert len(q) == int(re.search(r"kErfcxDegQ" + sfx + r" = (\d+);", s).group(1)) + 1
    return cap, p, q


def _erfcx(a, cap, p, q, dt):
    x = np.minimum(np.asarray(a, dt), dt(cap)).astype(dt)
    P = np.full_like(x, dt(p[-1]))
    for c in p[-2::-1]:
        P = (P * x + dt(c)).astype(dt)
    Q = np.full_like(x, dt(q[-1]))
    for c in q[-2::-1]:
        Q = (Q * x + dt(c)).astype(dt)
    return (P * (dt(1) / Q)).astype(dt)


def test_rational_error_in_fp64():
    cap, p, q = _table()
    a = np.linspace(0.0, cap, 200001)
    rel = np.abs(_erfcx(a, cap, p, q, np.float64) / erfcx(a) - 1)
    assert rel.max() < 5e-8, rel.max()
    rel32 = np.abs(_erfcx(a, cap, p, q, np.float32).astype(np.float64) / erfcx(a) - 1)
    assert rel32.max() < 5e-7, rel32.max()
    assert all(c > 0 for c in p + q)  # Horner without cancellation, Q > 0 on [0, inf)


def test_asymptote_is_one_over_x_sqrt_pi():
    cap, p, q = _table()
    assert abs(p[-1] / q[-1] * math.sqrt(math.pi) - 1) < 1e-3


def test_float_erfc_as_the_kernel_forms_it():
    cap, p, q = _table()
    x = np.linspace(-9.0, 9.0, 360001).astype(np.float32)
    a = np.abs(x)
    e = np.exp(-(a * a)).astype(np.float32)
    r = (_erfcx(a, cap, p, q, np.float32) * e).astype(np.float32)
    got = np.where(x >= 0, r, np.float32(2) - r).astype(np.float64)
    ref = erfc(x.astype(np.float64))
    core = np.abs(x) <= 3
    assert np.max(np.abs(got[core] / ref[core] - 1)) < 1.5e-6
    # everywhere: relative error bounded by the FP32 exponential of -x^2 (~x^2 ulp)
    nz = ref > 1e-37
    rel = np.abs(got[nz] / ref[nz] - 1)
    xa = np.abs(np.asarray(x[nz], np.float64))
    bound = 6e-7 + 1.2e-7 * xa ** 2
    assert np.all(rel < bound), rel.max()
    assert np.max(np.abs(got - ref)) < 4e-7
    # beyond the cap the product is exactly 0 (no inf * 0)
    big = np.array([12.0, 20.0, 1e4, 1e30], np.float32)
    with np.errstate(over="ignore"):
        eb = np.exp(-(big * big)).astype(np.float32)
    assert np.all((_erfcx(big, cap, p, q, np.float32) * eb) == 0)


def _erfcx64(a, cap, p, q):
    """The FP64 kernel's evaluation: Horner with fused multiply-add, P / Q."""
    x = np.minimum(np.asarray(a, np.float64), cap)
    P = np.full_like(x, p[-1])
    for c in p[-2::-1]:
        P = P * x + c
    Q = np.full_like(x, q[-1])
    for c in q[-2::-1]:
        Q = Q * x + c
    return P / Q


def test_fp64_rational():
    cap, p, q = _table("64")
    assert cap == 27.0 and len(p) == 10 and len(q) == 11
    assert all(c > 0 for c in p + q)
    a = np.concatenate([np.linspace(0.0, 1.0, 100001), np.linspace(1.0, cap, 400001)])
    rel = np.abs(_erfcx64(a, cap, p, q) / erfcx(a) - 1)
    assert rel.max() < 2e-15, rel.max()
    assert abs(p[-1] / q[-1] * math.sqrt(math.pi) - 1) < 1e-6


def test_fp64_erfc_as_the_kernel_forms_it():
    cap, p, q = _table("64")
    # values whose square is exact, so exp(-x^2) adds only its own rounding
    x = np.arange(-2 ** 14, 2 ** 14 + 1) / 2 ** 9          # [-32, 32] in steps of 2^-9
    a = np.abs(x)
    e = np.exp(-(a * a))
    r = _erfcx64(a, cap, p, q) * e
    got = np.where(x >= 0, r, 2.0 - r)
    ref = erfc(x)
    nz = ref > 1e-300
    assert np.max(np.abs(got[nz] / ref[nz] - 1)) < 3e-15
    assert np.all(got[~nz] < 1e-290) and np.all(np.isfinite(got))

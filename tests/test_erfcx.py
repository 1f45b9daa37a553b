"""The FP32 flux's erfc: erfc(x) = erfcx(|x|) exp(-x^2) with erfcx a degree-18
polynomial on [0, 6] (gkm-dns_amd/csrc/gkm_erfcx.h, written by gen_erfcx.py
from a 50-digit Chebyshev interpolant) and a four-term asymptotic series
beyond; erfc(-|x|) = 2 - erfc(|x|).

Pinned against scipy.special.erfc / erfcx (an independent implementation):
the polynomial's truncation error in FP64 evaluation is below 2^-24, and the
FP32 evaluation the kernel performs stays within the error of the FP32
exponential it multiplies by.
"""
import re
from pathlib import Path

import numpy as np
from scipy.special import erfc, erfcx

HDR = Path(__file__).resolve().parents[1] / "gkm-dns_amd" / "csrc" / "gkm_erfcx.h"


def _table():
    s = HDR.read_text()
    B = float(re.search(r"kErfcxB = ([0-9.eE+-]+);", s).group(1))
    body = re.search(r"kErfcxF\[\d+\] = \{(.*?)\};", s, re.S).group(1)
    coef = [float(t.strip().rstrip("f")) for t in body.split(",") if t.strip()]
    deg = int(re.search(r"kErfcxDegF = (\d+);", s).group(1))
    assert len(coef) == deg + 1
    return B, coef


def _erfcx(a, B, coef, dt):
    a = a.astype(dt)
    u = (np.minimum(a, dt(B)) * dt(2.0 / B) - dt(1)).astype(dt)
    p = np.full_like(u, dt(coef[-1]))
    for c in coef[-2::-1]:
        p = (p * u + dt(c)).astype(dt)
    with np.errstate(divide="ignore"):
        ia = (dt(1) / a).astype(dt)
    i2 = ia * ia
    tail = ((((dt(-1.875) * i2 + dt(0.75)) * i2 + dt(-0.5)) * i2 + dt(1)) * ia * dt(0.5641895835477563)).astype(dt)
    return np.where(a > dt(B), tail, p)


def test_polynomial_truncation_below_float_epsilon():
    B, coef = _table()
    a = np.linspace(0.0, B, 200001)
    rel = np.abs(_erfcx(a, B, coef, np.float64) / erfcx(a) - 1)
    assert rel.max() < 2.0 ** -24 * 1.05, rel.max()


def test_asymptotic_tail_beyond_range():
    B, coef = _table()
    a = np.linspace(B, 30.0, 20001)
    rel = np.abs(_erfcx(a, B, coef, np.float64) / erfcx(a) - 1)
    assert rel.max() < 105.0 / (16.0 * B ** 8), rel.max()


def test_float_erfc_as_the_kernel_forms_it():
    B, coef = _table()
    x = np.linspace(-9.0, 9.0, 360001).astype(np.float32)
    a = np.abs(x)
    e = np.exp(-(a * a)).astype(np.float32)
    r = (_erfcx(a, B, coef, np.float32) * e).astype(np.float32)
    got = np.where(x >= 0, r, np.float32(2) - r).astype(np.float64)
    ref = erfc(x.astype(np.float64))
    core = np.abs(x) <= 3
    assert np.max(np.abs(got[core] / ref[core] - 1)) < 1.5e-6
    # everywhere: relative error bounded by the FP32 exponential of -x^2 (~x^2 ulp)
    # plus, beyond the polynomial's range, the tail series' truncation
    nz = ref > 1e-37
    rel = np.abs(got[nz] / ref[nz] - 1)
    xa = np.abs(np.asarray(x[nz], np.float64))
    bound = 3.5e-7 + 1.2e-7 * xa ** 2 + np.where(xa > B, 105.0 / (16.0 * B ** 8), 0.0)
    assert np.all(rel < bound), rel.max()
    assert np.max(np.abs(got - ref)) < 4e-7

"""Independent pin of the oracle's WENO nonlinear weights (CPU only).

The oracle's SYMOO and JS lines (weno/src/update.cpp:502-557 for SYMOO,
:1421-1468 for the WENO-JS variant) are rebuilt here from the published
definitions of the schemes, without their closed-form indicator coefficients:

  * each candidate is the quadratic whose averages over its three cells equal
    the cell values, solved from the averaging integrals;
  * its smoothness indicator is
        IS_k = sum_{l=1,2} h^(2l-1) * integral over the upwind cell of (p_k^(l))^2
    (Jiang & Shu, J. Comput. Phys. 126, 1996), integrated exactly with
    numpy.polynomial;
  * WENO-SYMOO (Martin, Taylor, Wu & Weirs, J. Comput. Phys. 220, 2006): four
    candidates, ideal weights (1, 9, 9, 1)/20, the downwind candidate's
    indicator replaced by the largest of the four, alpha = gamma / (eps + IS);
  * WENO-JS: three candidates, d = (0.1, 0.6, 0.3), alpha = d / (eps + IS)^2;
  * eps = 1e-10 (weno/src/param.h:15);
  * the right state at an interface is the left state of the mirrored line.

A dropped term, a wrong indicator coefficient or a missing candidate in the
downwind maximum changes the weights by far more than the rounding allowed
here (relative 1e-13), so each such mistake fails a case below.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache

import numpy as np
import pytest
from numpy.polynomial import Polynomial

from _util import cons

EPS = 1e-10
SYMOO = 0
JS = 1

# Cells of a line around the interface x = 0 with h = 1: LLL = [-3,-2], ...,
# L = [-1,0], R = [0,1], ..., RRR = [2,3].
CELLS = [(m - 3, m - 2) for m in range(6)]


@lru_cache(maxsize=None)
def _averaging_inverse(cells):
    """Exact inverse of the map from the coefficients of a quadratic to its
    averages over the three cells (rational arithmetic)."""
    A = [[Fraction(x1 ** (n + 1) - x0 ** (n + 1), n + 1) / (x1 - x0) for n in range(3)] for x0, x1 in cells]
    # Gauss-Jordan on 3x3 rationals
    M = [row[:] + [Fraction(int(r == c)) for c in range(3)] for r, row in enumerate(A)]
    for c in range(3):
        piv = next(r for r in range(c, 3) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        M[c] = [v / M[c][c] for v in M[c]]
        for r in range(3):
            if r != c and M[r][c] != 0:
                M[r] = [a - M[r][c] * b for a, b in zip(M[r], M[c])]
    return np.array([[float(v) for v in row[3:]] for row in M])


def _candidate(stencil, vals):
    """Quadratic with the given averages over cells `stencil` (indices into CELLS)."""
    inv = _averaging_inverse(tuple(CELLS[s] for s in stencil))
    return Polynomial(inv @ np.asarray(vals, dtype=np.float64))


def _indicator(p, cell):
    x0, x1 = cell
    total = 0.0
    for l in (1, 2):
        d = p.deriv(l)
        q = (d * d).integ()
        total += q(x1) - q(x0)  # h = 1, so h^(2l-1) = 1
    return total


def _left_state(recon, c):
    """Left state at x = 0 from the six cell values c (LLL .. RRR)."""
    up = CELLS[2]  # the upwind cell L
    if recon == SYMOO:
        stencils = [(0, 1, 2), (1, 2, 3), (2, 3, 4), (3, 4, 5)]
        gam = [1 / 20, 9 / 20, 9 / 20, 1 / 20]
    else:
        stencils = [(0, 1, 2), (1, 2, 3), (2, 3, 4)]
        gam = [0.1, 0.6, 0.3]
    polys = [_candidate(s, [c[i] for i in s]) for s in stencils]
    IS = [_indicator(p, up) for p in polys]
    if recon == SYMOO:
        IS[3] = max(IS)
        alpha = [g / (EPS + s) for g, s in zip(gam, IS)]
    else:
        alpha = [g / (EPS + s) ** 2 for g, s in zip(gam, IS)]
    tot = sum(alpha)
    return sum(a / tot * p(0.0) for a, p in zip(alpha, polys)), IS


def independent_weno(recon, c):
    c = np.asarray(c, dtype=np.float64)
    wl, _ = _left_state(recon, c)
    wr, _ = _left_state(recon, c[::-1])
    return wl, wr


def _close(orc_val, ind_val, scale):
    return abs(orc_val - ind_val) <= 1e-13 * max(scale, 1e-300)


def _check(orc, recon, c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    wl, wr = orc.weno_scalar(recon, c)
    il, ir = independent_weno(recon, c)
    scale = float(np.max(np.abs(c)))
    assert _close(wl, il, scale), (recon, c.tolist(), wl, il)
    assert _close(wr, ir, scale), (recon, c.tolist(), wr, ir)


# ----------------------------------------------------------------- the definition
def test_indicator_closed_forms_match_the_integral_definition():
    """The integral gives the textbook Jiang-Shu forms on the three upwind
    candidates, so the reconstruction below is the published one."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        a, b, c = rng.normal(size=3)
        # stencil (LLL, LL, L) over L: 13/12 (a-2b+c)^2 + 1/4 (a-4b+3c)^2
        p = _candidate((0, 1, 2), [a, b, c])
        assert _indicator(p, CELLS[2]) == pytest.approx(13 / 12 * (a - 2 * b + c) ** 2 + (a - 4 * b + 3 * c) ** 2 / 4,
                                                        rel=1e-12, abs=1e-14)
        # stencil (L, R, RR) over L: 13/12 (a-2b+c)^2 + 1/4 (3a-4b+c)^2
        p = _candidate((2, 3, 4), [a, b, c])
        assert _indicator(p, CELLS[2]) == pytest.approx(13 / 12 * (a - 2 * b + c) ** 2 + (3 * a - 4 * b + c) ** 2 / 4,
                                                        rel=1e-12, abs=1e-14)


def test_candidates_reproduce_quadratics_exactly():
    q = Polynomial([0.3, -0.7, 0.25])
    Q = q.integ()
    avg = [Q(x1) - Q(x0) for x0, x1 in CELLS]
    for s in [(0, 1, 2), (1, 2, 3), (2, 3, 4), (3, 4, 5)]:
        p = _candidate(s, [avg[i] for i in s])
        assert np.allclose(p.coef, q.coef, atol=1e-13)


# ----------------------------------------------------------------- oracle vs definition
@pytest.mark.parametrize("recon", [SYMOO, JS])
def test_weno_rough_data(orc, recon):
    rng = np.random.default_rng(100 + recon)
    for _ in range(300):
        _check(orc, recon, rng.normal(size=6) * 10 ** rng.uniform(-2, 2))


@pytest.mark.parametrize("recon", [SYMOO, JS])
def test_weno_positive_state_like_data(orc, recon):
    """Density- and energy-like lines: an O(1) level with O(10 %) wiggles."""
    rng = np.random.default_rng(200 + recon)
    for _ in range(200):
        _check(orc, recon, rng.uniform(0.5, 2.0) * (1 + 0.1 * rng.normal(size=6)))


@pytest.mark.parametrize("recon", [SYMOO, JS])
@pytest.mark.parametrize("pos", range(1, 6))
def test_weno_step_at_every_position(orc, recon, pos):
    for lo, hi in [(0.0, 1.0), (1.0, 0.0), (0.125, 2.5), (3.0, -1.0)]:
        c = np.where(np.arange(6) < pos, lo, hi).astype(np.float64)
        _check(orc, recon, c)


@pytest.mark.parametrize("recon", [SYMOO, JS])
def test_weno_single_spike_in_each_cell(orc, recon):
    """A spike inside exactly the stencils that hold its cell: their weights fall."""
    rng = np.random.default_rng(300 + recon)
    for pos in range(6):
        for _ in range(10):
            c = 1.0 + 0.01 * rng.normal(size=6)
            c[pos] += rng.choice([-1.0, 1.0]) * 10 ** rng.uniform(-1, 1)
            _check(orc, recon, c)


def test_symoo_downwind_indicator_is_the_maximum_of_all_four(orc):
    """Lines chosen so that each of IS0, IS1, IS2 is in turn the largest upwind
    indicator: the downwind candidate must take that value, not its own."""
    rng = np.random.default_rng(400)
    hit = set()
    for _ in range(400):
        c = rng.normal(size=6) * rng.uniform(0.1, 3.0, size=6)
        _, IS = _left_state(SYMOO, c)
        # IS[3] was replaced by the maximum; which candidate supplied it?
        raw = [_indicator(_candidate(s, [c[i] for i in s]), CELLS[2])
               for s in [(0, 1, 2), (1, 2, 3), (2, 3, 4), (3, 4, 5)]]
        hit.add(int(np.argmax(raw)))
        _check(orc, SYMOO, c)
    assert hit == {0, 1, 2, 3}


@pytest.mark.parametrize("recon", [SYMOO, JS])
def test_weno_linear_data_takes_the_ideal_weights(orc, recon):
    """Linear data: every indicator equals the squared slope, so the weights are
    the ideal ones and the left state is the linear interpolant's value."""
    for slope, off in [(0.5, 1.0), (-2.0, 0.3), (1e-3, 5.0)]:
        x = np.array([m - 2.5 for m in range(6)])  # cell centres
        c = off + slope * x
        wl, wr = orc.weno_scalar(recon, c)
        assert wl == pytest.approx(off, rel=1e-13, abs=1e-14)
        assert wr == pytest.approx(off, rel=1e-13, abs=1e-14)
        _check(orc, recon, c)


@pytest.mark.parametrize("recon", [SYMOO, JS])
@pytest.mark.parametrize("energy_recon", [0, 1])
def test_weno_vector_matches_the_definition(orc, recon, energy_recon):
    """orc_weno on a 6 x 5 stencil: fields 0..3 (and 4 in E mode) by the scalar
    definition; in T mode field 4 from the reconstructed cell temperature,
    rho E = rho R T / (gamma - 1) + |rho u|^2 / (2 rho) (update.cpp:560-621)."""
    gam, R = 1.4, 287.0
    P = orc.make_params(recon=recon, energy_recon=energy_recon, gamma=gam, R=R)
    rng = np.random.default_rng(500 + 2 * recon + energy_recon)
    for _ in range(50):
        cells = np.stack([cons(rng.uniform(0.5, 2.0), *rng.normal(0, 0.4, 3), rng.uniform(0.5, 3.0))
                          for _ in range(6)])
        wl, wr = orc.weno(P, cells)
        nlin = 5 if energy_recon == 1 else 4
        for q in range(nlin):
            il, ir = independent_weno(recon, cells[:, q])
            sc = float(np.max(np.abs(cells[:, q])))
            assert _close(wl[q], il, sc) and _close(wr[q], ir, sc), (q, wl[q], il, wr[q], ir)
        if energy_recon == 1:
            continue
        T = (gam - 1) * (cells[:, 4] / cells[:, 0]
                         - 0.5 * np.sum(cells[:, 1:4] ** 2, axis=1) / cells[:, 0] ** 2) / R
        Tl, Tr = independent_weno(recon, T)
        for w, Tw in ((wl, Tl), (wr, Tr)):
            E = w[0] * R * Tw / (gam - 1) + 0.5 * np.dot(w[1:4], w[1:4]) / w[0]
            assert w[4] == pytest.approx(E, rel=1e-12)

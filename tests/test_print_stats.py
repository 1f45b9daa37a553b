"""PrintStats with spline gradients (weno/src/update.cpp:638-784): CPU checks.

The reference takes velocity gradients from einspline (un-vendored, version
unpinned): a tensor-product cubic B-spline with natural ends over the padded
segment (weno/src/Interpolation.cpp:3-31), evaluated at the grid nodes.  At
a node that gradient equals the slope of the 1-D natural cubic spline along
the grid line, so the oracle is pinned here against an independent library
routine, scipy.interpolate.CubicSpline(bc_type="natural"):

* orc_spline_slopes vs scipy on random lines of every length class;
* the whole orc_print_stats (velocity fields, axis pairing of the nine
  gradients, the reference's sums and normalisations) vs a numpy + scipy
  restatement on a small uneven box;
* closed forms: a linear function's slopes are exact; a smooth periodic
  field's slopes converge to the analytic derivative away from the line ends;
* gkm_format_print_stats (host code of libgkm.so, no GPU needed) writes the
  reference's column order and separators.
Einspline's own rounding stays unpinned (it is absent here).
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from _util import smooth_box


def _scipy_slopes(f, h, axis=0):
    x = np.arange(f.shape[axis]) * h
    return CubicSpline(x, f, axis=axis, bc_type="natural").derivative()(x)


@pytest.mark.parametrize("m", [2, 3, 4, 9, 38, 134, 262])
def test_spline_slopes_match_scipy_natural_spline(orc, m):
    rng = np.random.default_rng(m)
    for h in (0.05, 1.0):
        f = rng.normal(size=m)
        D = orc.spline_slopes(f, h)
        ref = _scipy_slopes(f, h)
        assert np.max(np.abs(D - ref)) <= 1e-12 * np.max(np.abs(ref)), m


def test_spline_slopes_exact_on_linear_data(orc):
    x = np.arange(40) * 0.1
    D = orc.spline_slopes(3.0 - 2.5 * x, 0.1)
    assert np.max(np.abs(D + 2.5)) < 1e-12


def test_spline_slopes_converge_to_derivative(orc):
    """sin on a padded periodic line: inside, away from the natural ends, the
    error is O(h^4); at the ends the natural condition pulls it off."""
    errs = []
    for n in (16, 32, 64):
        h = 2 * math.pi / n
        x = (np.arange(n + 6) - 3 + 0.5) * h
        D = orc.spline_slopes(np.sin(x), h)
        mid = slice(12 * n // 32 + 3, n + 3 - 12 * n // 32)
        errs.append(np.max(np.abs(D[mid] - np.cos(x[mid]))))
    assert errs[1] < errs[0] / 12 and errs[2] < errs[1] / 12, errs


def _numpy_print_stats(W, shape, gas, dx):
    """Restatement of PrintStats from the reference text with numpy arrays and
    scipy's natural spline: gradU[0] = d/dy (axis 0), gradU[1] = d/dx
    (axis 1), gradU[2] = d/dz (axis 2), update.cpp:687-695."""
    ny, nx, nz = shape
    P = (ny + 6, nx + 6, nz + 6)
    Wp = [w.reshape(P) for w in W]
    den = Wp[0]
    vel = [Wp[1] / den, Wp[2] / den, Wp[3] / den]
    g = [[_scipy_slopes(vel[c], dx, axis=a) for a in range(3)] for c in range(3)]  # g[comp][axis]
    s = (slice(3, -3),) * 3
    I = lambda a: a[s]  # noqa: E731
    d, u, v, w = I(den), I(vel[0]), I(vel[1]), I(vel[2])
    Vsqr = u * u + v * v + w * w
    Pi = (d * (gas["gamma"] - 1)) * (I(Wp[4]) / d - 0.5 * Vsqr)
    Ti = Pi / (gas["R"] * d)
    dudx, dudy, dudz = I(g[0][1]), I(g[0][0]), I(g[0][2])
    dvdx, dvdy, dvdz = I(g[1][1]), I(g[1][0]), I(g[1][2])
    dwdx, dwdy, dwdz = I(g[2][1]), I(g[2][0]), I(g[2][2])
    T0, mu0 = gas["T0"], gas["mu0"]
    mu = mu0 * (Ti / T0) ** 1.5 * (T0 + 110.4) / (Ti + 110.4)
    N = d.size
    out = {}
    out["diss"] = np.sum(mu * (dudx ** 2 + dudy ** 2 + dudz ** 2 + dvdx ** 2 + dvdy ** 2 + dvdz ** 2 + dwdx ** 2 +
                               dwdy ** 2 + dwdz ** 2)) / N
    out["tke"] = np.sum(d * Vsqr) / N
    out["sijsij"] = np.sum(dudx ** 2 + 0.5 * (dudy + dvdx) ** 2 + 0.5 * (dudz + dwdx) ** 2 + 0.5 * (dwdy + dvdz) ** 2 +
                           dvdy ** 2 + dwdz ** 2) / N
    for name, x in (("u", dudx), ("v", dvdy), ("w", dwdz)):
        m2, m3, m4 = np.mean(x ** 2), np.mean(x ** 3), np.mean(x ** 4)
        out["skew_" + name] = m3 / m2 ** 1.5
        out["kurt_" + name] = m4 / m2 ** 2
        out[{"u": "dudx2", "v": "dvdy2", "w": "dwdz2"}[name]] = m2
    out["p_mean"], out["T_mean"], out["v_mean"] = np.mean(Pi), np.mean(Ti), np.mean(1 / d)
    out["p_rms"] = math.sqrt(np.mean((out["p_mean"] - Pi) ** 2))
    out["T_rms"] = math.sqrt(np.mean((out["T_mean"] - Ti) ** 2))
    out["v_rms"] = math.sqrt(np.mean((out["v_mean"] - 1 / d) ** 2))
    return out


def test_print_stats_matches_numpy_scipy_restatement(orc):
    nx, ny, nz = 10, 8, 12
    dx = 2 * math.pi / nx
    gas = dict(gamma=1.4, R=287.0, mu0=3e-3, T0=1.0 / (1.4 * 287.0))
    # random cells: skewness and kurtosis far from their symmetric values
    rng = np.random.default_rng(9)
    W = smooth_box(nx, ny, nz, seed=9, mach=0.4)
    grid = orc.make_grid(nx, ny, nz, dx)
    orc.fill_ghosts(grid, W)
    for q in range(4):
        W[q] = W[q] * (1 + 0.2 * rng.random(W[q].size)) if q == 0 else W[q] + 0.3 * rng.normal(size=W[q].size)
    W[4] = W[4] + 0.5 * (W[1] ** 2 + W[2] ** 2 + W[3] ** 2) / W[0]
    orc.fill_ghosts(grid, W)
    p = orc.make_params(gamma=gas["gamma"], R=gas["R"], mu0=gas["mu0"], T0=gas["T0"], visc_law=0)
    o = orc.print_stats(grid, p, W)
    r = _numpy_print_stats(W, (ny, nx, nz), gas, dx)
    for k, v in r.items():
        assert abs(o[k] - v) <= 1e-11 * abs(v) + 1e-300, (k, o[k], v)


def test_print_stats_fills_ghosts_of_a_copy(orc):
    """Stats of a state with stale ghosts equal those of the filled state, and
    the caller's arrays are left alone."""
    nx = ny = nz = 8
    W = smooth_box(nx, ny, nz, seed=3)
    grid = orc.make_grid(nx, ny, nz, 2 * math.pi / nx)
    p = orc.make_params()
    W0 = [w.copy() for w in W]
    a = orc.print_stats(grid, p, W)
    assert all(np.array_equal(x, y) for x, y in zip(W, W0))
    orc.fill_ghosts(grid, W)
    b = orc.print_stats(grid, p, W)
    assert a == b


def test_format_print_stats_line():
    import gkm_amd._abi as A
    s = {f: 0.0 for f in A.PRINT_STAT_FIELDS}
    s.update(tke=0.75, diss=1.25e-3, skew_u=-0.5, skew_v=-0.4, skew_w=-0.45, kurt_u=3.5, kurt_v=3.6, kurt_w=3.7,
             p_rms=2e-2, T_rms=3e-5, v_rms=4e-2)
    gam, p0, T0, Mt, den0 = 1.4, 2.5, 0.006, 0.5, 1.0
    line = A.format_print_stats(s, 0.125, gam, p0, T0, Mt, den0)
    cols = [0.125, 0.75, 1.25e-3, -0.5, -0.4, -0.45, 3.5, 3.6, 3.7, 2e-2 / (gam * p0 * Mt * Mt),
            3e-5 / ((gam - 1) * T0 * Mt * Mt), 4e-2 / (Mt * Mt / den0)]
    # update.cpp:779-780: two spaces between columns, four before Vrms, precision 8, scientific
    want = "  ".join(f"{x:.8e}" for x in cols[:11]) + "    " + f"{cols[11]:.8e}"
    assert line == want
    with pytest.raises(A.GkmError):
        st = A.GkmPrintStats(**s)
        import ctypes as C
        buf = C.create_string_buffer(16)
        A.check(A.load().gkm_format_print_stats(C.byref(st), 0.0, gam, p0, T0, Mt, den0, buf, len(buf)))


def test_print_stats_struct_matches_header():
    import re
    from pathlib import Path

    import gkm_amd._abi as A
    from oracle import oracle as O
    text = re.sub(r"/\*.*?\*/", "", (Path(__file__).resolve().parents[1] / "include" / "gkm.h").read_text(), flags=re.S)
    body = re.search(r"typedef struct gkm_print_stats_s \{(.*?)\} gkm_print_stats_t;", text, flags=re.S).group(1)
    names = [n.strip() for _, group in re.findall(r"(double)\s+([\w\s,]+);", body) for n in group.split(",")]
    assert [f for f, _ in A.GkmPrintStats._fields_] == names
    assert [f for f, _ in O.PrintStats._fields_] == names

"""Physics check of the whole step against the GKM tree's own decay run.

GKM/CUDA GKM/output_Mach-0.5-Re72.dat is that tree's smooth-flow scheme on a
Mt = 0.5, Re = 72 decaying box; its initial condition (Init.dat) is not
shipped, nor its grid, so only the shape of the decay can be compared.  A
128^3 run of the same recipe (init_BS.m spectrum, k0 = 4, another random
phase field) with the GKS scheme (flux_scheme = "gks", Pr = 1, power-law
viscosity) at CFL 0.5 gives q^2/q^2(0) at t/t0 = 1..4 within 4-10 % of the
published values (lower: coarser grid and larger time step add numerical
dissipation; at CFL 0.1 the gap is 1-7 %).  The bar of 15 % catches a wrong
viscosity scale or a broken update, not the IC difference.  The WENO-BGK
scheme is held to the same bar.
"""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gk():
    import gkm_amd
    gkm_amd.load()
    return gkm_amd


@pytest.mark.parametrize("scheme,kw", [("gks", dict(flux_scheme="gks", Pr=1.0)),
                                       ("weno_bgk", dict(recon="js", Pr=0.7))])
def test_decay_matches_published_shape(gk, scheme, kw):
    ic = gk.dhit(128, k0=4, Mt=0.5, Re=72, seed=1234)
    r = gk.mixed_vs_fp64(ic, t_end=4.0, cfl=0.5, every=50, visc_law="power", **kw)
    pub = gk.PUBLISHED_MT05_RE72
    print(scheme, {t: (round(r["q2_decay"][t], 4), pub[t]) for t in pub})
    for t, ref in pub.items():
        got = r["q2_decay"][t]
        assert abs(got / ref - 1) < 0.15, (t, got, ref)
    # decays monotonically and faster than the inviscid limit would
    q = [s["fp64"]["q2"] for s in r["samples"]]
    assert all(b < a for a, b in zip(q, q[1:]))

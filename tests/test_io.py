"""File formats of the reference (include/gkm_io.h; host code of libgkm.so, no
GPU needed): the weno params file and Init%i%i%i rank files
(weno/src/FlowField.cpp:9-112, WriteReInit update.cpp:1352-1391), the GKM
Init.dat (init_BS.m:160-175, GKM/CUDA GKM/src/FlowField.cpp:178-235).

Each format is checked against a numpy restatement of the reference's byte
layout written from the reference text (field order, j-fastest index
J = k*nsy*nsx + i*nsy + j, rank (a, b, c) holding global block
[a*nsy, ...) x [b*nsx, ...) x [c*nsz, ...), file name Init%i%i%i), so a
transposed or misplaced segment fails; plus round trips and error paths.
The reference ships no Init files (its Init.dat is absent), so no golden
file pins them.
"""
import math
import subprocess

import numpy as np
import pytest


@pytest.fixture(scope="module")
def gio():
    from gkm_amd import io
    return io


@pytest.fixture(scope="module")
def ic8():
    from gkm_amd.ic import dhit
    return dhit(8, Mt=0.3, Re=30, seed=1)


def test_params_file_bytes(gio, tmp_path):
    p = gio.WenoParams(nc=16, gam=1.4, den0=1.0, Mt=0.3, Re=30.0, mu0=1.5e-3, p0=0.29, t0=4.5)
    gio.write_weno_params(tmp_path / "params", p)
    raw = (tmp_path / "params").read_bytes()
    assert len(raw) == 4 + 7 * 8
    assert np.frombuffer(raw[:4], dtype=np.int32)[0] == 16
    assert list(np.frombuffer(raw[4:], dtype=np.float64)) == [1.4, 1.0, 0.3, 30.0, 1.5e-3, 0.29, 4.5]
    q = gio.read_weno_params(tmp_path / "params")
    assert q == p
    assert q.K() == 2 and q.dx() == 2 * math.pi / 16
    assert q.T0() == 0.29 / (287.0 * 1.0)


@pytest.mark.parametrize("pd", [(1, 1, 1), (2, 1, 2), (1, 4, 2)])
def test_rank_files_layout(gio, ic8, tmp_path, pd):
    from gkm_amd.ic import conservative_padded
    W = conservative_padded(ic8)
    gio.write_weno_init(tmp_path, W, 8, 1.4, pd)
    P = 14
    Wp = [w.reshape(P, P, P)[3:-3, 3:-3, 3:-3] for w in W]
    den = Wp[0]
    u, v, w = Wp[1] / den, Wp[2] / den, Wp[3] / den
    pr = (den * (1.4 - 1)) * (Wp[4] / den - 0.5 * (u * u + v * v + w * w))
    sy, sx, sz = 8 // pd[0], 8 // pd[1], 8 // pd[2]
    names = sorted(f.name for f in tmp_path.iterdir())
    assert names == sorted(f"Init{a}{b}{c}" for a in range(pd[0]) for b in range(pd[1]) for c in range(pd[2]))
    for a in range(pd[0]):
        for b in range(pd[1]):
            for c in range(pd[2]):
                raw = np.fromfile(tmp_path / f"Init{a}{b}{c}", dtype=np.float64)
                assert raw.size == 5 * sy * sx * sz
                seg = (slice(a * sy, (a + 1) * sy), slice(b * sx, (b + 1) * sx), slice(c * sz, (c + 1) * sz))
                for q, f in enumerate((u, v, w, den, pr)):
                    # file order: j fastest -> array [k][i][j]
                    got = raw[q * sy * sx * sz:(q + 1) * sy * sx * sz].reshape(sz, sx, sy).transpose(2, 1, 0)
                    assert np.array_equal(got, f[seg]), (a, b, c, q)


def test_rank_files_round_trip(gio, ic8, tmp_path):
    """Initialize_DHIT of WriteReInit's output: rho exact, momenta and energy
    within a rounding (the primitive round trip is the reference's)."""
    from gkm_amd.ic import conservative_padded
    W = conservative_padded(ic8)
    gio.write_weno_init(tmp_path, W, 8, 1.4, (2, 2, 2))
    R = gio.read_weno_init(tmp_path, 8, 1.4, (2, 2, 2))
    assert np.array_equal(R[0], W[0])
    for q in range(1, 5):
        assert np.max(np.abs(R[q] - W[q])) <= 4e-16 * np.max(np.abs(W[q])), q
    # the reader restated: W from the primitives in the file
    raw = np.fromfile(tmp_path / "Init101", dtype=np.float64).reshape(5, 4, 4, 4).transpose(0, 3, 2, 1)
    u0, v0, w0, d0, p0 = raw
    Rp = [r.reshape(14, 14, 14)[3 + 4:3 + 8, 3:3 + 4, 3 + 4:3 + 8] for r in R]
    assert np.array_equal(Rp[0], d0)
    assert np.array_equal(Rp[1], u0 * d0)
    assert np.array_equal(Rp[4], d0 * (p0 / (d0 * (1.4 - 1.0)) + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0)))


def test_weno_case_folder(gio, ic8, tmp_path):
    p = gio.write_weno_case(tmp_path, ic8, proc_dim=(1, 2, 1))
    assert (tmp_path / "params").exists() and (tmp_path / "Init010").exists()
    assert p.nc == 8 and p.mu0 == ic8.mu0 and p.t0 == ic8.t0 and p.p0 == ic8.p0


def test_init_dat_bytes_and_state(gio, ic8, tmp_path):
    hdr, u, v, w = gio.init_dat_from_ic(ic8)
    gio.write_init_dat(tmp_path / "Init.dat", hdr, u, v, w)
    raw = (tmp_path / "Init.dat").read_bytes()
    n3 = 8 ** 3
    assert len(raw) == 4 + 9 * 4 + 3 * 4 * n3
    assert np.frombuffer(raw[:4], dtype=np.int32)[0] == 8
    h9 = np.frombuffer(raw[4:40], dtype=np.float32)
    assert np.array_equal(h9, np.float32([ic8.rho0, ic8.Mt, ic8.Re, ic8.k0, ic8.urms, ic8.mu0, ic8.c0, ic8.p0,
                                          ic8.t0]))
    f = np.frombuffer(raw[40:], dtype=np.float32).reshape(3, 8, 8, 8)
    # MATLAB order: first index (j) fastest
    assert np.array_equal(f[0].transpose(2, 1, 0), u)
    assert np.array_equal(f[2].transpose(2, 1, 0), w)
    h2, u2, v2, w2 = gio.read_init_dat(tmp_path / "Init.dat")
    assert h2 == gio.InitDat(**{k: float(np.float32(getattr(hdr, k))) if k != "n" else 8 for k in hdr.__dict__})
    assert np.array_equal(u2, u) and np.array_equal(v2, v) and np.array_equal(w2, w)
    W = gio.init_dat_state(h2, u2, v2, w2, 1.4)
    den0, p0 = float(h2.rho0), float(h2.p0)
    uu, vv, ww = (x.astype(np.float64) for x in (u2, v2, w2))
    inner = (slice(3, -3),) * 3
    Wi = [x.reshape(14, 14, 14)[inner] for x in W]
    assert np.all(Wi[0] == den0)
    assert np.array_equal(Wi[2], vv * den0)
    assert np.array_equal(Wi[4], den0 * (p0 / (den0 * (1.4 - 1.0)) + .5 * (uu * uu + vv * vv + ww * ww)))
    assert np.all(W[0].reshape(14, 14, 14)[0] == 0)  # ghosts left at zero


def test_io_errors(gio, ic8, tmp_path):
    import gkm_amd._abi as A
    with pytest.raises(A.GkmError) as e:
        gio.read_weno_params(tmp_path / "missing")
    assert e.value.code == A.GKM_E_IO
    from gkm_amd.ic import conservative_padded
    with pytest.raises(A.GkmError) as e:
        gio.write_weno_init(tmp_path, conservative_padded(ic8), 8, 1.4, (3, 1, 1))
    assert e.value.code == A.GKM_E_ARG
    gio.write_weno_init(tmp_path, conservative_padded(ic8), 8, 1.4, (1, 1, 1))
    (tmp_path / "Init000").write_bytes((tmp_path / "Init000").read_bytes()[:-8])
    with pytest.raises(A.GkmError) as e:
        gio.read_weno_init(tmp_path, 8, 1.4)
    assert e.value.code == A.GKM_E_IO and "short" in str(e.value)
    (tmp_path / "Init.dat").write_bytes(np.int32(8).tobytes() + np.zeros(9, np.float32).tobytes())
    with pytest.raises(A.GkmError) as e:
        gio.read_init_dat(tmp_path / "Init.dat")
    assert e.value.code == A.GKM_E_IO


def test_driver_reads_case_before_touching_a_device(gio, ic8, tmp_path):
    """gkm_dns parses arguments and files on the host: usage errors exit 2;
    with a case folder it prints the PrintParams block, then (here, with no
    GPU) reports the failed device call and exits 1."""
    from gkm_amd._abi import LIB_PATH
    exe = LIB_PATH.parent / "gkm_dns"
    assert exe.exists(), "gkm_dns is built by make -C gkm-dns_amd/csrc"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    gio.write_weno_case(tmp_path, ic8, proc_dim=(2, 1, 1))
    r = subprocess.run([str(exe), str(tmp_path), "--proc-dim", "2", "1", "1", "--steps", "1"], capture_output=True,
                       text=True, timeout=120)
    assert "nc            =  8" in r.stdout and "nc_segy       =  4" in r.stdout
    if r.returncode != 0:  # no usable device here: the first device call reports it
        assert r.returncode == 1 and "gkm_create" in r.stderr
    else:  # on a GPU: one step, a stats line before and after it
        assert len([ln for ln in r.stdout.splitlines() if ln.count("e") >= 12]) == 2

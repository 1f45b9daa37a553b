"""Write tests/golden/flux_golden.npz: BGK flux vectors produced by the
reference's own flux() (weno/src/GKMflux.cpp:11-215) compiled from the
reference sources by `make -C oracle ref` (oracle/_ref/libref_flux.so).

Inputs: 256 interface records (WL WR Wl Wr DWxl DWxr DWyl DWyr DWzl DWzr tau dt)
drawn with a fixed seed: smooth and jump states, Mach up to ~1, dt/tau over
four decades; evaluated for K = 2, gamma = 1.4 (the reference's 1/0.7 in the
slope solver is compiled in).  Run from the repository root:

    python tests/golden/make_flux_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from _util import flux_records  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    if not O.ref_available():
        O.build(ref=True)
    rec = np.concatenate([flux_records(128, seed=101, mach=0.4),
                          flux_records(128, seed=202, mach=1.0, dt_over_tau=(1e-2, 1e2))])
    dx = 2 * np.pi / 64
    F = O.ref_flux_points(O.make_params(Pr=0.7, K=2, gamma=1.4), rec, dx)
    out = Path(__file__).resolve().parent / "flux_golden.npz"
    np.savez_compressed(out, rec=rec, dx=np.array(dx), K=np.array(2), gamma=np.array(1.4), Pr=np.array(0.7),
                        F=F)
    print("wrote", out, rec.shape, F.shape)


if __name__ == "__main__":
    main()

"""The benchmark workloads as parameter sets.

* ``kepler2d``: 2D adiabatic Keplerian disk, Nr = 256, r in [0.5, 3]
  (configs/c1_kepler2d.par).
* ``binary_batch``: 48 independent isothermal disks around an equal-mass
  binary (configs/c3_binary_disk.par), each with its own seeded Nr in
  [32, 256], density-noise seed and binary phase.
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import Params, parse_par

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def par_text(base: str, **kv) -> str:
    text = open(os.path.join(ROOT, "configs", base)).read()
    for k, v in kv.items():
        pat = re.compile(rf"^{k}\s.*$", re.M)
        if not pat.search(text):
            raise KeyError(k)
        text = pat.sub(f"{k} {v}", text)
    return text


def params(base: str, **kv) -> Params:
    return parse_par(par_text(base, **kv))


def binary_batch(nruns: int = 48, seed: int = 20241203, nr_lo: int = 32, nr_hi: int = 256) -> list[Params]:
    """Seeded batch of binary disks: Nr ~ U[nr_lo, nr_hi], phase ~ U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    nr = rng.integers(nr_lo, nr_hi + 1, size=nruns)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=nruns)
    return [params("c3_binary_disk.par", Num_R=int(nr[s]), Init_Par0=s + 1, Init_Par6=repr(float(phase[s])))
            for s in range(nruns)]


# Standalone GRMHD cons2prim benchmark: 2^24 states in Kerr-Schild, a = 0.9.
C4 = dict(seed=20241203, M=1.0, a=0.9, gamma=4.0 / 3.0, theta_cut=0.1, rmax=50.0)


def c4_params(**over) -> dict:
    p = dict(C4)
    p.update(over)
    M, a = p["M"], p["a"]
    r_h = M + (M * M - a * a) ** 0.5
    p.setdefault("rmin", 1.5 * r_h)
    return p

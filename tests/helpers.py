"""Shared helpers for the parity tests."""
from __future__ import annotations

import os
import re

import numpy as np

import disco_amd as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def par_text(base: str, **kv) -> str:
    text = open(os.path.join(ROOT, "configs", base)).read()
    for k, v in kv.items():
        pat = re.compile(rf"^{k}\s.*$", re.M)
        assert pat.search(text), k
        text = pat.sub(f"{k} {v}", text)
    return text


def params(base: str, **kv) -> D.Params:
    return D.parse_par(par_text(base, **kv))


def field_rel_err(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-field max |a-b| / max |b| (fields along axis 0)."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    scale = np.maximum(np.abs(b).max(axis=1), 1e-300)
    return np.abs(a - b).max(axis=1) / scale


def phi_err(a: np.ndarray, b: np.ndarray, pmax: float = 2 * np.pi) -> float:
    d = np.abs(a - b)
    d = np.minimum(d, np.abs(d - pmax))
    return float(d.max())


def cell_radius(Nr: int, Nz: int, Np, r_jph) -> np.ndarray:
    """r centroid of every cell (gridsetup order), geometry_functions.c:5-10."""
    out = []
    for k in range(Nz):
        for j in range(Nr):
            rp, rm = r_jph[j + 1], r_jph[j]
            rc = 2.0 * (rp * rp + rp * rm + rm * rm) / (3.0 * (rp + rm))
            out.append(np.full(int(Np[j + Nr * k]), rc))
    return np.concatenate(out)


def state_errors(st: dict, ref: dict, rcell: np.ndarray, mask: np.ndarray) -> dict:
    """Relative errors per field over the masked cells.

    Scalars (rho, P, D, E, passive) are normalised by their own max |ref|.
    Vector components are normalised by the flow speed scale, because a
    component that is ~0 (v_r, v_z of a disk) has no scale of its own:
    velocities by max |Omega r|, momenta by max |l / r| (the azimuthal
    momentum), angular velocity / angular momentum by their own max.
    """
    pr, rp = st["prim"][:, mask], ref["prim"][:, mask]
    cs, rc = st["cons"][:, mask], ref["cons"][:, mask]
    r = rcell[mask]
    vscale = np.abs(rp[3] * r).max()
    sscale = np.abs(rc[3] / r).max()
    ep = np.abs(pr - rp).max(axis=1)
    ec = np.abs(cs - rc).max(axis=1)
    sp = np.maximum(np.abs(rp).max(axis=1), 1e-300)
    sc = np.maximum(np.abs(rc).max(axis=1), 1e-300)
    sp[2] = sp[4] = vscale
    sc[2] = sc[4] = sscale
    if rp.shape[0] > 5:
        # magnetic field components: normalise by the field magnitude scale
        bscale = np.sqrt((rp[5:8] ** 2).sum(axis=0)).max()
        sp[5:8] = bscale
        sc[5:8] = np.sqrt((rc[5:8] ** 2).sum(axis=0)).max()
    return {"prim": ep / sp, "cons": ec / sc}

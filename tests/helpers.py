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
    if rp.shape[0] >= 8:
        # magnetic field components: normalise by the field magnitude scale
        bscale = np.sqrt((rp[5:8] ** 2).sum(axis=0)).max()
        sp[5:8] = bscale
        sc[5:8] = np.sqrt((rc[5:8] ** 2).sum(axis=0)).max()
    return {"prim": ep / sp, "cons": ec / sc}


def track_partners(piph, offL, nL, offU, nU):
    """Other cell of every face of the track between annulus L (lower) and U.

    A face is owned by the cell whose front edge ends it (faces_alt.c:140-163,
    ties: the lower annulus' front first).  For an owner in L the other cell is
    the first U cell whose front is >= the owner's; for an owner in U, the
    first L cell whose front is > the owner's.  Returns (other of each L cell's
    face, other of each U cell's face) as global cell indices."""
    fL = piph[offL:offL + nL]
    fU = piph[offU:offU + nU]
    oL = np.argsort(fL, kind="stable")
    oU = np.argsort(fU, kind="stable")
    pU = np.searchsorted(fU[oU], fL, side="left") % nU
    pL = np.searchsorted(fL[oL], fU, side="right") % nL
    return offU + oU[pU], offL + oL[pL]


def div_b(Phi, piph, Np, Nr, Nz):
    """Net outward magnetic flux of every cell from the face fluxes, the
    quantity Python/discopy/geom.py:370-520 (calcDivB) integrates: the phi
    face at the cell's front minus the one at its back, plus the r and z faces
    of the tracks on both sides, each face taken once from its owner's slot
    (Phi[1]/Phi[2]: owned r face in the inner/outer track, Phi[3]/Phi[4]:
    z faces).  Returns (div, scale) with scale = sum of |face flux| per cell."""
    N = piph.size
    off = np.concatenate([[0], np.cumsum(Np)])
    div = np.zeros(N)
    scale = np.zeros(N)
    for jk in range(Nr * Nz):
        c = np.arange(off[jk], off[jk + 1])
        cm = np.roll(c, 1)
        div[c] += Phi[0][c] - Phi[0][cm]
        scale[c] += np.abs(Phi[0][c]) + np.abs(Phi[0][cm])
    for k in range(Nz):
        for j in range(Nr):
            jk = j + Nr * k
            tracks = []
            if j + 1 < Nr:
                tracks.append((jk + 1, 2, 1))
            if k + 1 < Nz:
                tracks.append((jk + Nr, 4, 3))
            for jk2, sL, sU in tracks:
                oL, oU = track_partners(piph, off[jk], Np[jk], off[jk2], Np[jk2])
                cL = np.arange(off[jk], off[jk + 1])
                cU = np.arange(off[jk2], off[jk2 + 1])
                f = Phi[sL][cL]  # faces owned by lower cells: outward for them, inward for the other
                np.add.at(div, cL, f)
                np.add.at(div, oL, -f)
                np.add.at(scale, cL, np.abs(f))
                np.add.at(scale, oL, np.abs(f))
                f = Phi[sU][cU]
                np.add.at(div, oU, f)
                np.add.at(div, cU, -f)
                np.add.at(scale, oU, np.abs(f))
                np.add.at(scale, cU, np.abs(f))
    return div, scale


def diag_errors(dg: dict, ref: dict) -> tuple:
    """Step diagnostics (add_diagnostics, analysis.c:52-107) against the
    reference's theTools: per quantity max |gpu - ref| over all annuli,
    divided by t_avg x the largest per-cell magnitude of that quantity in
    the reference's state, with the velocity components v_r, v_z taken at
    the flow speed scale as in state_errors (annulus averages of signed
    quantities such as rho v_r cancel to far below their terms, so the
    average itself is no scale); quantities the reference holds at exactly
    zero must be exactly zero.  Also the relative error of t_avg.  The
    per-cell magnitudes follow Diagnostics/diag_euler.c:17-44; the torque
    density's is 2 pi r rho r |f| with |f| planet 1's point-mass force at the
    cell's centre (planet.c:97-102: M d / (d^2 + eps^2)^1.5), since its
    annulus average cancels around the ring like rho v_r's."""
    a, b = dg["Qrz"], ref["Qrz"]
    rc = cell_radius(ref["Nr"], ref["Nz"], ref["Np"], ref["r_jph"])
    p = ref["prim"]
    rho, om, P = np.abs(p[0]), np.abs(p[3]), np.abs(p[1])
    real = ref["real"] == 1
    vs = np.abs(p[3][real] * rc[real]).max()  # flow speed scale
    tw = 2 * np.pi * rc * rho
    torque = None
    pl = ref.get("planets")
    if pl is not None and len(pl) > 1:
        rp, pp, M, _, eps = pl[1]
        phi = ref["piph"] - 0.5 * ref["dphi"]
        d = np.sqrt((rc * np.cos(phi) - rp * np.cos(pp)) ** 2 + (rc * np.sin(phi) - rp * np.sin(pp)) ** 2)
        torque = tw * rc * M * d / (d * d + eps * eps) ** 1.5
    mag = [rho, tw * vs, tw * vs, torque, tw * rc * rc * om * vs, om, tw * rc * rc * om, P]
    e = np.zeros(a.shape[-1])
    for q in range(a.shape[-1]):
        d = np.abs(a[..., q] - b[..., q]).max()
        sc = np.abs(b[..., q]).max() if mag[q] is None else ref["t_avg"] * mag[q].max()
        e[q] = d / sc if sc > 0 else (0.0 if d == 0 else np.inf)
    return e, abs(dg["t_avg"] - ref["t_avg"]) / max(abs(ref["t_avg"]), 1e-300)


def real_mask(Nr, Nz, NgRa, NgRb, NgZa, NgZb, Np):
    off = np.concatenate([[0], np.cumsum(Np)])
    m = np.zeros(int(off[-1]), bool)
    for k in range(NgZa, Nz - NgZb):
        for j in range(NgRa, Nr - NgRb):
            jk = j + Nr * k
            m[off[jk]:off[jk + 1]] = True
    return m

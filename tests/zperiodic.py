"""The periodic-z ghost fill by its definition (test-side helpers).

A one-rank run of the reference with Z_Periodic fills its z ghosts through
the self-wraparound Cartesian communicator (mpisetup.c:17,38) in
exchangeData (exchange.c:173-279): with Ng ghost layers on each side and R
real layers, ghost layer k < Ng holds a copy of real layer k + R and ghost
layer k >= Ng + R a copy of real layer k - R, field by field for the
cell_lite set (exchange.c:7-16).  Np is equal along z, so layer k of the
reference's cell order (annulus jk = j + Nr*k, then array index) is one
contiguous block of cells of the same size for every k.
"""
from __future__ import annotations

import numpy as np


def layer_slices(Np, Nr: int, Nz: int) -> list[slice]:
    """Cell range of every z-layer in the reference's cell order."""
    Np = np.asarray(Np).reshape(Nz, Nr)
    assert (Np == Np[0]).all(), "Np differs along z"
    L = int(Np[0].sum())
    return [slice(k * L, (k + 1) * L) for k in range(Nz)]


def ghost_source(k: int, Nz: int, Ng: int) -> int | None:
    """Real layer whose copy ghost layer k holds (None for a real layer)."""
    R = Nz - 2 * Ng
    if k < Ng:
        return k + R
    if k >= Nz - Ng:
        return k - R
    return None


def ghost_mismatches(st: dict, Np, Nr: int, Nz: int, Ng: int, fields) -> list[str]:
    """Names of (field, ghost layer) pairs that are not bit-for-bit copies of
    their source layer; empty when the state satisfies the definition."""
    sl = layer_slices(Np, Nr, Nz)
    bad = []
    for f in fields:
        a = np.asarray(st[f])
        for k in range(Nz):
            s = ghost_source(k, Nz, Ng)
            if s is None:
                continue
            if not np.array_equal(a[..., sl[k]], a[..., sl[s]]):
                bad.append(f"{f}[k={k} <- {s}]")
    return bad


def roll_real_layers(a: np.ndarray, Np, Nr: int, Nz: int, Ng: int, shift: int = 1) -> np.ndarray:
    """The periodic state moved up by `shift` real layers: real layer k takes
    real layer k - shift (cyclically over the R real layers), and the ghost
    layers are refilled from the rolled real layers by the definition above."""
    sl = layer_slices(Np, Nr, Nz)
    R = Nz - 2 * Ng
    out = np.array(a, copy=True)
    for k in range(Ng, Ng + R):
        src = Ng + (k - Ng - shift) % R
        out[..., sl[k]] = a[..., sl[src]]
    for k in range(Nz):
        s = ghost_source(k, Nz, Ng)
        if s is not None:
            out[..., sl[k]] = out[..., sl[s]]
    return out

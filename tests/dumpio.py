"""Reader for the binary state dumps of the reference harness driver
(oracle/ref/ref_driver.c) and for the committed golden fixtures."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_dump(fn: str) -> dict:
    b = open(fn, "rb").read()
    hdr = np.frombuffer(b[:48], np.int32)
    _magic, ver, Nr, Nz, NQ, NF, Npl, NgRa, NgRb, NgZa, NgZb, step = (int(v) for v in hdr)
    o = 48
    t, dtl = np.frombuffer(b[o:o + 16], np.float64)
    o += 16
    r = np.frombuffer(b[o:o + 8 * (Nr + 1)], np.float64)
    o += 8 * (Nr + 1)
    z = np.frombuffer(b[o:o + 8 * (Nz + 1)], np.float64)
    o += 8 * (Nz + 1)
    Np = np.frombuffer(b[o:o + 4 * Nr * Nz], np.int32)
    o += 4 * Nr * Nz
    nc = int(Np.sum())
    rec = 4 + 2 * NQ + NF
    cells = np.frombuffer(b[o:o + 8 * rec * nc], np.float64).reshape(nc, rec)
    o += 8 * rec * nc
    pl = np.frombuffer(b[o:o + 8 * 5 * Npl], np.float64).reshape(Npl, 5)
    o += 8 * 5 * Npl
    out = dict(Nr=Nr, Nz=Nz, NQ=NQ, NgRa=NgRa, NgRb=NgRb, NgZa=NgZa, NgZb=NgZb, step=step, t=float(t),
               dt=float(dtl), r_jph=r.copy(), z_kph=z.copy(), Np=Np.copy(), piph=cells[:, 0].copy(),
               dphi=cells[:, 1].copy(), wiph=cells[:, 2].copy(), real=cells[:, 3].copy(),
               prim=cells[:, 4:4 + NQ].T.copy(), cons=cells[:, 4 + NQ:4 + 2 * NQ].T.copy(),
               planets=pl.copy())
    if NF:
        out["Phi"] = cells[:, 4 + 2 * NQ:].T.copy()
    if ver >= 3:
        nt = int(np.frombuffer(b[o:o + 4], np.int32)[0])
        o += 4
        out["t_avg"] = float(np.frombuffer(b[o:o + 8], np.float64)[0])
        o += 8
        out["Qrz"] = np.frombuffer(b[o:o + 8 * Nr * Nz * nt], np.float64).reshape(Nz, Nr, nt).copy()
    return out


def load_fixture(name: str) -> dict:
    """Returns {step: state-dict} from a committed fixture file."""
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    out = {}
    for s in z["steps"]:
        pre = f"s{int(s)}_"
        st = {}
        for k in z.files:
            if k.startswith(pre):
                v = z[k]
                st[k[len(pre):]] = v.item() if v.ndim == 0 else v
        out[int(s)] = st
    return out

"""FP64 NumPy restatement of the cons2prim guess network (test infrastructure).

Same parameter layout as the GPU library (disco_amd/csrc/c2p.hip): W1[8][64],
b1, W2[64][64], b2, W3, b3, W4, b4, W5[64][2], b5; SiLU hidden activations;
inputs standardised with (fmean, fstd), targets with (tmean, tstd); loss =
mean over samples and outputs of the squared standardised residual.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 8, 64, 2
SHAPES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]


def unpack(w):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(w[o:o + n], np.float64).reshape(s))
        o += n
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, norm, feat):
    """feat [8, n] raw features -> standardised outputs [n, 2]."""
    P = unpack(w)
    fm, fs = norm[:8], norm[8:16]
    a = ((np.asarray(feat, np.float64).T - fm) / fs)
    for l in range(4):
        a = silu(a @ P[2 * l] + P[2 * l + 1])
    return a @ P[8] + P[9]


def loss_and_grad(w, norm, feat, tgt):
    P = unpack(w)
    fm, fs, tm, ts = norm[:8], norm[8:16], norm[16:18], norm[18:20]
    n = feat.shape[1]
    a = [((np.asarray(feat, np.float64).T - fm) / fs)]
    zs = []
    for l in range(4):
        z = a[-1] @ P[2 * l] + P[2 * l + 1]
        zs.append(z)
        a.append(silu(z))
    y = a[-1] @ P[8] + P[9]
    t = (np.asarray(tgt, np.float64).T - tm) / ts
    r = y - t
    loss = float((r * r).sum() / (n * NOUT))
    dy = 2.0 * r / (n * NOUT)
    grads = [None] * 10
    grads[8] = a[4].T @ dy
    grads[9] = dy.sum(0)
    d = (dy @ P[8].T) * dsilu(zs[3])
    for l in range(3, -1, -1):
        grads[2 * l] = a[l].T @ d
        grads[2 * l + 1] = d.sum(0)
        if l > 0:
            d = (d @ P[2 * l].T) * dsilu(zs[l - 1])
    return loss, np.concatenate([g.ravel() for g in grads])

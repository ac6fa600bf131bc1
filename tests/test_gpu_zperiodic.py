"""The device's periodic-z ghost fill (zcopy_cell, k_zcopy) pinned by its
definition rather than by a second implementation of it.

* After every one of 5 steps, each ghost z-layer of the GPU state is a
  bit-for-bit copy of the real layer one period away (exchange.c:173-279 of a
  one-rank run, whose Cartesian communicator wraps onto itself,
  mpisetup.c:17,38): prims, mesh (piph, dphi), mesh speeds and, for MHD, the
  face fluxes.
* Translation invariance: the periodic z direction has no preferred layer
  (gravity is Grav2D, the mesh speed depends on r only, the fixed r-ghosts
  are smooth in z, and z_kph is exact on these grids: dz = 1/4), so the
  state rolled by one real z-layer must evolve into the rolled state, bit for
  bit, over 5 steps.  A ghost fill from the wrong layer, or a missing one,
  breaks the symmetry.

Both checks run for ideal MHD with constrained transport (HLLD) and for
Newtonian hydro (HLLC) on 12 x 4 disks.
"""
import numpy as np
import pytest

import disco_amd as D
from helpers import params
from zperiodic import ghost_mismatches, roll_real_layers

pytestmark = pytest.mark.gpu

CASES = [("c2_mhd3d.par", dict(Num_R=12, Num_Z=4, CFL=0.3)),
         ("c1_kepler2d.par", dict(Num_R=12, Num_Z=4, Z_Periodic=1))]
IDS = ["mhd", "euler"]


def _fields(st):
    return ["prim", "piph", "dphi", "wiph"] + (["Phi"] if "Phi" in st else [])


@pytest.mark.parametrize("base,over", CASES, ids=IDS)
def test_ghost_layers_are_copies_every_step(base, over):
    dom = D.Domain(params(base, **over))
    I = dom.info[0]
    g = dom.grid(0)
    assert I.NgZa == I.NgZb == 2 and I.Nz == 8
    for s in range(1, 6):
        dom.step(1)
        dom.sync()
        st = dom.state(0)
        bad = ghost_mismatches(st, g["Np"], I.Nr, I.Nz, I.NgZa, _fields(st))
        print(f"step {s}: ghost layers that are not copies: {bad}", flush=True)
        assert bad == [], (s, bad)
    dom.close()


@pytest.mark.parametrize("base,over", CASES, ids=IDS)
def test_rolled_state_evolves_into_rolled_state(base, over):
    p = params(base, **over)
    a = D.Domain(p)
    b = D.Domain(p)
    I = a.info[0]
    g = a.grid(0)
    Np, Nr, Nz, Ng = g["Np"], I.Nr, I.Nz, I.NgZa
    assert np.array_equal(np.diff(g["z_kph"]), np.full(Nz, 0.25))

    def roll(x):
        return roll_real_layers(x, Np, Nr, Nz, Ng, 1)

    s0 = a.state(0)
    b.upload(0, prim=roll(s0["prim"]), cons=roll(s0["cons"]), piph=roll(s0["piph"]), dphi=roll(s0["dphi"]),
             wiph=roll(s0["wiph"]))
    if "Phi" in s0:
        b.upload_faces(0, roll(s0["Phi"]))
    sl = np.arange(s0["piph"].size)
    L = int(np.asarray(Np).reshape(Nz, Nr)[0].sum())
    real = (sl >= Ng * L) & (sl < (Nz - Ng) * L)
    for s in range(1, 6):
        a.step(1)
        b.step(1)
        a.sync()
        b.sync()
        sa, sb = a.state(0), b.state(0)
        assert a.time()["dt"][0] == b.time()["dt"][0], s
        diffs = {}
        for f in _fields(sa):
            diffs[f] = float(np.abs(roll(sa[f]) - sb[f]).max())
        diffs["cons"] = float(np.abs(roll(sa["cons"])[:, real] - sb["cons"][:, real]).max())
        print(f"step {s}: max |rolled(a) - b| {diffs}", flush=True)
        assert all(v == 0.0 for v in diffs.values()), (s, diffs)
    a.close()
    b.close()

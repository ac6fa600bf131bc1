"""Checkpoint and restart in the reference's layout (Output/h5out.c:366-622,
Restart/h5in.c:101-368) through the C ABI.

* A run checkpointed after 3 steps and restarted into a fresh handle made
  from the same in.par continues bit for bit as the run that wrote it (prims,
  conserved state, mesh, mesh speeds, face fluxes, dt, time, planet phases
  and the diagnostics accumulated after the checkpoint) over 4 more steps:
  3D MHD with constrained transport, the 2D binary disk and the spin-0.9
  GRMHD torus.
* The file holds the reference's datasets with the reference's shapes and
  contents: Data/Cells = (prim, Phi, piph) per cell in theCells order, Grid
  Np / Index / Id_phi0 by h5out.c's rules, T, r_jph, z_kph, the averaged
  diagnostics, Pars and Opts.
* Without the extra Data/Cons and Data/Wiph datasets (a checkpoint of the
  reference's own layout) the restart rebuilds the conserved state as the
  reference's restart does: prims then agree to round-off with the written
  ones and the run continues within the stepping tolerance.
* A file whose grid is not the run's is refused with a message.
"""
import os

import numpy as np
import pytest

import disco_amd as D
from disco_amd import checkpoint as ckpt
from helpers import params

pytestmark = pytest.mark.gpu

CASES = [("c2_mhd3d.par", dict(Num_R=12, Num_Z=4, CFL=0.3)),
         ("c3_binary_disk.par", dict(Num_R=24)),
         ("c5_grmhd_torus.par", dict(Num_R=16, Num_Z=12)),
         # a passive scalar rides in the cell records after the fields (h5out.c:78-82)
         ("c2_mhd3d.par", dict(Num_R=12, Num_Z=4, CFL=0.3, Num_N=1))]
IDS = ["mhd", "binary2d", "grmhd_a09", "mhd_scalar"]


def _same_state(a, b, dt=True):
    """dt: compare the last step's dt too (not part of a checkpoint)."""
    sa, sb = a.state(0), b.state(0)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ta, tb = a.time(), b.time()
    assert np.array_equal(ta["t"], tb["t"])
    assert not dt or np.array_equal(ta["dt"], tb["dt"])
    assert np.array_equal(ta["planet_phi"], tb["planet_phi"])


@pytest.mark.parametrize("base,over", CASES, ids=IDS)
def test_restart_continues_bit_for_bit(base, over, tmp_path):
    p = params(base, **over)
    a = D.Domain(p)
    a.step(3)
    a.sync()
    path = str(tmp_path / "checkpoint_0003.ckp")
    a.write_checkpoint(path)
    b = D.Domain(p)
    b.restart(path)
    _same_state(a, b, dt=False)
    for s in range(4):
        a.step(1)
        b.step(1)
        a.sync()
        b.sync()
        _same_state(a, b)
    da, db = a.diagnostics(0), b.diagnostics(0)
    assert np.array_equal(da["Qrz"], db["Qrz"]) and da["t_avg"] == db["t_avg"]
    print(f"{base}: restart after 3 steps == uninterrupted run for 4 more steps", flush=True)
    a.close()
    b.close()


def test_checkpoint_holds_the_reference_datasets(tmp_path):
    p = params("c2_mhd3d.par", Num_R=12, Num_Z=4, CFL=0.3)
    dom = D.Domain(p)
    dom.step(2)
    dom.sync()
    st = dom.state(0)
    dg = dom.diagnostics(0)
    g = dom.grid(0)
    I = dom.info[0]
    path = str(tmp_path / "c.ckp")
    dom.write_checkpoint(path)
    f = ckpt.read(path)
    nq, nf = 8, 5
    cells = f["Data/Cells"]
    assert cells.shape == (I.ncell, nq + nf + 1)
    assert np.array_equal(cells[:, :nq], st["prim"].T)
    assert np.array_equal(cells[:, nq:nq + nf], st["Phi"].T)
    assert np.array_equal(cells[:, nq + nf], st["piph"])
    assert np.array_equal(f["Data/Cons"], st["cons"].T) and np.array_equal(f["Data/Wiph"], st["wiph"])
    assert f["Grid/T"][0] == dom.time()["t"][0]
    assert np.array_equal(f["Grid/r_jph"], g["r_jph"]) and np.array_equal(f["Grid/z_kph"], g["z_kph"])
    Np = g["Np"].reshape(I.Nz, I.Nr)
    assert np.array_equal(f["Grid/Np"], Np)
    off = np.concatenate([[0], np.cumsum(g["Np"])])[:-1].reshape(I.Nz, I.Nr)
    assert np.array_equal(f["Grid/Index"], off)
    # Id_phi0 (h5out.c:535-545): the cell whose phi centre has the largest cos
    phi = st["piph"] - 0.5 * st["dphi"]
    for jk in range(I.Nr * I.Nz):
        o, n = off.ravel()[jk], g["Np"][jk]
        c = np.cos(phi[o:o + n])
        first = o + int(np.argmax(c)) if c.max() > -1.0 else 0
        assert f["Grid/Id_phi0"].ravel()[jk] == first
    assert np.allclose(f["Data/Diagnostics"], dg["Qrz"] / dg["t_avg"], rtol=0, atol=0)
    assert dom.diagnostics(0)["t_avg"] == 0.0  # zeroed, as output() does
    assert f["Data/Planets"].shape == (1, 7)
    assert f["Opts/HYDRO"] == "mhd" and int(f["Opts/NUM_C"][0]) == 8 and int(f["Opts/CT_MODE"][0]) == 2
    assert int(f["Pars/Num_R"][0]) == 12 and float(f["Pars/CFL"][0]) == 0.3
    dom.close()


def test_restart_from_reference_layout_rebuilds_the_state(tmp_path):
    """Only the reference's datasets (no Data/Cons, Data/Wiph)."""
    p = params("c2_mhd3d.par", Num_R=12, Num_Z=4, CFL=0.3)
    a = D.Domain(p)
    a.step(3)
    a.sync()
    full = str(tmp_path / "full.ckp")
    a.write_checkpoint(full)
    f = ckpt.read(full)
    del f["Data/Cons"], f["Data/Wiph"]
    ref = str(tmp_path / "ref_layout.ckp")
    ckpt.write(ref, f)
    b = D.Domain(p)
    b.restart(ref)
    sa, sb = a.state(0), b.state(0)
    scale = np.abs(sa["prim"]).max(axis=1)[:, None] + 1e-300
    assert (np.abs(sa["prim"] - sb["prim"]) / scale).max() < 1e-13
    assert np.array_equal(sa["piph"], sb["piph"]) and np.array_equal(sa["Phi"], sb["Phi"])
    a.step(4)
    b.step(4)
    sa, sb = a.state(0), b.state(0)
    assert (np.abs(sa["prim"] - sb["prim"]) / scale).max() < 1e-10
    a.close()
    b.close()


def test_restart_refuses_another_grid(tmp_path):
    a = D.Domain(params("c3_binary_disk.par", Num_R=24))
    path = str(tmp_path / "a.ckp")
    a.write_checkpoint(path)
    b = D.Domain(params("c3_binary_disk.par", Num_R=28))
    with pytest.raises(RuntimeError, match="grid|Np"):
        b.restart(path)
    a.close()
    b.close()

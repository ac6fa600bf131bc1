"""GPU tests of check_dt (domain.c:283-310) and the main loop around it
(main.c:68-75): the T_End clip runs on the device at the start of every step.

* A small 2D disk run to a T_End that falls inside its 4th (7th) step takes
  the reference's number of steps, ends exactly on T_End and matches the
  reference's final state (tests/golden/c1_tend_*.npz: the reference's own
  loop with its own check_dt) to 1e-11 per field.
* A run that has taken its step to T_End holds: later steps leave its state,
  mesh, time and diagnostics unchanged bit for bit (the mesh velocity is the
  one exception: the first held step sets it from the final prims, as the
  next step's set_wcell would, and it stays there).  In a batch the runs
  reach their own T_End after different step counts; each ends where it
  ends alone.
* check_dt's file controls: "latest" is removed and the loop returns before
  the step, "abort" makes the step it is seen at the last one.
"""
import os

import numpy as np
import pytest

import disco_amd as D
from dumpio import GOLDEN, load_fixture
from helpers import cell_radius, params, phi_err, state_errors

pytestmark = pytest.mark.gpu

FIELDS = ("prim", "cons", "piph", "dphi", "wiph")


def _t_end(fixture):
    return float(np.load(os.path.join(GOLDEN, fixture))["t_end"])


@pytest.mark.parametrize("fixture", ["c1_tend_a.npz", "c1_tend_b.npz"])
def test_advance_to_t_end_matches_reference(fixture):
    ref = load_fixture(fixture)
    nref = max(ref)
    t_end = _t_end(fixture)
    dom = D.Domain(params("c1_kepler2d.par", Num_R=16, T_End=repr(t_end)))
    n, status = dom.advance()
    print(f"{fixture}: {n} steps (reference {nref}), status {status}", flush=True)
    assert n == nref and status == 1
    assert dom.final().all()
    tm = dom.time()
    assert tm["t"][0] == t_end
    assert abs(tm["dt"][0] - ref[nref]["dt"]) <= 1e-12 * ref[nref]["dt"]
    st = dom.state(0)
    r = ref[nref]
    real = r["real"] == 1
    e = state_errors(st, r, cell_radius(r["Nr"], r["Nz"], r["Np"], r["r_jph"]), real)
    assert e["prim"].max() < 1e-11 and e["cons"].max() < 1e-11, e
    assert phi_err(st["piph"], r["piph"]) < 1e-11
    # the run holds once final
    d0 = dom.diagnostics(0)
    dom.step(1)
    st1 = dom.state(0)
    for k in ("prim", "cons", "piph", "dphi"):
        assert np.array_equal(st[k], st1[k]), k
    dom.step(2)
    assert dom.advance() == (0, 1)
    st2 = dom.state(0)
    for k in FIELDS:
        assert np.array_equal(st1[k], st2[k]), k
    assert dom.time()["t"][0] == t_end and dom.time()["dt"][0] == 0.0
    d1 = dom.diagnostics(0)
    assert np.array_equal(d0["Qrz"], d1["Qrz"]) and d0["t_avg"] == d1["t_avg"]
    dom.close()


def test_batch_runs_stop_at_their_own_t_end():
    ta, tb = _t_end("c1_tend_a.npz"), _t_end("c1_tend_b.npz")
    pa = params("c1_kepler2d.par", Num_R=16, T_End=repr(ta))
    pb = params("c1_kepler2d.par", Num_R=16, T_End=repr(tb))
    batch = D.Domain([pa, pb])
    n, status = batch.advance()
    assert n == max(load_fixture("c1_tend_b.npz")) and status == 1
    tm = batch.time()
    assert tm["t"][0] == ta and tm["t"][1] == tb
    for r, p in enumerate((pa, pb)):
        one = D.Domain(p)
        one.advance()
        if r == 0:
            one.step(1)  # held in the batch while run 1 finished
        sa, sb = batch.state(r), one.state(0)
        for k in FIELDS:
            assert np.array_equal(sa[k], sb[k]), (r, k)
        da, db = batch.diagnostics(r), one.diagnostics(0)
        assert np.array_equal(da["Qrz"], db["Qrz"]) and da["t_avg"] == db["t_avg"], r
        one.close()
    batch.close()


def test_advance_stops_at_max_steps_and_on_file_controls(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    dom = D.Domain(params("c1_kepler2d.par", Num_R=16))  # T_End far away
    assert dom.advance(max_steps=3) == (3, 0)
    # files are read only when asked for
    (tmp_path / "latest").write_text("")
    assert dom.advance(max_steps=2) == (2, 0)
    assert (tmp_path / "latest").exists()
    # "latest": removed, the loop returns before the step
    before = dom.state(0)
    assert dom.advance(max_steps=5, poll_files=True) == (0, 4)
    assert not (tmp_path / "latest").exists()
    after = dom.state(0)
    for k in FIELDS:
        assert np.array_equal(before[k], after[k]), k
    # "abort": the step it is seen at is the last
    (tmp_path / "abort").write_text("")
    assert dom.advance(max_steps=5, poll_files=True) == (1, 2)
    assert not dom.final().any()
    # the loop goes on once the file is gone, and counts its steps
    (tmp_path / "abort").unlink()
    assert dom.advance(max_steps=2, poll_files=True) == (2, 0)
    # 3 + 2 + 0 + 1 + 2 steps in all, each of the same kind as step()
    one = D.Domain(params("c1_kepler2d.par", Num_R=16))
    one.step(8)
    sa, sb = dom.state(0), one.state(0)
    for k in FIELDS:
        assert np.array_equal(sa[k], sb[k]), k
    one.close()
    dom.close()

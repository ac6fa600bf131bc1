"""The C host (disco_amd/host/disco_main.c) on the GPU: the reference's main
loop to T_End through the C ABI reproduces the reference's own run to a
T_End inside its 4th step (tests/golden/c1_tend_a.npz) and the Python
binding's advance() bit for bit, writes its checkpoints, and a restart
from one of them ("Restart 1", input.ckp) ends on the same final state."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import disco_amd as D
from disco_amd import checkpoint as ckpt
from dumpio import GOLDEN, load_fixture
from helpers import par_text, params
from test_host_cpu import build_host

pytestmark = pytest.mark.gpu


def test_c_host_runs_the_reference_loop(tmp_path):
    exe = str(tmp_path / "disco_main")
    build_host(exe)
    t_end = float(np.load(os.path.join(GOLDEN, "c1_tend_a.npz"))["t_end"])
    text = par_text("c1_kepler2d.par", Num_R=16, T_End=repr(t_end))
    (tmp_path / "in.par").write_text(text)
    out = subprocess.run([exe, "in.par", "2"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    ref = load_fixture("c1_tend_a.npz")
    n = max(ref)
    assert f"steps {n}" in out.stdout
    fin = ckpt.read(str(tmp_path / "checkpoint_final.ckp"))
    assert fin["Grid/T"][0] == t_end
    dom = D.Domain(params("c1_kepler2d.par", Num_R=16, T_End=repr(t_end)))
    dom.advance()
    st = dom.state(0)
    assert np.array_equal(fin["Data/Cells"][:, :5], st["prim"].T)
    assert np.array_equal(fin["Data/Cons"], st["cons"].T)
    real = ref[n]["real"] == 1
    rel = np.abs(fin["Data/Cells"][real, 0] - ref[n]["prim"][0, real]).max() / np.abs(ref[n]["prim"][0, real]).max()
    assert rel < 1e-11, rel
    # restart from the first checkpoint and run to T_End again
    run2 = tmp_path / "restart"
    run2.mkdir()
    shutil.copy(tmp_path / "checkpoint_0000.ckp", run2 / "input.ckp")
    (run2 / "in.par").write_text(par_text("c1_kepler2d.par", Num_R=16, T_End=repr(t_end), Restart=1))
    out2 = subprocess.run([exe, "in.par", "2"], cwd=run2, capture_output=True, text=True, timeout=120)
    assert out2.returncode == 0, out2.stdout
    fin2 = ckpt.read(str(run2 / "checkpoint_final.ckp"))
    assert np.array_equal(fin2["Data/Cells"], fin["Data/Cells"]) and fin2["Grid/T"][0] == t_end
    dom.close()

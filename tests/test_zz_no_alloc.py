"""disco_gpu_report_noalloc -- the report entry point that allocates nothing
(its annulus sums live in a per-GPU static device scratch) -- against
disco_gpu_report on the same states, bit for bit, on consecutive calls whose
second call has a smaller input (a run with fewer annuli), and its clear
refusal of a run with more annuli than the scratch holds."""
import numpy as np
import pytest

import disco_amd as D
from helpers import params

pytestmark = pytest.mark.gpu


def test_noalloc_report_equals_report_big_then_small():
    big = params("c3_binary_disk.par", Num_R=64, Sink_Par3=0.2)
    small = params("c3_binary_disk.par", Num_R=24, Sink_Par3=0.2, Init_Par0=2)
    dom = D.Domain([big, small])
    assert dom.info[0].Nr * dom.info[0].Nz > dom.info[1].Nr * dom.info[1].Nz
    for s in range(3):
        dom.step(2)
        dom.sync()
        for run in (0, 1):  # the second call has the smaller input
            a = dom.report(run)
            b = dom.report_noalloc(run)
            print(f"step {2 * (s + 1)} run {run}: mass {b[2 + 16]:.6e}", flush=True)
            assert a.shape == b.shape and np.array_equal(a, b), (s, run, a, b)
    # the torques are not trivially zero
    assert np.abs(dom.report_noalloc(0)[2 + 16 + 1:]).max() > 0
    dom.close()


def test_noalloc_report_3d_mhd():
    dom = D.Domain(params("c2_mhd3d.par", Num_R=12, Num_Z=4, CFL=0.3))
    dom.step(2)
    dom.sync()
    assert np.array_equal(dom.report(0), dom.report_noalloc(0))
    dom.close()


def test_noalloc_report_refuses_too_many_annuli():
    # 256 x 256 real annuli of 4 cells each: 260 x 260 = 67,600 annuli with ghosts
    dom = D.Domain(params("c1_kepler2d.par", Num_R=256, Num_Z=256, Z_Periodic=1, aspect=1000.0))
    assert dom.info[0].Nr * dom.info[0].Nz > 65536
    with pytest.raises(RuntimeError, match="more annuli than the scratch"):
        dom.report_noalloc(0)
    dom.report(0)  # the allocating entry point still serves it
    dom.close()

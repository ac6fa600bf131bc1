"""The flux-flip replay (ct_cells.h flip_chain_cell, compiled by the host
compiler from the source the GPU kernel uses) against a literal sequential
restatement of the reference's flip_fluxes loop (bfields.c:760-814, explicit
face lists and LRtype toggles) on the overlap faces of the small MHD disk,
for random flag patterns from sparse to every face: the face fluxes agree
bit for bit, whatever the pattern (adjacent flips, flips two apart, chains
over the end of a face list, chains closed around a track).  A planted
mistake in the restatement (the toggle of the face before the flip left
out) is detected."""
import os
import re
import subprocess

import pytest

from helpers import par_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "disco_amd", "csrc")


def _build(tmp, extra=()):
    exe = str(tmp / "flip_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-I", CSRC, "-I", os.path.join(ROOT, "include"), *extra,
                    os.path.join(ROOT, "tests", "probes", "flip_probe.cpp"), os.path.join(CSRC, "host_setup.cpp"),
                    "-o", exe], check=True, timeout=300)
    return exe


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    return _build(tmp_path_factory.mktemp("flip"))


@pytest.fixture(scope="module")
def par(tmp_path_factory):
    p = tmp_path_factory.mktemp("par") / "in.par"
    p.write_text(par_text("c2_mhd3d.par", Num_R=12, Num_Z=4, CFL=0.3))
    return str(p)


@pytest.mark.parametrize("prob", [0.05, 0.3, 0.6, 0.9, 1.0])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_replay_equals_reference_loop(probe, par, prob, seed):
    out = subprocess.run([probe, par, str(seed), str(prob)], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    m = re.search(r"flags (\d+) mismatches (\d+) err (\d+)", out)
    assert m, out
    nflags, mism, err = map(int, m.groups())
    assert nflags > 0 and mism == 0 and err == 0, out


def test_planted_mistake_in_the_restatement_is_detected(par, tmp_path):
    exe = _build(tmp_path, ("-DPLANT_NO_PREV_TOGGLE",))
    out = subprocess.run([exe, par, "1", "0.6"], check=True, capture_output=True, text=True, timeout=120).stdout
    mism = int(re.search(r"mismatches (\d+)", out).group(1))
    assert mism > 0, out

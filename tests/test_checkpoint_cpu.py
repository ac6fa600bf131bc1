"""The checkpoint container (disco_amd/csrc/checkpoint.cpp, read and written
in Python by disco_amd/checkpoint.py) round-trips both ways: a file the C++
writer produces reads back in Python with the same names, types, shapes and
bits, and a file Python writes reads back in C++.  A truncated file is
refused.  (The GPU tests cover the checkpoint contents and restarts.)"""
import os
import subprocess

import numpy as np
import pytest

from disco_amd import checkpoint as ckpt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "disco_amd", "csrc")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ck") / "ckpt_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-I", CSRC, os.path.join(ROOT, "tests", "probes", "ckpt_probe.cpp"),
                    os.path.join(CSRC, "checkpoint.cpp"), "-o", exe], check=True, timeout=120)
    return exe


def test_cpp_writes_python_reads(probe, tmp_path):
    f = str(tmp_path / "a.ckp")
    subprocess.run([probe, "write", f], check=True, timeout=30)
    d = ckpt.read(f)
    assert list(d) == ["Data/Cells", "Grid/Np", "Opts/HYDRO"]
    cells = d["Data/Cells"]
    assert cells.dtype == np.float64 and cells.shape == (2, 3)
    assert np.array_equal(cells.ravel(), np.array([1.5, -2.25, 3.0, 1e-300, 7.0, -0.0]))
    assert np.signbit(cells[1, 2])
    assert d["Grid/Np"].dtype == np.int32 and d["Grid/Np"].tolist() == [[-7, 123456789]]
    assert d["Opts/HYDRO"] == "mhd"


def test_python_writes_cpp_reads(probe, tmp_path):
    f = str(tmp_path / "b.ckp")
    x = np.arange(24, dtype=np.float64).reshape(2, 3, 4) * 0.5
    ckpt.write(f, {"Data/Diagnostics": x, "Grid/Index": np.array([[1, 2], [3, 4]], np.int32), "Opts/INITIAL": "gr_torus"})
    out = subprocess.run([probe, "read", f], check=True, capture_output=True, text=True, timeout=30).stdout.splitlines()
    assert out[0] == "Data/Diagnostics 0 2 3 4 : 138"
    assert out[1] == "Grid/Index 1 2 2 : 10"
    assert out[2].startswith("Opts/INITIAL 2 8 : ")
    assert ckpt.read(f)["Opts/INITIAL"] == "gr_torus"


def test_truncated_file_is_refused(probe, tmp_path):
    f = str(tmp_path / "c.ckp")
    subprocess.run([probe, "write", f], check=True, timeout=30)
    data = open(f, "rb").read()
    open(f, "wb").write(data[:-5])
    r = subprocess.run([probe, "read", f], capture_output=True, text=True, timeout=30)
    assert r.returncode != 0 and "not a readable checkpoint" in r.stderr

"""Per-leaf parity of the physics the GPU kernels compile, against the
reference's own functions.

tests/golden/leaf_vectors.npz holds random face states and what the
reference build returns for them (oracle/ref/ref_leaf.c, linked with the
reference's riemann.c, hlld.c and Hydro/euler.c, mhd.c, grmhd.c, grmhd2.c
or greuler.c + schw_ks_rpz.c): solve_riemann's conserved increments (F - w U*) dA dt and
flux_to_E's EMFs for HLL, HLLC, HLLD and HLLC_DAMPCENTER on phi, r and z faces, prim2cons,
cons2prim back, mindt and source.  The library's leaves (the step_cells.h
policies over phys_euler.h, phys_mhd.h, phys_grmhd_fv.h and phys_greuler.h -- the source the
GPU kernels compile) are built here with the host compiler into
tests/probes/leaf_probe.cpp and fed the same records.  Both sides are
compiled without fused multiply-adds and call the same libm, so every value
must agree to 1e-14 of its group's scale; most agree to the bit.
"""
import os
import subprocess

import numpy as np
import pytest

from dumpio import GOLDEN
from helpers import par_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "disco_amd", "csrc")

# (config, overrides) of each case: the same as tests/golden/make_fixtures.py
CASES = {
    "euler_hllc": ("c1_kepler2d.par", {"Num_R": 16}),
    "euler_hll": ("c1_kepler2d.par", {"Num_R": 16, "Riemann_Solver": 0}),
    "mhd_hlld": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4}),
    "mhd_hllc": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4, "Riemann_Solver": 1}),
    "mhd_hll": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4, "Riemann_Solver": 0}),
    "grmhd_hll": ("c5_grmhd_torus.par", {"Num_R": 16, "Num_Z": 12, "Metric_Par3": 0.0}),
    "grmhd_hllc": ("c5_grmhd_torus.par", {"Num_R": 16, "Num_Z": 12, "Metric_Par3": 0.0, "Riemann_Solver": 1}),
    "greuler_hll": ("c5_grmhd_torus.par", {"Hydro": "greuler", "Num_C": 5, "Constrained_Transport": 0, "Num_R": 16,
                                           "Num_Z": 12, "Metric_Par3": 0.0}),
    "greuler_hllc": ("c5_grmhd_torus.par", {"Hydro": "greuler", "Num_C": 5, "Constrained_Transport": 0, "Num_R": 16,
                                            "Num_Z": 12, "Metric_Par3": 0.0, "Riemann_Solver": 1}),
    "grmhd2_hll": ("c5_grmhd_torus.par", {"Hydro": "grmhd2", "Num_R": 16, "Num_Z": 12, "Metric_Par3": 0.0}),
    "grmhd2_hllc": ("c5_grmhd_torus.par", {"Hydro": "grmhd2", "Num_R": 16, "Num_Z": 12, "Metric_Par3": 0.0,
                                           "Riemann_Solver": 1}),
    "grmhd2_hlld": ("c5_grmhd_torus.par", {"Hydro": "grmhd2", "Num_R": 16, "Num_Z": 12, "Metric_Par3": 0.0,
                                           "Riemann_Solver": 2}),
    # strong jumps and fields: the relativistic HLLD's retried starting
    # pressures and its HLL fallback
    "grmhd2_hlld_strong": ("c5_grmhd_torus.par", {"Hydro": "grmhd2", "Num_R": 16, "Num_Z": 12, "Metric_Par3": 0.0,
                                                  "Riemann_Solver": 2}),
    # HLLC_DAMPCENTER (riemann.c:253) on faces either side of r = 0.1: the
    # records inside take HLL's state, those outside HLLC's
    "euler_dampc": ("c1_kepler2d.par", {"Num_R": 16, "Riemann_Solver": 3}),
    "mhd_dampc": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4, "Riemann_Solver": 3}),
    # one passive scalar (NUM_N = 1): prim2cons, cons2prim, flux and getUstar
    # carry it as prim[5] * rho
    "euler_hllc_n1": ("c1_kepler2d.par", {"Num_R": 16, "Num_N": 1}),
    "euler_hll_n1": ("c1_kepler2d.par", {"Num_R": 16, "Num_N": 1, "Riemann_Solver": 0}),
    # ... and for MHD, HLLD's rule too (the scalar rides on the density of the
    # contact's side, hlld.c:345-348)
    "mhd_hlld_n1": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4, "Num_N": 1}),
    "mhd_hllc_n1": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4, "Num_N": 1, "Riemann_Solver": 1}),
    "mhd_hll_n1": ("c2_mhd3d.par", {"Num_R": 12, "Num_Z": 4, "Num_N": 1, "Riemann_Solver": 0}),
}


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("leaf") / "leaf_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-I", CSRC, "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "probes", "leaf_probe.cpp"),
                    os.path.join(CSRC, "host_setup.cpp"), "-o", exe], check=True, timeout=300)
    return exe


@pytest.fixture(scope="module")
def vectors():
    return np.load(os.path.join(GOLDEN, "leaf_vectors.npz"))


@pytest.mark.parametrize("case", list(CASES))
def test_leaves_match_reference(case, probe, vectors, tmp_path):
    base, over = CASES[case]
    rec = vectors[f"{case}_in"]
    ref = vectors[f"{case}_out"]
    n = rec.shape[0]
    nq = (rec.shape[1] - 9) // 2
    par = tmp_path / "in.par"
    par.write_text(par_text(base, **over))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.ascontiguousarray(rec, np.float64).tobytes())
    subprocess.run([probe, str(par), str(fin), str(fout)], check=True, timeout=60)
    out = np.fromfile(fout, np.float64).reshape(n, 4 * nq + 5)
    groups = {"dU": slice(0, nq), "emf": slice(nq, nq + 4), "cons": slice(nq + 4, 2 * nq + 4),
              "prim": slice(2 * nq + 4, 3 * nq + 4), "mindt": slice(3 * nq + 4, 3 * nq + 5),
              "source": slice(3 * nq + 5, 4 * nq + 5)}
    worst = {}
    for g, sl in groups.items():
        a, b = out[:, sl], ref[:, sl]
        assert np.isfinite(a).all() and np.isfinite(b).all(), g
        sc = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
        e = (np.abs(a - b) / sc).max()
        worst[g] = e
        assert e <= 1e-14, (case, g, e)
    exact = float((out == ref).mean())
    print(f"{case}: worst {worst}, bit-identical values {exact:.3f}")


@pytest.mark.parametrize("case", ["euler_dampc", "mhd_dampc"])
def test_dampcenter_switches_at_r_0p1(case, probe, vectors, tmp_path):
    """The HLLC_DAMPCENTER vectors are not HLLC's in disguise: run as HLL the
    library gives the reference's values bit for bit on the faces inside
    r = 0.1, run as HLLC those on the faces outside, and each differs from the
    other solver on its own side."""
    base, over = CASES[case]
    rec = vectors[f"{case}_in"]
    ref = vectors[f"{case}_out"]
    n = rec.shape[0]
    nq = (rec.shape[1] - 9) // 2
    inner = rec[:, 2 * nq] < 0.1
    assert 0 < inner.sum() < n
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.ascontiguousarray(rec, np.float64).tobytes())
    outs = {}
    for rs in (0, 1):
        par, fout = tmp_path / f"in{rs}.par", tmp_path / f"out{rs}.bin"
        par.write_text(par_text(base, **dict(over, Riemann_Solver=rs)))
        subprocess.run([probe, str(par), str(fin), str(fout)], check=True, timeout=60)
        outs[rs] = np.fromfile(fout, np.float64).reshape(n, 4 * nq + 5)
    dU = slice(0, nq)
    assert np.array_equal(outs[0][inner], ref[inner])
    assert np.array_equal(outs[1][~inner], ref[~inner])
    moving = ~np.all(rec[:, :nq] == rec[:, nq:2 * nq], axis=1)  # distinct left and right states
    assert (outs[1][inner & moving, dU] != ref[inner & moving, dU]).any(axis=1).all()
    assert (outs[0][~inner & moving, dU] != ref[~inner & moving, dU]).any(axis=1).all()

"""CPU tests: the oracle against the reference's own states, the overlap-face
merge against brute force, the host set-up against the reference's step-0
state, the parameter parser, and the C ABI's exported symbols."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import disco_amd as D
from dumpio import load_fixture
from helpers import ROOT, cell_radius, params, phi_err, state_errors
from oracle_py import Oracle, track_faces

TOL = {1: 1e-12, 10: 1e-10, 50: 1e-9, 100: 1e-9}
CASES = [
    ("c1_small.npz", "c1_kepler2d.par", dict(Num_R=16)),
    ("c1_small_hll_mm4.npz", "c1_kepler2d.par", dict(Num_R=16, Riemann_Solver=0, Mesh_Motion=4)),
    ("c3_small.npz", "c3_binary_disk.par", dict(Num_R=24)),
]


def _grid(ref0):
    return dict(Nr=ref0["Nr"], Nz=ref0["Nz"], NgRa=ref0["NgRa"], NgRb=ref0["NgRb"], NgZa=ref0["NgZa"],
                NgZb=ref0["NgZb"], Np=ref0["Np"].astype(np.int32), r_jph=ref0["r_jph"], z_kph=ref0["z_kph"])


@pytest.mark.parametrize("fixture,base,over", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_reference(fixture, base, over):
    """The restatement, started from the reference's own step-0 state, tracks
    the reference to round-off (1e-12 after one step)."""
    ref = load_fixture(fixture)
    p = params(base, **over)
    r0 = ref[0]
    pl = r0["planets"][:, 1] if r0["planets"].size else None
    o = Oracle(p, _grid(r0), r0, planet_phi=pl, t=r0["t"])
    rcell = cell_radius(r0["Nr"], r0["Nz"], r0["Np"], r0["r_jph"])
    done = 0
    for s in sorted(ref):
        if s == 0:
            continue
        for _ in range(s - done):
            dt = o.getmindt()
            o.timestep(dt)
        done = s
        st = o.state()
        real = ref[s]["real"] == 1
        assert abs(dt - ref[s]["dt"]) <= 1e-12 * ref[s]["dt"]
        e = state_errors(st, ref[s], rcell, real)
        assert e["prim"].max() < TOL[s], (s, e)
        assert e["cons"].max() < TOL[s], (s, e)
        assert phi_err(st["piph"], ref[s]["piph"]) < TOL[s]
        if r0["planets"].size:
            assert np.allclose(st["planet_phi"][:len(ref[s]["planets"])], ref[s]["planets"][:, 1], rtol=0,
                               atol=1e-12)
    o.close()


def test_oracle_detects_a_planted_sign_error():
    """The pin is sensitive: flipping the sign of the pressure in one cell's
    primitive state moves the oracle's next step far outside tolerance."""
    ref = load_fixture("c1_small.npz")
    p = params("c1_kepler2d.par", Num_R=16)
    r0 = ref[0]
    bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in r0.items()}
    c = int(np.nonzero(r0["real"] == 1)[0][100])
    bad["prim"][1, c] *= 1.001
    o = Oracle(p, _grid(r0), bad, planet_phi=r0["planets"][:, 1], t=0.0)
    o.timestep(o.getmindt())
    rcell = cell_radius(r0["Nr"], r0["Nz"], r0["Np"], r0["r_jph"])
    e = state_errors(o.state(), ref[1], rcell, ref[1]["real"] == 1)
    assert e["prim"].max() > 1e-9


@pytest.mark.parametrize("fixture", ["c1_tend_a.npz", "c1_tend_b.npz"])
def test_oracle_check_dt_matches_reference(fixture):
    """check_dt (domain.c:283-292): the reference's own loop, run to a T_End
    inside a step, takes the same number of steps as the oracle's, its last
    step cut to end on T_End (same final time, and a state that a step of
    any other length would miss by far more than round-off)."""
    z = np.load(os.path.join(ROOT, "tests", "golden", fixture))
    t_end = float(z["t_end"])
    ref = load_fixture(fixture)
    nref = max(ref)
    p = params("c1_kepler2d.par", Num_R=16, T_End=repr(t_end))
    r0 = ref[0]
    o = Oracle(p, _grid(r0), r0, planet_phi=r0["planets"][:, 1], t=r0["t"])
    n = o.run_to_end()
    st = o.state()
    assert n == nref
    assert st["t"] == ref[nref]["t"] == t_end
    rcell = cell_radius(r0["Nr"], r0["Nz"], r0["Np"], r0["r_jph"])
    e = state_errors(st, ref[nref], rcell, ref[nref]["real"] == 1)
    assert e["prim"].max() < 1e-11 and e["cons"].max() < 1e-11, e
    o.close()
    # a full-length last step (no clip) misses the reference's state
    o = Oracle(p, _grid(r0), r0, planet_phi=r0["planets"][:, 1], t=r0["t"])
    o.step(nref)
    e = state_errors(o.state(), ref[nref], rcell, ref[nref]["real"] == 1)
    assert e["prim"].max() > 1e-6
    o.close()


REPORT_CASES = [("c3_report.npz", "c3_binary_disk.par", dict(Num_R=24, Sink_Par3=0.2)),
                ("c1_report.npz", "c1_kepler2d.par", dict(Num_R=16))]


@pytest.mark.parametrize("fixture,base,over", REPORT_CASES, ids=[c[0] for c in REPORT_CASES])
def test_oracle_report_matches_reference(fixture, base, over):
    """report.c on the reference's own state after 10 steps: the oracle's
    full-precision values against the line the reference printed (%le: 7
    significant digits, so 5e-7 relative), zeros exactly; and the mass
    against an independent sum of rho dV."""
    rep = np.load(os.path.join(ROOT, "tests", "golden", fixture))["report"]
    ref = load_fixture(fixture)
    st = ref[max(ref)]
    pl = st["planets"][:, 1] if st["planets"].size else None
    o = Oracle(params(base, **over), _grid(st), st, planet_phi=pl, t=st["t"])
    got = o.report()
    o.close()
    assert got.size == rep.size
    assert np.all(np.abs(got - rep) <= 5.01e-7 * np.abs(rep)), np.abs(got - rep) / np.maximum(np.abs(rep), 1e-300)
    # the gas mass: sum of rho x rbar dr dphi dz over the real cells
    rj, zk = st["r_jph"], st["z_kph"]
    off = np.concatenate([[0], np.cumsum(st["Np"])])
    mass = 0.0
    for k in range(st["NgZa"], st["Nz"] - st["NgZb"]):
        for j in range(st["NgRa"], st["Nr"] - st["NgRb"]):
            c = slice(off[j + st["Nr"] * k], off[j + st["Nr"] * k + 1])
            dV = .5 * (rj[j + 1] + rj[j]) * (rj[j + 1] - rj[j]) * st["dphi"][c] * (zk[k + 1] - zk[k])
            mass += (st["prim"][0, c] * dV).sum()
    npl = (rep.size - 15) // 8
    assert abs(got[2 + 8 * npl] - mass) <= 1e-13 * mass
    if npl > 1:
        # with a cut length of zero every cell lies outside every cut: each
        # cut sum is the whole torque
        o = Oracle(params(base, **{**over, "Sink_Par3": 0.0}), _grid(st), st, planet_phi=pl, t=st["t"])
        z = o.report()
        o.close()
        i = 3 + 8 * npl
        assert np.array_equal(z[i + 2::2], np.full(5, z[i])) and np.array_equal(z[i + 3::2], np.full(5, z[i + 1]))


def _brute_overlaps(piA, dpA, piB, dpB, P=2 * np.pi):
    """All (a, b) pairs whose phi intervals overlap (positive length), by
    brute force on the circle."""
    out = set()
    for a in range(len(piA)):
        a0, a1 = piA[a] - dpA[a], piA[a]
        for b in range(len(piB)):
            b0, b1 = piB[b] - dpB[b], piB[b]
            for s in (-P, 0.0, P):
                lo, hi = max(a0, b0 + s), min(a1, b1 + s)
                if hi - lo > 1e-12:
                    out.add((a, b))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_track_faces_equal_brute_force(seed):
    rng = np.random.default_rng(seed)
    P = 2 * np.pi
    NA, NB = rng.integers(4, 40), rng.integers(4, 40)

    def ring(n):
        d = rng.uniform(0.5, 1.5, n)
        d *= P / d.sum()
        p0 = rng.uniform(0, P)
        pi = (p0 + np.cumsum(d)) % P
        return pi, d

    piA, dpA = ring(NA)
    piB, dpB = ring(NB)
    iL, iR, LR, pm, pp = track_faces(piA, dpA, piB, dpB, P)
    assert len(iL) == NA + NB
    faces = {(int(a), int(b)) for a, b, lo, hi in zip(iL, iR, pm, pp) if (hi - lo) % P > 1e-12}
    assert faces == _brute_overlaps(piA, dpA, piB, dpB, P)
    # total face extent covers the circle exactly once
    width = np.array([(hi - lo) % P if abs(hi - lo) > 0 else 0.0 for lo, hi in zip(pm, pp)])
    assert abs(width.sum() - P) < 1e-9


SETUP_CASES = CASES + [("c2e_small.npz", "c1_kepler2d.par", dict(Num_R=12, Num_Z=6, Z_Periodic=1)),
                       ("c1_zerograd.npz", "c1_kepler2d.par", dict(Num_R=16, Boundary="zerograd")),
                       ("c1_outflow.npz", "c1_kepler2d.par", dict(Num_R=16, Boundary="outflow3d")),
                       ("c2e_outflow.npz", "c1_kepler2d.par", dict(Num_R=12, Num_Z=6, Boundary="outflow3d")),
                       ("c2_mhd_outflow.npz", "c2_mhd3d.par",
                        dict(Num_R=12, Num_Z=4, CFL=0.3, Z_Periodic=0, Boundary="outflow3d")),
                       ("c5_grmhd_outflow.npz", "c5_grmhd_torus.par",
                        dict(Num_R=16, Num_Z=12, Metric_Par3=0.0, Boundary="outflow3d")),
                       # Noise_Type 1-3 (noise.c): the glibc rand() stream continued after the phases
                       ("c1_noise1.npz", "c1_kepler2d.par", dict(Num_R=16, Noise_Type=1, Noise_Abs=1e-3,
                                                                 Noise_Rel=1e-2)),
                       ("c1_noise2.npz", "c1_kepler2d.par", dict(Num_R=16, Noise_Type=2, Noise_Abs=1e-4,
                                                                 Noise_Rel=1e-2)),
                       ("c2e_noise3.npz", "c1_kepler2d.par", dict(Num_R=12, Num_Z=6, Z_Periodic=1, Noise_Type=3,
                                                                  Noise_Abs=1e-3, Noise_Rel=1e-2)),
                       # geometric boundaries without ghost annuli
                       ("c3_nobc_r0.npz", "c3_binary_disk.par", dict(Num_R=24, R_Min=0.0, NoBC_Rmin=1, Softening=0.3)),
                       ("c1_nobc_rmax.npz", "c1_kepler2d.par", dict(Num_R=16, NoBC_Rmax=1)),
                       ("c2e_nobc_z.npz", "c1_kepler2d.par", dict(Num_R=12, Num_Z=6, NoBC_Zmin=1, NoBC_Zmax=1)),
                       # GR hydrodynamics (greuler.c) on the torus at spin 0
                       ("c6_greuler_small.npz", "c5_grmhd_torus.par",
                        dict(Hydro="greuler", Num_C=5, Constrained_Transport=0, Num_R=16, Num_Z=12,
                             Metric_Par3=0.0)),
                       ("c7_grmhd2_small.npz", "c5_grmhd_torus.par",
                        dict(Hydro="grmhd2", Num_R=16, Num_Z=12, Metric_Par3=0.0)),
                       # Mesh_Motion 3: the annulus mean of the initial face speeds
                       ("c1_mm3.npz", "c1_kepler2d.par", dict(Num_R=16, Mesh_Motion=3)),
                       # the rotating frame's initial and fixed-ghost data
                       ("c1_rot_fixed.npz", "c1_kepler2d.par", dict(Num_R=16, RotFrame=1, RotOmega=0.3, RotD=0.05))]


@pytest.mark.parametrize("fixture,base,over", SETUP_CASES, ids=[c[0] for c in SETUP_CASES])
def test_host_setup_is_bit_exact(fixture, base, over):
    """setupGrid/setupDomain/setupCells + first ghost fill on the host give the
    reference's step-0 state bit for bit (same seeded phases, same IC)."""
    r0 = load_fixture(fixture)[0]
    s = D.setup_run(params(base, **over))
    real = r0["real"] == 1
    assert np.array_equal(s["Np"], r0["Np"])
    assert np.array_equal(s["r_jph"], r0["r_jph"])
    assert np.array_equal(s["z_kph"], r0["z_kph"])
    assert np.array_equal(s["piph"], r0["piph"])
    assert np.array_equal(s["dphi"], r0["dphi"])
    assert np.array_equal(s["prim"][:, real], r0["prim"][:, real])
    assert np.array_equal(s["cons"][:, real], r0["cons"][:, real])
    assert np.array_equal(s["prim"][:, ~real], r0["prim"][:, ~real])
    assert np.array_equal(s["wiph"][real], r0["wiph"][real])


def test_parser_semantics():
    text = open(os.path.join(ROOT, "configs", "c1_kepler2d.par")).read()
    p = D.parse_par(text)
    assert p.Num_R == 256 and p.Num_Z == 1
    assert abs(p.phimax - 2 * np.pi) < 1e-15          # P_Times_2pi applied
    assert p.Riemann_Solver == 1 and p.CFL == 0.5
    # separators "\t :=>_" between key and value (readpar.c:23)
    q = D.parse_par(text.replace("Num_R           256", "Num_R = 300"))
    assert q.Num_R == 300
    # ints are read as doubles and truncated
    q = D.parse_par(text.replace("Num_R           256", "Num_R 12.9"))
    assert q.Num_R == 12
    # a missing key is an error, like "Read Failed"
    with pytest.raises(ValueError):
        D.parse_par(re.sub(r"^CFL .*$", "", text, flags=re.M))


def test_unsupported_configuration_is_rejected():
    p = params("c1_kepler2d.par", R_Periodic=1)
    with pytest.raises(RuntimeError, match="R_Periodic"):
        D.setup_run(p)
    # every Timestep value sets up (timestep.c: 1 and 3 have their own substeps, the rest run RK2)
    for ts in (0, 1, 2, 3):
        D.setup_run(params("c1_kepler2d.par", Timestep=ts))
    # viscosity is euler.c's (mhd.c defines no visc_source) and needs nu > 0
    with pytest.raises(RuntimeError, match="euler only"):
        D.setup_run(params("c2_mhd3d.par", Num_R=12, Num_Z=4, Use_Viscosity=1, Viscosity=0.01))
    with pytest.raises(RuntimeError, match="Viscosity > 0"):
        D.setup_run(params("c1_kepler2d.par", Use_Viscosity=1))
    # the atmosphere is set-up work of the Newtonian modules (domain.c:194-200)
    with pytest.raises(RuntimeError, match="Include_Atmos"):
        D.setup_run(params("c5_grmhd_torus.par", Num_R=16, Num_Z=12, Include_Atmos=1))
    # get_cs2's profile 5 needs the planets inside cons2prim: refused
    with pytest.raises(RuntimeError, match="Cs2_Profile 5"):
        D.setup_run(params("c3_binary_disk.par", Num_R=24, Cs2_Profile=5))
    # sink.c's types beyond the built ones are refused, not ignored
    for key, msg in (("Sink_Type", "Sink_Type"), ("Nozzle_Type", "Nozzle_Type"), ("DampInner_Type", "Damp"),
                     ("Cool_Type", "Cool_Type")):
        with pytest.raises(RuntimeError, match=msg):
            D.setup_run(params("c1_kepler2d.par", **{key: 4}))
    with pytest.raises(RuntimeError, match="Cool_Par1"):
        D.setup_run(params("c1_kepler2d.par", Cool_Type=1))
    with pytest.raises(RuntimeError, match="alone in 2D"):
        D.setup_run(params("c1_kepler2d.par", DampUpper_Type=1))
    with pytest.raises(RuntimeError, match="Nozzle_Par2"):
        D.setup_run(params("c1_kepler2d.par", Nozzle_Type=1))


@pytest.mark.parametrize("over,msg", [
    (dict(Riemann_Solver=2), "HLLD with grmhd"),
    (dict(Metric_Par3=1.2), "Kerr needs"),
    (dict(Mesh_Motion=1), "grmhd: 0 or 4"),
    (dict(Constrained_Transport=0), "Constrained_Transport"),
    (dict(Sink_Type=1), "Sink_Type with grmhd"),
    (dict(NoBC_Rmax=1), "euler only"),
    (dict(RotFrame=1), "RotFrame with grmhd"),
    (dict(Use_Viscosity=1, Viscosity=0.1), "euler only"),
])
def test_unsupported_grmhd_configuration_is_rejected(over, msg):
    """GRMHD set-ups the step does not run are refused with the key named."""
    with pytest.raises(RuntimeError, match=msg):
        D.setup_run(params("c5_grmhd_torus.par", Num_R=8, Num_Z=6, **over))


def test_grmhd_setup_is_bit_exact():
    """Host set-up of the torus (GRMHD prim2cons, set_B_fields, the Noble
    cons2prim from the initial prims, fixed r and z ghosts) gives the
    reference's step-0 state bit for bit."""
    r0 = load_fixture("c5_grmhd_small.npz")[0]
    s = D.setup_run(params("c5_grmhd_torus.par", Num_R=16, Num_Z=12, Metric_Par3=0.0))
    real = r0["real"] == 1
    for k in ("Np", "r_jph", "z_kph", "piph", "dphi"):
        assert np.array_equal(s[k], r0[k]), k
    assert np.array_equal(s["prim"][:, real], r0["prim"][:, real])
    assert np.array_equal(s["cons"][:, real], r0["cons"][:, real])
    assert np.array_equal(s["prim"][:, ~real], r0["prim"][:, ~real])


def test_library_exports_every_declared_symbol():
    """The C ABI library loads and exports every function include/*.h declares."""
    L = C.CDLL(D.LIB_PATH)
    decl = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            decl |= set(re.findall(r"^\s*(?:int|void|const char\*|double)\s+\**(\w+)\s*\(", src, re.M))
    assert len(decl) >= 15
    for name in sorted(decl):
        assert hasattr(L, name), name

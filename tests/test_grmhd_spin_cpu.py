"""The spinning-metric GRMHD leaves pinned from the mathematics.

The reference has no spinning metric, so its fixtures pin the GRMHD module
only at a = 0, where every beta^phi and gamma_{R phi} term vanishes.  Here
the library's leaves (phys_grmhd.h, phys_grmhd_fv.h: the source the GPU
kernels compile, built with the host compiler into
tests/probes/grmhd_spin_probe.cpp) are compared at a = 0.5 and 0.9 with a
covariant restatement (oracle/grmhd_covariant.py) that shares no code with
them: the Cartesian Kerr-Schild metric pulled back to cylindrical
coordinates, numerical inverse and 3+1 split, complex-step derivatives, the
stress tensor T^{mu nu}, and the characteristic quadratic for the wave
speeds.  Compared on random states (to 1e-12 of each quantity's scale):
the 3+1 metric, prim2cons, the flux through r, phi and z faces, the HLL wave
speeds, solve_riemann's HLL increment and EMFs, the geometric sources and
the CFL step.  The angular-momentum source must vanish exactly (phi is a
Killing direction), and the oracle's own phi-derivative of the metric is
zero to rounding.

The oracle itself is pinned first: at a = 0 against the reference's own
grmhd.c + schw_ks_rpz.c leaf vectors (tests/golden/leaf_vectors.npz, HLL
rows; tests/golden/grmhd_leaf_a0.npz), and at a = 0.9 by identities
(sqrt(-g) = R, u.u = -1, u.b = 0, a planted sign error in the spin term
caught by the flux comparison).
"""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

from dumpio import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "disco_amd", "csrc")
GAMMA = 1.6666666666666667
M = 1.0


def _load_oracle():
    spec = importlib.util.spec_from_file_location("grmhd_covariant", os.path.join(ROOT, "oracle", "grmhd_covariant.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


GC = _load_oracle()


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("spin") / "grmhd_spin_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "probes", "grmhd_spin_probe.cpp"), "-o", exe], check=True, timeout=300)

    def run(a, recs, mode=0):
        lines = []
        for p1, p2, x, d in recs:
            vals = [*p1, *p2, *x]
            lines.append(f"{M!r} {a!r} {GAMMA!r} {mode} " + " ".join(repr(float(v)) for v in vals) + f" {int(d)}")
        out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True,
                             timeout=120).stdout
        return np.array([[float(v) for v in ln.split()] for ln in out.strip().splitlines()])

    return run


def _states(n, seed):
    """Random primitive pairs and positions outside the horizon."""
    rng = np.random.default_rng(seed)
    recs = []
    for k in range(n):
        R = rng.uniform(2.0, 30.0)
        z = rng.uniform(-10.0, 10.0)
        phi = rng.uniform(0.0, 2 * np.pi)
        p = np.zeros(8)
        p[0] = 10 ** rng.uniform(-4, 0)
        p[1] = p[0] * 10 ** rng.uniform(-3, 0)
        p[2] = rng.normal(0, 0.4)
        p[3] = rng.normal(0, 0.4) / R  # u_phi / R^2 with u_phi ~ R
        p[4] = rng.normal(0, 0.4)
        p[5:8] = rng.normal(0, 0.3, 3) * np.sqrt(p[0])
        q = p * (1.0 + rng.uniform(-0.1, 0.1, 8))
        recs.append((p, q, np.array([R, phi, z]), k % 3))
    return recs


def _rel(a, b, axis=0):
    scale = np.maximum(np.abs(b).max(axis=axis), 1e-300)
    return np.abs(a - b).max(axis=axis) / scale


# ------------------------------------------------------------ oracle pins
def test_oracle_matches_reference_leaves_at_a0():
    """The covariant oracle at a = 0 against the reference's own HLL leaf
    outputs (solve_riemann increment and EMFs, prim2cons with dV = 1.37,
    mindt, source) on its random face states."""
    z = np.load(os.path.join(GOLDEN, "leaf_vectors.npz"))
    rec, ref = z["grmhd_hll_in"], z["grmhd_hll_out"]
    worst = {}
    got = {"dU": [], "E": [], "cons": [], "dt": [], "src": []}
    for r in rec:
        pL, pR, x, n = r[0:8], r[8:16], r[16:19], r[19:22]
        w, dAdt, dim = r[22], r[23], int(r[24])
        k = int(np.argmax(n))
        dU, E, _ = GC.riemann_hll(x, pL, pR, k, w, dAdt, M, 0.0, GAMMA)
        got["dU"].append(dU)
        got["E"].append(E[:3] if dim else E)
        pt = GC.Point(x, M, 0.0)
        got["cons"].append(GC.prim2cons(pt, GC.State(pt, pL, GAMMA), 1.37))
        xp = x + np.array([0.05, 0.02, 0.03])
        xm = x - np.array([0.05, 0.02, 0.03])
        got["dt"].append(GC.mindt(pL, w, xp, xm, M, 0.0, GAMMA))
        got["src"].append(GC.source(pL, xp, xm, 0.7, M, 0.0, GAMMA))
    worst["dU"] = _rel(np.array(got["dU"]), ref[:, 0:8]).max()
    eref = [row[8:12] if int(r[24]) == 0 else row[8:11] for row, r in zip(ref, rec)]
    worst["E"] = max(_rel(np.array([g[i] for g in got["E"]]), np.array([e[i] for e in eref])) for i in range(3))
    worst["cons"] = _rel(np.array(got["cons"]), ref[:, 12:20]).max()
    worst["dt"] = _rel(np.array(got["dt"]), ref[:, 28]).max()
    src = np.array(got["src"])
    worst["src"] = _rel(src[:, [1, 2, 4]], ref[:, [29 + 1, 29 + 2, 29 + 4]]).max()
    # L: the reference skips the Killing direction (exact zero); the oracle's
    # phi-derivative of the metric gives zero to rounding
    assert (ref[:, 29 + 3] == 0.0).all()
    worst["src_L"] = np.abs(src[:, 3]).max() / np.abs(src[:, 2]).max()
    print(worst)
    assert max(worst.values()) < 1e-12, worst


def test_oracle_matches_reference_metric_and_cons_at_a0():
    z = np.load(os.path.join(GOLDEN, "grmhd_leaf_a0.npz"))
    gam = float(z["gamma"])
    em = ec = 0.0
    for x, p, mt, cons in zip(z["x"][:128], z["prim"][:128], z["metric"][:128], z["cons"][:128]):
        pt = GC.Point(x, float(z["M"]), 0.0)
        mine = np.concatenate([[pt.al], pt.be, pt.gam.ravel(), pt.igam.ravel(), [pt.jac * pt.R]])
        em = max(em, np.abs(mine - mt).max() / np.abs(mt).max())
        c = GC.prim2cons(pt, GC.State(pt, p, gam))
        ec = max(ec, (np.abs(c - cons) / np.maximum(np.abs(cons), 1e-300)).max())
    assert em < 1e-13 and ec < 1e-11, (em, ec)


def test_oracle_identities_at_spin():
    for p, _, x, _ in _states(64, 7):
        pt = GC.Point(x, M, 0.9)
        assert abs(pt.sqrtmg - x[0]) < 1e-13 * x[0]
        st = GC.State(pt, p, GAMMA)
        assert abs(st.uu @ st.ud + 1.0) < 1e-12
        assert abs(st.bu @ (pt.g @ st.uu)) < 1e-12 * max(1.0, np.sqrt(abs(st.b2)) * abs(st.uu[0]))
        # axisymmetry: the phi-derivative of the metric is zero to rounding
        assert np.abs(GC.dmetric(x[0], x[1], x[2], M, 0.9)[1]).max() < 1e-15 * np.abs(pt.g).max()


# ------------------------------------------------------------ library pins
@pytest.mark.parametrize("a", [0.5, 0.9])
def test_library_leaves_equal_covariant_oracle(probe, a):
    recs = _states(240, 11 if a == 0.9 else 12)
    lib = probe(a, recs)
    groups = {"metric": [], "cons": [], "flux": [], "speeds": [], "riemann": [], "src": [], "dt": []}
    for (p1, p2, x, d) in recs:
        pt = GC.Point(x, M, a)
        s1 = GC.State(pt, p1, GAMMA)
        groups["metric"].append(np.concatenate([[pt.al], pt.be, pt.gam.ravel(), pt.igam.ravel()]))
        groups["cons"].append(GC.prim2cons(pt, s1))
        groups["flux"].append(GC.flux(pt, s1, d))
        dU, E, (Sl, Sr) = GC.riemann_hll(x, p1, p2, d, 0.0, 1.0, M, a, GAMMA)
        groups["speeds"].append(np.array([Sl, Sr]))
        groups["riemann"].append(np.concatenate([dU, E]))
        xp = x + np.array([0.05, 0.02, 0.03])
        xm = x - np.array([0.05, 0.02, 0.03])
        groups["src"].append(GC.source(p1, xp, xm, 0.7, M, a, GAMMA))
        groups["dt"].append(GC.mindt(p1, 0.1, xp, xm, M, a, GAMMA))
    cols = {"metric": slice(0, 22), "cons": slice(22, 30), "flux": slice(30, 38), "speeds": slice(38, 40),
            "riemann": slice(41, 53), "src": slice(53, 61), "dt": slice(61, 62)}
    worst = {}
    for k, sl in cols.items():
        ora = np.array(groups[k]).reshape(len(recs), -1)
        mine = lib[:, sl]
        if k == "flux":  # per direction: each has its own scale
            worst[k] = max(_rel(mine[[i for i, r in enumerate(recs) if r[3] == d]],
                                ora[[i for i, r in enumerate(recs) if r[3] == d]]).max() for d in range(3))
        elif k == "src":
            worst[k] = _rel(mine[:, [1, 2, 4]], ora[:, [1, 2, 4]]).max()
            # phi is Killing: the library skips it (exact zero); the oracle's
            # derivative gives zero to rounding
            assert (mine[:, 3] == 0.0).all()
            assert np.abs(ora[:, 3]).max() < 1e-13 * np.abs(ora[:, 2]).max()
        else:
            worst[k] = _rel(mine, ora).max()
    print(f"a = {a}: {worst}")
    assert max(worst.values()) < 1e-12, worst


def test_planted_spin_sign_error_is_caught(probe):
    """A mutation of the oracle's metric (the spin term of l_phi with the
    wrong sign, i.e. a -> -a in the frame dragging only) must fail the flux
    comparison: the comparison sees the a-dependent terms."""
    recs = _states(60, 13)
    lib = probe(0.9, recs)
    orig = GC.metric_cart

    def mutated(x, y, z, Mm, a):
        g = orig(x, y, z, Mm, a)
        gneg = orig(x, y, z, Mm, -a)
        out = g.copy()
        out[0, 1:] = gneg[0, 1:]
        out[1:, 0] = gneg[1:, 0]
        return out

    GC.metric_cart = mutated
    try:
        errs = []
        for (p1, p2, x, d), row in zip(recs, lib):
            pt = GC.Point(x, M, 0.9)
            F = GC.flux(pt, GC.State(pt, p1, GAMMA), d)
            errs.append(np.abs(row[30:38] - F).max() / max(np.abs(F).max(), 1e-300))
    finally:
        GC.metric_cart = orig
    assert max(errs) > 1e-6, max(errs)


# ------------------------------------------------------------ greuler.c
@pytest.mark.parametrize("a", [0.5, 0.9])
def test_greuler_leaves_equal_covariant_oracle(probe, a):
    """greuler.c's GR hydrodynamics (phys_greuler.h) is the GRMHD stress
    tensor without field: its prim2cons, fluxes, HLL wave speeds and HLL
    increment must equal the covariant oracle's at B = 0 (greuler keeps
    u_phi in prim[UPP], the oracle u_phi / R^2), to 1e-12 of each quantity's
    scale; and its cons2prim (the quartic for W - 1) must invert prim2cons
    to 1e-10 from another start.  Its geometric source and CFL step differ
    from grmhd.c's by construction (greuler.c drops the z row of the frame
    terms and uses centroids) and are pinned at a = 0 by the reference's leaf
    vectors only."""
    recs = []
    for p1, p2, x, d in _states(240, 21 if a == 0.9 else 22):
        q1, q2 = p1[:5].copy(), p2[:5].copy()
        q1[3] *= x[0] ** 2
        q2[3] *= x[0] ** 2
        recs.append((np.concatenate([q1, np.zeros(3)]), np.concatenate([q2, np.zeros(3)]), x, d))
    lib = probe(a, recs, mode=10)
    worst = {"cons": 0.0, "flux": 0.0, "speeds": 0.0, "riemann": 0.0, "roundtrip": 0.0}
    for row, (p1, p2, x, d) in zip(lib, recs):
        g1, g2 = p1.copy(), p2.copy()
        g1[3] /= x[0] ** 2
        g2[3] /= x[0] ** 2
        pt = GC.Point(x, M, a)
        s1 = GC.State(pt, g1, GAMMA)
        for k, mine, ora in (("cons", row[22:27], GC.prim2cons(pt, s1)[:5]),
                             ("flux", row[30:35], GC.flux(pt, s1, d)[:5])):
            worst[k] = max(worst[k], np.abs(mine - ora).max() / max(np.abs(ora).max(), 1e-300))
        dU, _, (Sl, Sr) = GC.riemann_hll(x, g1, g2, d, 0.0, 1.0, M, a, GAMMA)
        worst["speeds"] = max(worst["speeds"], np.abs(row[38:40] - [Sl, Sr]).max() / max(abs(Sl), abs(Sr)))
        worst["riemann"] = max(worst["riemann"], np.abs(row[41:46] - dU[:5]).max() / max(np.abs(dU[:5]).max(), 1e-300))
        back = row[53:58]
        sc = np.array([p1[0], p1[1], *([max(np.abs(p1[2:5]).max(), 1e-300)] * 3)])
        worst["roundtrip"] = max(worst["roundtrip"], (np.abs(back - p1[:5]) / sc).max())
    print(f"greuler a = {a}: {worst}")
    assert max(worst["cons"], worst["flux"], worst["speeds"], worst["riemann"]) < 1e-12, worst
    assert worst["roundtrip"] < 1e-10, worst


# ------------------------------------------------------------ grmhd2.c
@pytest.mark.parametrize("a", [0.5, 0.9])
def test_grmhd2_leaves_equal_covariant_oracle(probe, a):
    """grmhd2.c is grmhd.c with prim[UPP] = u_phi: the 3+1 metric, prim2cons,
    the fluxes, the HLL wave speeds and the HLL increment with its EMFs must
    equal the covariant oracle's (given u_phi / R^2) to 1e-12 of each
    quantity's scale; the CFL step too (the same formula).  Its source drops
    grmhd.c's cell-average corrections and is pinned at a = 0 by the
    reference's leaf vectors only."""
    recs = []
    for p1, p2, x, d in _states(240, 31 if a == 0.9 else 32):
        q1, q2 = p1.copy(), p2.copy()
        q1[3] *= x[0] ** 2
        q2[3] *= x[0] ** 2
        recs.append((q1, q2, x, d))
    lib = probe(a, recs, mode=20)
    worst = {}
    for k in ("metric", "cons", "flux", "speeds", "riemann", "dt"):
        worst[k] = 0.0
    for row, (p1, p2, x, d) in zip(lib, recs):
        g1, g2 = p1.copy(), p2.copy()
        g1[3] /= x[0] ** 2
        g2[3] /= x[0] ** 2
        pt = GC.Point(x, M, a)
        s1 = GC.State(pt, g1, GAMMA)
        dU, E, (Sl, Sr) = GC.riemann_hll(x, g1, g2, d, 0.0, 1.0, M, a, GAMMA)
        xp = x + np.array([0.05, 0.02, 0.03])
        xm = x - np.array([0.05, 0.02, 0.03])
        for k, sl, ora in (("metric", slice(0, 22), np.concatenate([[pt.al], pt.be, pt.gam.ravel(), pt.igam.ravel()])),
                           ("cons", slice(22, 30), GC.prim2cons(pt, s1)),
                           ("flux", slice(30, 38), GC.flux(pt, s1, d)),
                           ("speeds", slice(38, 40), np.array([Sl, Sr])),
                           ("riemann", slice(41, 53), np.concatenate([dU, E])),
                           ("dt", slice(61, 62), np.array([GC.mindt(g1, 0.1, xp, xm, M, a, GAMMA)]))):
            worst[k] = max(worst[k], np.abs(row[sl] - ora).max() / max(np.abs(ora).max(), 1e-300))
    print(f"grmhd2 a = {a}: {worst}")
    assert max(worst.values()) < 1e-12, worst

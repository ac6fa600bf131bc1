"""The reference-side periodic-z ghost fill (oracle/ref/ref_exchange_serial.c,
linked into the reference harness build in place of exchange.c) held to the
definition of exchange.c:173-279 with mpisetup.c:17,38: in every stored state
of every periodic-z fixture, each ghost z-layer is a bit-for-bit copy of the
real layer one period away (prim, cons, piph, dphi, wiph and, for MHD, the
face fluxes Phi).  A second check makes sure the definition is not met by
accident: no ghost layer equals the layer next to its source (the off-by-one
fills) in any fixture state past step 0.
"""
import numpy as np
import pytest

from dumpio import load_fixture
from zperiodic import ghost_mismatches, ghost_source, layer_slices, roll_real_layers

PERIODIC = ["c2_mhd_small.npz", "c2_mhd_small_hllc.npz", "c2_mhd_small_hll.npz", "c2_mhd_sinkterms.npz",
            "c2e_small.npz", "c2e_nozzle.npz", "c3z_damp.npz"]


@pytest.mark.parametrize("fixture", PERIODIC)
def test_reference_ghost_layers_follow_the_definition(fixture):
    ref = load_fixture(fixture)
    for s, st in sorted(ref.items()):
        Nr, Nz, Ng = st["Nr"], st["Nz"], st["NgZa"]
        assert st["NgZb"] == Ng == 2
        fields = ["prim", "cons", "piph", "dphi", "wiph"] + (["Phi"] if "Phi" in st else [])
        assert ghost_mismatches(st, st["Np"], Nr, Nz, Ng, fields) == [], (fixture, s)
        if s == 0:
            continue
        # discriminating: an off-by-one source layer would not match
        sl = layer_slices(st["Np"], Nr, Nz)
        for k in range(Nz):
            src = ghost_source(k, Nz, Ng)
            if src is None:
                continue
            for wrong in (src - 1, src + 1):
                if 0 <= wrong < Nz and wrong != k:
                    assert not np.array_equal(st["prim"][:, sl[k]], st["prim"][:, sl[wrong]]), (fixture, s, k, wrong)


def test_roll_helper_round_trip():
    """roll_real_layers by R real layers is the identity, by 1 then -1 too,
    and a rolled state satisfies the ghost definition."""
    Nr, Nz, Ng = 3, 8, 2
    Np = np.tile(np.array([4, 5, 6]), Nz)
    n = int(Np.sum())
    a = np.arange(2 * n, dtype=float).reshape(2, n)
    b = roll_real_layers(a, Np, Nr, Nz, Ng, 1)
    assert ghost_mismatches({"x": b}, Np, Nr, Nz, Ng, ["x"]) == []
    assert np.array_equal(roll_real_layers(b, Np, Nr, Nz, Ng, -1)[:, layer_slices(Np, Nr, Nz)[Ng].start:
                                                                   layer_slices(Np, Nr, Nz)[Nz - Ng - 1].stop],
                          a[:, layer_slices(Np, Nr, Nz)[Ng].start:layer_slices(Np, Nr, Nz)[Nz - Ng - 1].stop])
    c = roll_real_layers(a, Np, Nr, Nz, Ng, Nz - 2 * Ng)
    sl = layer_slices(Np, Nr, Nz)
    assert np.array_equal(c[:, sl[Ng].start:sl[Nz - Ng - 1].stop], a[:, sl[Ng].start:sl[Nz - Ng - 1].stop])

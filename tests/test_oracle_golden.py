"""The oracle against the reference Fortran's own dumps (golden fixtures).

The fixtures were written by tests/golden/make_golden.py from the reference
compiled from its sources (oracle/_ref).  Bitwise equality is required for
every state field and for the pair list (indices, types, single-precision
w and gradients, in linked-list order).
"""
import hashlib

import numpy as np
import pytest

from conftest import bits_equal, deck_dir, describe_diff, golden, load_pkg

SMALL = {"t1_tiny": [1, 2, 3, 10, 50, 200], "t2_cspm": [1, 2, 50, 200], "t3_elastic": [1, 50, 200],
         "t4_xsph": [1, 2, 3, 10, 50, 200]}


def _sim(deck_name, oracle_lib):
    sps = load_pkg()
    return sps.Simulation.from_deck(sps.read_deck(deck_dir(deck_name)), lib_path=oracle_lib)


@pytest.mark.parametrize("deck_name", sorted(SMALL))
def test_setup_matches_reference(deck_name):
    """Read_2DMesh / set_up_dummy_nodes / Read_2DMesh D-matrix restated."""
    sps = load_pkg()
    g = golden(deck_name)
    parts = sps.build_particles(sps.read_deck(deck_dir(deck_name)))
    nn, ns, nd, nd2 = (int(v) for v in g["counts"])
    assert (parts.nnode, parts.nstress, parts.ndummy, parts.ndummy2) == (nn, ns, nd, nd2)
    for k in ("x", "mass", "rho", "hsml", "itype", "wall_position", "horizontal_or_not"):
        assert bits_equal(g[f"s0_{k}"], getattr(parts, k)), k
    for k in ("D11", "D22", "D12", "D33", "D41", "D42"):
        assert float(g[f"s0_{k}"]) == parts.D[k], k
    assert float(g["s0_pi"]) == parts.pi


@pytest.mark.parametrize("deck_name", sorted(SMALL))
def test_oracle_bitwise_small_decks(deck_name, oracle_lib):
    g = golden(deck_name)
    sim = _sim(deck_name, oracle_lib)
    done = 0
    for st in SMALL[deck_name]:
        sim.step(st - done)
        done = st
        s = sim.get_state()
        s["ivar1"] = s["internal_vars"][:, 0]
        for k in ("x", "vel", "stress", "ivar1", "f_drucker", "ddev_strn", "grad_u", "displ", "f_bound",
                  "art_visc"):
            assert bits_equal(g[f"s{st}_{k}"], s[k]), f"step {st} {k}: {describe_diff(g[f's{st}_{k}'], s[k])}"
        p = sim.get_pairs()
        assert len(p["i"]) == int(g[f"s{st}_npairs"])
        for k in ("i", "j", "type", "w", "dwdx", "dwdy"):
            assert bits_equal(g[f"s{st}_pairs_{k}"], p[k]), f"step {st} pairs {k}"
    sim.close()


def test_oracle_c1_collapse_1000_steps(oracle_lib):
    """BASELINE config 1 (5,151 + 5,000 particles, 1000 steps of 1e-5 s)."""
    g = golden("c1_collapse")
    sps = load_pkg()
    parts = sps.build_particles(sps.read_deck(deck_dir("c1_collapse")))
    for k in ("x", "mass", "rho", "hsml", "itype", "wall_position", "horizontal_or_not"):
        digest = hashlib.sha256(np.ascontiguousarray(getattr(parts, k)).tobytes()).hexdigest()
        assert digest == str(g[f"s0_sha_{k}"]), f"setup {k}"
    sim = _sim("c1_collapse", oracle_lib)
    done = 0
    for st in (1, 2, 10, 100, 1000):
        sim.step(st - done)
        done = st
        s = sim.get_state()
        s["ivar1"] = np.ascontiguousarray(s["internal_vars"][:, 0])
        for k in ("x", "vel", "stress", "ivar1", "f_drucker", "ddev_strn", "grad_u", "displ", "f_bound",
                  "art_visc"):
            digest = hashlib.sha256(np.ascontiguousarray(s[k]).tobytes()).hexdigest()
            assert digest == str(g[f"s{st}_sha_{k}"]), f"step {st} field {k}"
        p = sim.get_pairs()
        rec = np.zeros(len(p["i"]), dtype=[("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("w", "<f4"),
                                           ("dwdx", "<f4"), ("dwdy", "<f4")])
        for k in rec.dtype.names:
            rec[k] = p[k]
        assert hashlib.sha256(rec.tobytes()).hexdigest() == str(g[f"s{st}_sha_pairs"]), f"step {st} pairs"
    sim.close()

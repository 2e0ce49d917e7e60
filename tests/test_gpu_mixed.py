"""The mixed-precision form (SPS_MODE_FAST) against the double-precision oracle.

Single-precision pair arithmetic with compensated sums differs from the
reference by rounding only; these tests bound that difference:

* elastic decks (no yield switching) over long runs: the error stays at the
  level of single-precision rounding of the pair terms (1e-5 relative);
* yielding decks before yield starts (plasticity is off in the first step,
  time_sph = 0): 1e-5;
* yielding decks over more steps, where the reference's own results move by
  far more than 1e-5 under a perturbation of the initial positions at single
  precision (the yield tests f >= 0 and df >= 0 switch on it): the mixed
  form's error is bounded by ten times the double-precision oracle's spread
  under such a perturbation (the conditioning of the problem, not of the
  form);
* the pair set and statistics are those of the double-precision form;
* with the pair lists kept over steps (a skin), fewer searches than steps and
  the same bounds.
"""
import importlib
import sys

import numpy as np
import pytest

from conftest import ORACLE_LIB, deck_dir, load_pkg

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_pair(deck_name, steps, skin=0.0):
    sps = load_pkg()
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    deck = sps.read_deck(deck_dir(deck_name))
    mix = sps.Simulation.from_deck(deck, mode=sps.MODE_FAST)
    if skin:
        mix.set_list_skin(skin * mix.params.dx)
    ora = sps.Simulation.from_deck(deck, lib_path=ORACLE_LIB)
    assert mix.backend == "hip-gfx950", mix.backend
    x0 = ora.get_state(["x"])["x"]
    say(f"mixed {deck_name} {steps} steps")
    mix.step(steps)
    ora.step(steps)
    return acc.compare_states(mix.get_state(), ora.get_state(), x0), mix, ora


@pytest.mark.parametrize("deck_name,steps,tol", [
    ("t3_elastic", 200, 1e-5),     # elastic: no yield switching
    ("t2_cspm", 1, 1e-5),          # yielding decks, before yield starts
    ("t1_tiny", 1, 1e-5),
    ("c1_collapse", 1, 5e-5),      # stresses start at zero: first-step sums cancel
    ("v1_vslope_sp1", 200, 1e-5),  # prescribed velocities and stresses, elastic
    ("l1_localisation", 50, 1e-5),  # BCs with a time ramp, Perzyna von Mises, before yield
    ("b1_bui_sp1", 50, 1e-5),      # XSPH position update (Bui fork slope)
    ("b4_bui_outside", 50, 1e-5),  # the outside approach, fixed shifts
    ("b8_bui_out3", 50, 1e-5),
    ("t4_xsph", 1, 1e-5),
])
def test_mixed_close_to_oracle(deck_name, steps, tol):
    err, _, _ = run_pair(deck_name, steps)
    say(f"  {err}")
    for k in ("disp_rel_max", "vel_rel_max", "stress_rel_max"):
        assert err[k] <= tol, f"{deck_name}: {k} {err[k]:.3e} > {tol}"


def test_mixed_pairs_and_stats_equal_double_form():
    _, mix, ora = run_pair("t1_tiny", 1)
    sm, so = mix.stats(), ora.stats()
    for k in ("n_pairs", "n_by_type", "ndivx", "ndivy"):
        assert sm[k] == so[k], f"stats {k}: {sm[k]} vs {so[k]}"


@pytest.mark.parametrize("deck_name,steps,tol", [("t3_elastic", 200, 1e-5), ("t2_cspm", 1, 1e-5),
                                                ("b1_bui_sp1", 50, 1e-5), ("b8_bui_out3", 50, 1e-5)])
def test_mixed_with_list_skin_close_to_oracle(deck_name, steps, tol):
    err, mix, _ = run_pair(deck_name, steps, skin=0.25)
    n = mix.search_count()
    say(f"  {n} searches: {err}")
    assert 1 <= n < max(steps, 2), f"{n} searches in {steps} steps"
    for k in ("disp_rel_max", "vel_rel_max", "stress_rel_max"):
        assert err[k] <= tol, f"{deck_name}: {k} {err[k]:.3e} > {tol}"


def test_list_skin_needs_the_mixed_form():
    sps = load_pkg()
    sim = sps.Simulation.from_deck(sps.read_deck(deck_dir("t1_tiny")))
    with pytest.raises(sps.SpsError):
        sim.set_list_skin(0.001)


@pytest.mark.parametrize("deck_name,steps,skin", [("t2_cspm", 60, 0.0), ("t1_tiny", 3, 0.0), ("t1_tiny", 60, 0.0),
                                                   ("t2_cspm", 60, 0.25), ("t1_tiny", 60, 0.25)])
def test_mixed_error_within_conditioning_of_yielding_deck(deck_name, steps, skin):
    sps = load_pkg()
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    setup = importlib.import_module("stress-particle-sph_amd.setup")
    err, _, ora = run_pair(deck_name, steps, skin)
    ref = ora.get_state()
    deck = sps.read_deck(deck_dir(deck_name))
    parts = setup.build_particles(deck)
    st = sps.initial_state(parts)
    x0 = st["x"].copy()
    n = ora.n
    rng = np.random.default_rng(7)
    st["x"][:n] = st["x"][:n] * (1 + 1e-7 * rng.standard_normal((n, 2)))
    pert = sps.Simulation(sps.make_params(deck, parts), ORACLE_LIB)
    pert.set_state(st)
    pert.deck, pert.parts = deck, parts
    pert.step(steps)
    spread = acc.compare_states(pert.get_state(), ref, x0)
    say(f"  mixed {err}\n  oracle perturbed at 1e-7: {spread}")
    for k in ("stress_rel_rms", "vel_rel_max"):
        assert err[k] <= 10 * spread[k] + 1e-6, f"{k}: mixed {err[k]:.3e} vs conditioning {spread[k]:.3e}"

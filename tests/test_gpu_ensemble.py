"""The batched form: instances concatenated into one particle set, each with
its own material, and the mixed form's step replayed as a graph.

* An ensemble of slope instances (cohesive and cohesionless, vertical and
  inclined) in the double-precision form equals the oracle running the same
  ensemble (same per-particle materials) bit for bit, through yield.
* Graph replay of the mixed form's list-reusing step equals direct launches
  bit for bit (a deck, and an ensemble).
* Each instance of a mixed-form ensemble stays within single-precision
  rounding of the same instance run alone (the concatenation adds no
  interaction).
* The autotuner returns one of its candidates.
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


def ensemble_mod():
    load_pkg()
    return importlib.import_module("stress-particle-sph_amd.ensemble")


def small_instances(ens):
    return [ens.Instance(6.0, 60.0, 20.0, 0.0, 1e7), ens.Instance(5.0, 45.0, 35.0, 500.0, 3e7),
            ens.Instance(7.0, 90.0, 25.0, 0.0, 5e6), ens.Instance(5.5, 75.0, 30.0, 2e3, 2e7)]


def bits_equal_states(a, b):
    return [k for k in a if not np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))]


def test_ensemble_double_form_equals_oracle():
    sps = load_pkg()
    ens = ensemble_mod()
    base = sps.read_deck(deck_dir("c2_slope_dx01"))
    inst = small_instances(ens)
    gpu = ens.open_ensemble(sps, base, inst, dx=0.25, mode=sps.MODE_ORDERED)
    ora = ens.open_ensemble(sps, base, inst, dx=0.25, mode=sps.MODE_ORDERED, lib_path=ORACLE_LIB)
    say(f"ensemble of {len(inst)} instances, {gpu.n2} particles")
    gpu.step(60)
    ora.step(60)
    a, b = gpu.get_state(), ora.get_state()
    bad = bits_equal_states(a, b)
    io = gpu.ensemble[1][:gpu.n]
    plastic = [int(np.count_nonzero(b["internal_vars"][io == k, 0])) for k in range(len(inst))]
    say(f"  plastic stress points per instance {plastic}")
    assert not bad, bad
    assert sum(plastic) > 0


@pytest.mark.parametrize("which", ["deck", "ensemble", "bcs"])
def test_graph_replay_equals_direct_launches(which):
    sps = load_pkg()
    ens = ensemble_mod()
    states = []
    for graph in (False, True):
        if which in ("deck", "bcs"):
            # "bcs": prescribed velocities ramped in time, whose values a
            # replayed step must take from the step, not from its capture
            name = "t2_cspm" if which == "deck" else "l1_localisation"
            sim = sps.Simulation.from_deck(sps.read_deck(deck_dir(name)), mode=sps.MODE_FAST)
            sim.set_list_skin(0.25 * sim.params.dx)
        else:
            base = sps.read_deck(deck_dir("c2_slope_dx01"))
            sim = ens.open_ensemble(sps, base, small_instances(ens), dx=0.25, skin=0.25 * 0.25)
        sim.set_graph(graph)
        sim.step(40)
        states.append(sim.get_state())
        say(f"  {which} graph={graph}: {sim.search_count()} searches")
    assert not bits_equal_states(*states)


def test_ensemble_instances_match_single_runs():
    sps = load_pkg()
    ens = ensemble_mod()
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    base = sps.read_deck(deck_dir("c2_slope_dx01"))
    inst = small_instances(ens)[:2]
    steps = 30
    batch = ens.open_ensemble(sps, base, inst, dx=0.25)
    batch.step(steps)
    B = batch.get_state()
    _, inst_of, offs = batch.ensemble
    for k in range(len(inst)):
        one = ens.open_ensemble(sps, base, [inst[k]], dx=0.25)
        x0 = one.get_state(["x"])["x"]
        one.step(steps)
        S = one.get_state()
        # the instance's particles keep their order within each kind range
        sel = np.nonzero(inst_of == k)[0]
        sub = {name: B[name][sel[sel < len(B[name])]] for name in ("x", "vel", "stress", "internal_vars")}
        sub["x"] = sub["x"] - np.array([offs[k], 0.0])
        ref = {name: S[name] for name in ("x", "vel", "stress", "internal_vars")}
        ref["x"] = ref["x"][:len(sub["x"])]
        e = acc.compare_states(sub, ref, x0 - 0.0)
        say(f"  instance {k}: {e}")
        assert e["stress_rel_max"] < 1e-4 and e["vel_rel_max"] < 1e-4


def test_autotuner_picks_a_candidate():
    sps = load_pkg()
    ens = ensemble_mod()
    tune = importlib.import_module("stress-particle-sph_amd.autotune")
    base = sps.read_deck(deck_dir("c2_slope_dx01"))
    inst = small_instances(ens)
    best, table = tune.tune(lambda: ens.open_ensemble(sps, base, inst, dx=0.25), 0.25, skins=(0.1, 0.3),
                            warmup=2, steps=4)
    say(f"  {best}; {table}")
    assert len(table) == 4
    assert (best["skin_dx"], best["graph"]) in {(r["skin_dx"], r["graph"]) for r in table}


def test_combined_form_on_an_ensemble():
    """The combined form end to end on a small ensemble: the plain FP64 run
    logs its plastic tuples with each point's material, a network trained on
    them (and on an augmentation over drawn materials) corrects the mixed,
    batched, list-reusing, graph-replayed run, and that run stays close to the
    FP64 one."""
    import importlib
    sps = load_pkg()
    ens = ensemble_mod()
    sur = importlib.import_module("stress-particle-sph_amd.surrogate")
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    base = sps.read_deck(deck_dir("c2_slope_dx01"))
    inst = small_instances(ens)
    steps = 60
    ref = ens.open_ensemble(sps, base, inst, dx=0.25, mode=sps.MODE_ORDERED)
    x0 = ref.get_state(["x"])["x"]
    sur.log_plastic(ref, 1 << 20)
    ref.step(steps)
    S_ref = ref.get_state()
    sn, st, sr, mat = sur.plastic_log(ref, 1 << 20)
    ref.close()
    table = ens.build_ensemble(inst, 0.25, base.sml)[1]
    assert len(sn) > 0
    # each tuple carries one of the instances' materials
    assert set(map(tuple, np.round(mat, 12))) <= set(map(tuple, np.round(table[:, :2], 12)))
    tf, coh = mat[:, 0], mat[:, 1]
    ok = sur.correctable(st, tf, coh)
    sn, st, sr, tf, coh = sn[ok], st[ok], sr[ok], tf[ok], coh[ok]
    lib = __import__("ctypes").CDLL(sps.PRODUCT_LIB)
    S, Sc = float(np.abs(st).max()), max(float(coh.max()), 1e3)
    rng = np.random.default_rng(0)
    a = sur.augment(lib, sn, st, tf, coh, rng, copies=8, scale_max=10.0, material=((0.0, 2e3), (20.0, 35.0)))
    x = np.vstack([sur.features(sn, st, coh, tf, S, Sc), sur.features(a[0], a[1], a[3], a[4], S, Sc)])
    t = np.vstack([sur.targets(st, sr, tf, coh), sur.targets(a[1], a[2], a[4], a[3])])
    net = sur.Network(lib, seed=3, batch_rows=4096)
    losses = net.train(x, t, epochs=40, batch=512, lr=2e-3)
    comb = ens.open_ensemble(sps, base, inst, dx=0.25, skin=0.25 * 0.25)
    comb.set_graph(True)
    sur.attach(comb, net, S, Sc, 1e-5)
    comb.step(steps)
    e = acc.compare_states(comb.get_state(), S_ref, x0)
    ev, kept = sur.stats(comb)
    comb.close()
    net.close()
    say(f"  {len(sn)} correctable tuples, loss {losses[0]:.3e} -> {losses[-1]:.3e}; evaluated {ev}, kept {kept}; {e}")
    assert ev > 0 and 0 <= kept <= ev
    assert e["stress_rel_rms"] < 1e-2 and e["disp_rel_max"] < 1e-2

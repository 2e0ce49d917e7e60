"""The BASELINE configurations at their stated sizes, on the GPU.

* C2 (the cohesive slope at dx 0.02 m, 251,001 + 250,000 particles, and at
  dx 0.01 m, 1,002,001 + 1,000,000, the bench's deck): 1000 FP64 steps,
  every state field and the pair list bitwise against the oracle's digests
  at steps 1, 2, 10, 100, 500 and 1000 (tests/golden/<deck>_oracle.npz,
  written in the build container by tests/golden/make_oracle_digests.py;
  the oracle itself is pinned to the compiled reference on the smaller
  decks).
* C3 (64 slope instances from seed 2024 at dx 0.05 m, 7.7 M particles): the
  batch in the FP64 form bitwise against the oracle running the same batch
  over a step (seconds per step for the oracle on the host's threads); then
  every instance of the batch against the same instance run alone over 20
  steps (alone at the x offset it has in the batch, so that its coordinates
  are the same bits).  The list order of a pair depends on the whole
  particle set -- its cell grid and its pair counts (the high-water mark of
  2_SPH_main_2018.f90:1361-1369) -- so an instance's particles sum their
  terms in another order in the batch than alone and need not agree to the
  bit.  The difference is
  summation-order rounding (1.6e-12 of the largest velocity after 20 steps
  on MI355X); the instances are held to 1e-10 relative, and the count of
  bitwise-equal ones and the worst difference are reported.  A wall
  dummy's stored stress is a copy of the stress of its last node partner in
  list order, taken in the last derivative pass (2_SPH_main_2018.f90:580):
  which node that is depends on the list order, so the dummies' stored
  stresses are not compared (the nodes' stresses they copy are).  Mass is
  constant and every instance's momentum is the one it has alone.
* C4 (the learned return map on the 1 M slope, c2_slope_dx01): a network
  trained on the plain method's tuples of the FP64 run is attached to the
  mixed form; the yield-function check keeps or rejects each correction,
  and the run stays within a stated bound of the FP64 run.
* C5 (biaxial compression, Drucker-Prager, rigid platens, friction angle
  30 deg +- 2 % per particle from a seed): at dx 1 mm (40 k particles),
  bitwise against the oracle with the same per-particle materials over 50
  steps; at the full dx 0.05 mm (8 M + 8 M particles), the platen, mass and
  pair-count properties over 5 steps.
"""
import hashlib
import importlib
import sys
import time

import numpy as np
import pytest

from conftest import ORACLE_LIB, bits_equal, deck_dir, describe_diff, golden, load_pkg

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vel", "stress", "internal_vars", "f_drucker", "ddev_strn", "grad_u", "displ", "f_bound",
          "art_visc", "if_out_domain")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pair_digest(p):
    rec = np.zeros(len(p["i"]), dtype=[("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("w", "<f4"),
                                       ("dwdx", "<f4"), ("dwdy", "<f4")])
    for k in rec.dtype.names:
        rec[k] = p[k]
    return hashlib.sha256(rec.tobytes()).hexdigest()


@pytest.mark.parametrize("deck_name", ["c2_slope_dx02", "c2_slope_dx01"])
def test_c2_1000_steps_vs_oracle_digests(deck_name):
    sps = load_pkg()
    g = golden(f"{deck_name}_oracle")
    steps = sorted(int(k[1:].split("_")[0]) for k in g.files if k.endswith("_sha_x"))
    say(f"[{deck_name}] checkpoints {steps}")
    gpu = sps.Simulation.from_deck(sps.read_deck(deck_dir(deck_name)))
    done, t0 = 0, time.time()
    for st in steps:
        gpu.step(st - done)
        done = st
        s = gpu.get_state(list(FIELDS))
        bad = [k for k in FIELDS if sha(s[k]) != str(g[f"s{st}_sha_{k}"])]
        assert not bad, f"step {st}: {bad} differ from the oracle"
        p = gpu.get_pairs()
        assert len(p["i"]) == int(g[f"s{st}_npairs"]), f"step {st} pair count"
        assert pair_digest(p) == str(g[f"s{st}_sha_pairs"]), f"step {st} pairs"
        say(f"  step {st}: bitwise ({time.time() - t0:.1f} s)")
    gpu.close()


def ensemble_mod():
    load_pkg()
    return importlib.import_module("stress-particle-sph_amd.ensemble")


def test_c3_batch_vs_oracle_and_instances_alone():
    sps = load_pkg()
    ens = ensemble_mod()
    base = sps.read_deck(deck_dir("c2_slope_dx01"))
    inst = ens.draw_instances(64, 2024)
    gpu = ens.open_ensemble(sps, base, inst, dx=0.05, mode=sps.MODE_ORDERED)
    ora = ens.open_ensemble(sps, base, inst, dx=0.05, mode=sps.MODE_ORDERED, lib_path=ORACLE_LIB)
    say(f"[c3] {len(inst)} instances, {gpu.n2} particles")
    gpu.step(1)
    t0 = time.time()
    ora.step(1)
    say(f"  oracle step in {time.time() - t0:.1f} s")
    a, b = gpu.get_state(list(FIELDS)), ora.get_state(list(FIELDS))
    bad = [k for k in FIELDS if not bits_equal(b[k], a[k])]
    assert not bad, f"batch vs oracle after a step: {bad}"
    pg, po = gpu.get_pairs(), ora.get_pairs()
    assert pair_digest(pg) == pair_digest(po), "pair list"
    ora.close()
    gpu.step(19)
    s = gpu.get_state(["x", "vel", "stress", "mass"])
    table, inst_of, offs = gpu.ensemble
    nt = gpu.n                      # nodes and stress points; dummies after them
    nn_all = gpu.nn
    worst, equal = 0.0, 0
    for k in range(len(inst)):
        # alone at the offset it has in the batch: the same coordinates
        one = ens.open_ensemble(sps, base, [inst[k]], dx=0.05, mode=sps.MODE_ORDERED, x_start=float(offs[k]))
        one.step(20)
        o = one.get_state(["x", "vel", "stress", "mass"])
        sel = inst_of == k
        # the instance's particles keep their relative order in the batch
        phys = np.flatnonzero(sel[:nt])
        same = True
        for f in ("x", "vel", "stress"):
            ref, got = o[f][:one.n], s[f][phys]
            assert got.shape == ref.shape, f"instance {k} {f} shape"
            scale = max(np.abs(ref).max(), 1e-300)
            err = np.abs(got - ref).max() / scale
            worst = max(worst, err)
            same = same and bits_equal(ref, got)
            assert err <= 1e-10, f"instance {k} {f}: {describe_diff(ref, got)}"
        assert bits_equal(o["mass"], s["mass"][sel])
        mom_b = (s["mass"][phys[phys < nn_all], None] * s["vel"][phys[phys < nn_all]]).sum(axis=0)
        mom_a = (o["mass"][:one.nn, None] * o["vel"][:one.nn]).sum(axis=0)
        mscale = (o["mass"][:one.nn, None] * np.abs(o["vel"][:one.nn])).sum(axis=0).max()
        assert np.abs(mom_b - mom_a).max() <= 1e-10 * max(mscale, 1e-300), f"instance {k} momentum"
        equal += same
        one.close()
    say(f"  20 steps: {equal} of {len(inst)} instances bitwise equal to their run alone, worst relative "
        f"difference {worst:.2e}")
    gpu.close()


def test_c4_learned_return_map_on_1m_slope():
    sps = load_pkg()
    sur = importlib.import_module("stress-particle-sph_amd.surrogate")
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    deck = sps.read_deck(deck_dir("c2_slope_dx01"))
    steps = 1500
    plain = sps.Simulation.from_deck(deck)
    x0 = plain.get_state(["x"])["x"]
    sur.log_plastic(plain, 1 << 20)
    plain.step(steps)
    ref = plain.get_state()
    sn, st, sr, _ = sur.plastic_log(plain, 1 << 20)
    plain.close()
    p = plain.params
    tanfi, coh = p.props[12], p.props[13]
    ok = sur.correctable(st, tanfi, coh)
    say(f"[c4] {len(sn)} plastic tuples logged over {steps} FP64 steps, {int(ok.sum())} correctable")
    assert ok.sum() > 256
    sn, st, sr = sn[ok], st[ok], sr[ok]
    lib = __import__("ctypes").CDLL(sps.PRODUCT_LIB)
    S, Sc = float(np.abs(st).max()), max(float(coh), 1e3)
    rng = np.random.default_rng(0)
    a = sur.augment(lib, sn, st, tanfi, coh, rng, copies=8, scale_max=10.0)
    x = np.vstack([sur.features(sn, st, coh, tanfi, S, Sc), sur.features(a[0], a[1], a[3], a[4], S, Sc)])
    t = np.vstack([sur.targets(st, sr, tanfi, coh), sur.targets(a[1], a[2], a[4], a[3])])
    net = sur.Network(lib, seed=1, batch_rows=8192)
    losses = net.train(x, t, epochs=100, batch=4096, lr=2e-3)
    say(f"  trained on {len(x)} rows: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.05 * losses[0]
    mix = sps.Simulation.from_deck(deck, mode=sps.MODE_FAST)
    mix.set_list_skin(0.25 * p.dx)
    sur.attach(mix, net, S, Sc, tol=1e-5)
    mix.step(steps)
    ev, kept = sur.stats(mix)
    e = acc.compare_states(mix.get_state(), ref, x0)
    say(f"  {ev} plastic states through the network, {kept} kept ({kept / max(ev, 1):.1%}), "
        f"{ev - kept} fell back to the exact return map; vs FP64: {e}")
    assert ev > 0 and 0 < kept <= ev
    assert e["stress_rel_max"] < 1e-2 and e["disp_rel_max"] < 1e-2
    net.close()
    mix.close()


def biaxial_mod():
    load_pkg()
    return importlib.import_module("stress-particle-sph_amd.biaxial")


def test_c5_reduced_with_perturbed_friction_vs_oracle():
    sps = load_pkg()
    bx = biaxial_mod()
    deck = sps.read_deck(deck_dir("c5_biaxial_dx1"))
    gpu = bx.open_biaxial(sps, deck, seed=2024)
    ora = bx.open_biaxial(sps, deck, seed=2024, lib_path=ORACLE_LIB)
    say(f"[c5 dx 1 mm] {gpu.n2} particles, {len(np.unique(gpu.biaxial[1]))} friction levels in use")
    for st in (1, 10, 50):
        gpu.step(st - gpu.itimestep_sph.value)
        ora.step(st - ora.itimestep_sph.value)
        a, b = gpu.get_state(list(FIELDS)), ora.get_state(list(FIELDS))
        bad = [f"{k}: {describe_diff(b[k], a[k])}" for k in FIELDS if not bits_equal(b[k], a[k])]
        assert not bad, f"step {st}: " + "; ".join(bad)
        assert gpu.stats()["n_pairs"] == ora.stats()["n_pairs"]
    plastic = int(np.count_nonzero(b["internal_vars"][gpu.nn:, 0]))
    say(f"  bitwise through step 50; {plastic} plastic stress points")
    gpu.close()
    ora.close()


def test_c5_full_size_properties():
    """8 M + 8 M particles (one CPU oracle step takes minutes there: the
    bitwise check is the reduced deck's above, with the same kernels)."""
    sps = load_pkg()
    bx = biaxial_mod()
    deck = sps.read_deck(deck_dir("c5_biaxial"))
    t0 = time.time()
    gpu = bx.open_biaxial(sps, deck, seed=2024)
    nn, n = gpu.nn, gpu.n
    say(f"[c5 full] {nn} velocity + {n - nn} stress particles, set up in {time.time() - t0:.1f} s")
    assert nn > 8_000_000 and n - nn == 8_000_000
    gpu.step(1)
    a = gpu.get_state(["mass"])
    st = gpu.stats()
    # pair count of the lattices: 10 node-node pairs per node, 16 node-point
    # and 10 point-point pairs per stress point in the interior; types 3, 1, 2
    expect = 10 * nn + 26 * (n - nn)
    say(f"  {st['n_pairs']} pairs (interior lattice count {expect}); by type {st['n_by_type'][:4]}")
    assert 0.99 * expect < st["n_pairs"] <= expect
    assert st["n_by_type"][3] <= 10 * nn and st["n_by_type"][1] <= 16 * (n - nn)
    assert st["n_by_type"][2] <= 10 * (n - nn) and st["n_by_type"][0] == 0
    gpu.step(4)
    s = gpu.get_state(["x", "vel", "stress", "mass"])
    x0 = gpu.parts.x
    top = np.isclose(x0[:nn, 1], 0.2)
    bottom = np.isclose(x0[:nn, 1], 0.0)
    assert np.all(s["vel"][:nn][top, 1] == -0.05) and np.all(s["vel"][:nn][top, 0] == 0.0)
    assert np.all(s["vel"][:nn][bottom] == 0.0)
    assert np.all(np.isfinite(s["stress"])) and np.all(np.isfinite(s["x"]))
    assert bits_equal(a["mass"], s["mass"])
    say("  platens prescribed, state finite, mass unchanged after 5 steps")
    gpu.close()

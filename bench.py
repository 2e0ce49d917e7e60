#!/usr/bin/env python3
"""Throughput of the Stress-Particle SPH hot path on one MI355X.

Workload: the cohesive vertical slope at full size -- 10 m x 10 m block on a
0.01 m lattice (1,002,001 velocity particles + 1,000,000 stress particles)
with a left wall and a 30 m bottom wall (12,177 dummies), Drucker-Prager
c = 10 kPa, phi = 20 deg, E = 10 MPa, nu = 0.3, rho = 1850, dt = 1e-5 s, from
rest under gravity (deck stress-particle-sph_amd/decks/c2_slope_dx01).  The
form measured is the plain FP64 one whose results are bitwise equal to the
reference (ordered summation).

One step = one `time_integration` of the reference: neighbour search and the
four RK4 stages with all their pair loops, plastic update, return mapping and
position update.  The metric is pair interactions/s = (pairs found by each
step's search, summed over the timed steps) / wall time.  Inputs are resident
on the GPU before the timed region; nothing is copied to the host inside it.

Beside it, the mixed-precision form (single-precision pair arithmetic with
compensated sums, double-precision state and return map) is timed on the
same workload, and its error against the double-precision form (whose
results are the reference's bit for bit) is measured after --acc-steps steps
from the same initial state: displacement, stress and failure-surface
location, next to the spread of the double-precision form itself under a
single-precision perturbation of the initial positions (the conditioning of
the problem).

The learned stress update (the surrogate form) is trained on the GPU on the
plastic tuples the double-precision run logged, then used in the
mixed-precision form with list reuse; its speed, its error after
--acc-steps steps and the share of its corrections kept by the yield-function
check are reported.  The batched form runs --instances slope instances drawn
from a seed (heights 5-20 m, angles 30-90 deg, dx 0.05 m, each with its own
material) as one particle set, with the configuration the autotuner picks,
beside the same instances run one after the other.  The combined form runs
that ensemble mixed, batched, with list reuse and the learned stress update
(trained on the plain method's tuples of the ensemble, with their
materials), against the plain FP64 form one instance at a time and as one
particle set; its stress error is measured after 500 and --ens-acc-steps
steps, beside the mixed form without the network and the FP64 form's own
spread under a 1e-7 perturbation.

The CPU baseline is the oracle (a C restatement of the reference, 1 thread)
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair interactions/s and steps/s at 1M particles; surrogate+FP32 vs plain FP64"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def load_pkg():
    import importlib
    return importlib.import_module("stress-particle-sph_amd")


def algorithmic_bytes(kernel, st, counts, params):
    """Minimal single-touch bytes of one launch of `kernel` (the average over
    its launches in a step where they differ).

    Pair-list entries are read once each (16-bit window index 2 B, w 4 B,
    gradient 2 x 4 B, as far as the kernel uses them); every particle's state
    that the launch reads or writes is counted once, however many receivers
    gather it.  P_t = pairs of type t of the last search."""
    nn, ns, nd = counts
    P = st["n_by_type"]
    P1, P3, P6, P9 = P[1], P[3], P[6], P[9]
    cspm = params.cspm
    av = params.alpha > 0 or params.beta > 0
    if kernel == "k_deriv_nodes":
        lists = 10 * (P1 + P6) + (10 * 2 * P3 if av else 0)
        node = 16 + 16 + 32 + 8 + 8 + 8 + 16 + 16 + 16 + 16 + 16   # x v sigma m rho h rk_vel(rw) rhs2 av fb
        sp = 24 + 8 + 8 + (16 if cspm else 0)                        # sigma(xx,yy,xy) m rho (x)
        dummy = 16 + 8 + 8 + 8 + 4                                   # x m rho h orig
        return lists + nn * node + ns * sp + nd * dummy
    if kernel == "k_sp_update":
        # four launches per step with the derivative, one (after the last stage) without
        deriv = 14 * (P1 + P9) + ns * (16 + 16 + 32 + 32 + 8 + 32 + 32 + 64 + 16) \
            + nn * (16 + 16 + 8 + 8) + nd * (16 + 8 + 8 + 4 + 4)
        update = 6 * P1 + ns * (16 + 64) + nn * (16 + 8 + 8)
        return (4 * deriv + update) / 5
    if kernel == "k_spu_nodes":
        return 6 * (P1 + P6) + nn * (32 + 32 + 8 + 16) + ns * (32 + 8 + 8 + 8)
    if kernel == "k_count":
        # three launches per step (node, stress-point and dummy receivers)
        entries_a = (P1 + P6) + (P1 + P9)
        lists = 14 * entries_a + 10 * 2 * P3 + 2 * (nn + ns) + 2 * nn
        return ((nn + ns + nd) * (16 + 8 + 4 + 4 + 4) + lists) / 3
    return None


def measured_traffic(deck_name, kernel):
    """HBM bytes per launch of `kernel` from the calibrated FETCH_SIZE and
    WRITE_SIZE counter runs summarised in profiles/ (None if not profiled)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{deck_name}.json")
    if not os.path.exists(path):
        return None
    k = json.load(open(path))["kernels"].get(kernel)
    return None if k is None else k["bytes"]


def run_gpu(args, sps, deck, mode=0, skin=0.0, net=None, graph=False, profile=True):
    import torch
    sim = sps.Simulation.from_deck(deck, mode=mode)
    if skin:
        sim.set_list_skin(skin)
    if graph:
        sim.set_graph(True)
    if net is not None:
        import importlib
        importlib.import_module("stress-particle-sph_amd.surrogate").attach(sim, net, *net.scales)
    counts = (sim.parts.nnode, sim.parts.nstress, sim.parts.ndummy)
    print(f"[bench] deck {args.deck}: {counts[0]} nodes, {counts[1]} stress points, {counts[2]} dummies",
          file=sys.stderr, flush=True)
    for k in range(args.warmup):
        sim.step(1)
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    sim.set_profiling(profile)
    sim.profile(reset=True)
    torch.cuda.synchronize()
    pairs = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.step(1)
        pairs += sim.stats()["n_pairs"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = sim.profile(reset=True)
    st = sim.stats()
    elapsed = t1 - t0
    return sim, counts, pairs, elapsed, prof, st


def form_errors(args, sps, deck):
    """Errors of the mixed-precision form against the double-precision one
    after args.acc_steps steps from rest, and the double-precision form's own
    spread under a 1e-7 relative perturbation of the initial positions."""
    import importlib
    import numpy as np
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    setup = importlib.import_module("stress-particle-sph_amd.setup")
    parts = setup.build_particles(deck)

    def run(mode, perturb=0.0, skin=0.0, log=0, net=None):
        st = sps.initial_state(parts)
        if perturb:
            n = parts.nnode + parts.nstress
            rng = np.random.default_rng(2024)
            st["x"][:n] = st["x"][:n] * (1 + perturb * rng.standard_normal((n, 2)))
        sim = sps.Simulation(sps.make_params(deck, parts, mode=mode), sps.PRODUCT_LIB)
        sim.set_state(st)
        sim.deck, sim.parts = deck, parts
        if skin:
            sim.set_list_skin(skin)
        if log:
            sur.log_plastic(sim, log)
        if net is not None:
            sur.attach(sim, net, *net.scales)
        done = 0
        while done < args.acc_steps:
            k = min(500, args.acc_steps - done)
            sim.step(k)
            done += k
            print(f"[bench] accuracy run mode {mode} perturb {perturb:g}: {done} steps", file=sys.stderr,
                  flush=True)
        out = sim.get_state()
        if log:
            logged[0] = sur.plastic_log(sim, log)
        if net is not None:
            logged[1] = sur.stats(sim)
        sim.close()
        return out

    sur = importlib.import_module("stress-particle-sph_amd.surrogate")
    logged = [None, None]
    x0 = sps.initial_state(parts)["x"]
    ref = run(sps.MODE_ORDERED, log=args.log_cap)
    mixed = acc.compare_states(run(sps.MODE_FAST), ref, x0)
    skinned = acc.compare_states(run(sps.MODE_FAST, skin=args.skin * deck_dx(sps, deck)), ref, x0)
    spread = acc.compare_states(run(sps.MODE_ORDERED, 1e-7), ref, x0)
    rnd = lambda d: {k: float(f"{v:.4g}") for k, v in d.items()}
    out = {"steps": args.acc_steps, "mixed_vs_fp64": rnd(mixed), "mixed_skin_vs_fp64": rnd(skinned),
           "fp64_perturbed_1e-7_vs_fp64": rnd(spread)}
    surrogate = None
    net, info = train_surrogate(args, sps, sur, logged[0], parts, deck)
    if net is not None:
        e = acc.compare_states(run(sps.MODE_FAST, skin=args.skin * deck_dx(sps, deck), net=net), ref, x0)
        ev, kept = logged[1]
        info.update({"mixed_skin_surrogate_vs_fp64": rnd(e), "plastic_states_evaluated": ev, "kept": kept,
                     "kept_share": round(kept / max(ev, 1), 4)})
        surrogate = (net, info)
    else:
        surrogate = (None, info)
    return out, surrogate


def train_surrogate(args, sps, sur, log, parts, deck):
    """The network trained on the logged plastic tuples of the double-precision
    run (and an augmentation of them), with its training record."""
    import ctypes
    import numpy as np
    sn, st, sr, _ = log
    p = sps.make_params(deck, parts)
    tanfi, coh = p.props[12], p.props[13]
    ok = sur.correctable(st, tanfi, coh)
    info = {"plastic_tuples_logged": int(len(sn)), "correctable": int(ok.sum())}
    if ok.sum() < 256:
        info["note"] = "too few plastic states in the logged steps to train on"
        return None, info
    sn, st, sr = sn[ok], st[ok], sr[ok]
    lib = ctypes.CDLL(sps.PRODUCT_LIB)
    S = float(np.abs(st).max())
    Sc = max(float(coh), 1e3)
    rng = np.random.default_rng(0)
    hold = rng.random(len(sn)) < 0.1
    a = sur.augment(lib, sn[~hold], st[~hold], tanfi, coh, rng, copies=8, scale_max=10.0)
    x = np.vstack([sur.features(sn[~hold], st[~hold], coh, tanfi, S, Sc), sur.features(a[0], a[1], a[3], a[4], S, Sc)])
    t = np.vstack([sur.targets(st[~hold], sr[~hold], tanfi, coh), sur.targets(a[1], a[2], a[4], a[3])])
    net = sur.Network(lib, seed=1, batch_rows=8192)
    t0 = time.perf_counter()
    losses = net.train(x, t, epochs=args.epochs, batch=4096, lr=2e-3)
    train_s = time.perf_counter() - t0
    y = net.forward(sur.features(sn[hold], st[hold], coh, tanfi, S, Sc)).astype(np.float64)
    held = float(np.linalg.norm(sur.corrected(st[hold], y, tanfi, coh) - sr[hold]) /
                 max(np.linalg.norm(st[hold] - sr[hold]), 1e-300))
    net.scales = (S, Sc, args.surrogate_tol)
    info.update({"training_rows": int(len(x)), "epochs": args.epochs, "train_s": round(train_s, 2),
                 "loss_first_last": [float(f"{losses[0]:.4g}"), float(f"{losses[-1]:.4g}")],
                 "heldout_error_vs_correction": float(f"{held:.4g}"), "tolerance": args.surrogate_tol})
    print(f"[bench] surrogate trained on {len(x)} rows in {train_s:.1f} s, held-out error {held:.3g} of the "
          "correction", file=sys.stderr, flush=True)
    return net, info


def deck_dx(sps, deck):
    import importlib
    setup = importlib.import_module("stress-particle-sph_amd.setup")
    return sps.make_params(deck, setup.build_particles(deck)).dx


def run_batched(args, sps, base_deck):
    """The batched form: args.instances slope instances as one particle set
    (autotuned), against the same instances one after the other."""
    import importlib
    import torch
    ens = importlib.import_module("stress-particle-sph_amd.ensemble")
    tune = importlib.import_module("stress-particle-sph_amd.autotune")
    dx = 0.05
    inst = ens.draw_instances(args.instances, 2024)
    sync = torch.cuda.synchronize
    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    best, table = tune.tune(lambda: ens.open_ensemble(sps, base_deck, inst, dx=dx), dx, warmup=3, steps=8,
                            sync=sync, log=log)
    tune_s = time.perf_counter() - t0
    sim = ens.open_ensemble(sps, base_deck, inst, dx=dx, skin=best["skin_dx"] * dx)
    sim.set_graph(best["graph"])
    sim.step(args.warmup)
    sync()
    nsteps = max(10, args.steps // 2)
    pairs = 0
    t0 = time.perf_counter()
    for _ in range(nsteps):
        sim.step(1)
        pairs += sim.stats()["n_pairs"]
    sync()
    el = time.perf_counter() - t0
    n2, searches = sim.n2, sim.search_count()
    sizes = np.bincount(sim.ensemble[1], minlength=len(inst))
    sim.close()
    log(f"[bench] batched: {len(inst)} instances, {n2} particles, {1e3 * el / nsteps:.3f} ms/step")
    # the same instances one at a time (same configuration), a few steps each
    seq = 0.0
    for k in range(len(inst)):
        one = ens.open_ensemble(sps, base_deck, [inst[k]], dx=dx, skin=best["skin_dx"] * dx)
        one.set_graph(best["graph"])
        one.step(2)
        sync()
        t1 = time.perf_counter()
        one.step(4)
        sync()
        seq += (time.perf_counter() - t1) / 4
        one.close()
    log(f"[bench] instances one at a time: {1e3 * seq:.3f} ms per step of all")
    return {"instances": len(inst), "seed": 2024, "dx": dx, "particles": int(n2),
            "particles_per_instance_min_max": [int(sizes.min()), int(sizes.max())],
            "ms_per_step": round(1e3 * el / nsteps, 4), "steps_per_s": round(nsteps / el, 3),
            "pairs_per_s": round(pairs / el, 1), "searches": f"{searches} in {args.warmup + nsteps} steps",
            "one_at_a_time_ms_per_step": round(1e3 * seq, 4), "speedup_vs_one_at_a_time": round(seq * nsteps / el, 3),
            "autotuned": best, "autotune_table": table, "autotune_s": round(tune_s, 2)}


def run_combined(args, sps, base_deck, tuned):
    """The combined form on the slope ensemble: mixed precision, the
    instances batched into one particle set, list reuse and graph replay as
    autotuned, and the learned stress update trained on the plain method's
    tuples of the same ensemble (each with its point's material) -- against
    the plain FP64 form on the same instances, one at a time (as one would
    run the reference) and as one particle set.  Error of the combined run
    against the FP64 run after args.ens_acc_steps steps from rest."""
    import importlib
    import torch
    ens = importlib.import_module("stress-particle-sph_amd.ensemble")
    sur = importlib.import_module("stress-particle-sph_amd.surrogate")
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    dx, sync = 0.05, torch.cuda.synchronize
    inst = ens.draw_instances(args.instances, 2024)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    # plain FP64, one instance at a time
    seq = 0.0
    for k in range(len(inst)):
        one = ens.open_ensemble(sps, base_deck, [inst[k]], dx=dx, mode=sps.MODE_ORDERED)
        one.step(1)
        sync()
        t1 = time.perf_counter()
        one.step(3)
        sync()
        seq += (time.perf_counter() - t1) / 3
        one.close()
    log(f"[bench] combined: FP64 one instance at a time {1e3 * seq:.3f} ms per step of all")
    # plain FP64 on the ensemble as one particle set: timing, plastic tuples, reference state
    nacc, ntime = args.ens_acc_steps, 10
    ref = ens.open_ensemble(sps, base_deck, inst, dx=dx, mode=sps.MODE_ORDERED)
    x0 = ref.get_state(["x"])["x"]
    sur.log_plastic(ref, args.log_cap)
    ref.step(5)
    sync()
    t1 = time.perf_counter()
    ref.step(ntime)
    sync()
    fp64_batched = (time.perf_counter() - t1) / ntime
    early = min(500, nacc)
    ref.step(early - 5 - ntime)
    S_ref_early = ref.get_state()
    ref.step(nacc - early)
    S_ref = ref.get_state()
    sn, st, sr, mat = sur.plastic_log(ref, args.log_cap)
    ref.close()
    log(f"[bench] combined: FP64 batched {1e3 * fp64_batched:.3f} ms/step, {len(sn)} plastic tuples logged")
    out = {"instances": len(inst), "seed": 2024, "steps": nacc,
           "fp64_one_at_a_time_ms_per_step": round(1e3 * seq, 4),
           "fp64_batched_ms_per_step": round(1e3 * fp64_batched, 4), "plastic_tuples_logged": int(len(sn))}
    tf, coh = mat[:, 0], mat[:, 1]
    ok = sur.correctable(st, tf, coh)
    if ok.sum() < 256:
        out["note"] = "too few plastic states in the logged steps to train on"
        return out
    sn, st, sr, tf, coh = sn[ok], st[ok], sr[ok], tf[ok], coh[ok]
    lib = __import__("ctypes").CDLL(sps.PRODUCT_LIB)
    S, Sc = float(np.abs(st).max()), max(float(coh.max()), 1e3)
    rng = np.random.default_rng(0)
    a = sur.augment(lib, sn, st, tf, coh, rng, copies=4, scale_max=10.0, material=((0.0, 20e3), (15.0, 40.0)))
    x = np.vstack([sur.features(sn, st, coh, tf, S, Sc), sur.features(a[0], a[1], a[3], a[4], S, Sc)])
    t = np.vstack([sur.targets(st, sr, tf, coh), sur.targets(a[1], a[2], a[4], a[3])])
    net = sur.Network(lib, seed=1, batch_rows=8192)
    t1 = time.perf_counter()
    losses = net.train(x, t, epochs=max(args.epochs // 2, 1), batch=4096, lr=2e-3)
    train_s = time.perf_counter() - t1
    log(f"[bench] combined: surrogate trained on {len(x)} rows in {train_s:.1f} s")
    def ens_run(mode, skin=0.0, graph=False, net_=None, perturb=0.0):
        """A run of the ensemble from rest over nacc steps, timed over a window."""
        sim = ens.open_ensemble(sps, base_deck, inst, dx=dx, mode=mode, skin=skin)
        if perturb:
            st0 = sim.get_state()
            n = sim.n
            st0["x"][:n] = st0["x"][:n] * (1 + perturb * np.random.default_rng(2024).standard_normal((n, 2)))
            sim.set_state({k: st0[k] for k in ("x", "vel", "stress", "mass", "rho", "hsml", "itype",
                                               "if_out_domain", "internal_vars", "f_drucker", "x00", "displ",
                                               "wall_position", "horizontal_or_not")})
            sim.set_materials(sim.ensemble[0], sim.ensemble[1])
        sim.set_graph(graph)
        if net_ is not None:
            sur.attach(sim, net_, S, Sc, args.surrogate_tol)
        sim.step(5)
        sync()
        t1 = time.perf_counter()
        sim.step(ntime)
        sync()
        ms = (time.perf_counter() - t1) / ntime
        sim.step(early - 5 - ntime)
        state_early = sim.get_state()
        sim.step(nacc - early)
        state = sim.get_state()
        stats = sur.stats(sim) if net_ is not None else None
        sim.close()
        return (state_early, state), ms, stats

    skin = tuned["skin_dx"] * dx
    S_comb, comb_ms, (ev, kept) = ens_run(sps.MODE_FAST, skin, tuned["graph"], net)
    e, e_early = acc.compare_states(S_comb[1], S_ref, x0), acc.compare_states(S_comb[0], S_ref_early, x0)
    net.close()
    S_mix, _, _ = ens_run(sps.MODE_FAST, skin, tuned["graph"])
    e_mix, e_mix_early = acc.compare_states(S_mix[1], S_ref, x0), acc.compare_states(S_mix[0], S_ref_early, x0)
    S_pert, _, _ = ens_run(sps.MODE_ORDERED, perturb=1e-7)
    e_pert, e_pert_early = (acc.compare_states(S_pert[1], S_ref, x0),
                            acc.compare_states(S_pert[0], S_ref_early, x0))
    rnd = lambda d: {k: float(f"{v:.4g}") for k, v in d.items()}
    out.update({"ms_per_step": round(1e3 * comb_ms, 4), "steps_per_s": round(1.0 / comb_ms, 3),
                "speedup_vs_fp64_one_at_a_time": round(seq / comb_ms, 3),
                "speedup_vs_fp64_batched": round(fp64_batched / comb_ms, 3),
                "autotuned": tuned, "training_rows": int(len(x)), "train_s": round(train_s, 2),
                "loss_first_last": [float(f"{losses[0]:.4g}"), float(f"{losses[-1]:.4g}")],
                "plastic_states_evaluated": ev, "kept": kept, "accuracy_vs_fp64": rnd(e),
                "mixed_batched_no_surrogate_vs_fp64": rnd(e_mix),
                "fp64_perturbed_1e-7_vs_fp64": rnd(e_pert),
                f"at_step_{early}": {"accuracy_vs_fp64": rnd(e_early),
                                     "mixed_batched_no_surrogate_vs_fp64": rnd(e_mix_early),
                                     "fp64_perturbed_1e-7_vs_fp64": rnd(e_pert_early)}})
    log(f"[bench] combined: {1e3 * comb_ms:.3f} ms/step, {out['speedup_vs_fp64_one_at_a_time']}x FP64 one at a "
        f"time, {out['speedup_vs_fp64_batched']}x FP64 batched; stress error {e['stress_rel_max']:.3g} max, "
        f"{e['stress_rel_rms']:.3g} rms after {nacc} steps (mixed without the network "
        f"{e_mix['stress_rel_max']:.3g} max; FP64 perturbed 1e-7 {e_pert['stress_rel_max']:.3g} max); after "
        f"{early} steps {e_early['stress_rel_max']:.3g} max (FP64 perturbed {e_pert_early['stress_rel_max']:.3g})")
    return out


def run_cpu(args, sps, deck, oracle_path):
    sim = sps.Simulation.from_deck(deck, lib_path=oracle_path)
    pairs = 0
    t0 = time.perf_counter()
    for k in range(args.cpu_steps):
        sim.step(1)
        pairs += sim.stats()["n_pairs"]
    t1 = time.perf_counter()
    sim.close()
    return pairs, t1 - t0


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--deck", default="c2_slope_dx01")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision form")
    ap.add_argument("--skin", type=float, default=0.25,
                    help="list skin of the mixed form's list-reuse leg, in lattice spacings")
    ap.add_argument("--log-cap", type=int, default=4_000_000, help="plastic tuples logged for the surrogate")
    ap.add_argument("--epochs", type=int, default=200, help="surrogate training epochs")
    ap.add_argument("--surrogate-tol", type=float, default=1e-5, help="yield-function residual kept")
    ap.add_argument("--instances", type=int, default=64, help="slope instances of the batched form (0: skip)")
    ap.add_argument("--ens-acc-steps", type=int, default=2000,
                    help="steps of the combined form's accuracy run on the ensemble (0: skip the combined form)")
    ap.add_argument("--acc-steps", type=int, default=2000,
                    help="steps of the mixed-vs-double accuracy runs (0: skip them)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 times) and exit; its "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic of the kernels")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 8, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)
    sps = load_pkg()
    deck = sps.read_deck(os.path.join(ROOT, "stress-particle-sph_amd", "decks", args.deck))
    sim, counts, pairs, elapsed, prof, st = run_gpu(args, sps, deck)
    ms_per_step = 1e3 * elapsed / args.steps
    value = pairs / elapsed
    # dominant kernel and its roofline
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dlaunch, dms) = dom
    avg_s = dms / dlaunch / 1e3
    ab = algorithmic_bytes(dname, st, counts, sim.params)
    roofline = None
    if ab is not None:
        achieved = ab / avg_s / 1e9
        roofline = {"kernel": dname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": measured_traffic(args.deck, dname),
                    "bytes_per_launch": int(ab), "avg_launch_ms": round(dms / dlaunch, 5),
                    "launches": dlaunch}
    breakdown = {k: {"launches": n, "ms_per_step": round(ms / args.steps, 4)} for k, (n, ms) in
                 sorted(prof.items(), key=lambda kv: -kv[1][1])}
    print("[bench] kernel breakdown (ms per step): " +
          ", ".join(f"{k} {v['ms_per_step']}" for k, v in breakdown.items()), file=sys.stderr, flush=True)
    # what a driver that takes the whole state back every step pays on top (PCIe)
    t0 = time.perf_counter()
    sim.get_state()
    host_copy_ms = 1e3 * (time.perf_counter() - t0)
    sim.close()
    forms = {"fp64": {"ms_per_step": round(ms_per_step, 4), "pairs_per_s": round(value, 1),
                      "host_copy_ms": round(host_copy_ms, 3)}}
    if not args.no_mixed:
        msim, _, mpairs, melapsed, mprof, _ = run_gpu(args, sps, deck, mode=sps.MODE_FAST)
        msim.close()
        forms["mixed"] = {
            "ms_per_step": round(1e3 * melapsed / args.steps, 4), "pairs_per_s": round(mpairs / melapsed, 1),
            "speedup_vs_fp64": round(elapsed / melapsed, 3),
            "kernels": {k: round(ms / args.steps, 4) for k, (n, ms) in sorted(mprof.items(), key=lambda kv: -kv[1][1])}}
        print(f"[bench] mixed form: {forms['mixed']['ms_per_step']} ms/step "
              f"({forms['mixed']['speedup_vs_fp64']}x)", file=sys.stderr, flush=True)
        skin = args.skin * sim.params.dx
        ssim, _, spairs, selapsed, sprof, _ = run_gpu(args, sps, deck, mode=sps.MODE_FAST, skin=skin)
        searches = ssim.search_count()
        ssim.close()
        forms["mixed_skin"] = {
            "ms_per_step": round(1e3 * selapsed / args.steps, 4), "pairs_per_s": round(spairs / selapsed, 1),
            "speedup_vs_fp64": round(elapsed / selapsed, 3), "skin": skin,
            "searches": f"{searches} in {args.warmup + args.steps} steps (pairs/s counts each step's pairs as "
                        "those of the last search)",
            "kernels": {k: round(ms / args.steps, 4) for k, (n, ms) in sorted(sprof.items(), key=lambda kv: -kv[1][1])}}
        print(f"[bench] mixed form with list reuse: {forms['mixed_skin']['ms_per_step']} ms/step "
              f"({forms['mixed_skin']['speedup_vs_fp64']}x, {searches} searches)", file=sys.stderr, flush=True)
        if args.acc_steps > 0:
            forms["mixed"]["accuracy"], (net, sinfo) = form_errors(args, sps, deck)
            if net is not None:
                nsim, _, npairs, nelapsed, _, _ = run_gpu(args, sps, deck, mode=sps.MODE_FAST, skin=skin, net=net)
                nsim.close()
                sinfo.update({"ms_per_step": round(1e3 * nelapsed / args.steps, 4),
                              "pairs_per_s": round(npairs / nelapsed, 1),
                              "speedup_vs_fp64": round(elapsed / nelapsed, 3)})
                net.close()
            forms["mixed_skin_surrogate"] = sinfo
    if args.instances > 0:
        forms["batched"] = run_batched(args, sps, deck)
        if args.ens_acc_steps > 0:
            forms["combined"] = run_combined(args, sps, deck, forms["batched"]["autotuned"])
    cpu = None
    if not args.no_cpu:
        oracle = os.path.join(ROOT, "oracle", "libsps_oracle.so")
        cpairs, ct = run_cpu(args, sps, deck, oracle)
        cpu = {"value": round(cpairs / ct, 1), "unit": "pair interactions/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of {args.deck} from rest ({sum(counts)} particles), "
                         f"C oracle on 1 thread, {ct:.1f} s",
               "steps_per_s": round(args.cpu_steps / ct, 4)}
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.deck}: cohesive slope 10 m x 10 m on a 0.01 m lattice, "
                               f"{counts[0]} velocity + {counts[1]} stress particles + {counts[2]} wall "
                               "dummies, plain FP64 form bitwise equal to the reference",
                   "deck": args.deck, "particles": sum(counts), "pairs_per_step": st["n_pairs"],
                   "steps_per_s": round(args.steps / elapsed, 3), "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": breakdown,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

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
(trained on the plain method's tuples of an FP64 run of the ensemble
perturbed by 1e-7, with their materials), against the plain FP64 form on
the same particle set (the baseline) and one instance at a time; every
form is timed over the windows its error is measured on (steps 0-500,
0-1000, 0-N of an N = --ens-acc-steps run), with its searches per step,
beside the mixed form without the network and the FP64 form's own spread.
The biaxial compression leg (--c5-steps, 16 M particles) compares the FP64
form with the mixed form with list reuse, with and without the learned
return map, over a stated window in the same way.

The CPU baseline is the oracle (a C restatement of the reference, bitwise
equal to it at any thread count) on a bounded sample of the same workload:
with OpenMP on every CPU this process may use (cpu_baseline), and on one
thread, the reference's own schedule (cpu_baseline.one_thread).
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


# Algorithmic bytes of the step: minimal single-touch bytes of the state
# each part of the step must read and write, with no stored pair list (the
# kernel values recomputed), per velocity particle (node), stress particle
# and wall dummy.  Per RK stage: a node reads x, v0, the RK accumulator, m,
# rho, h and writes v and the accumulator (104 B); a stress particle reads x,
# sigma0, the RK accumulator, m, rho, h, the yield-function memory and the
# plastic strain and writes sigma, the accumulator, the yield-function memory
# and grad u (224 B); a dummy reads x, m, rho (24 B); after the stages 100 B
# per particle for the final update and the position update, and 64 B per
# particle for the neighbour search.
STAGE_NODE_B, STAGE_SP_B, STAGE_DUMMY_B, POST_B, SEARCH_B = 104, 224, 24, 100, 64


def step_bytes(counts):
    nn, ns, nd = counts
    return 4 * (STAGE_NODE_B * nn + STAGE_SP_B * ns + STAGE_DUMMY_B * nd) + (POST_B + SEARCH_B) * (nn + ns)


def kernel_algorithmic_bytes(kernel, counts):
    """Algorithmic bytes of one launch of a stage kernel (the share of the
    stage bytes above that it does), with the unit they are counted in."""
    nn, ns, nd = counts
    if kernel == "k_deriv_nodes":      # the node half of a stage
        return STAGE_NODE_B * nn, f"{STAGE_NODE_B} B per velocity particle per stage x {nn}"
    if kernel == "k_sp_update":        # the stress-particle half of a stage
        return STAGE_SP_B * ns, f"{STAGE_SP_B} B per stress particle per stage x {ns}"
    if kernel == "k_count_lists":      # the search (two launches: node and stress-point receivers)
        return SEARCH_B * (nn + ns) / 2, f"{SEARCH_B} B per particle per step / 2 launches"
    return None, None


def implementation_bytes(kernel, st, counts, params):
    """Bytes one launch of `kernel` of this implementation moves at least
    (the average over its launches in a step where they differ): the
    algorithmic bytes plus its stored pair lists, read once each (16-bit
    window index 2 B, w 4 B, gradient 2 x 4 B, as far as the kernel uses
    them); every particle's state that the launch reads or writes is counted
    once, however many receivers gather it.  P_t = pairs of type t of the
    last search."""
    nn, ns, nd = counts
    P = st["n_by_type"]
    P1, P3, P6, P9 = P[1], P[3], P[6], P[9]
    cspm = params.cspm
    av = params.alpha > 0 or params.beta > 0
    if kernel == "k_deriv_nodes":
        lists = 10 * (P1 + P6)
        node = 16 + 16 + 32 + 8 + 8 + 8 + 16 + 16 + 16 + 8 + 16 + 16   # x v sigma m rho h rk_vel(rw) rhs2 av(in, out) fb
        sp = 24 + 8 + 8 + (16 if cspm else 0)                            # sigma(xx,yy,xy) m rho (x)
        dummy = 8 + 8 + 4                                                # m rho orig
        return lists + nn * node + ns * sp + nd * dummy
    if kernel == "k_av_nodes":
        return (10 * 2 * P3 + nn * (16 + 16 + 8 + 8 + 8 + 8)) if av else 0   # list B; x v m rho h, av out
    if kernel == "k_sp_update":
        # four launches per step with the derivative, one (after the last stage) without
        deriv = 14 * (P1 + P9) + ns * (16 + 16 + 32 + 32 + 8 + 32 + 32 + 64 + 16) \
            + nn * (16 + 16 + 8 + 8) + nd * (16 + 8 + 8 + 4 + 4)
        update = 6 * P1 + ns * (16 + 64) + nn * (16 + 8 + 8)
        return (4 * deriv + update) / 5
    if kernel == "k_spu_nodes":
        return 6 * (P1 + P6) + nn * (32 + 32 + 8 + 16) + ns * (32 + 8 + 8 + 8)
    if kernel == "k_count_lists":
        # two launches per step (node and stress-point receivers); each
        # stages x of the window (h too when the smoothing lengths differ)
        entries_a = (P1 + P6) + (P1 + P9)
        lists = 14 * entries_a + 10 * 2 * P3 + 2 * (nn + ns) + 2 * nn
        return ((nn + ns + nd) * (16 + 4 + 4 + 4) + lists) / 2
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
    """Warm-up, then args.steps timed steps in one sps_step call (the state
    resident on the GPU, one synchronisation at the end), then, with
    profile, the same number of steps again with every kernel bracketed by
    HIP events on the library's stream (the per-kernel times)."""
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
    sim.step(args.warmup)
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    p0 = sim.stats()["pairs_total"]
    t0 = time.perf_counter()
    sim.step(args.steps)                 # returns when the steps are done on the GPU
    t1 = time.perf_counter()
    st = sim.stats()
    pairs = st["pairs_total"] - p0
    if skin:
        # lists kept over steps: each step is counted with the pairs within
        # 2h that the last search found (list entries in the skin, beyond
        # 2h, contribute zero and are not counted)
        pairs = st["n_pairs"] * args.steps
    elapsed = t1 - t0
    prof = {}
    sim.events_s = None
    if profile:
        sim.set_profiling(True)
        sim.profile(reset=True)
        t2 = time.perf_counter()
        sim.step(args.steps)
        sim.events_s = time.perf_counter() - t2
        prof = sim.profile(reset=True)
        sim.set_profiling(False)
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
    ens = importlib.import_module("stress-particle-sph_amd.ensemble")
    tune = importlib.import_module("stress-particle-sph_amd.autotune")
    dx = 0.05
    inst = ens.draw_instances(args.instances, 2024)
    sync = lambda: None          # sps_step returns when its steps are done
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


def windows_of(n):
    """Window ends of an n-step run: 500, 1000 and n (those up to n)."""
    return sorted({w for w in (500, 1000) if w < n} | {n})


ACC_FIELDS = ["x", "vel", "stress", "internal_vars"]     # what accuracy.compare_states reads


def run_windows(sim, wins, keep=True):
    """Advance `sim` from its current step through each window end of `wins`,
    timing every segment (sps_step returns when its steps are done) and
    counting its pair searches; the state at each end when `keep`."""
    done, total, rows, states = 0, 0.0, [], []
    s0 = sim.search_count()
    for w in wins:
        t1 = time.perf_counter()
        sim.step(w - done)
        total += time.perf_counter() - t1
        done = w
        rows.append({"step": w, "s": total, "searches": sim.search_count() - s0})
        if keep:
            states.append(sim.get_state(ACC_FIELDS))
    return rows, states


def run_combined(args, sps, base_deck, tuned):
    """The combined form on the slope ensemble -- mixed precision, the
    instances batched into one particle set, list reuse and graph replay as
    autotuned, and the learned stress update -- against the plain FP64 form
    on the same particle set (the baseline) and one instance at a time.
    Every form runs from rest over args.ens_acc_steps steps; its time and its
    error against the FP64 run are taken over the same windows (steps 0-500,
    0-1000, 0-N).  The network is trained on the plastic tuples of a second
    FP64 run whose initial positions are perturbed by 1e-7 (the same
    problem, not the trajectory it is compared with), which also measures
    the FP64 form's own spread."""
    import importlib
    ens = importlib.import_module("stress-particle-sph_amd.ensemble")
    sur = importlib.import_module("stress-particle-sph_amd.surrogate")
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    dx = 0.05
    inst = ens.draw_instances(args.instances, 2024)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    rnd = lambda d: {k: float(f"{v:.4g}") for k, v in d.items()}
    nacc = args.ens_acc_steps
    wins = windows_of(nacc)
    # plain FP64, one instance at a time (sampled: steps 2-4 of each; this
    # form searches every step and reuses nothing, so its step cost is flat)
    seq = 0.0
    for k in range(len(inst)):
        one = ens.open_ensemble(sps, base_deck, [inst[k]], dx=dx, mode=sps.MODE_ORDERED)
        one.step(1)
        t1 = time.perf_counter()
        one.step(3)
        seq += (time.perf_counter() - t1) / 3
        one.close()
    log(f"[bench] combined: FP64 one instance at a time {1e3 * seq:.3f} ms per step of all")
    # the baseline: plain FP64 on the ensemble as one particle set
    ref = ens.open_ensemble(sps, base_deck, inst, dx=dx, mode=sps.MODE_ORDERED)
    x0 = ref.get_state(["x"])["x"]
    ref_rows, ref_states = run_windows(ref, wins)
    ref.close()
    log(f"[bench] combined: FP64 batched {1e3 * ref_rows[-1]['s'] / nacc:.3f} ms/step over {nacc} steps")

    def perturbed_run():
        sim = ens.open_ensemble(sps, base_deck, inst, dx=dx, mode=sps.MODE_ORDERED)
        st0 = sim.get_state()
        n = sim.n
        st0["x"][:n] = st0["x"][:n] * (1 + 1e-7 * np.random.default_rng(2024).standard_normal((n, 2)))
        sim.set_state({k: st0[k] for k in ("x", "vel", "stress", "mass", "rho", "hsml", "itype",
                                           "if_out_domain", "internal_vars", "f_drucker", "x00", "displ",
                                           "wall_position", "horizontal_or_not")})
        sim.set_materials(sim.ensemble[0], sim.ensemble[1])
        sur.log_plastic(sim, args.log_cap)
        _, states = run_windows(sim, wins)
        logged = sur.plastic_log(sim, args.log_cap)
        sim.close()
        return states, logged

    pert_states, (sn, st, sr, mat) = perturbed_run()
    e_pert = [acc.compare_states(a, b, x0) for a, b in zip(pert_states, ref_states)]
    del pert_states
    out = {"instances": len(inst), "seed": 2024, "steps": nacc, "windows": wins,
           "baseline": "FP64 batched (the plain form on the same particle set)",
           "fp64_one_at_a_time_ms_per_step": round(1e3 * seq, 4),
           "fp64_one_at_a_time_note": "sampled over steps 2-4 of each instance",
           "plastic_tuples_logged": int(len(sn)),
           "training_data": "plastic tuples of the FP64 run perturbed by 1e-7 (not the compared trajectory)"}

    def form(mode, skin=0.0, graph=False, net_=None):
        sim = ens.open_ensemble(sps, base_deck, inst, dx=dx, mode=mode, skin=skin)
        sim.set_graph(graph)
        if net_ is not None:
            sur.attach(sim, net_, S, Sc, args.surrogate_tol)
        rows, states = run_windows(sim, wins)
        stats = sur.stats(sim) if net_ is not None else None
        sim.close()
        return rows, states, stats

    def report(rows, states):
        res = []
        for r, ref_r, s_, ref_s, ep in zip(rows, ref_rows, states, ref_states, e_pert):
            e = acc.compare_states(s_, ref_s, x0)
            res.append({"steps": f"0-{r['step']}", "ms_per_step": round(1e3 * r["s"] / r["step"], 4),
                        "speedup_vs_fp64_batched": round(ref_r["s"] / r["s"], 3),
                        "speedup_vs_fp64_one_at_a_time": round(seq * r["step"] / r["s"], 3),
                        "searches_per_step": round(r["searches"] / r["step"], 4),
                        "error_vs_fp64": rnd(e), "fp64_perturbed_1e-7_vs_fp64": rnd(ep)})
        return res

    skin = tuned["skin_dx"] * dx
    mrows, mstates, _ = form(sps.MODE_FAST, skin, tuned["graph"])
    out["mixed_batched_skin"] = report(mrows, mstates)
    del mstates
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
    crows, cstates, (ev, kept) = form(sps.MODE_FAST, skin, tuned["graph"], net)
    net.close()
    comb = report(crows, cstates)
    del cstates
    last = comb[-1]
    holds = last["speedup_vs_fp64_batched"] >= 3.0 and last["error_vs_fp64"]["stress_rel_max"] <= 1e-3
    out.update({"combined": comb, "autotuned": tuned, "training_rows": int(len(x)), "train_s": round(train_s, 2),
                "loss_first_last": [float(f"{losses[0]:.4g}"), float(f"{losses[-1]:.4g}")],
                "plastic_states_evaluated": ev, "kept": kept,
                "target_3x_at_1e-3_over_all_steps": {
                    "holds": bool(holds), "speedup_vs_fp64_batched": last["speedup_vs_fp64_batched"],
                    "speedup_vs_fp64_one_at_a_time": last["speedup_vs_fp64_one_at_a_time"],
                    "stress_rel_max": last["error_vs_fp64"]["stress_rel_max"]}})
    for name, rows in (("mixed + batched + list reuse", out["mixed_batched_skin"]), ("combined", comb)):
        for r in rows:
            log(f"[bench] {name} steps {r['steps']}: {r['ms_per_step']} ms/step, {r['speedup_vs_fp64_batched']}x "
                f"FP64 batched, {r['speedup_vs_fp64_one_at_a_time']}x one at a time, "
                f"{r['searches_per_step']} searches/step, stress error {r['error_vs_fp64']['stress_rel_max']:.3g} "
                f"(FP64 perturbed {r['fp64_perturbed_1e-7_vs_fp64']['stress_rel_max']:.3g})")
    return out


def run_biaxial(args, sps):
    """Biaxial compression at full size (decks/c5_biaxial: 0.1 m x 0.2 m at
    dx 0.05 mm, 8 M velocity + 8 M stress particles, friction angle
    30 deg +- 2 % per particle from seed 2024, rigid platens at 0.05 m/s,
    100 kPa confinement) over a stated window of its 5000 steps
    (args.c5_steps from rest): the plain FP64 form against the mixed form
    with list reuse, with and without the learned return map (trained on
    the plastic tuples of an FP64 run perturbed by 1e-7, which also gives
    the FP64 form's own spread), each timed and compared over the same
    windows."""
    import importlib
    bx = importlib.import_module("stress-particle-sph_amd.biaxial")
    sur = importlib.import_module("stress-particle-sph_amd.surrogate")
    acc = importlib.import_module("stress-particle-sph_amd.accuracy")
    log = lambda m: print(m, file=sys.stderr, flush=True)
    rnd = lambda d: {k: float(f"{v:.4g}") for k, v in d.items()}
    deck = sps.read_deck(os.path.join(ROOT, "stress-particle-sph_amd", "decks", "c5_biaxial"))
    nsteps = args.c5_steps
    wins = sorted({nsteps // 2, nsteps} - {0})
    t0 = time.perf_counter()
    parts = sps.build_particles(deck)
    setup_s = time.perf_counter() - t0

    def open_(mode, skin=0.0, perturb=0.0):
        sim = bx.open_biaxial(sps, deck, seed=2024, mode=mode, parts=parts)
        if perturb:
            st0 = sim.get_state()
            n = sim.n
            st0["x"][:n] = st0["x"][:n] * (1 + perturb * np.random.default_rng(2024).standard_normal((n, 2)))
            sim.set_state({k: st0[k] for k in ("x", "vel", "stress", "mass", "rho", "hsml", "itype",
                                               "if_out_domain", "internal_vars", "f_drucker", "x00", "displ",
                                               "wall_position", "horizontal_or_not")})
            sim.set_materials(*sim.biaxial)
        if skin:
            sim.set_list_skin(skin)
        return sim

    ref = open_(sps.MODE_ORDERED)
    nn, n2, dx = ref.nn, ref.n2, ref.params.dx
    x0 = ref.get_state(["x"])["x"]
    ref_rows, ref_states = run_windows(ref, wins)
    pairs = ref.stats()["pairs_total"]
    ref.close()
    log(f"[bench] biaxial: {n2} particles, FP64 {1e3 * ref_rows[-1]['s'] / nsteps:.3f} ms/step over {nsteps} steps")
    pert = open_(sps.MODE_ORDERED, perturb=1e-7)
    sur.log_plastic(pert, args.log_cap)
    _, pstates = run_windows(pert, wins)
    sn, st, sr, mat = sur.plastic_log(pert, args.log_cap)
    pert.close()
    e_pert = [acc.compare_states(a, b, x0) for a, b in zip(pstates, ref_states)]
    del pstates
    out = {"deck": "c5_biaxial", "particles": int(n2), "velocity_particles": int(nn), "steps": nsteps,
           "window_of": "the first steps of the 5000 of the configuration", "windows": wins,
           "setup_s": round(setup_s, 2),
           "fp64": [{"steps": f"0-{r['step']}", "ms_per_step": round(1e3 * r["s"] / r["step"], 4)} for r in ref_rows],
           "fp64_pairs_per_s": round(pairs / ref_rows[-1]["s"], 1), "plastic_tuples_logged": int(len(sn))}

    def form(mode, skin=0.0, net_=None, S=1.0, Sc=1.0):
        sim = open_(mode, skin)
        if net_ is not None:
            sur.attach(sim, net_, S, Sc, args.surrogate_tol)
        rows, states = run_windows(sim, wins)
        stats = sur.stats(sim) if net_ is not None else None
        sim.close()
        res = []
        for r, ref_r, s_, ref_s, ep in zip(rows, ref_rows, states, ref_states, e_pert):
            res.append({"steps": f"0-{r['step']}", "ms_per_step": round(1e3 * r["s"] / r["step"], 4),
                        "speedup_vs_fp64": round(ref_r["s"] / r["s"], 3),
                        "searches_per_step": round(r["searches"] / r["step"], 4),
                        "error_vs_fp64": rnd(acc.compare_states(s_, ref_s, x0)),
                        "fp64_perturbed_1e-7_vs_fp64": rnd(ep)})
        return res, stats

    skin = args.skin * dx
    out["mixed_skin"], _ = form(sps.MODE_FAST, skin)
    for r in out["mixed_skin"]:
        log(f"[bench] biaxial mixed + list reuse steps {r['steps']}: {r['ms_per_step']} ms/step "
            f"({r['speedup_vs_fp64']}x), stress error {r['error_vs_fp64']['stress_rel_max']:.3g} "
            f"(FP64 perturbed {r['fp64_perturbed_1e-7_vs_fp64']['stress_rel_max']:.3g})")
    tf, coh = mat[:, 0], mat[:, 1]
    ok = sur.correctable(st, tf, coh) if len(sn) else np.zeros(0, bool)
    if ok.sum() < 256:
        out["surrogate_note"] = (f"{int(ok.sum())} correctable plastic tuples in the window: too few to train "
                                 "the learned return map on, so the form with it is not run")
        log(f"[bench] biaxial: {out['surrogate_note']}")
        return out
    sn, st, sr, tf, coh = sn[ok], st[ok], sr[ok], tf[ok], coh[ok]
    lib = __import__("ctypes").CDLL(sps.PRODUCT_LIB)
    S, Sc = float(np.abs(st).max()), max(float(coh.max()), 1e3)
    rng = np.random.default_rng(0)
    a = sur.augment(lib, sn, st, tf, coh, rng, copies=4, scale_max=10.0)
    x = np.vstack([sur.features(sn, st, coh, tf, S, Sc), sur.features(a[0], a[1], a[3], a[4], S, Sc)])
    t = np.vstack([sur.targets(st, sr, tf, coh), sur.targets(a[1], a[2], a[4], a[3])])
    net = sur.Network(lib, seed=1, batch_rows=8192)
    losses = net.train(x, t, epochs=max(args.epochs // 2, 1), batch=4096, lr=2e-3)
    out["mixed_skin_surrogate"], (ev, kept) = form(sps.MODE_FAST, skin, net, S, Sc)
    net.close()
    out.update({"training_rows": int(len(x)), "loss_first_last": [float(f"{losses[0]:.4g}"),
                                                                  float(f"{losses[-1]:.4g}")],
                "plastic_states_evaluated": ev, "kept": kept})
    for r in out["mixed_skin_surrogate"]:
        log(f"[bench] biaxial mixed + list reuse + learned return map steps {r['steps']}: {r['ms_per_step']} "
            f"ms/step ({r['speedup_vs_fp64']}x), stress error {r['error_vs_fp64']['stress_rel_max']:.3g}")
    return out


def run_cpu(args, sps, deck, oracle_path, threads):
    """The oracle over args.cpu_steps steps from rest on `threads` OpenMP
    threads (its result does not depend on the count): pairs, seconds."""
    import ctypes
    import ctypes.util
    sim = sps.Simulation.from_deck(deck, lib_path=oracle_path)
    omp = ctypes.CDLL(ctypes.util.find_library("gomp") or "libgomp.so.1")
    omp.omp_set_num_threads(threads)
    try:
        t0 = time.perf_counter()
        sim.step(args.cpu_steps)
        t1 = time.perf_counter()
    finally:
        omp.omp_set_num_threads(1)
    pairs = sim.stats()["pairs_total"]
    sim.close()
    return pairs, t1 - t0


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB, 5 times (the HIP runtime's
    own copy, through ctypes): the HBM counters of a known traffic."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(5):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:       # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes_each_way": n, "launches": 5}), flush=True)


def host_cpu():
    """The host's CPU model, logical CPU count and the CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0))}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--deck", default="c2_slope_dx01")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps timed for the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0: every CPU this process may use)")
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
    ap.add_argument("--c5-steps", type=int, default=200,
                    help="steps of the biaxial compression leg at 16 M particles (0: skip it)")
    ap.add_argument("--acc-steps", type=int, default=2000,
                    help="steps of the mixed-vs-double accuracy runs (0: skip them)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 times) and exit; its "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic of the kernels")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
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
    # dominant kernel and its roofline: the algorithmic bytes of one launch
    # over its average duration (HIP events on the library's stream, over
    # the timed steps)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dlaunch, dms) = dom
    avg_s = dms / dlaunch / 1e3
    ab, unit = kernel_algorithmic_bytes(dname, counts)
    ib = implementation_bytes(dname, st, counts, sim.params)
    roofline = None
    if ab is not None:
        achieved = ab / avg_s / 1e9
        traffic = measured_traffic(args.deck, dname)
        sb = step_bytes(counts)
        roofline = {"kernel": dname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                    "algorithmic_bytes_per_launch": int(ab), "algorithmic_unit": unit,
                    "implementation_bytes_per_launch": int(ib) if ib else None,
                    "traffic_over_algorithmic": round(traffic / ab, 2) if traffic else None,
                    "avg_launch_ms": round(dms / dlaunch, 5), "launches": dlaunch,
                    "step": {"algorithmic_bytes": int(sb), "achieved": round(sb / (elapsed / args.steps) / 1e9, 2),
                             "frac": round(sb / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 5),
                             "traffic": measured_traffic(args.deck, "step")}}
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
                      "ms_per_step_with_kernel_events": round(1e3 * sim.events_s / args.steps, 4),
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
    if args.c5_steps > 0:
        forms["biaxial_c5"] = run_biaxial(args, sps)
    cpu = None
    if not args.no_cpu:
        oracle = os.path.join(ROOT, "oracle", "libsps_oracle.so")
        hc = host_cpu()
        nth = min(hc["usable"], args.cpu_threads) if args.cpu_threads > 0 else hc["usable"]
        cpairs, ct = run_cpu(args, sps, deck, oracle, nth)
        spairs, sct = run_cpu(args, sps, deck, oracle, 1)
        print(f"[bench] CPU oracle: {cpairs / ct:.4g} pairs/s on {nth} threads, {spairs / sct:.4g} on one",
              file=sys.stderr, flush=True)
        cpu = {"value": round(cpairs / ct, 1), "unit": "pair interactions/s", "cores": nth, "kind": "port",
               "sample": f"{args.cpu_steps} steps of {args.deck} from rest ({sum(counts)} particles), "
                         f"the oracle (C restatement of the reference, ordered, bitwise equal to it) with "
                         f"OpenMP on {nth} threads, {ct:.1f} s",
               "steps_per_s": round(args.cpu_steps / ct, 4), "cpu_model": hc["model"], "nproc": hc["nproc"],
               "usable_cpus": hc["usable"],
               "one_thread": {"value": round(spairs / sct, 1), "cores": 1, "seconds": round(sct, 2),
                              "steps_per_s": round(args.cpu_steps / sct, 4),
                              "note": "the reference's own schedule (it is serial)"}}
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

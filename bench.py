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

The CPU baseline is the oracle (a C restatement of the reference, 1 thread)
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

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


def run_gpu(args, sps, deck, mode=0, skin=0.0):
    import torch
    sim = sps.Simulation.from_deck(deck, mode=mode)
    if skin:
        sim.set_list_skin(skin)
    counts = (sim.parts.nnode, sim.parts.nstress, sim.parts.ndummy)
    print(f"[bench] deck {args.deck}: {counts[0]} nodes, {counts[1]} stress points, {counts[2]} dummies",
          file=sys.stderr, flush=True)
    for k in range(args.warmup):
        sim.step(1)
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    sim.set_profiling(True)
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

    def run(mode, perturb=0.0, skin=0.0):
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
        done = 0
        while done < args.acc_steps:
            k = min(500, args.acc_steps - done)
            sim.step(k)
            done += k
            print(f"[bench] accuracy run mode {mode} perturb {perturb:g}: {done} steps", file=sys.stderr,
                  flush=True)
        out = sim.get_state()
        sim.close()
        return out

    x0 = sps.initial_state(parts)["x"]
    ref = run(sps.MODE_ORDERED)
    mixed = acc.compare_states(run(sps.MODE_FAST), ref, x0)
    skinned = acc.compare_states(run(sps.MODE_FAST, skin=args.skin * deck_dx(sps, deck)), ref, x0)
    spread = acc.compare_states(run(sps.MODE_ORDERED, 1e-7), ref, x0)
    rnd = lambda d: {k: float(f"{v:.4g}") for k, v in d.items()}
    return {"steps": args.acc_steps, "mixed_vs_fp64": rnd(mixed), "mixed_skin_vs_fp64": rnd(skinned),
            "fp64_perturbed_1e-7_vs_fp64": rnd(spread)}


def deck_dx(sps, deck):
    import importlib
    setup = importlib.import_module("stress-particle-sph_amd.setup")
    return sps.make_params(deck, setup.build_particles(deck)).dx


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
    sim.close()
    forms = {"fp64": {"ms_per_step": round(ms_per_step, 4), "pairs_per_s": round(value, 1)}}
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
            forms["mixed"]["accuracy"] = form_errors(args, sps, deck)
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

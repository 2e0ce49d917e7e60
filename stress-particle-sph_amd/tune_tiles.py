"""Build-time autotuner of the tiled pair kernels' tile width and window
capacities.

No counterpart in the reference.  The tiled kernels give each workgroup
TILE consecutive receivers and stage the union of their stencils (the
window) in LDS, at most CAP partners of a main kind and CAP_D dummies; a
tile whose window does not fit walks global memory instead, with the same
order and results.  These are compile-time constants of the LDS layouts
(csrc/sps_tiles.h), so tuning them is a build step in three parts:

  python stress-particle-sph_amd/tune_tiles.py build        # any machine with hipcc
  python stress-particle-sph_amd/tune_tiles.py measure      # on the GPU
  python stress-particle-sph_amd/tune_tiles.py write T.json

`build` compiles the library once per candidate into _build/tiles/<name>/.
`measure` runs each candidate in a process of its own on the workloads --
the plain FP64 form on the 1 M slope (c2_slope_dx01) and the batched
ensemble (64 slopes from seed 2024 at dx 0.05 m, mixed form with list
reuse) -- over `--steps` steps after a warm-up, checks that each
candidate's state after the FP64 steps is bitwise the first candidate's
(the results may not depend on the tile), and prints a JSON table.
`write` puts the candidate with the smallest geometric mean of times
relative to the first candidate into csrc/sps_tiles.h.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "tiles")

# name: (TILE, TROWS, CAP, CAP_D, STAGE_WAVES).  A window of a TILE-receiver
# tile spans about TILE / 8 + 2 rows of a strip 10 cells wide: about 490
# partners of a kind at TILE 256 on the 1 M slope, about 260 at 128.  LDS per
# workgroup on that deck (no CSPM, one smoothing length): the list build
# 37.3 KiB at t256, 33.6 at t256s, 22.9 at t128, 19.1 at t128s; the stage
# kernels at most 35 KiB at t256.  With one smoothing length the list build
# leaves h out of its window: 27.9 KiB with its plans at t256, five
# workgroups per CU; 26.9 KiB at t256m (608), six.  The stage kernels k_sp_update and
# k_deriv_nodes are held to 3 waves per SIMD by their registers unless
# compiled for 4 (STAGE_WAVES), which spills a few of them.
CANDIDATES = {
    "t256": (256, 24, 640, 192, 0),
    "t256m": (256, 24, 608, 192, 0),
    "t256w4": (256, 24, 640, 192, 4),
    "t256s": (256, 24, 576, 160, 0),
    "t256sw4": (256, 24, 576, 160, 4),
    "t128": (128, 16, 384, 128, 0),
    "t128s": (128, 16, 320, 96, 0),
}


def lib_path(name):
    return os.path.join(BUILD, name, "libsps_hip.so")


def build(names):
    for name in names:
        t, r, c, d, w = CANDIDATES[name]
        os.makedirs(os.path.dirname(lib_path(name)), exist_ok=True)
        tiles = f"-DSPS_TILE={t} -DSPS_TROWS={r} -DSPS_CAP={c} -DSPS_CAP_D={d} -DSPS_STAGE_WAVES={w}"
        print(f"[tune_tiles] building {name}: {tiles}", flush=True)
        subprocess.run(["make", "-B", "-C", CSRC, f"OUT={lib_path(name)}", f"TILES={tiles}"], check=True,
                       stdout=subprocess.DEVNULL)


def _load_pkg():
    import importlib
    sys.path.insert(0, ROOT)
    return importlib.import_module("stress-particle-sph_amd")


def measure_one(name, steps, warmup, instances):
    """Times of candidate `name` (run in its own process)."""
    import numpy as np
    sps = _load_pkg()
    lib = lib_path(name)
    deck = sps.read_deck(os.path.join(HERE, "decks", "c2_slope_dx01"))
    sim = sps.Simulation.from_deck(deck, lib_path=lib)
    sim.step(warmup)
    t0 = time.perf_counter()
    sim.step(steps)
    fp64 = (time.perf_counter() - t0) / steps
    s = sim.get_state(["x", "vel", "stress"])
    digest = hashlib.sha256(b"".join(np.ascontiguousarray(s[k]).tobytes() for k in ("x", "vel", "stress")))
    sim.close()
    out = {"name": name, "tile": CANDIDATES[name], "fp64_1m_ms_per_step": round(1e3 * fp64, 4),
           "fp64_state_sha256": digest.hexdigest()}
    if instances > 0:
        import importlib
        ens = importlib.import_module("stress-particle-sph_amd.ensemble")
        inst = ens.draw_instances(instances, 2024)
        e = ens.open_ensemble(sps, deck, inst, dx=0.05, mode=sps.MODE_FAST, lib_path=lib, skin=0.25 * 0.05)
        e.step(warmup)
        t0 = time.perf_counter()
        e.step(steps)
        out["batched_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 4)
        e.close()
    print(json.dumps(out), flush=True)


def measure(names, steps, warmup, instances, timeout):
    rows = []
    for name in names:
        if not os.path.exists(lib_path(name)):
            raise SystemExit(f"{lib_path(name)} not built: run `tune_tiles build` first")
        print(f"[tune_tiles] measuring {name} {CANDIDATES[name]}", file=sys.stderr, flush=True)
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "_one", name, str(steps), str(warmup),
                            str(instances)], capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise SystemExit(f"candidate {name} failed:\n{p.stderr[-3000:]}")
        rows.append(json.loads(p.stdout.strip().splitlines()[-1]))
        print(f"[tune_tiles] {rows[-1]}", file=sys.stderr, flush=True)
    ref = rows[0]["fp64_state_sha256"]
    for r in rows:
        r["bitwise_equal_to_first"] = r["fp64_state_sha256"] == ref
    keys = [k for k in ("fp64_1m_ms_per_step", "batched_ms_per_step") if k in rows[0]]
    for r in rows:
        r["relative_time"] = round(math.exp(sum(math.log(r[k] / rows[0][k]) for k in keys) / len(keys)), 4)
    best = min((r for r in rows if r["bitwise_equal_to_first"]), key=lambda r: r["relative_time"])
    table = {"candidates": rows, "best": best["name"], "steps": steps, "warmup": warmup, "instances": instances}
    print(json.dumps(table), flush=True)
    if not all(r["bitwise_equal_to_first"] for r in rows):
        raise SystemExit("a candidate's results differ from the first candidate's")
    return table


def write(table_path):
    table = json.load(open(table_path))
    best = next(r for r in table["candidates"] if r["name"] == table["best"])
    t, r, c, d, w = best["tile"]
    lines = ["// Tile width and window capacities of the tiled pair kernels.",
             "// tune_tiles.py rewrites this file with the fastest candidate it measured;",
             "// a compiler definition (-DSPS_TILE=... and the others) overrides it.",
             "// Measured (ms per step: FP64 on the 1 M slope; batched ensemble, mixed):"]
    for row in table["candidates"]:
        lines.append(f"//   {row['name']:6s} {tuple(row['tile'])}: {row['fp64_1m_ms_per_step']}"
                     + (f", {row['batched_ms_per_step']}" if "batched_ms_per_step" in row else "")
                     + f" (relative {row['relative_time']})")
    lines += ["#pragma once"]
    for macro, val, what in (("SPS_TILE", t, "receivers per workgroup"), ("SPS_TROWS", r, "most window rows staged"),
                             ("SPS_CAP", c, "staged partners of a main kind per window"),
                             ("SPS_CAP_D", d, "staged dummies per window"),
                             ("SPS_STAGE_WAVES", w, "waves per SIMD of the FP64 stage kernels (0: the compiler's choice)")):
        lines += [f"#ifndef {macro}", f"#define {macro} {val}  // {what}", "#endif"]
    open(os.path.join(CSRC, "sps_tiles.h"), "w").write("\n".join(lines) + "\n")
    print(f"[tune_tiles] wrote {best['name']} {tuple(best['tile'])} into csrc/sps_tiles.h")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("what", choices=["build", "measure", "write", "_one"])
    ap.add_argument("rest", nargs="*")
    ap.add_argument("--only", nargs="*", default=None, help="candidate names (default: all, the first is the baseline)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--timeout", type=int, default=300, help="seconds per candidate")
    a = ap.parse_args()
    names = a.only or list(CANDIDATES)
    if a.what == "build":
        build(names)
    elif a.what == "measure":
        measure(names, a.steps, a.warmup, a.instances, a.timeout)
    elif a.what == "write":
        write(a.rest[0])
    else:
        measure_one(a.rest[0], int(a.rest[1]), int(a.rest[2]), int(a.rest[3]))


if __name__ == "__main__":
    main()

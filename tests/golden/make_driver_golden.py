"""Digests of every file the reference's own driver writes (code/1_SPH_2018.f90
with its CSV / GiD writers, OutputRes 3_SPH_material_2018.f90:2750-3010, and
out_print_sph / the neighbour statistics in output.txt), run as written on
decks with a short plot and print period: oracle/_ref/driver/sph_ref, built
by `make -C oracle driver` from the reference's sources.

  driver_<deck>.json   {"plot_step": p, "print_step": q, "steps": n,
                        "files": {name: sha256, ...}}

tests/test_driver_dropin.py runs the same driver with the drop-in patch
(time_integration -> the glue over include/sps.h) and compares every file.
Only this container can run it (the GPU box has no /root/reference).
Usage: python tests/golden/make_driver_golden.py [deck ...]
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "driver", "sph_ref")
DECKS = os.path.join(ROOT, "stress-particle-sph_amd", "decks")

# deck: (steps, plot_step, print_step)
PLAN = {
    "t1_tiny": (200, 10, 20),           # the position update, no BCs: surface_points every plot
    "t4_xsph": (200, 25, 50),           # XSPH
    "v1_vslope_sp1": (100, 10, 25),     # prescribed velocities and stresses, no position update
    "c1_collapse": (1000, 100, 250),
}


def prepare(deck, work, steps, plot, prnt):
    for f in os.listdir(os.path.join(DECKS, deck)):
        shutil.copy(os.path.join(DECKS, deck, f), work)
    path = os.path.join(work, "input.txt")
    lines = open(path).read().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("time step"))
    dt = lines[k + 1].split(",")[0]
    lines[k + 1] = f"{dt},100.0,{steps}"
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("print, save, plot"))
    lines[k + 1] = f"{prnt},1000000,{plot}"
    open(path, "w").write("\n".join(lines) + "\n")
    return set(os.listdir(work))


def outputs(work, inputs):
    out = {}
    for f in sorted(os.listdir(work)):
        if f in inputs or f == "run.log":
            continue
        out[f] = hashlib.sha256(open(os.path.join(work, f), "rb").read()).hexdigest()
    return out


def run(deck):
    steps, plot, prnt = PLAN[deck]
    work = tempfile.mkdtemp(prefix=f"drv_{deck}_")
    try:
        inputs = prepare(deck, work, steps, plot, prnt)
        subprocess.run(["bash", "-c", f"ulimit -s unlimited && {DRIVER} > run.log 2>&1"], cwd=work, check=True)
        rec = {"plot_step": plot, "print_step": prnt, "steps": steps, "files": outputs(work, inputs)}
        json.dump(rec, open(os.path.join(HERE, f"driver_{deck}.json"), "w"), indent=1)
        print(deck, len(rec["files"]), "files", flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle driver")
    for d in (sys.argv[1:] or list(PLAN)):
        run(d)

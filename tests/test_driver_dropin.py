"""The reference's own driver with the drop-in patch, against the same driver
as written: every output file byte for byte.

oracle/_ref/driver/sph_dropin_{oracle,hip} are code/1_SPH_2018.f90 (the
reference's driver program) with `call time_integration` replaced by the
glue's `call sps_time_integration` (plus its `use` line and the closing
`call sps_glue_finish`), linked with the reference's modules, the glue and
the CPU oracle or the HIP library; `make -C oracle driver` builds them and
deletes the patched copy of the source.  The reference's writers run
unchanged: nodes.csv.*, stress_points.csv.*, surface_points.csv.*, the GiD
.post.res / .post.msh, geom.csv, the .chk echo and output.txt (out_print_sph
and the neighbour statistics, which the glue writes with the reference's
statements).  tests/golden/driver_<deck>.json holds the digests of the files
the unpatched driver wrote (tests/golden/make_driver_golden.py).

SPS_GLUE_SYNC=output keeps the state on the device except after the steps
the driver writes; the files must be the same.
"""
import hashlib
import json
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT, deck_dir

DRV = os.path.join(ROOT, "oracle", "_ref", "driver")


def run_driver(binary, deck_name, tmp_path, sync=None):
    exe = os.path.join(DRV, binary)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference sources at build time)")
    rec = json.load(open(os.path.join(GOLDEN, f"driver_{deck_name}.json")))
    work = tmp_path / f"{deck_name}_{sync or 'step'}"
    shutil.copytree(deck_dir(deck_name), work)
    inputs = set(os.listdir(work))
    path = work / "input.txt"
    lines = path.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("time step"))
    lines[k + 1] = f"{lines[k + 1].split(',')[0]},100.0,{rec['steps']}"
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("print, save, plot"))
    lines[k + 1] = f"{rec['print_step']},1000000,{rec['plot_step']}"
    path.write_text("\n".join(lines) + "\n")
    env = dict(os.environ)
    if sync:
        env["SPS_GLUE_SYNC"] = sync
    print(f"[driver] {binary} deck={deck_name} steps={rec['steps']} sync={sync}", flush=True)
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && '{exe}' > run.log 2>&1"], cwd=work, check=True,
                   timeout=600, env=env)
    got = {f: hashlib.sha256((work / f).read_bytes()).hexdigest() for f in sorted(os.listdir(work))
           if f not in inputs and f != "run.log"}
    return rec["files"], got


def check(want, got):
    assert sorted(got) == sorted(want), f"files differ: {sorted(set(got) ^ set(want))[:6]}"
    bad = [f for f in want if got[f] != want[f]]
    assert not bad, f"{len(bad)} of {len(want)} files differ from the reference's, e.g. {bad[:4]}"


@pytest.mark.parametrize("deck_name", ["t1_tiny", "t4_xsph", "v1_vslope_sp1"])
@pytest.mark.parametrize("sync", [None, "output"])
def test_reference_driver_dropin_oracle(deck_name, sync, tmp_path):
    check(*run_driver("sph_dropin_oracle", deck_name, tmp_path, sync))


@pytest.mark.gpu
@pytest.mark.parametrize("deck_name", ["t1_tiny", "t4_xsph", "v1_vslope_sp1", "c1_collapse"])
@pytest.mark.parametrize("sync", [None, "output"])
def test_reference_driver_dropin_hip(deck_name, sync, tmp_path):
    check(*run_driver("sph_dropin_hip", deck_name, tmp_path, sync))

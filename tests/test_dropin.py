"""The Fortran drop-in: the reference's own driver and readers with
`call time_integration` replaced by the iso_c_binding glue
(stress-particle-sph_amd/fortran/sps_glue.f90) over the C-ABI.

oracle/_ref/code/sps_dropin_{oracle,hip} are linked by `make -C oracle ref`
in the build container (they contain the reference's compiled modules and
travel with the tree as built files).  Their state dumps must equal the
reference's own, bit for bit.
"""
import hashlib
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, bits_equal, deck_dir, golden

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from refdump import read_state  # noqa: E402

BIN = os.path.join(ROOT, "oracle", "_ref", "code")
FIELDS = ("x", "vel", "stress", "ivar1", "f_drucker", "ddev_strn", "grad_u", "displ", "f_bound", "art_visc")


def run_dropin(binary, deck_name, steps, tmp_path):
    exe = os.path.join(BIN, binary)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference sources at build time)")
    work = tmp_path / deck_name
    shutil.copytree(deck_dir(deck_name), work)
    out = work / "out"
    out.mkdir()
    print(f"[dropin] {binary} deck={deck_name} steps={steps}", flush=True)
    cmd = f"ulimit -s unlimited && '{exe}' out {max(steps)} " + " ".join(map(str, steps))
    subprocess.run(["bash", "-c", cmd], cwd=work, check=True, timeout=240,
                   stdout=subprocess.DEVNULL)
    return {st: read_state(str(out / f"state_{st:06d}.bin")) for st in steps}


def check(states, deck_name):
    g = golden(deck_name)
    for st, s in states.items():
        for k in FIELDS:
            if f"s{st}_{k}" in g.files:
                assert bits_equal(g[f"s{st}_{k}"], s[k]), f"step {st} {k}"
            else:                   # fixtures that keep only a digest of this field
                digest = hashlib.sha256(np.ascontiguousarray(s[k]).tobytes()).hexdigest()
                assert digest == str(g[f"s{st}_sha_{k}"]), f"step {st} {k}"


@pytest.mark.parametrize("deck_name,steps", [("t1_tiny", [1, 50, 200]), ("t2_cspm", [1, 50, 200]),
                                             ("v1_vslope_sp1", [1, 10, 100])])
def test_fortran_dropin_oracle(deck_name, steps, tmp_path):
    check(run_dropin("sps_dropin_oracle", deck_name, steps, tmp_path), deck_name)


@pytest.mark.gpu
@pytest.mark.parametrize("deck_name,steps", [("t1_tiny", [1, 50, 200]), ("t2_cspm", [1, 50, 200]),
                                             ("t3_elastic", [1, 50, 200]), ("v1_vslope_sp1", [1, 10, 100])])
def test_fortran_dropin_hip(deck_name, steps, tmp_path):
    check(run_dropin("sps_dropin_hip", deck_name, steps, tmp_path), deck_name)

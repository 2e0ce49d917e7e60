"""The product's HIP kernels, compiled for the CPU with AddressSanitizer
(tests/emu: a host stand-in for the HIP runtime and a fiber scheduler for
barriers and wave exchanges), against the oracle.

This checks on the CPU what the GPU parity tests check on MI355X -- the
kernel sources reproduce the reference bit for bit -- and, through the
sanitizer, that no kernel reads or writes outside its buffers, its LDS or
its stack on these decks.
"""
import os
import subprocess
import sys

import pytest

from conftest import ORACLE_LIB, ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))
import build_emu  # noqa: E402


def _asan_runtime():
    try:
        p = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def emu_lib():
    if _asan_runtime() is None:
        pytest.skip("g++ with AddressSanitizer runtime not available")
    if not os.path.exists(ORACLE_LIB):
        pytest.skip("oracle library not built")
    return build_emu.build()


# several cases per process: each process pays the start-up of the
# AddressSanitizer build once
EXACT_GROUPS = {
    "tiny": [("t1_tiny", 20), ("t2_cspm", 20), ("t3_elastic", 20), ("t2_cspm:hvar", 20), ("t2_cspm:graph", 20),
             ("t4_xsph", 20), ("p1_vm", 20), ("p4_mc", 20), ("t2_cspm:nan", 2)],
    "bui": [("b3_bui_sp3", 8), ("b4_bui_outside", 8), ("b6_bui_velvec_shift:kick", 8)],
    "examples": [("c1_collapse", 2), ("v3_vslope_sp3", 10), ("l1_localisation", 6), ("l1_localisation:graph", 6)],
    "surface": [("l2_localisation_free", 4), ("l3_localisation_xsph", 4)],
}


def run_cases(emu_lib, cases):
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0:halt_on_error=1")
    argv = [a for c in cases for a in map(str, c)]
    return subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emu", "run_case.py"), emu_lib,
                           ORACLE_LIB] + argv, env=env, capture_output=True, text=True, timeout=1500)


@pytest.mark.parametrize("group", sorted(EXACT_GROUPS))
def test_cpu_build_of_kernels_matches_oracle(emu_lib, group):
    p = run_cases(emu_lib, EXACT_GROUPS[group])
    print(p.stdout[-3000:], p.stderr[-4000:])
    assert p.returncode == 0, p.stderr[-4000:] or p.stdout


# the mixed-precision kernels: in bounds (AddressSanitizer) and within single-
# precision rounding of the oracle where the problem does not amplify it (an
# elastic deck; a yielding deck before yield starts)
MIXED_CASES = [("t3_elastic:mixed", 20, "1e-5"), ("t2_cspm:mixed", 1, "1e-5"), ("c1_collapse:mixed", 1, "5e-5"),
               ("v3_vslope_sp3:mixed", 10, "1e-5"), ("l1_localisation:mixed", 6, "1e-5"),
               ("t4_xsph:mixed", 1, "1e-5"), ("b1_bui_sp1:skin", 8, "1e-5"), ("b8_bui_out3:skin", 8, "1e-5")]


def test_cpu_build_of_mixed_kernels_near_oracle(emu_lib):
    p = run_cases(emu_lib, MIXED_CASES)
    print(p.stdout[-3000:], p.stderr[-4000:])
    assert p.returncode == 0, p.stderr[-4000:] or p.stdout


# the network kernels (matrix-core GEMM emulated from its lane layout, loss,
# backward, Adam) against the numpy restatement
def test_cpu_build_of_network_kernels_matches_numpy_oracle(emu_lib):
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0:halt_on_error=1")
    p = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emu", "run_mlp.py"), emu_lib], env=env,
                       capture_output=True, text=True, timeout=900)
    print(p.stdout[-2000:], p.stderr[-4000:])
    assert p.returncode == 0, p.stderr[-4000:] or p.stdout

"""The product's HIP kernels, compiled for the CPU with AddressSanitizer
(tests/emu: a host stand-in for the HIP runtime and a fiber scheduler for
barriers and wave exchanges), against the oracle.

This checks on the CPU what the GPU parity tests check on MI355X -- the
kernel sources reproduce the reference bit for bit -- and, through the
sanitizer, that no kernel reads or writes outside its buffers, its LDS or
its stack on these decks.

Every group of cases runs in a process of its own (each pays the start-up
of the sanitizer build once), and the processes are started together when
the module's first test asks for them, so the groups run side by side.
"""
import os
import subprocess
import sys
import threading

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


# exact cases: bitwise against the oracle
EXACT_GROUPS = {
    "tiny": [("t1_tiny", 20), ("t2_cspm", 20), ("t3_elastic", 20)],
    "tiny2": [("t2_cspm:hvar", 20), ("t2_cspm:graph", 20), ("t4_xsph", 20)],
    "models": [("p1_vm", 10), ("p4_mc", 10), ("t2_cspm:nan", 2), ("t2_cspm:lf", 10), ("v1_vslope_sp1:lf", 10)],
    "bui": [("b3_bui_sp3", 8), ("b4_bui_outside", 8), ("b6_bui_velvec_shift:kick", 8)],
    "collapse": [("c1_collapse", 2), ("v3_vslope_sp3", 10)],
    "localisation": [("l1_localisation", 3), ("l1_localisation:graph", 3)],
    "surface": [("l2_localisation_free", 4), ("l3_localisation_xsph", 4)],
    "bcs": [("l4_localisation_outside", 3), ("c3_collapse_free", 3)],
}

# the mixed-precision kernels: in bounds (AddressSanitizer) and within single-
# precision rounding of the oracle where the problem does not amplify it (an
# elastic deck; a yielding deck before yield starts)
MIXED_CASES = [("t3_elastic:mixed", 20, "1e-5"), ("t2_cspm:mixed", 1, "1e-5"), ("c1_collapse:mixed", 1, "5e-5"),
               ("v3_vslope_sp3:mixed", 10, "1e-5"), ("l1_localisation:mixed", 6, "1e-5"),
               ("t4_xsph:mixed", 1, "1e-5"), ("b1_bui_sp1:skin", 8, "1e-5"), ("b8_bui_out3:skin", 8, "1e-5")]

# the tile width and window capacities are build parameters (csrc/sps_tiles.h,
# tune_tiles.py): the kernels built with another candidate -- narrower tiles,
# smaller windows, so more tiles take the global walk -- give the same bits
TILE_VARIANTS = {"t128s": ("SPS_TILE=128", "SPS_TROWS=16", "SPS_CAP=320", "SPS_CAP_D=96")}
TILE_CASES = [("c1_collapse", 2), ("b4_bui_outside", 4), ("l2_localisation_free", 3), ("t2_cspm:graph", 10),
              ("t4_xsph", 10), ("c1_collapse:mixed", 1, "5e-5")]


class Runs:
    """The case processes, started together; result(name) waits for one."""

    def __init__(self):
        self.env = dict(os.environ, LD_PRELOAD=_asan_runtime(),
                        ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0:halt_on_error=1")
        self.procs, self.out = {}, {}

    def start(self, name, argv):
        self.procs[name] = subprocess.Popen([sys.executable] + argv, env=self.env, stdout=subprocess.PIPE,
                                            stderr=subprocess.PIPE, text=True)

    def cases(self, name, lib, cases):
        self.start(name, [os.path.join(ROOT, "tests", "emu", "run_case.py"), lib, ORACLE_LIB] +
                   [a for c in cases for a in map(str, c)])

    def result(self, name, timeout=2400):
        if name not in self.out:
            p = self.procs[name]
            try:
                so, se = p.communicate(timeout=timeout)
            except subprocess.TimeoutExpired:
                p.kill()
                so, se = p.communicate()
            self.out[name] = (p.returncode, so, se)
        rc, so, se = self.out[name]
        print(so[-3000:], se[-4000:])
        return rc, so, se

    def stop(self):
        for p in self.procs.values():
            if p.poll() is None:
                p.kill()
                p.wait()


@pytest.fixture(scope="module")
def runs():
    if _asan_runtime() is None:
        pytest.skip("g++ with AddressSanitizer runtime not available")
    if not os.path.exists(ORACLE_LIB):
        pytest.skip("oracle library not built")
    # the default build and the tile variants' builds side by side
    libs, errs = {}, []

    def build(key, defines):
        try:
            libs[key] = build_emu.build(defines=defines)
        except Exception as e:          # reported by the tests that need the build
            errs.append(f"{key}: {e}")

    threads = [threading.Thread(target=build, args=(None, ()))]
    threads += [threading.Thread(target=build, args=(v, d)) for v, d in TILE_VARIANTS.items()]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if None not in libs:
        pytest.fail("emulated build failed: " + "; ".join(errs))
    r = Runs()
    for g, cases in EXACT_GROUPS.items():
        r.cases(g, libs[None], cases)
    r.cases("mixed", libs[None], MIXED_CASES)
    r.start("mlp", [os.path.join(ROOT, "tests", "emu", "run_mlp.py"), libs[None]])
    for v in TILE_VARIANTS:
        if v in libs:
            r.cases("tile_" + v, libs[v], TILE_CASES)
    yield r
    r.stop()


@pytest.mark.parametrize("group", sorted(EXACT_GROUPS))
def test_cpu_build_of_kernels_matches_oracle(runs, group):
    rc, so, se = runs.result(group)
    assert rc == 0, se[-4000:] or so


@pytest.mark.parametrize("variant", sorted(TILE_VARIANTS))
def test_cpu_build_of_other_tile_width_matches_oracle(runs, variant):
    assert "tile_" + variant in runs.procs, f"the {variant} build failed"
    rc, so, se = runs.result("tile_" + variant)
    assert rc == 0, se[-4000:] or so


def test_cpu_build_of_mixed_kernels_near_oracle(runs):
    rc, so, se = runs.result("mixed")
    assert rc == 0, se[-4000:] or so


# the network kernels (matrix-core GEMM emulated from its lane layout, loss,
# backward, Adam) against the numpy restatement
def test_cpu_build_of_network_kernels_matches_numpy_oracle(runs):
    rc, so, se = runs.result("mlp")
    assert rc == 0, se[-4000:] or so

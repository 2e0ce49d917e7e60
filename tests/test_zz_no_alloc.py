"""sps_return_map_noalloc, the entry point that allocates nothing: its
results equal sps_return_map's bit for bit (the exact return map,
adapt_stress2 of code/3_SPH_material_2018.f90:2066-2140, on the GPU) on two
calls in a row, the second smaller than the first, so that the shared
device staging array keeps the first call's rows past the second's end;
sizes past the stated maximum are refused."""
import ctypes

import numpy as np
import pytest

from conftest import ORACLE_LIB, PRODUCT_LIB

MAX_ROWS = 262144   # SPS_RETURN_MAP_NOALLOC_MAX


def stresses(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.normal(0.0, 5e4, (n, 4))
    s[:, 3] = 0.3 * (s[:, 0] + s[:, 1])
    s[::7, :2] += 8e4                         # some states beyond the tension cutoff
    tanfi = np.tan(np.radians(rng.uniform(15.0, 40.0, n)))
    coh = rng.uniform(0.0, 2e4, n)
    return np.ascontiguousarray(s), tanfi, coh


def call(lib, name, s, tanfi, coh):
    out = np.full_like(s, np.nan)
    d = ctypes.POINTER(ctypes.c_double)
    fn = getattr(lib, name)
    fn.argtypes = [ctypes.c_int64, d, d, d, d]
    fn.restype = ctypes.c_int
    rc = fn(len(s), s.ctypes.data_as(d), tanfi.ctypes.data_as(d), coh.ctypes.data_as(d), out.ctypes.data_as(d))
    return rc, out


def test_noalloc_refuses_beyond_its_maximum():
    lib = ctypes.CDLL(ORACLE_LIB)
    s, tf, c = stresses(MAX_ROWS + 1, 3)
    rc, _ = call(lib, "sps_return_map_noalloc", s, tf, c)
    assert rc == -1


@pytest.mark.gpu
def test_noalloc_return_map_equals_allocating_one():
    lib = ctypes.CDLL(PRODUCT_LIB)
    for n, seed in ((200_000, 1), (4_099, 2)):      # the second call smaller than the first
        s, tf, c = stresses(n, seed)
        rc_a, a = call(lib, "sps_return_map", s, tf, c)
        rc_b, b = call(lib, "sps_return_map_noalloc", s, tf, c)
        assert rc_a == 0 and rc_b == 0, lib.sps_last_error
        assert a.tobytes() == b.tobytes(), f"n={n}"
        assert not np.array_equal(a, s)          # the map changed some states
    s, tf, c = stresses(MAX_ROWS + 1, 3)
    rc, _ = call(lib, "sps_return_map_noalloc", s, tf, c)
    assert rc == -1
    rc, _ = call(lib, "sps_return_map_noalloc", s[:0], tf[:0], c[:0])
    assert rc == 0

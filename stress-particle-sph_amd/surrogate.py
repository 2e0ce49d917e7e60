"""Host side of the learned Drucker-Prager stress update.

The network (include/sps.h, sps_mlp_*) is created, trained and run on the GPU
by the library; this module builds its training set from the plain method's
logged tuples, scales features and targets, and attaches a trained network
to a simulation of the mixed-precision form.

Features of a plastic trial state (stress components xx, yy, xy, zz):
    x[0:4] = sigma_n / S,  x[4:8] = (sigma_trial - sigma_n) / S,
    x[8] = c / Sc,  x[9] = tan(phi)
Target: (R(sigma_trial) - sigma_trial) / F(sigma_trial), R the exact return
map (adapt_stress2, code/3_SPH_material_2018.f90:2066-2140) and F = sqrt(J2)
+ 3 alpha p - k_c its yield function, positive at every plastic trial state:
the correction per unit overstress.
"""
import ctypes
import math

import numpy as np

_P = ctypes.c_void_p


def _declare(lib):
    if getattr(lib, "_sps_mlp_declared", False):
        return lib
    lib.sps_mlp_param_count.restype = ctypes.c_int64
    lib.sps_mlp_create.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(_P)]
    lib.sps_mlp_destroy.argtypes = [_P]
    lib.sps_mlp_destroy.restype = None
    lib.sps_mlp_get_params.argtypes = [_P, _P]
    lib.sps_mlp_set_params.argtypes = [_P, _P]
    lib.sps_mlp_forward.argtypes = [_P, ctypes.c_int64, _P, _P]
    lib.sps_mlp_gradient.argtypes = [_P, ctypes.c_int64, _P, _P, _P, _P]
    lib.sps_mlp_train.argtypes = [_P, ctypes.c_int64, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                  ctypes.c_uint64, _P]
    lib.sps_return_map.argtypes = [ctypes.c_int64, _P, _P, _P, _P]
    lib.sps_set_surrogate.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.sps_get_surrogate_stats.argtypes = [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    lib.sps_set_plastic_log.argtypes = [_P, ctypes.c_int64]
    lib.sps_get_plastic_log.argtypes = [_P, ctypes.c_int64, _P, ctypes.POINTER(ctypes.c_int64)]
    for f in ("sps_mlp_create", "sps_mlp_get_params", "sps_mlp_set_params", "sps_mlp_forward", "sps_mlp_gradient",
              "sps_mlp_train", "sps_return_map", "sps_set_surrogate", "sps_get_surrogate_stats",
              "sps_set_plastic_log", "sps_get_plastic_log"):
        getattr(lib, f).restype = ctypes.c_int
    lib._sps_mlp_declared = True
    return lib


def _ptr(a):
    return a.ctypes.data


class Network:
    """The GPU network of a library implementing include/sps.h."""

    def __init__(self, lib, seed=1, batch_rows=8192):
        from . import SpsError
        self._err = SpsError
        self.lib = _declare(lib)
        self.handle = _P()
        self._check(self.lib.sps_mlp_create(int(batch_rows), int(seed), ctypes.byref(self.handle)))
        self.nparam = int(self.lib.sps_mlp_param_count())

    def _check(self, rc):
        if rc != 0:
            raise self._err(rc, self.lib.sps_last_error().decode())

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        self._check(self.lib.sps_mlp_get_params(self.handle, _ptr(out)))
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        assert theta.shape == (self.nparam,)
        self._check(self.lib.sps_mlp_set_params(self.handle, _ptr(theta)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 4), dtype=np.float32)
        self._check(self.lib.sps_mlp_forward(self.handle, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.lib.sps_mlp_gradient(self.handle, x.shape[0], _ptr(x), _ptr(t), _ptr(g), _ptr(loss)))
        return float(loss[0]), g

    def train(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=7):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        self._check(self.lib.sps_mlp_train(self.handle, x.shape[0], _ptr(x), _ptr(t), int(epochs), int(batch),
                                           float(lr), int(seed), _ptr(losses)))
        return losses

    def close(self):
        if self.handle:
            self.lib.sps_mlp_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def return_map(lib, stress, tanfi, coh):
    """The exact return map of stresses (n x 4) on the GPU; tanfi and coh are
    scalars or one value per row."""
    lib = _declare(lib)
    s = np.ascontiguousarray(stress, dtype=np.float64)
    n = s.shape[0]
    tf = np.ascontiguousarray(np.broadcast_to(np.asarray(tanfi, dtype=np.float64), (n,)))
    c = np.ascontiguousarray(np.broadcast_to(np.asarray(coh, dtype=np.float64), (n,)))
    out = np.zeros_like(s)
    rc = lib.sps_return_map(n, _ptr(s), _ptr(tf), _ptr(c), _ptr(out))
    if rc != 0:
        from . import SpsError
        raise SpsError(rc, lib.sps_last_error().decode())
    return out


def features(s_n, s_tr, coh, tanfi, S, Sc):
    """Feature rows (n x 10) of trial states."""
    n = s_n.shape[0]
    x = np.empty((n, 10), dtype=np.float64)
    x[:, 0:4] = s_n / S
    x[:, 4:8] = (s_tr - s_n) / S
    x[:, 8] = np.broadcast_to(coh, (n,)) / Sc
    x[:, 9] = np.broadcast_to(tanfi, (n,))
    return x.astype(np.float32)


def yield_function(s, tanfi, coh):
    """F = sqrt(J2) + 3 alpha p - k_c of adapt_stress2 (alpha = tan(phi) /
    sqrt(9 + 12 tan^2(phi)), k_c = 3 c / sqrt(9 + 12 tan^2(phi)))."""
    tanfi = np.asarray(tanfi, dtype=np.float64)
    coh = np.asarray(coh, dtype=np.float64)
    root = np.sqrt(9 + 12 * tanfi * tanfi)
    alpha2, kc = tanfi / root, 3 * coh / root
    p = (s[:, 0] + s[:, 1] + s[:, 3]) / 3.0
    d0, d1, d3 = s[:, 0] - p, s[:, 1] - p, s[:, 3] - p
    sj2 = np.sqrt(s[:, 2] ** 2 + 0.5 * (d0 * d0 + d1 * d1 + d3 * d3))
    return sj2 + 3 * alpha2 * p - kc


OVERSTRESS_MIN = 1e-9   # as the library's: smaller overstresses are left to the exact return map


def state_size(s, tanfi, coh):
    """sqrt(J2) + |3 alpha p| + k_c of stresses s."""
    tanfi = np.asarray(tanfi, dtype=np.float64)
    coh = np.asarray(coh, dtype=np.float64)
    root = np.sqrt(9 + 12 * tanfi * tanfi)
    p = (s[:, 0] + s[:, 1] + s[:, 3]) / 3.0
    d0, d1, d3 = s[:, 0] - p, s[:, 1] - p, s[:, 3] - p
    sj2 = np.sqrt(s[:, 2] ** 2 + 0.5 * (d0 * d0 + d1 * d1 + d3 * d3))
    return sj2 + np.abs(3 * (tanfi / root) * p) + 3 * coh / root


def correctable(s_tr, tanfi, coh):
    """Trial states the network corrects: overstress above rounding level and
    sqrt(J2) above the 1e-5 below which the return map zeroes the deviator
    whatever the overstress (3_SPH_material_2018.f90:2122)."""
    p = (s_tr[:, 0] + s_tr[:, 1] + s_tr[:, 3]) / 3.0
    d0, d1, d3 = s_tr[:, 0] - p, s_tr[:, 1] - p, s_tr[:, 3] - p
    sj2 = np.sqrt(s_tr[:, 2] ** 2 + 0.5 * (d0 * d0 + d1 * d1 + d3 * d3))
    return (yield_function(s_tr, tanfi, coh) > OVERSTRESS_MIN * state_size(s_tr, tanfi, coh)) & \
        (sj2 > float(np.float32(10e-06)))


def targets(s_tr, s_ret, tanfi, coh):
    """(R - sigma_trial) / F(sigma_trial) of correctable plastic trial states."""
    f = yield_function(s_tr, tanfi, coh)
    return ((s_ret - s_tr) / f[:, None]).astype(np.float32)


def corrected(s_tr, y, tanfi, coh):
    """The stress the network's output y gives for trial states s_tr."""
    return s_tr + yield_function(s_tr, tanfi, coh)[:, None] * y


def log_plastic(sim, cap):
    """Keep the plain method's plastic tuples of the following steps."""
    lib = _declare(sim.lib)
    sim._check(lib.sps_set_plastic_log(sim.handle, int(cap)))


PLOG_ROW = 14


def plastic_log(sim, cap):
    """The logged tuples: sigma_n, sigma_trial, R(sigma_trial) (each n x 4)
    and the material (tan phi, c) of each (n x 2)."""
    lib = _declare(sim.lib)
    out = np.zeros((cap, PLOG_ROW), dtype=np.float64)
    n = ctypes.c_int64(0)
    sim._check(lib.sps_get_plastic_log(sim.handle, int(cap), _ptr(out), ctypes.byref(n)))
    out = out[:n.value]
    return out[:, 0:4], out[:, 4:8], out[:, 8:12], out[:, 12:14]


def augment(lib, s_n, s_tr, tanfi, coh, rng, copies=4, scale_max=100.0, material=None):
    """More plastic trial states around the logged ones: the elastic
    increment sigma_trial - sigma_n scaled by log-uniform factors in
    [1, scale_max] (larger loading steps) and, when `material` = ((c_lo, c_hi),
    (phi_lo_deg, phi_hi_deg)) is given, drawn materials (otherwise each
    row's own: `tanfi` and `coh` are one value or one per row); labelled
    with the exact return map and kept where plastic.  Returns (s_n, s_tr,
    R, c, tanfi)."""
    rows = []
    n = s_n.shape[0]
    for _ in range(copies):
        k = np.exp(rng.uniform(0.0, math.log(scale_max), size=(n, 1)))
        tr = s_n + k * (s_tr - s_n)
        if material is not None:
            (c0, c1), (p0, p1) = material
            c = rng.uniform(c0, c1, size=n)
            tf = np.tan(np.radians(rng.uniform(p0, p1, size=n)))
        else:
            c = np.broadcast_to(np.asarray(coh, dtype=np.float64), (n,)).copy()
            tf = np.broadcast_to(np.asarray(tanfi, dtype=np.float64), (n,)).copy()
        ret = return_map(lib, tr, tf, c)
        plastic = np.any(ret != tr, axis=1) & correctable(tr, tf, c)
        rows.append((s_n[plastic], tr[plastic], ret[plastic], c[plastic], tf[plastic]))
    return tuple(np.concatenate([r[i] for r in rows]) for i in range(5))


def attach(sim, net, stress_scale, coh_scale, tol=1e-4):
    """Use the network for the plastic trial states of `sim` (mixed form)."""
    lib = _declare(sim.lib)
    sim._check(lib.sps_set_surrogate(sim.handle, net.handle if net is not None else None, float(stress_scale),
                                     float(coh_scale), float(tol)))


def stats(sim):
    lib = _declare(sim.lib)
    ev, kept = ctypes.c_int64(0), ctypes.c_int64(0)
    sim._check(lib.sps_get_surrogate_stats(sim.handle, ctypes.byref(ev), ctypes.byref(kept)))
    return int(ev.value), int(kept.value)

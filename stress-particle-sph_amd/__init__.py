"""Stress-Particle SPH hot path on MI355X: Python host side.

This package mirrors, on the host, the interface of the reference's hot path
(`call time_integration`, code/2_SPH_main_2018.f90:78-184, driven once per
step by code/1_SPH_2018.f90:72-102) over the C-ABI declared in
include/sps.h.  The compute runs in libsps_hip.so (hand-written HIP kernels
for gfx950, sources under csrc/); this module only reads decks, lays out the
initial state the way the reference does and moves arrays across the ABI.

    import importlib; sps = importlib.import_module("stress-particle-sph_amd")
    deck = sps.read_deck("decks/c1_collapse")
    sim = sps.Simulation.from_deck(deck)          # product library
    sim.step(10)
    st = sim.get_state()
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .deck import Deck, DeckError, read_deck  # noqa: F401
from .setup import Particles, bcs_on_particles, build_particles, elastic_matrix, reference_pi  # noqa: F401

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
PRODUCT_LIB = os.path.join(PKG_DIR, "libsps_hip.so")
DECK_DIR = os.path.join(PKG_DIR, "decks")

SPS_ABI_VERSION = 4
INTEGRATOR_RK4, INTEGRATOR_LEAPFROG = 0, 1
SPS_MAX_TCURVE = 64
SPS_MAX_BCS = 32
SPS_MAX_CURVES = 16
MODE_ORDERED = 0
MODE_FAST = 1

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ECAPACITY", -5: "ESTATE", -6: "ENUMERIC"}


class SpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sps error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("nnode", ctypes.c_int32),
                ("nstress", ctypes.c_int32), ("ndummy", ctypes.c_int32),
                ("ndummy2", ctypes.c_int32), ("nint_vars", ctypes.c_int32),
                ("ndimn", ctypes.c_int32), ("nstre", ctypes.c_int32), ("skf", ctypes.c_int32),
                ("sp_sph", ctypes.c_int32), ("inside_approach", ctypes.c_int32),
                ("cspm", ctypes.c_int32), ("update_x", ctypes.c_int32), ("xsph", ctypes.c_int32),
                ("cont_density", ctypes.c_int32), ("art_stress", ctypes.c_int32),
                ("dummy_nodes", ctypes.c_int32), ("sph_shift", ctypes.c_int32),
                ("no_bcs", ctypes.c_int32), ("bforce_gate", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("ic_grav", ctypes.c_int32),
                ("tcurve_npts", ctypes.c_int32),
                ("props", ctypes.c_double * 20), ("pi", ctypes.c_double), ("dx", ctypes.c_double),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("damping", ctypes.c_double),
                ("cspm_det_tol", ctypes.c_double),
                ("D11", ctypes.c_double), ("D22", ctypes.c_double), ("D12", ctypes.c_double),
                ("D33", ctypes.c_double), ("D41", ctypes.c_double), ("D42", ctypes.c_double),
                ("xmin_domain", ctypes.c_double * 2), ("xmax_domain", ctypes.c_double * 2),
                ("cgrav", ctypes.c_double * 2), ("ft_grav", ctypes.c_double),
                ("tcurve_t", ctypes.c_double * SPS_MAX_TCURVE),
                ("tcurve_f", ctypes.c_float * SPS_MAX_TCURVE),
                ("npoints", ctypes.c_int32), ("vel_vector", ctypes.c_int32),
                ("shift_update", ctypes.c_int32), ("reserved0", ctypes.c_int32),
                ("r_x", ctypes.c_double), ("r_y", ctypes.c_double), ("disp_tol", ctypes.c_double)]


_P = ctypes.c_void_p


class State(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("x", "vel", "stress", "mass", "rho", "hsml", "itype",
                                   "if_out_domain", "internal_vars", "f_drucker", "x00",
                                   "displ", "wall_position", "horizontal_or_not", "grad_u",
                                   "ddev_strn", "f_bound", "art_visc", "x_10", "disp_10")]


class Bcs(ctypes.Structure):
    _fields_ = [("no_bcs", ctypes.c_int32), ("ifsigman", ctypes.c_int32), ("ncurves", ctypes.c_int32),
                ("reserved0", ctypes.c_int32)] + [(n, _P) for n in ("bc_list", "bc_info", "bc_or_not", "curve_npts",
                                                                     "curve_t", "curve_f")]


class Stats(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_int64), ("m_pairs", ctypes.c_int64),
                ("n_by_type", ctypes.c_int64 * 10), ("ndivx", ctypes.c_int32),
                ("ndivy", ctypes.c_int32), ("steps", ctypes.c_int64), ("pairs_total", ctypes.c_int64)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double)]


def load_library(path):
    """Load a library implementing include/sps.h and declare its signatures."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    h = ctypes.c_void_p
    lib.sps_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(h)]
    lib.sps_set_state.argtypes = [h, ctypes.POINTER(State)]
    lib.sps_step.argtypes = [h, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_int32)]
    lib.sps_get_state.argtypes = [h, ctypes.POINTER(State)]
    lib.sps_get_pairs.argtypes = [h, ctypes.c_int64] + [_P] * 6 + [ctypes.POINTER(ctypes.c_int64)]
    lib.sps_get_stats.argtypes = [h, ctypes.POINTER(Stats)]
    lib.sps_get_surface.argtypes = [h, _P, _P]
    lib.sps_set_profiling.argtypes = [h, ctypes.c_int32]
    lib.sps_get_profile.argtypes = [h, ctypes.c_int32, ctypes.POINTER(KernelTime),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    lib.sps_set_list_skin.argtypes = [h, ctypes.c_double]
    lib.sps_set_materials.argtypes = [h, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    lib.sps_set_graph.argtypes = [h, ctypes.c_int32]
    lib.sps_get_search_count.argtypes = [h, ctypes.POINTER(ctypes.c_int64)]
    lib.sps_set_bcs.argtypes = [h, ctypes.POINTER(Bcs)]
    lib.sps_destroy.argtypes = [h]
    lib.sps_destroy.restype = None
    lib.sps_last_error.restype = ctypes.c_char_p
    lib.sps_backend.restype = ctypes.c_char_p
    for f in ("sps_create", "sps_set_state", "sps_step", "sps_get_state", "sps_get_pairs",
              "sps_get_stats", "sps_set_profiling", "sps_get_profile", "sps_set_list_skin", "sps_set_materials", "sps_set_graph",
              "sps_get_search_count", "sps_set_bcs"):
        getattr(lib, f).restype = ctypes.c_int
    return lib


def make_params(deck: Deck, parts: Particles, mode=MODE_ORDERED, bforce_gate=None,
                cspm_det_tol=None, nint_vars=10):
    """Pack the sps_params POD from a deck and its particle setup."""
    p = Params()
    p.abi_version = SPS_ABI_VERSION
    p.nnode, p.nstress, p.ndummy, p.ndummy2 = parts.nnode, parts.nstress, parts.ndummy, parts.ndummy2
    p.nint_vars = nint_vars
    p.ndimn, p.nstre, p.skf = deck.ndimn, deck.nstre, deck.skf
    p.sp_sph, p.inside_approach = int(deck.sp_sph), int(deck.inside_approach)
    p.cspm, p.update_x, p.xsph = int(deck.cspm), int(deck.update_x), int(deck.xsph)
    p.cont_density, p.art_stress = int(deck.cont_density), int(deck.art_stress)
    p.dummy_nodes, p.no_bcs = int(deck.dummy_nodes), deck.no_bcs
    p.npoints = deck.npoints
    if not deck.inside_approach:                # 3_...:371-378, 609
        p.sph_shift, p.vel_vector, p.shift_update = int(deck.sph_shift), int(deck.vel_vector), deck.shift_update
        p.r_x, p.r_y, p.disp_tol = deck.dx * deck.rx_factor, deck.dx * deck.ry_factor, deck.disp_tol
    bui = deck.layout == "bui"
    p.bforce_gate = int(bui) if bforce_gate is None else int(bforce_gate)
    p.mode = mode
    p.ic_grav = deck.ic_grav
    for k in range(20):
        p.props[k] = float(deck.props[k])
    p.pi = parts.pi
    p.dx = deck.dx
    p.alpha, p.beta, p.damping = deck.alpha, deck.beta, deck.damping
    if cspm_det_tol is None:
        cspm_det_tol = float(np.float32(1e-7 if bui else 1e-3))
    p.cspm_det_tol = cspm_det_tol
    for k, v in parts.D.items():
        setattr(p, k, v)
    for a in range(2):
        p.xmin_domain[a], p.xmax_domain[a] = deck.xmin_domain[a], deck.xmax_domain[a]
        p.cgrav[a] = deck.cgrav[a]
    p.ft_grav = deck.ft_grav
    if deck.ic_grav:
        tt = deck.ttcurves[deck.tcurve_grav - 1]
        ff = deck.ftcurves[deck.tcurve_grav - 1]
        if not 2 <= len(tt) <= SPS_MAX_TCURVE:
            raise DeckError("gravity time curve needs 2..64 points")
        p.tcurve_npts = len(tt)
        for k in range(len(tt)):
            p.tcurve_t[k], p.tcurve_f[k] = tt[k], ff[k]
    return p


def initial_state(parts: Particles, nint_vars=10):
    """Arrays of the reference state after Setup_Global_Arrays and
    Initial_conditions with zero stress and velocity (3_...:1005-1022)."""
    n2, n = parts.ntotal2, parts.ntotal
    return dict(
        x=parts.x.copy(), vel=np.zeros((n2, 2)), stress=np.zeros((n2, 4)),
        mass=parts.mass.copy(), rho=parts.rho.copy(), hsml=parts.hsml.copy(),
        itype=parts.itype.copy(), if_out_domain=np.zeros(n2, dtype=np.int32),
        internal_vars=np.zeros((n, nint_vars)), f_drucker=np.zeros(n),
        x00=parts.x.copy(), displ=np.zeros((parts.nnode, 2)),
        wall_position=parts.wall_position.copy(),
        horizontal_or_not=parts.horizontal_or_not.copy(),
        x_10=parts.x[:parts.nnode].copy(), disp_10=np.zeros(parts.nnode))


_STATE_SHAPES = {  # name: (dtype, per-particle shape, count key)
    "x": (np.float64, (2,), "n2"), "vel": (np.float64, (2,), "n2"),
    "stress": (np.float64, (4,), "n2"), "mass": (np.float64, (), "n2"),
    "rho": (np.float64, (), "n2"), "hsml": (np.float64, (), "n2"),
    "itype": (np.int32, (), "n2"), "if_out_domain": (np.int32, (), "n2"),
    "internal_vars": (np.float64, ("niv",), "n"), "f_drucker": (np.float64, (), "n"),
    "x00": (np.float64, (2,), "n2"), "displ": (np.float64, (2,), "nn"),
    "wall_position": (np.float32, (), "n2"), "horizontal_or_not": (np.float32, (), "n2"),
    "grad_u": (np.float64, (4,), "n"), "ddev_strn": (np.float64, (), "n"),
    "f_bound": (np.float64, (2,), "nn"), "art_visc": (np.float64, (2,), "nn"),
    "x_10": (np.float64, (2,), "nn"), "disp_10": (np.float64, (), "nn"),
}


class Simulation:
    """One handle of a library implementing include/sps.h."""

    def __init__(self, params: Params, lib_path=PRODUCT_LIB):
        self.lib = load_library(lib_path)
        self.params = params
        self.handle = ctypes.c_void_p()
        self._check(self.lib.sps_create(ctypes.byref(params), ctypes.byref(self.handle)))
        self.nn = params.nnode
        self.n = params.nnode + params.nstress
        self.n2 = self.n + params.ndummy
        self.niv = params.nint_vars
        self.time_sph = ctypes.c_double(0.0)
        self.itimestep_sph = ctypes.c_int32(0)

    @classmethod
    def from_deck(cls, deck: Deck, lib_path=PRODUCT_LIB, mode=MODE_ORDERED, parts=None, **kw):
        """A simulation of the deck from rest (its particles built here, or
        `parts`, the same deck's, built before)."""
        if parts is None:
            parts = build_particles(deck)
        sim = cls(make_params(deck, parts, mode=mode, **kw), lib_path)
        sim.set_state(initial_state(parts, sim.niv))
        sim.deck, sim.parts = deck, parts
        if deck.no_bcs > 0:
            sim.set_bcs(deck, parts)
        return sim

    def set_bcs(self, deck: Deck, parts: Particles):
        """The deck's prescribed BCs (sps_set_bcs): bc_list, the nodes' rows of
        Get_BCs_on_node at the initial positions, BC_or_not of every particle,
        every time curve, ifsigman."""
        info, bc_or_not = bcs_on_particles(deck, parts.x[:parts.ntotal])
        nc = len(deck.ttcurves)
        if nc > SPS_MAX_CURVES:
            raise DeckError(f"more than {SPS_MAX_CURVES} time curves")
        ct = np.zeros((max(nc, 1), SPS_MAX_TCURVE))
        cf = np.zeros((max(nc, 1), SPS_MAX_TCURVE), dtype=np.float32)
        npts = np.zeros(max(nc, 1), dtype=np.int32)
        for c in range(nc):
            npts[c] = len(deck.ttcurves[c])
            ct[c, :npts[c]], cf[c, :npts[c]] = deck.ttcurves[c], deck.ftcurves[c]
        keep = [np.ascontiguousarray(deck.bc_list, dtype=np.float64), np.ascontiguousarray(info[:self.nn]),
                np.ascontiguousarray(bc_or_not[:self.n]), npts, ct, cf]
        b = Bcs(no_bcs=deck.no_bcs, ifsigman=deck.ifsigman, ncurves=nc, reserved0=0)
        for name, a in zip(("bc_list", "bc_info", "bc_or_not", "curve_npts", "curve_t", "curve_f"), keep):
            setattr(b, name, a.ctypes.data)
        self._check(self.lib.sps_set_bcs(self.handle, ctypes.byref(b)))

    def _check(self, rc):
        if rc != 0:
            raise SpsError(rc, self.lib.sps_last_error().decode())

    @property
    def backend(self):
        return self.lib.sps_backend().decode()

    def _shape(self, name):
        dt, tail, key = _STATE_SHAPES[name]
        cnt = {"n2": self.n2, "n": self.n, "nn": self.nn}[key]
        tail = tuple(self.niv if t == "niv" else t for t in tail)
        return dt, (cnt,) + tail

    def set_state(self, arrays: dict):
        st = State()
        keep = []
        for name, arr in arrays.items():
            dt, shape = self._shape(name)
            a = np.ascontiguousarray(arr, dtype=dt)
            if a.shape != shape:
                raise ValueError(f"{name}: shape {a.shape}, expected {shape}")
            keep.append(a)
            setattr(st, name, a.ctypes.data)
        self._check(self.lib.sps_set_state(self.handle, ctypes.byref(st)))

    def get_state(self, names=None):
        names = names or list(_STATE_SHAPES)
        out, st = {}, State()
        for name in names:
            dt, shape = self._shape(name)
            out[name] = np.zeros(shape, dtype=dt)
            setattr(st, name, out[name].ctypes.data)
        self._check(self.lib.sps_get_state(self.handle, ctypes.byref(st)))
        return out

    def step(self, nsteps=1, dt=None):
        if dt is None:
            dt = self.deck.dt
        self._check(self.lib.sps_step(self.handle, nsteps, dt, ctypes.byref(self.time_sph),
                                      ctypes.byref(self.itimestep_sph)))

    def get_surface(self):
        """BC_or_not (ntotal; 2 on the free surface) and the unit normals
        (ntotal, 2) of the last step's free-surface detection."""
        bc = np.zeros(self.n, dtype=np.int32)
        nrm = np.zeros((self.n, 2))
        self._check(self.lib.sps_get_surface(self.handle, bc.ctypes.data, nrm.ctypes.data))
        return bc, nrm

    def get_pairs(self):
        n = ctypes.c_int64(0)
        self.lib.sps_get_pairs(self.handle, 0, None, None, None, None, None, None, ctypes.byref(n))
        cnt = n.value
        out = dict(i=np.zeros(cnt, np.int32), j=np.zeros(cnt, np.int32), type=np.zeros(cnt, np.int32),
                   w=np.zeros(cnt, np.float32), dwdx=np.zeros(cnt, np.float32),
                   dwdy=np.zeros(cnt, np.float32))
        self._check(self.lib.sps_get_pairs(self.handle, cnt, *(out[k].ctypes.data for k in
                                                                ("i", "j", "type", "w", "dwdx", "dwdy")),
                                           ctypes.byref(n)))
        return out

    def stats(self):
        s = Stats()
        self._check(self.lib.sps_get_stats(self.handle, ctypes.byref(s)))
        return dict(n_pairs=s.n_pairs, m_pairs=s.m_pairs, n_by_type=list(s.n_by_type),
                    ndivx=s.ndivx, ndivy=s.ndivy, steps=s.steps, pairs_total=s.pairs_total)

    def set_materials(self, table, mat):
        """A material per particle: table rows (tan phi, c, E, nu, D11, D22,
        D12, D33, D41, D42), mat the row of each particle (after set_state)."""
        table = np.ascontiguousarray(table, dtype=np.float64)
        mat = np.ascontiguousarray(mat, dtype=np.int32)
        assert table.ndim == 2 and table.shape[1] == 10 and mat.shape == (self.n2,)
        self._check(self.lib.sps_set_materials(self.handle, table.shape[0], table.ctypes.data, mat.ctypes.data))

    def set_integrator(self, which):
        """RK4 (the reference's, default) or leapfrog: INTEGRATOR_RK4 / INTEGRATOR_LEAPFROG."""
        self._check(self.lib.sps_set_integrator(self.handle, int(which)))

    def set_graph(self, on=True):
        """Replay the mixed form's list-reusing steps as a captured graph."""
        self._check(self.lib.sps_set_graph(self.handle, int(bool(on))))

    def set_list_skin(self, skin):
        """Keep the pair lists over steps, rebuilt when a particle has moved
        skin / 2 (SPS_MODE_FAST only; 0 searches every step)."""
        self._check(self.lib.sps_set_list_skin(self.handle, float(skin)))

    def search_count(self):
        n = ctypes.c_int64(0)
        self._check(self.lib.sps_get_search_count(self.handle, ctypes.byref(n)))
        return int(n.value)

    def set_profiling(self, on=True):
        self._check(self.lib.sps_set_profiling(self.handle, int(bool(on))))

    def profile(self, reset=True):
        """Per-kernel launch counts and total milliseconds (HIP events on the
        library's stream) since the last reset."""
        buf = (KernelTime * 64)()
        n = ctypes.c_int32(0)
        self._check(self.lib.sps_get_profile(self.handle, 64, buf, ctypes.byref(n), int(bool(reset))))
        return {buf[k].name.decode(): (int(buf[k].launches), float(buf[k].total_ms)) for k in range(n.value)}

    def close(self):
        if self.handle:
            self.lib.sps_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

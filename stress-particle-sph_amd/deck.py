"""Readers for the reference's input decks: input.txt, <name>.dat, <name>.pts.

The reference reads all three with Fortran list-directed input in a fixed
sequence (code/3_SPH_material_2018.f90:40-321 for the .dat file and the head
of input.txt, :346-405 and :758-770 for the .pts file, code/1_SPH_2018.f90
:52-59 for the time-step block of input.txt).  `ListReader` reproduces the
list-directed rules those reads rely on: every READ starts a new record,
values may run over several records, the rest of the last record is dropped,
a character item takes the first blank- or comma-delimited token, logicals
accept T/F/.true./.false.

Two .dat layouts exist: code/ reads `nstre nmats` on the material-type line
(code/3_...:101), the Bui et al. (2008) fork reads `ntype_solid nstre nmats`
(soil_failure_bui_et_al_2008/3_...:99).
"""
from __future__ import annotations

import dataclasses
import os
import re
from fractions import Fraction

import numpy as np

_TOKEN = re.compile(r"[^\s,]+")


class DeckError(ValueError):
    pass


class ListReader:
    """Fortran list-directed READ over one text file."""

    def __init__(self, path):
        self.path = path
        with open(path, "r") as f:
            self.lines = f.read().splitlines()
        self.pos = 0

    def _record(self):
        if self.pos >= len(self.lines):
            raise DeckError(f"{self.path}: unexpected end of file")
        line = self.lines[self.pos]
        self.pos += 1
        return _TOKEN.findall(line)

    def values(self, n):
        """READ(u,*) of n numeric/logical items: returns n raw tokens."""
        out = []
        while len(out) < n:
            out.extend(self._record())
        return out[:n]

    def peek(self):
        """The tokens of the next non-blank record, without consuming it."""
        k = self.pos
        while k < len(self.lines):
            toks = _TOKEN.findall(self.lines[k])
            if toks:
                return toks
            k += 1
        return []

    def text(self):
        """READ(u,*) text: a character item (first token; a blank record
        supplies nothing and the read moves on, as for numbers)."""
        while True:
            toks = self._record()
            if toks:
                return toks[0]

    def ints(self, n):
        return [int(t) if re.fullmatch(r"[+-]?\d+", t) else int(float(t)) for t in self.values(n)]

    def int1(self):
        return self.ints(1)[0]

    def reals(self, n):
        return [_real(t) for t in self.values(n)]

    def logicals(self, n):
        return [_logical(t, self.path) for t in self.values(n)]


def _real(tok):
    """A real(DP) item: decimal to binary64, correctly rounded."""
    return float(tok.replace("d", "e").replace("D", "e"))


def _real32(tok):
    """A default-real item: decimal to binary32, correctly rounded directly
    (going through binary64 first could round twice)."""
    exact = Fraction(tok.replace("d", "e").replace("D", "e"))
    c = np.float32(float(exact))
    best = c
    for cand in (np.nextafter(c, np.float32(-np.inf)), np.nextafter(c, np.float32(np.inf))):
        e_c, e_b = abs(Fraction(float(cand)) - exact), abs(Fraction(float(best)) - exact)
        if e_c < e_b or (e_c == e_b and (int(cand.view(np.uint32)) & 1) == 0):
            best = cand
    return np.float32(best)


def _logical(tok, path):
    t = tok.lstrip(".").upper()
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise DeckError(f"{path}: bad logical {tok!r}")


@dataclasses.dataclass
class Wall:
    wall_id: int
    position: np.float32
    x1: np.float32
    x2: np.float32


@dataclasses.dataclass
class Deck:
    """Everything the reference reads from the three input files."""
    directory: str
    problem_name: str
    layout: str
    # input.txt
    sp_sph: bool = True
    art_stress: bool = False
    particle_shift: bool = False
    inside_approach: bool = True
    npoints: int = 1
    # the outside approach (3_...:371-373)
    sph_shift: bool = False
    vel_vector: bool = False
    shift_update: int = 1
    rx_factor: float = 0.0
    ry_factor: float = 0.0
    disp_tol: float = 0.0
    sml: float = 1.2
    dt: float = 0.0
    time_end: float = 0.0
    maxtimestep: int = 0
    print_step: int = 1
    save_step: int = 1
    plot_step: int = 1
    # .dat
    ndimn: int = 2
    ntype_solid: int = 2
    nstre: int = 4
    nmats: int = 1
    props: np.ndarray = None          # (20,) props(1,:)
    icunk_s: int = 0
    no_bcs: int = 0
    ifsigman: int = 0
    no_segments_bc: int = 0
    no_nodes_bc: int = 0
    bc_list: np.ndarray = None        # (no_bcs, 8): number, variable, curve, a1, a0, w, phi*twopi, Tf
    segment_bcs: np.ndarray = None    # (no_segments_bc, 5): x1 y1 x2 y2 bc
    nodal_bcs: np.ndarray = None      # (no_nodes_bc, 2): node, bc
    ic_tcurve: int = 0
    ttcurves: list = None             # per curve: float64 array
    ftcurves: list = None             # per curve: float32 array
    xmin_domain: np.ndarray = None
    xmax_domain: np.ndarray = None
    icunkno: int = 2
    pa_sph: int = 2
    nnps: int = 2
    sle: int = 0
    skf: int = 1
    cspm: bool = False
    update_x: bool = True
    xsph: bool = False
    summation_density: bool = False
    cont_density: bool = False
    damping: float = 0.0
    alpha: float = 0.0
    beta: float = 0.0
    ic_grav: int = 0
    cgrav: np.ndarray = None
    tcurve_grav: int = 1
    ft_grav: float = 1.0
    # .pts
    geom_type: int = 1
    xs: tuple = None                  # x1 x2 x3 x4
    ys: tuple = None                  # y1 y2 y3 y4
    dx: float = 0.0
    dy: float = 0.0
    dummy_nodes: bool = False
    walls: list = None


def read_deck(directory, layout="auto"):
    """Read a deck directory the way the reference does.

    layout: "code" (code/ and the vertical-slope / strain-localisation forks),
    "bui" (the Bui et al. 2008 fork with its extra ntype_solid field), or
    "auto": "bui" when the material-type record holds three values, else
    "code".  The layout also selects the fork's hot-path variant
    (make_params: wall-force gate and CSPM determinant threshold).
    """
    if layout not in ("code", "bui", "auto"):
        raise DeckError("layout must be 'code', 'bui' or 'auto'")
    inp = ListReader(os.path.join(directory, "input.txt"))
    inp.text()                                   # 3_...:62
    # character(15) problem_name; character(10) in the Bui fork (its 5_...:115),
    # whose decks' names "auto" therefore also needs within 10 characters
    name = inp.text()[:10 if layout == "bui" else 15]
    d = Deck(directory=directory, problem_name=name, layout=layout)
    dat = ListReader(os.path.join(directory, name + ".dat"))
    pts = ListReader(os.path.join(directory, name + ".pts"))

    # .dat head (3_...:74-139)
    nline = dat.int1()
    for _ in range(nline):
        dat.text()
    dat.text()
    d.ndimn = dat.int1()
    dat.text()
    if layout == "auto":
        layout = d.layout = "bui" if len(dat.peek()) >= 3 else "code"
    if layout == "code":
        d.nstre, d.nmats = dat.ints(2)
    else:
        d.ntype_solid, d.nstre, d.nmats = dat.ints(3)
    props = np.zeros((max(d.nmats, 1) + 1, 20), dtype=np.float64)
    dat.text()
    for _ in range(d.nmats):
        vals = dat.values(13)
        lmat = int(float(vals[0]))
        props[lmat, 0:12] = [_real(t) for t in vals[1:13]]
        ncrit = int(props[lmat, 1])
        if ncrit in (5, 12):
            dat.text()
            props[lmat, 12:20] = dat.reals(8)
    d.props = props[1].copy()

    # setup_particles (3_...:332-335) then Read_2DMesh input.txt part (:364-375)
    dat.text()
    d.icunk_s = dat.int1()
    inp.text()
    d.sp_sph, d.art_stress, d.particle_shift = inp.logicals(3)
    if d.sp_sph:
        inp.text()
        d.inside_approach = inp.logicals(1)[0]
        inp.text()
        d.npoints = inp.int1()
        if not d.inside_approach:                # 3_...:371-374
            inp.text()
            v = inp.values(6)
            d.sph_shift, d.vel_vector = _logical(v[0], inp.path), _logical(v[1], inp.path)
            d.shift_update = int(v[2])
            d.rx_factor, d.ry_factor, d.disp_tol = _real(v[3]), _real(v[4]), _real(v[5])
    if d.vel_vector:                             # 3_...:377-378
        d.npoints = 2
    if not d.sp_sph:
        d.npoints = 1
    # .pts geometry (3_...:383-405)
    pts.text()
    d.geom_type = pts.int1()
    pts.text()
    v = pts.reals(5)
    d.xs, d.dx = tuple(v[:4]), v[4]
    pts.text()
    v = pts.reals(5)
    d.ys, d.dy = tuple(v[:4]), v[4]
    pts.text()
    d.dummy_nodes = pts.logicals(1)[0]
    inp.text()
    d.sml = inp.reals(1)[0]                      # 3_...:447-448
    d.walls = []
    if d.dummy_nodes:                            # set_up_dummy_nodes 3_...:768-783
        pts.text()
        n_walls = pts.int1()
        for _ in range(n_walls):
            pts.text()
            pts.text()
            v = pts.values(4)
            d.walls.append(Wall(int(float(v[0])), _real32(v[1]), _real32(v[2]), _real32(v[3])))

    # .dat remainder (3_...:160-317)
    dat.text()
    d.no_bcs, d.ifsigman = dat.ints(2)
    d.bc_list = np.zeros((max(d.no_bcs, 0), 8))
    if d.no_bcs > 0:                             # 3_...:168-181
        # twopi = acos(0.0)*4.0/360. in default real: degrees to radians
        twopi = np.float32(np.float32(np.float32(np.pi / 2) * np.float32(4.0)) / np.float32(360.0))
        dat.text()
        for i in range(d.no_bcs):
            d.bc_list[i] = dat.reals(8)
            d.bc_list[i, 6] = d.bc_list[i, 6] * float(twopi)
    dat.text()
    d.no_segments_bc = dat.int1()
    d.segment_bcs = np.zeros((max(d.no_segments_bc, 0), 5))
    if d.no_segments_bc > 0:                     # 3_...:185-196
        dat.text()
        for i in range(d.no_segments_bc):
            d.segment_bcs[i] = dat.reals(5)
    dat.text()
    d.no_nodes_bc = dat.int1()
    d.nodal_bcs = np.zeros((max(d.no_nodes_bc, 0), 2))
    if d.no_nodes_bc > 0:                        # 3_...:199-210
        dat.text()
        for i in range(d.no_nodes_bc):
            d.nodal_bcs[i] = dat.reals(2)
    dat.text()
    d.ic_tcurve = dat.int1()
    d.ttcurves, d.ftcurves = [], []
    if d.ic_tcurve == 1:
        dat.text()
        ntc, _mpts = dat.ints(2)
        for _ in range(ntc):
            dat.text()
            npt = dat.int1()
            dat.text()
            d.ttcurves.append(np.array(dat.reals(npt), dtype=np.float64))
            d.ftcurves.append(np.array([_real32(t) for t in dat.values(npt)], dtype=np.float32))
    dat.text()
    v = dat.reals(2 * d.ndimn)
    d.xmin_domain = np.array(v[:d.ndimn], dtype=np.float64)
    d.xmax_domain = np.array(v[d.ndimn:], dtype=np.float64)
    dat.text()                                   # Initial_conditions 3_...:1424-1436
    d.icunkno = dat.int1()
    if d.icunkno == 2:                           # Get_Init2D 3_...:1458-1484
        for _ in range(d.nstre + d.ndimn):
            dat.text()
            if dat.int1() != 0:
                raise DeckError("only zero initial conditions (ICtype 0) are supported")
    else:
        raise DeckError("only icunkno = 2 initial conditions are supported")
    dat.text()
    v = dat.values(7)                            # one READ, 3_...:286
    d.pa_sph, d.nnps, d.sle, d.skf = (int(float(t)) for t in v[:4])
    d.cspm, d.update_x, d.xsph = (_logical(t, dat.path) for t in v[4:])
    dat.text()
    d.summation_density, d.cont_density = dat.logicals(2)
    dat.text()
    d.damping = dat.reals(1)[0]
    dat.text()
    d.alpha, d.beta = dat.reals(2)
    dat.text()
    d.ic_grav = dat.int1()
    d.cgrav = np.zeros(2)
    if d.ic_grav == 1:
        dat.text()
        v = dat.values(d.ndimn + 2)
        d.cgrav = np.array([_real(t) for t in v[:d.ndimn]], dtype=np.float64)
        d.tcurve_grav = int(float(v[d.ndimn]))
        d.ft_grav = _real(v[d.ndimn + 1])

    # input.txt time block (code/1_SPH_2018.f90:52-59)
    inp.text()
    v = inp.values(3)
    d.dt, d.time_end, d.maxtimestep = _real(v[0]), _real(v[1]), int(float(v[2]))
    inp.text()
    d.print_step, d.save_step, d.plot_step = (int(float(t)) for t in inp.values(3))
    return d

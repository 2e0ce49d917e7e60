"""Ensembles of slope instances run as one particle set (the batched form).

No counterpart in the reference, which runs one problem with one material
and rectangular blocks.  An instance is a slope of height H whose face
descends at angle beta from a crest of width W = H / 2 to the ground, on a
lattice of spacing dx (nodes on the lattice, stress points at the cell
centres, both inside the slope), with a left wall and a bottom wall that
extends two heights beyond the toe (three dummy rows each, as the
reference's set_up_dummy_nodes lays them, 3_SPH_material_2018.f90:736-878).
The instances sit side by side along x, far enough apart that no pair
crosses between them, and are concatenated into one particle set with the
reference's index ranges (all nodes, all stress points, all innermost dummy
rows, all outer rows), each particle carrying its instance's material
(include/sps.h, sps_set_materials).

Instances are drawn from a seed with SplitMix64 (Steele, Lea & Flood 2014),
in a fixed order per instance: H ~ U[5, 20] m, beta ~ U[30, 90] deg,
phi ~ U[15, 40] deg, c ~ U[0, 20] kPa, E ~ U[5, 50] MPa (nu = 0.3,
rho = 1850 kg/m^3).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .deck import Wall
from .setup import Particles, elastic_matrix, reference_pi, wall_dummies

F32 = np.float32
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))


@dataclasses.dataclass
class Instance:
    height: float
    angle_deg: float
    phi_deg: float
    coh: float
    young: float
    poiss: float = 0.3
    rho: float = 1850.0


def draw_instances(n, seed):
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        out.append(Instance(height=rng.uniform(5.0, 20.0), angle_deg=rng.uniform(30.0, 90.0),
                            phi_deg=rng.uniform(15.0, 40.0), coh=rng.uniform(0.0, 20e3),
                            young=rng.uniform(5e6, 50e6)))
    return out


def _slope_lattice(inst, dx):
    """Nodes and stress points of one slope (origin at its crest's left end on
    the ground)."""
    H, W = inst.height, 0.5 * inst.height
    cot = 0.0 if inst.angle_deg >= 90.0 else 1.0 / math.tan(math.radians(inst.angle_deg))
    nx = int((W + H * cot) / dx + 1)
    ny = int(H / dx + 1)

    def inside(x, y):   # 0 <= y <= H, 0 <= x <= W + (H - y) cot(beta)
        return (y <= H + 1e-9 * dx) & (x <= W + (H - y) * cot + 1e-9 * dx)

    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    xn, yn = (i * dx).ravel(), (j * dx).ravel()
    node_in = inside(xn, yn)
    ci, cj = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    xs, ys = ((ci + 0.5) * dx).ravel(), ((cj + 0.5) * dx).ravel()
    cell_in = inside(xs, ys)
    # a node's mass: 2 / 8 of rho dx^2 per adjacent cell inside the slope (the
    # reference's no_int rule for a rectangle: 8 inside, 4 on an edge, 2 at a corner)
    cin = np.zeros((nx + 1, ny + 1), dtype=np.int32)
    cin[1:nx, 1:ny] = cell_in.reshape(nx - 1, ny - 1)
    no_int = 2 * (cin[:-1, :-1] + cin[1:, :-1] + cin[:-1, 1:] + cin[1:, 1:]).ravel()
    keep = node_in & (no_int > 0)
    return (xn[keep], yn[keep], no_int[keep]), (xs[cell_in], ys[cell_in]), W + H * cot


def build_ensemble(instances, dx=0.05, sml=1.2, x_start=0.0):
    """One particle set of the instances, side by side in x from x_start;
    returns (Particles, material table (n_instances x 10), particle ->
    instance, instance x offsets, domain box).  An instance built alone at
    the offset it has in a batch has the same coordinates, bit for bit."""
    nodes, sps, walls_all, offs = [], [], [], []
    x0 = x_start
    for k, inst in enumerate(instances):
        (xn, yn, no_int), (xs, ys), xtoe = _slope_lattice(inst, dx)
        H = inst.height
        offs.append(x0)
        nodes.append((xn + x0, yn, no_int, k))
        sps.append((xs + x0, ys, k))
        wl = Wall(1, F32(x0 - dx), F32(-dx), F32(H + 0.05 * H))
        wb = Wall(2, F32(-dx), F32(x0 - 6 * dx), F32(x0 + xtoe + 2 * H))
        walls_all.append(([wl, wb], k))
        x0 = x0 + xtoe + 2 * H + 20 * dx        # the next instance's left wall: cells apart
    # dummies of every instance: innermost rows first
    inner, outer = [], []
    for walls, k in walls_all:
        wx, wy, hor, wpos, n2 = wall_dummies(walls, dx, dx)
        inner.append((wx[:n2], wy[:n2], hor[:n2], wpos[:n2], k))
        outer.append((wx[n2:], wy[n2:], hor[n2:], wpos[n2:], k))
    nnode = sum(len(a[0]) for a in nodes)
    nstress = sum(len(a[0]) for a in sps)
    ndummy2 = sum(len(a[0]) for a in inner)
    ndummy = ndummy2 + sum(len(a[0]) for a in outer)
    n2 = nnode + nstress + ndummy
    x = np.empty((n2, 2))
    mass = np.empty(n2)
    rho = np.empty(n2)
    inst_of = np.empty(n2, dtype=np.int32)
    wall_position = np.zeros(n2, dtype=np.float32)
    horizontal_or_not = np.zeros(n2, dtype=np.float32)
    o = 0
    for xn, yn, no_int, k in nodes:
        r = instances[k].rho
        sl = slice(o, o + len(xn))
        x[sl, 0], x[sl, 1] = xn, yn
        mass[sl] = no_int * ((r * dx * dx) / 8)
        rho[sl], inst_of[sl] = r, k
        o += len(xn)
    for xs, ys, k in sps:
        r = instances[k].rho
        sl = slice(o, o + len(xs))
        x[sl, 0], x[sl, 1] = xs, ys
        mass[sl] = 8 * ((r * dx * dx) / 8)
        rho[sl], inst_of[sl] = r, k
        o += len(xs)
    for wx, wy, hor, wpos, k in inner + outer:
        r = instances[k].rho
        sl = slice(o, o + len(wx))
        x[sl, 0], x[sl, 1] = wx.astype(np.float64), wy.astype(np.float64)
        mass[sl] = float(F32(dx) * F32(dx)) * r
        rho[sl], inst_of[sl] = r, k
        horizontal_or_not[sl], wall_position[sl] = hor, wpos
        o += len(wx)
    itype = np.empty(n2, dtype=np.int32)
    itype[:nnode], itype[nnode:nnode + nstress], itype[nnode + nstress:] = 2, 1, 25
    D0 = elastic_matrix(instances[0].young, instances[0].poiss)
    parts = Particles(nnode=nnode, nstress=nstress, ndummy=ndummy, ndummy2=ndummy2, ndivx=0, ndivy=0, x=x,
                      mass=mass, rho=rho, hsml=np.full(n2, sml * dx), itype=itype, wall_position=wall_position,
                      horizontal_or_not=horizontal_or_not, D=D0, pi=reference_pi())
    table = np.empty((len(instances), 10))
    for k, inst in enumerate(instances):
        D = elastic_matrix(inst.young, inst.poiss)
        table[k] = [math.tan(math.radians(inst.phi_deg)), inst.coh, inst.young, inst.poiss,
                    D["D11"], D["D22"], D["D12"], D["D33"], D["D41"], D["D42"]]
    margin = 0.5 * max(i.height for i in instances)
    box = ((float(x[:, 0].min()) - margin, float(x[:, 1].min()) - margin),
           (float(x[:, 0].max()) + margin, float(x[:, 1].max()) + 4 * margin))
    return parts, table, inst_of, np.array(offs), box


def open_ensemble(sps, base_deck, instances, dx=0.05, mode=1, lib_path=None, skin=0.0, x_start=0.0):
    """A Simulation of the ensemble with the base deck's controls (time step,
    gravity curve, damping, viscosity, CSPM), the instances' geometry and
    materials; the first instance at x offset x_start."""
    parts, table, inst_of, offs, box = build_ensemble(instances, dx, base_deck.sml, x_start)
    p = sps.make_params(base_deck, parts, mode=mode)
    p.dx = dx
    for a in range(2):
        p.xmin_domain[a], p.xmax_domain[a] = box[0][a], box[1][a]
    # props(1,:) of the first instance (the table overrides them particle by particle)
    t0 = table[0]
    p.props[2], p.props[3], p.props[12], p.props[13] = t0[2], t0[3], t0[0], t0[1]
    sim = sps.Simulation(p, lib_path or sps.PRODUCT_LIB)
    sim.set_state(sps.initial_state(parts, sim.niv))
    sim.set_materials(table, inst_of)
    if skin:
        sim.set_list_skin(skin)
    sim.deck, sim.parts = base_deck, parts
    sim.ensemble = (table, inst_of, offs)
    return sim

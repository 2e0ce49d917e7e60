"""Initial particle state from a deck, as the reference builds it.

Mirrors Read_2DMesh (code/3_SPH_material_2018.f90:346-732),
set_up_dummy_nodes (:736-878) and Setup_Global_Arrays (:883-1026) for the
inside approach with one to four stress points per lattice cell and the
outside approach with one to three per node, including their
single-precision arithmetic (the wall geometry is default real) and the
lattice-count truncations.  Vectorised with numpy; every expression keeps the
reference's operation order so the result is bit-identical (checked against
dumps of the reference in tests/test_oracle_golden.py,
test_setup_matches_reference and test_setup_matches_reference_other_forks).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .deck import Deck, DeckError

F32 = np.float32


@dataclasses.dataclass
class Particles:
    nnode: int
    nstress: int
    ndummy: int
    ndummy2: int
    ndivx: int
    ndivy: int
    x: np.ndarray            # (ntotal2, 2) float64
    mass: np.ndarray
    rho: np.ndarray
    hsml: np.ndarray
    itype: np.ndarray        # int32
    wall_position: np.ndarray  # (ntotal2,) float32
    horizontal_or_not: np.ndarray
    D: dict                  # D11..D42
    pi: float

    @property
    def ntotal(self):
        return self.nnode + self.nstress

    @property
    def ntotal2(self):
        return self.ntotal + self.ndummy


def reference_pi():
    """pi = 4*atan(1.0) evaluated in default real (3_...:144): atan(1.0) is
    the binary32 value nearest pi/4, and scaling by 4 is exact, so this is
    the binary32 value nearest pi, widened."""
    return float(F32(4) * F32(np.arctan(F32(1))))


def elastic_matrix(young, poiss):
    """3_...:719-730 (DP with the single-precision literals 4./3., 2./3.)."""
    K_mod = young / (3 * (1 - 2 * poiss))
    G_mod = young / (2 * (1 + poiss))
    D11 = float(F32(4) / F32(3)) * G_mod + K_mod
    D12 = -(float(F32(2) / F32(3)) * G_mod) + K_mod
    return dict(D11=D11, D22=D11, D12=D12, D33=G_mod, D41=D12, D42=D12)


def wall_dummies(walls, dx, dy):
    """Dummy particles of the walls (set_up_dummy_nodes, 3_...:736-878) in
    default-real arithmetic: three rows per wall, the innermost rows of all
    walls first.  Returns x, y, horizontal_or_not, wall_position (float32)
    and the innermost-row count ndummy2."""
    nrow = 3
    dx2, dy2 = F32(dx), F32(dy)
    ndiv_wall = [int(F32((w.x2 - w.x1) / dx2 + F32(1))) for w in walls]
    ndummy2 = sum(ndiv_wall)
    dxs, dys, hor, wpos = [], [], [], []

    def layer(w, nd, m):
        j = np.arange(nd, dtype=np.float32)
        if w.wall_id in (1, 3):
            xs_ = np.full(nd, F32(w.position - F32(m - 1) * dx2), dtype=np.float32) if m > 1 else \
                np.full(nd, w.position, dtype=np.float32)
            ys_ = (w.x1 + j * dy2).astype(np.float32)
            h = 2.0 if w.wall_id == 1 else 22.0
        elif w.wall_id == 2:
            xs_ = (w.x1 + j * dx2).astype(np.float32)
            ys_ = np.full(nd, F32(w.position - F32(m - 1) * dy2), dtype=np.float32) if m > 1 else \
                np.full(nd, w.position, dtype=np.float32)
            h = 1.0
        else:
            raise DeckError(f"unknown wall id {w.wall_id}")
        dxs.append(xs_); dys.append(ys_)
        hor.append(np.full(nd, h, dtype=np.float32))
        wpos.append(np.full(nd, w.position, dtype=np.float32))

    for w, nd in zip(walls, ndiv_wall):
        layer(w, nd, 1)
    for w, nd in zip(walls, ndiv_wall):
        for m in range(2, nrow + 1):
            layer(w, nd, m)
    if not dxs:
        e = np.zeros(0, dtype=np.float32)
        return e, e, e, e, 0
    return np.concatenate(dxs), np.concatenate(dys), np.concatenate(hor), np.concatenate(wpos), ndummy2


def inside_stress_points(xn, yn, ndivx, ndivy, npoints, x1, y1, dx, dy):
    """Stress-point positions of the inside approach (3_...:541-605).

    One point per cell sits at the cell centre.  With 2 to 4 points, each is
    the mean of three corners of the cell (a lower-left triangle and its
    complements), summed left to right and divided by 3 in double precision;
    cells run x column outer, y row inner, and a cell's points are
    consecutive.  The corner choices are the reference's, including the
    repeated corner of the second point of 2 and of the first point of 4.
    """
    if npoints == 1:
        si = np.repeat(np.arange(1, ndivx, dtype=np.float64), ndivy - 1)
        sj = np.tile(np.arange(1, ndivy, dtype=np.float64), ndivx - 1)
        return x1 - dx / 2. + si * dx, y1 - dx / 2. + sj * dy
    X = xn.reshape(ndivx, ndivy)
    Y = yn[:ndivy]                     # y of a node depends on its row only
    a, b = X[:-1, :-1], X[:-1, 1:]     # column j: rows i, i + 1
    c, d = X[1:, :-1], X[1:, 1:]       # column j + 1: rows i, i + 1
    ya = np.broadcast_to(Y[:-1], a.shape)
    yb = np.broadcast_to(Y[1:], a.shape)
    corners = {
        2: [((a, b, c), (ya, yb, ya)), ((b, c, c), (yb, ya, yb))],
        3: [((a, b, c), (ya, yb, ya)), ((a, c, d), (ya, ya, yb)), None],
        4: [((a, a, d), (ya, yb, ya)), ((b, c, d), (yb, ya, yb)), ((a, b, d), (ya, yb, yb)),
            ((a, c, d), (ya, ya, yb))],
    }[npoints]
    xs = np.empty((ndivx - 1, ndivy - 1, npoints))
    ys = np.empty_like(xs)
    for p, cor in enumerate(corners):
        if cor is None:                # third point of 3: two corners and the left edge's midpoint
            xs[:, :, p] = (b + d + (c + a) / 2.) / 3.
            ys[:, :, p] = (yb + yb + (ya + yb) / 2.) / 3.
            continue
        (p0, p1, p2), (q0, q1, q2) = cor
        xs[:, :, p] = (p0 + p1 + p2) / 3.
        ys[:, :, p] = (q0 + q1 + q2) / 3.
    return xs.ravel(), ys.ravel()


def outside_stress_points(xn, yn, npoints, r_x, r_y):
    """Stress points of the outside approach (3_...:607-660): npoints per
    node at the offsets (+-r_x, +-r_y) of the reference's patterns, a node's
    points consecutive."""
    pat = {1: [(1, 1)], 2: [(1, 1), (-1, -1)], 3: [(0, 1), (-1, -1), (1, -1)],
           4: [(-1, -1), (-1, 1), (1, 1), (1, -1)]}[npoints]
    xs = np.empty((len(xn), npoints))
    ys = np.empty_like(xs)
    for p, (sx, sy) in enumerate(pat):
        xs[:, p] = xn if sx == 0 else (xn + r_x if sx > 0 else xn - r_x)
        ys[:, p] = yn + r_y if sy > 0 else yn - r_y
    return xs.ravel(), ys.ravel()


def bcs_on_particles(deck: Deck, x):
    """Get_BCs_on_node (3_...:1030-1113) over the particles 1..ntotal at
    their positions x (ntotal, 2): bc_info (ntotal, 8) int32 (column 0 the
    1-based particle number, 1 how many BCs, 2.. the BC numbers) and
    bc_or_not (ntotal) int32 (1 where a BC applies).  A particle is on a
    segment when (y1-y2)(x-x1) + (x2-x1)(y-y1) is exactly zero in double
    precision and it lies in the segment's box; segments add their BC in the
    next column, nodal BCs put theirs in the column of its variable (the
    reference asks on standard input when that column already holds another
    BC: here that is an error)."""
    n = len(x)
    info = np.zeros((n, 8), dtype=np.int32)
    info[:, 0] = np.arange(1, n + 1)
    bc_or_not = np.zeros(n, dtype=np.int32)
    xx, yy = x[:, 0], x[:, 1]
    for x1, y1, x2, y2, nb in deck.segment_bcs:
        z = (y1 - y2) * (xx - x1) + (x2 - x1) * (yy - y1)
        on = (z == 0.0) & (xx >= min(x1, x2)) & (xx <= max(x1, x2)) & (yy >= min(y1, y2)) & (yy <= max(y1, y2))
        idx = np.nonzero(on)[0]
        info[idx, 1] += 1
        if np.any(info[idx, 1] > 6):
            raise DeckError("more than 6 boundary conditions on one particle")
        info[idx, info[idx, 1] + 1] = int(nb)
        bc_or_not[idx] = 1
    for node, nb in deck.nodal_bcs:
        i, nb = int(node) - 1, int(nb)
        k = int(deck.bc_list[nb - 1, 1]) + 1
        info[i, 1] += 1
        if info[i, k] not in (0, nb):
            raise DeckError(f"node {i + 1}: two boundary conditions for variable {k - 1}")
        info[i, k] = nb
        bc_or_not[i] = 1
    return info, bc_or_not


def build_particles(deck: Deck) -> Particles:
    if deck.geom_type != 1:
        raise DeckError("only geom_type 1 (rectangle) is implemented, as in the reference (3_...:388-390)")
    if not (deck.sp_sph and 1 <= deck.npoints <= 4):
        raise DeckError("only stress-particle SPH with 1 to 4 stress points per cell or node is supported")
    x1, _x2, _x3, x4 = deck.xs
    y1, _y2, _y3, y4 = deck.ys
    dx, dy = deck.dx, deck.dy
    lx = x4 - x1
    ly = y4 - y1
    ndivx = int(lx / dx + 1)          # integer truncation, 3_...:415-416
    ndivy = int(ly / dy + 1)
    nnode = ndivx * ndivy
    if deck.inside_approach:                    # 3_...:424-438
        nstress = deck.npoints * (ndivx - 1) * (ndivy - 1)
    else:
        nstress = deck.npoints * nnode

    # nodes: i (x) outer, j (y) inner, 3_...:498-514
    ii = np.repeat(np.arange(ndivx, dtype=np.float64), ndivy)
    jj = np.tile(np.arange(ndivy, dtype=np.float64), ndivx)
    xn = x1 + ii * dx
    yn = y1 + jj * dy
    no_int = np.where((xn == x1) | (xn == x4) | (yn == y4), 4, 8)
    corner = ((xn == x1) & ((yn == y1) | (yn == y4))) | ((xn == x4) & ((yn == y1) | (yn == y4)))
    no_int = np.where(corner, 2, no_int)

    if deck.inside_approach:
        xs, ys = inside_stress_points(xn, yn, ndivx, ndivy, deck.npoints, x1, y1, dx, dy)
    else:
        xs, ys = outside_stress_points(xn, yn, deck.npoints, dx * deck.rx_factor, dx * deck.ry_factor)

    rho0 = deck.props[5]
    hs = deck.sml * dx
    area = (rho0 * dx * dy) / 8
    mass_n = no_int * area
    if deck.inside_approach:                    # 3_...:680-705
        mass_s = np.full(nstress, (8 * area) / deck.npoints)
    else:
        mass_s = np.repeat(mass_n / deck.npoints, deck.npoints)

    # dummy walls, default-real arithmetic (3_...:736-878)
    walls = deck.walls if deck.dummy_nodes else []
    wx, wy, hor_all, wpos_all, ndummy2 = wall_dummies(walls, dx, dy)
    ndummy = len(wx)
    dx2 = F32(dx)

    ntotal = nnode + nstress
    ntotal2 = ntotal + ndummy
    x = np.empty((ntotal2, 2), dtype=np.float64)
    x[:nnode, 0], x[:nnode, 1] = xn, yn
    x[nnode:ntotal, 0], x[nnode:ntotal, 1] = xs, ys
    if ndummy:
        x[ntotal:, 0] = wx.astype(np.float64)
        x[ntotal:, 1] = wy.astype(np.float64)
    mass = np.empty(ntotal2)
    mass[:nnode] = mass_n
    mass[nnode:ntotal] = mass_s
    mass[ntotal:] = float(dx2 * dx2) * rho0
    rho = np.full(ntotal2, rho0)
    hsml = np.full(ntotal2, hs)
    itype = np.empty(ntotal2, dtype=np.int32)
    itype[:nnode], itype[nnode:ntotal], itype[ntotal:] = 2, 1, 25
    wall_position = np.zeros(ntotal2, dtype=np.float32)
    horizontal_or_not = np.zeros(ntotal2, dtype=np.float32)
    if ndummy:
        wall_position[ntotal:] = wpos_all
        horizontal_or_not[ntotal:] = hor_all
    return Particles(nnode=nnode, nstress=nstress, ndummy=ndummy, ndummy2=ndummy2,
                     ndivx=ndivx, ndivy=ndivy, x=x, mass=mass, rho=rho, hsml=hsml,
                     itype=itype, wall_position=wall_position,
                     horizontal_or_not=horizontal_or_not,
                     D=elastic_matrix(deck.props[2], deck.props[3]), pi=reference_pi())

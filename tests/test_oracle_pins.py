"""Independent pins of the oracle, beside the golden fixtures.

- the pair set equals an O(N^2) brute force with the reference's acceptance
  test (code/2_SPH_main_2018.f90:1345-1353);
- the list order equals a simulation of the reference's linked-list
  bookkeeping (allocate and prepend on growth, reuse in place, truncate and
  re-attach: 2_...:1286-1401) applied to the discovery order of its cell
  traversal, over steps where the pair count rises and falls;
- the kernel values of isolated pairs satisfy the cubic spline's
  normalisation, its value at the origin, its compact support and
  W'(r) = dW/dr (finite differences), which a dropped factor, a wrong power
  or a sign error would all break;
- the error behaviour of the C-ABI.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import deck_dir, load_pkg

F32 = np.float32


def _sim(deck_name, lib):
    sps = load_pkg()
    return sps.Simulation.from_deck(sps.read_deck(deck_dir(deck_name)), lib_path=lib)


def brute_force_pairs(x, h, ood):
    """All unordered pairs with sqrt(dx^2 + dy^2) < 2 * (h_i + h_j) / 2."""
    idx = np.nonzero(ood == 0)[0]
    xs, hs = x[idx], h[idx]
    dx0 = xs[:, None, 0] - xs[None, :, 0]
    dx1 = xs[:, None, 1] - xs[None, :, 1]
    d = dx0 * dx0
    d = d + dx1 * dx1
    mh = (hs[:, None] + hs[None, :]) / 2.0
    acc = np.sqrt(d) < 2 * mh
    iu = np.triu_indices(len(idx), 1)
    sel = acc[iu]
    return {(int(idx[a]), int(idx[b])) for a, b in zip(iu[0][sel], iu[1][sel])}


@pytest.mark.parametrize("steps_before", [0, 150])
def test_pair_set_is_brute_force(oracle_lib, steps_before):
    sim = _sim("t1_tiny", oracle_lib)
    if steps_before:
        sim.step(steps_before)
    st = sim.get_state()
    sim.step(1)
    p = sim.get_pairs()
    got = {(min(a, b) - 1, max(a, b) - 1) for a, b in zip(p["i"], p["j"])}
    assert len(got) == len(p["i"])
    assert got == brute_force_pairs(st["x"], st["hsml"], st["if_out_domain"])


def discovery_order(x, h, ood):
    """Pairs in the order the reference's cell traversal discovers them."""
    n = len(x)
    live = [i for i in range(n) if ood[i] != 1]
    xmin = [1e10, 1e10]
    xmax = [-1e10, -1e10]
    hmax = 0.0
    for i in live:
        for d in range(2):
            xmin[d] = min(xmin[d], x[i, d])
            xmax[d] = max(xmax[d], x[i, d])
        hmax = max(hmax, h[i])
    deltx = hmax * 2
    nd, lo = [0, 0], [0.0, 0.0]
    for d in range(2):
        length = xmax[d] - xmin[d]
        nd[d] = int(length / deltx + 1)
        lo[d] = xmin[d] - (nd[d] * deltx - length) / 2 - float(F32(0.001)) * length
    cells = {}
    for i in live:
        c = [min(int((x[i, d] - lo[d]) / deltx + 1), nd[d]) for d in range(2)]
        cells.setdefault((c[0], c[1]), []).append(i)
    out = []
    for cy in range(1, nd[1] + 1):
        for cx in range(1, nd[0] + 1):
            mine = cells.get((cx, cy), [])
            for a, i in enumerate(mine):
                for jy in range(cy, min(cy + 1, nd[1]) + 1):
                    for jx in range(max(1, cx - 1), min(nd[0], cx + 1) + 1):
                        if (jy, jx) < (cy, cx):
                            continue
                        other = cells.get((jx, jy), [])
                        start = a + 1 if (jx, jy) == (cx, cy) else 0
                        for j in other[start:]:
                            d0, d1 = x[i, 0] - x[j, 0], x[i, 1] - x[j, 1]
                            if math.sqrt(d0 * d0 + d1 * d1) < 2 * ((h[i] + h[j]) / 2.0):
                                out.append((min(i, j), max(i, j)))
    return out


class LinkedList:
    """The reference's pair list bookkeeping, node by node (2_...:1286-1401)."""

    def __init__(self):
        self.head = None          # `last`
        self.m_pairs = 0
        self.n_pairs = 0
        self.keep0 = self.keep1 = None

    def search(self, payloads):
        class Node:
            __slots__ = ("val", "next")

        if self.n_pairs < self.m_pairs:       # re-attach the truncated tail
            self.keep0.next = self.keep1
        cur = self.head
        self.n_pairs = 0
        for val in payloads:
            self.n_pairs += 1
            if self.n_pairs > self.m_pairs:   # grow: prepend a fresh node
                node = Node()
                node.next = self.head
                self.head = node
                cur = node
                self.m_pairs += 1
            cur.val = val
            if self.n_pairs < self.m_pairs:
                self.keep0 = cur
                cur = cur.next
        if self.n_pairs < self.m_pairs:       # truncate
            self.keep1 = cur
            self.keep0.next = None

    def traverse(self):
        out, node = [], self.head
        while node is not None:
            out.append(node.val)
            node = node.next
        return out


def test_list_order_matches_linked_list_simulation(oracle_lib):
    sim = _sim("t2_cspm", oracle_lib)
    ll = LinkedList()
    counts = []
    for step in range(1, 61):
        st = sim.get_state()
        sim.step(1)
        disc = discovery_order(st["x"], st["hsml"], st["if_out_domain"])
        ll.search(disc)
        p = sim.get_pairs()
        got = [(min(a, b) - 1, max(a, b) - 1) for a, b in zip(p["i"], p["j"])]
        assert got == ll.traverse(), f"step {step}"
        counts.append(len(disc))
    # the run must exercise growth past the high-water mark and shrinking
    assert max(counts) > counts[0] and min(counts) < counts[0]


def _isolated_pairs(oracle_lib, rs, h=1.0):
    """One node and one stress point per distance r, groups 10 h apart."""
    sps = load_pkg()
    deck = sps.read_deck(deck_dir("t3_elastic"))
    n = len(rs)
    parts = sps.build_particles(deck)
    parts.nnode, parts.nstress, parts.ndummy, parts.ndummy2 = n, n, 0, 0
    x = np.zeros((2 * n, 2))
    x[:n, 0] = np.arange(n) * 10.0 * h
    x[n:, 0] = x[:n, 0] + np.asarray(rs)
    parts.x = x
    parts.mass = np.full(2 * n, 1.0)
    parts.rho = np.full(2 * n, 1.0)
    parts.hsml = np.full(2 * n, h)
    parts.itype = np.array([2] * n + [1] * n, dtype=np.int32)
    parts.wall_position = np.zeros(2 * n, dtype=np.float32)
    parts.horizontal_or_not = np.zeros(2 * n, dtype=np.float32)
    deck.dummy_nodes = False
    deck.xmin_domain = np.array([-1e3, -1e3])
    deck.xmax_domain = np.array([1e5, 1e3])
    prm = sps.make_params(deck, parts)
    sim = sps.Simulation(prm, oracle_lib)
    sim.deck = deck
    sim.set_state(sps.initial_state(parts, sim.niv))
    sim.step(1, dt=1e-9)
    p = sim.get_pairs()
    # map back to r order: pair_i is the stress point (type 1)
    order = np.argsort(p["i"])
    return {k: p[k][order] for k in p}, parts.pi


def test_cubic_spline_kernel_properties(oracle_lib):
    h = 1.0
    rs = np.concatenate([[0.0], np.linspace(1e-3, 2.0 - 1e-3, 4001)])
    p, pi32 = _isolated_pairs(oracle_lib, rs, h)
    assert len(p["w"]) == len(rs)
    assert np.all(p["type"] == 1)
    w = p["w"].astype(np.float64)
    # value at the origin: 10 / (7 pi h^2) with the single-precision pi
    assert abs(w[0] - 10.0 / (7.0 * pi32 * h * h)) <= 1e-6 * w[0]
    assert p["dwdx"][0] == 0.0 and p["dwdy"][0] == 0.0
    r, wr = rs[1:], w[1:]
    # normalisation: int_0^{2h} W(r) 2 pi r dr = 1 (pi of the factor is single precision)
    integral = np.trapezoid(wr * 2 * np.pi * r, r) + w[0] * np.pi * r[0] ** 2
    assert abs(integral - 1.0) < 2e-6
    # compact support: W and W' vanish at 2h
    assert wr[-1] < 1e-8 and abs(p["dwdx"][-1]) < 1e-5
    # gradient: dW/dx_receiver for dx = x_sp - x_node = +r along x is W'(r)
    dwdr = p["dwdx"][1:].astype(np.float64)
    fd = np.gradient(wr, r)
    inner = (r > 0.01) & (r < 1.99) & (np.abs(r - 1.0) > 0.01)
    assert np.allclose(dwdr[inner], fd[inner], rtol=2e-3, atol=2e-5)
    assert np.all(p["dwdy"] == 0.0)
    # continuity at q = 1 (branch change)
    k = np.argmin(np.abs(r - 1.0))
    assert abs(wr[k] - 15.0 / (7 * np.pi) / 6.0) < 1e-3


def test_error_behaviour(oracle_lib):
    sps = load_pkg()
    deck = sps.read_deck(deck_dir("t1_tiny"))
    parts = sps.build_particles(deck)
    for name, mode in (("cont_density", sps.MODE_FAST), ("art_stress", sps.MODE_ORDERED)):
        prm = sps.make_params(deck, parts, mode=mode)
        setattr(prm, name, 1)
        with pytest.raises(sps.SpsError) as e:
            sps.Simulation(prm, oracle_lib)
        assert e.value.code == -1
    prm = sps.make_params(deck, parts)
    sim = sps.Simulation(prm, oracle_lib)
    sim.deck = deck
    with pytest.raises(sps.SpsError) as e:
        sim.step(1)
    assert e.value.code == -5
    sim.set_state(sps.initial_state(parts, sim.niv))
    sim.step(1)
    n = ctypes.c_int64(0)
    rc = sim.lib.sps_get_pairs(sim.handle, 10, None, None, None, None, None, None, ctypes.byref(n))
    assert rc == -4 and n.value == sim.stats()["n_pairs"]


def test_non_finite_position_is_flagged_out_of_domain(oracle_lib):
    """A particle whose position is not finite is flagged out of domain (the
    reference's dxx > 0 test misses NaN) and never enters the search: the
    step completes and the particle has no pairs."""
    sps = load_pkg()
    deck = sps.read_deck(deck_dir("t1_tiny"))
    parts = sps.build_particles(deck)
    sim = sps.Simulation(sps.make_params(deck, parts), oracle_lib)
    st = sps.initial_state(parts, sim.niv)
    k = parts.nnode + 3                       # a stress point
    st["x"][k] = (np.nan, 0.005)
    sim.set_state(st)
    sim.deck = deck
    sim.step(2)
    s = sim.get_state(["if_out_domain"])
    assert s["if_out_domain"][k] == 1
    assert int(s["if_out_domain"].sum()) == 1
    p = sim.get_pairs()
    assert not np.any(p["i"] == k + 1) and not np.any(p["j"] == k + 1)

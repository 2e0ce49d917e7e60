"""The leapfrog integrator (sps_set_integrator; no counterpart in the
reference, which integrates with RK4 only, 2_SPH_main_2018.f90:653-802).

Pinned without the reference, as it has none:
* free fall: a block without walls under gravity keeps (to rounding) zero
  stress, and kick-drift leapfrog started with a half kick puts every
  particle at the closed form x0 + g t^2 / 2 at every whole step, with the
  velocity g (t - dt / 2) at the half steps;
* convergence to RK4: on a smooth elastic block the leapfrog displacement
  and stress approach a fine-step RK4 solution at second order (each
  halving of dt divides the error by 4);
* the GPU's leapfrog equals the oracle's bit for bit (the kick, the stress
  update, the drift and the first half kick are the same arithmetic).
Leapfrog is never used for the reference-parity claims.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import ORACLE_LIB, PRODUCT_LIB, bits_equal, deck_dir, describe_diff, load_pkg


def wall_free_block(tmp_path):
    """The tiny deck without its walls (dummy nodes F)."""
    d = tmp_path / "freefall"
    shutil.copytree(deck_dir("t1_tiny"), d)
    pts = d / "t1_tiny.pts"
    lines = pts.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("dummy nodes"))
    pts.write_text("\n".join(lines[:k + 1] + ["F"]) + "\n")
    return str(d)


def test_leapfrog_free_fall_closed_form(oracle_lib, tmp_path):
    sps = load_pkg()
    deck = sps.read_deck(wall_free_block(tmp_path))
    sim = sps.Simulation.from_deck(deck, lib_path=oracle_lib)
    assert sim.params.ndummy == 0
    sim.set_integrator(sps.INTEGRATOR_LEAPFROG)
    x0 = sim.get_state(["x"])["x"][:sim.n].copy()
    g, dt = -9.81, deck.dt
    for k in (1, 10, 100):
        sim.step(k - sim.itimestep_sph.value)
        s = sim.get_state(["x", "vel", "stress"])
        t = k * dt
        x_exact = x0[:, 1] + 0.5 * g * t * t
        v_half = g * (t - 0.5 * dt)
        assert np.abs(s["x"][:sim.n, 1] - x_exact).max() <= 1e-12 * max(1.0, np.abs(x0[:, 1]).max())
        assert np.abs(s["x"][:sim.n, 0] - x0[:, 0]).max() <= 1e-15
        assert np.abs(s["vel"][:sim.nn, 1] - v_half).max() <= 1e-12 * abs(v_half)
        assert np.abs(s["stress"][:sim.n]).max() <= 1e-6     # Pa: no stress builds up in free fall


def smooth_elastic_block(tmp_path):
    """The elastic block (t3_elastic) with nothing non-smooth in it: material
    law 1 (with eco 1: elastic and no return map; the deck's law 12 applies
    the Drucker-Prager return map at every stage even to the elastic
    material) and no artificial viscosity (its switch on the sign of v.x)."""
    d = tmp_path / "smooth"
    shutil.copytree(deck_dir("t3_elastic"), d)
    dat = d / "t3_elastic.dat"
    lines = dat.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("mater."))
    f = lines[k + 1].split()
    f[2] = "1"
    lines[k + 1] = "  ".join(f)
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("tan(phi)"))
    del lines[k:k + 2]
    k = next(i for i, ln in enumerate(lines) if ln.lower().startswith("alpha"))
    lines[k + 1] = "  0.0  0.0"
    dat.write_text("\n".join(lines) + "\n")
    return str(d)


def run_to(sps, deck, lib, integrator, dt, t_end):
    sim = sps.Simulation.from_deck(deck, lib_path=lib)
    if integrator:
        sim.set_integrator(integrator)
    sim.step(int(round(t_end / dt)), dt=dt)
    s = sim.get_state(["displ", "stress"])
    # the nodes and stress points (a wall dummy's stress is a copy of a node
    # stress taken at a different point of the step by each integrator)
    s["stress"] = s["stress"][:sim.n]
    return s


def test_leapfrog_converges_to_rk4_at_second_order(oracle_lib, tmp_path):
    sps = load_pkg()
    deck = sps.read_deck(smooth_elastic_block(tmp_path))
    dt0 = deck.dt
    t_end = 40 * dt0
    ref = run_to(sps, deck, oracle_lib, 0, dt0 / 64, t_end)   # RK4 is fourth order here
    errs = {"displ": [], "stress": []}
    for m in (2, 4, 8):
        s = run_to(sps, deck, oracle_lib, sps.INTEGRATOR_LEAPFROG, dt0 / m, t_end)
        for k in errs:
            errs[k].append(np.abs(s[k] - ref[k]).max() / np.abs(ref[k]).max())
    print(f"leapfrog errors vs RK4 at dt/64: {errs}")
    for k, e in errs.items():
        ratios = [e[0] / e[1], e[1] / e[2]]
        assert all(3.5 < r < 4.5 for r in ratios), (k, ratios)
    assert errs["displ"][-1] < 1e-3 and errs["stress"][-1] < 1e-4


def test_leapfrog_rejects_what_it_does_not_cover(oracle_lib):
    sps = load_pkg()
    sim = sps.Simulation.from_deck(sps.read_deck(deck_dir("t4_xsph")), lib_path=oracle_lib)
    with pytest.raises(sps.SpsError):
        sim.set_integrator(sps.INTEGRATOR_LEAPFROG)
    sim = sps.Simulation.from_deck(sps.read_deck(deck_dir("t1_tiny")), lib_path=oracle_lib)
    with pytest.raises(sps.SpsError):
        sim.set_integrator(7)


@pytest.mark.gpu
@pytest.mark.parametrize("deck_name,steps", [("t2_cspm", 60), ("c1_collapse", 20), ("v1_vslope_sp1", 50)])
def test_leapfrog_gpu_equals_oracle(deck_name, steps):
    sps = load_pkg()
    deck = sps.read_deck(deck_dir(deck_name))
    sims = []
    for lib in (PRODUCT_LIB, ORACLE_LIB):
        sim = sps.Simulation.from_deck(deck, lib_path=lib)
        sim.set_integrator(sps.INTEGRATOR_LEAPFROG)
        sim.step(steps)
        sims.append(sim.get_state())
    a, b = sims
    bad = [f"{k}: {describe_diff(b[k], a[k])}" for k in a if not bits_equal(b[k], a[k])]
    assert not bad, "; ".join(bad)

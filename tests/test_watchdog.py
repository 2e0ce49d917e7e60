"""The non-finite watchdog (failure detection; the reference has none and runs
on with NaNs, code/3_SPH_material_2018.f90:1793 being its only isnan).

After every step both libraries check that each particle in the domain box
has a finite position, velocity and stress; an sps_step call that met a
non-finite one returns SPS_ENUMERIC (-6).  The HIP library then refuses to
step until the state is set again.  A particle outside the domain box (which
a non-finite position makes it) interacts with nothing and is not checked.
The NaN is planted in the input state, on the host.
"""
import numpy as np
import pytest

from conftest import ORACLE_LIB, PRODUCT_LIB, deck_dir, load_pkg


def nan_state(sps, deck_name, lib, where="stress"):
    deck = sps.read_deck(deck_dir(deck_name))
    sim = sps.Simulation.from_deck(deck, lib_path=lib)
    st = sps.initial_state(sim.parts, sim.niv)
    k = sim.parts.nnode + 7                  # an in-domain stress point
    if where == "stress":
        st["stress"][k, 1] = np.nan
    else:
        st["vel"][3] = (np.inf, 0.0)         # a node
    sim.set_state(st)
    return sim, st


@pytest.mark.parametrize("where", ["stress", "vel"])
def test_oracle_flags_non_finite_state(oracle_lib, where):
    sps = load_pkg()
    sim, _ = nan_state(sps, "t2_cspm", oracle_lib, where)
    with pytest.raises(sps.SpsError) as e:
        sim.step(2)
    assert e.value.code == -6


def test_oracle_clean_state_passes(oracle_lib):
    sps = load_pkg()
    sim = sps.Simulation.from_deck(sps.read_deck(deck_dir("t2_cspm")), lib_path=oracle_lib)
    sim.step(3)


@pytest.mark.gpu
@pytest.mark.parametrize("where", ["stress", "vel"])
def test_hip_flags_non_finite_state(where):
    sps = load_pkg()
    sim, st = nan_state(sps, "t2_cspm", PRODUCT_LIB, where)
    with pytest.raises(sps.SpsError) as e:
        sim.step(2)
    assert e.value.code == -6
    with pytest.raises(sps.SpsError) as e:    # the handle needs a new state
        sim.step(1)
    assert e.value.code == -5
    clean = sps.initial_state(sim.parts, sim.niv)
    sim.set_state(clean)                        # the flag starts clear again
    sim.time_sph.value, sim.itimestep_sph.value = 0.0, 0
    sim.step(2)
    ora = sps.Simulation.from_deck(sps.read_deck(deck_dir("t2_cspm")), lib_path=ORACLE_LIB)
    ora.step(2)
    a, b = sim.get_state(["stress", "vel"]), ora.get_state(["stress", "vel"])
    for k in a:
        assert a[k].tobytes() == b[k].tobytes(), k

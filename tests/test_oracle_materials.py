"""The oracle's material path (sps_set_materials), pinned: with a one-row
table equal to the deck's own constants every particle's material is the
reference's, so the results must be those of the golden-pinned single-
material path, bit for bit -- through yield (c1_collapse) and with cohesion
and CSPM (t2_cspm).  A table of two rows with the particles split between
them, both equal to the deck's constants, must give the same again."""
import numpy as np
import pytest

from conftest import ORACLE_LIB, deck_dir, load_pkg


def table_of(sim):
    p = sim.params
    return np.array([[p.props[12], p.props[13], p.props[2], p.props[3], p.D11, p.D22, p.D12, p.D33, p.D41,
                      p.D42]])


@pytest.mark.parametrize("deck_name,steps,rows", [("c1_collapse", 20, 1), ("t2_cspm", 40, 1), ("t2_cspm", 40, 2)])
def test_material_table_of_the_deck_reduces_to_single_material(deck_name, steps, rows):
    sps = load_pkg()
    deck = sps.read_deck(deck_dir(deck_name))
    plain = sps.Simulation.from_deck(deck, lib_path=ORACLE_LIB)
    mat = sps.Simulation.from_deck(deck, lib_path=ORACLE_LIB)
    table = np.repeat(table_of(mat), rows, axis=0)
    mat.set_materials(table, np.arange(mat.n2, dtype=np.int32) % rows)
    plain.step(steps)
    mat.step(steps)
    a, b = plain.get_state(), mat.get_state()
    bad = [k for k in a if not np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))]
    assert not bad, bad
    assert np.count_nonzero(a["internal_vars"][:, 0]) > 0 or deck_name == "t2_cspm"


def test_material_table_changes_results():
    """A different friction angle in the table changes the yielding."""
    sps = load_pkg()
    deck = sps.read_deck(deck_dir("c1_collapse"))
    plain = sps.Simulation.from_deck(deck, lib_path=ORACLE_LIB)
    mat = sps.Simulation.from_deck(deck, lib_path=ORACLE_LIB)
    table = table_of(mat)
    table[0, 0] = np.tan(np.radians(40.0))
    mat.set_materials(table, np.zeros(mat.n2, dtype=np.int32))
    plain.step(20)
    mat.step(20)
    assert not np.array_equal(plain.get_state(["stress"])["stress"], mat.get_state(["stress"])["stress"])

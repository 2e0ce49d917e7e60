"""Biaxial compression with a perturbed friction angle.

The sample, its platens and its confined faces come from a deck in the
reference's format (decks/c5_biaxial: 0.1 m x 0.2 m, prescribed platen
velocities and face stresses as the reference's strain-localisation example
prescribes them, 3_SPH_material_2018.f90:1644-1753).  What the reference
cannot express is a friction angle that varies from particle to particle
(it has one material row, props(1,:)): here every particle draws its angle
from phi * (1 + rel * u), u uniform on [-1, 1] in LEVELS equal steps, from
a seeded SplitMix64 stream (Steele, Lea & Flood 2014; the i-th particle takes
the stream's i-th output), and the angles reach the kernels as a material
table (include/sps.h, sps_set_materials) of LEVELS rows.
"""
from __future__ import annotations

import math

import numpy as np

from .setup import elastic_matrix

LEVELS = 1024
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_stream(n, seed):
    """The first n outputs of SplitMix64 from `seed`, vectorised: the state
    advances by the golden gamma per output, each output mixes its state."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def phi_levels(n, seed, levels=LEVELS):
    """Each particle's level 0..levels-1, uniform."""
    u = (splitmix64_stream(n, seed) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.minimum((u * levels).astype(np.int32), levels - 1)


def material_table(tan_phi, coh, young, poiss, rel=0.02, levels=LEVELS):
    """Rows (tan phi, c, E, nu, D11, D22, D12, D33, D41, D42) of the levels:
    phi_k = phi (1 + rel (2 (k + 1/2) / levels - 1))."""
    phi = math.atan(tan_phi)
    D = elastic_matrix(young, poiss)
    k = np.arange(levels)
    phik = phi * (1.0 + rel * (2.0 * (k + 0.5) / levels - 1.0))
    table = np.empty((levels, 10))
    table[:, 0] = np.tan(phik)
    table[:, 1:4] = (coh, young, poiss)
    table[:, 4:] = (D["D11"], D["D22"], D["D12"], D["D33"], D["D41"], D["D42"])
    return table


def open_biaxial(sps, deck, seed=2024, rel=0.02, mode=0, lib_path=None, parts=None):
    """A Simulation of the deck with each particle's friction angle perturbed
    by up to rel (seeded); the deck's BCs are set.  `parts`: the deck's
    particles if already built."""
    sim = sps.Simulation.from_deck(deck, lib_path=lib_path or sps.PRODUCT_LIB, mode=mode, parts=parts)
    p = sim.params
    table = material_table(p.props[12], p.props[13], p.props[2], p.props[3], rel)
    mat = phi_levels(sim.n2, seed)
    sim.set_materials(table, mat)
    sim.biaxial = (table, mat)
    return sim

"""Readers for the binary dumps written by oracle/ref_dump_driver.f90.

The dump driver is linked against the reference Fortran modules and writes
state in the reference's own (Fortran column-major) layout.  These readers turn
the dumps into dicts of numpy arrays; `make_golden.py` reduces them to the
committed fixtures, and the tests read the fixtures, never the raw dumps.
"""
import numpy as np

STATE_MAGIC = 0x53505331
PAIRS_MAGIC = 0x53505032

SCALARS = ("time_sph", "dt_sph", "pi", "D11", "D22", "D12", "D33", "D41", "D42",
           "alpha", "beta", "sml", "dx", "dy", "ft_grav", "dampingTG")


def read_state(path):
    b = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(b, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a

    hdr = take("<i4", 8)
    assert hdr[0] == STATE_MAGIC, "not a state dump"
    nnode, nstress, ndummy, ntotal, ntotal2, step, ndummy2 = (int(v) for v in hdr[1:])
    d = dict(nnode=nnode, nstress=nstress, ndummy=ndummy, ntotal=ntotal,
             ntotal2=ntotal2, step=step, ndummy2=ndummy2)
    for k, v in zip(SCALARS, take("<f8", 16)):
        d[k] = float(v)
    d["xmin_dom"] = take("<f8", 2)
    d["xmax_dom"] = take("<f8", 2)
    d["cgrav"] = take("<f8", 2)
    d["props"] = take("<f8", 20)
    d["x"] = take("<f8", 2 * ntotal2).reshape(ntotal2, 2)
    d["vel"] = take("<f8", 2 * ntotal2).reshape(ntotal2, 2)
    d["stress"] = take("<f8", 4 * ntotal2).reshape(ntotal2, 4)
    d["mass"] = take("<f8", ntotal2)
    d["rho"] = take("<f8", ntotal2)
    d["hsml"] = take("<f8", ntotal2)
    d["itype"] = take("<i4", ntotal2)
    d["if_out_domain"] = take("<i4", ntotal2)
    d["ivar1"] = take("<f8", ntotal)
    d["f_drucker"] = take("<f8", ntotal)
    d["ddev_strn"] = take("<f8", ntotal)
    # grad_u(2,2,ntotal) column-major: per particle [g11, g21, g12, g22]
    d["grad_u"] = take("<f8", 4 * ntotal).reshape(ntotal, 4)
    d["displ"] = take("<f8", 2 * nnode).reshape(nnode, 2)
    d["f_bound"] = take("<f8", 2 * nnode).reshape(nnode, 2)
    d["art_visc"] = take("<f8", 2 * nnode).reshape(nnode, 2)
    d["x00"] = take("<f8", 2 * ntotal2).reshape(ntotal2, 2)
    if off < len(b):
        d["wall_position"] = take("<f4", ntotal2)
        d["horizontal_or_not"] = take("<f4", ntotal2)
    assert off == len(b), (off, len(b))
    return d


PAIR_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("type", "<i4"),
                       ("w", "<f4"), ("dwdx", "<f4"), ("dwdy", "<f4")])


def read_pairs(path):
    b = open(path, "rb").read()
    magic, n = np.frombuffer(b, "<i4", 2)
    assert magic == PAIRS_MAGIC
    p = np.frombuffer(b, PAIR_DTYPE, int(n), 8).copy()
    assert 8 + p.nbytes == len(b)
    return p

"""Autotuner of the mixed form's run-time configuration.

No counterpart in the reference.  For a simulation built by `factory` (a
callable returning a fresh Simulation of SPS_MODE_FAST at the same initial
state), every candidate configuration is timed over `steps` steps after
`warmup` steps, and the fastest is returned with the table of timings.  The
knobs are the ones the library exposes at run time: the pair-list skin
(lists kept while no particle has moved half of it: a larger skin searches
less often but walks longer lists) and graph replay of the list-reusing step.
The kernels' tile (256 receivers, one workgroup) and window capacities are
compile-time constants of the LDS layouts and are not tuned.
"""
import itertools
import time

DEFAULT_SKINS = (0.1, 0.25, 0.5)     # in lattice spacings


def tune(factory, dx, skins=DEFAULT_SKINS, graphs=(False, True), warmup=3, steps=10, sync=None, log=None):
    results = []
    best = None
    for skin, graph in itertools.product(skins, graphs):
        sim = factory()
        sim.set_list_skin(skin * dx)
        sim.set_graph(graph)
        sim.step(warmup)
        if sync:
            sync()
        t0 = time.perf_counter()
        sim.step(steps)
        if sync:
            sync()
        dt = (time.perf_counter() - t0) / steps
        searches = sim.search_count()
        sim.close()
        row = {"skin_dx": skin, "graph": graph, "ms_per_step": round(1e3 * dt, 4), "searches": searches}
        results.append(row)
        if log:
            log(f"[autotune] skin {skin} dx, graph {graph}: {1e3 * dt:.3f} ms/step ({searches} searches)")
        if best is None or dt < best[0]:
            best = (dt, skin, graph)
    return {"skin_dx": best[1], "graph": best[2], "ms_per_step": round(1e3 * best[0], 4)}, results

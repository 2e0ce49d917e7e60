"""Error of a state against a double-precision reference state.

The measures reported for the mixed-precision form (and the surrogate
return map): displacement, stress and failure-surface location, each
relative to the reference field's own scale so that they read the same on
any deck.  Both states come from Simulation.get_state() of the same deck
after the same steps; x0 is the reference's initial positions.
"""
import numpy as np


def _rel_max(a, b, scale):
    return float(np.abs(a - b).max() / scale) if a.size else 0.0


def plastic_band(ivars, frac=0.1):
    """Stress points in the failure surface: plastic strain (internal
    variable 1, as the reference writes it out) above `frac` of its maximum.
    Empty when nothing has yielded."""
    ep = ivars[:, 0]
    top = np.abs(ep).max() if ep.size else 0.0
    if top <= 0.0:
        return np.zeros(ep.shape, dtype=bool)
    return np.abs(ep) >= frac * top


def compare_states(test, ref, x0, nn=None):
    """Relative errors of `test` against `ref`:

    disp_rel_max   max |x_t - x_r| / max |x_r - x0|  (displacement scale;
                   positions when nothing has moved)
    vel_rel_max    max |v_t - v_r| / max |v_r|
    stress_rel_max max |s_t - s_r| / max |s_r| over nodes and stress points
    stress_rel_rms rms |s_t - s_r| / rms |s_r|
    band_mismatch  1 - |A and B| / |A or B| of the two failure surfaces
                   (plastic_band), 0 when neither has yielded
    """
    n = ref["internal_vars"].shape[0]      # nodes + stress points
    xr, xt = ref["x"][:n], test["x"][:n]
    u = np.linalg.norm(xr - x0[:n], axis=1)
    du = np.linalg.norm(xt - xr, axis=1)
    uscale = u.max() if u.max() > 0 else np.abs(xr).max()
    vr, vt = ref["vel"][:n], test["vel"][:n]
    vscale = max(np.abs(vr).max(), 1e-300)
    sr, st = ref["stress"][:n], test["stress"][:n]
    sscale = max(np.abs(sr).max(), 1e-300)
    a, b = plastic_band(test["internal_vars"]), plastic_band(ref["internal_vars"])
    union = np.count_nonzero(a | b)
    return {
        "disp_rel_max": float(du.max() / uscale) if du.size else 0.0,
        "vel_rel_max": _rel_max(vt, vr, vscale),
        "stress_rel_max": _rel_max(st, sr, sscale),
        "stress_rel_rms": float(np.sqrt(np.mean((st - sr) ** 2)) / max(np.sqrt(np.mean(sr ** 2)), 1e-300)),
        "band_mismatch": float(1.0 - np.count_nonzero(a & b) / union) if union else 0.0,
    }

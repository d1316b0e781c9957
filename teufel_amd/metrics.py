"""Error and throughput measures of the radiation step.

trace_rel_error    eps_E: max over pixels of max_i |E - E_ref| / max_i |E_ref|,
                   taken separately for the E and the B triplet of each pixel
                   (vector norms per sample), then the larger of the two.
energy_rel_error   eps_W: max_p |W_p - W_p^ref| / |W_p^ref| over pixels with
                   |W_p^ref| >= floor * max |W^ref|; also returns the fraction
                   of pixels excluded by the floor.
"""
import numpy as np


def trace_rel_error(traces, ref):
    traces = np.asarray(traces, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    worst = 0.0
    for sl in (slice(0, 3), slice(3, 6)):
        d = np.linalg.norm(traces[..., sl] - ref[..., sl], axis=-1).max(axis=-1)
        m = np.linalg.norm(ref[..., sl], axis=-1).max(axis=-1)
        ok = m > 0
        if np.any(~ok) and np.any(d[~ok] > 0):
            return float("inf")
        if np.any(ok):
            worst = max(worst, float((d[ok] / m[ok]).max()))
    return worst


def energy_rel_error(W, W_ref, floor=1e-3):
    W = np.asarray(W, dtype=np.float64)
    W_ref = np.asarray(W_ref, dtype=np.float64)
    a = np.abs(W_ref)
    if a.size == 0 or a.max() == 0:
        return 0.0, 0.0
    sel = a >= floor * a.max()
    err = np.abs(W[sel] - W_ref[sel]) / a[sel]
    return float(err.max()), float(1.0 - sel.mean())

"""Shared helpers for the tests: problem builders and comparison metrics."""
import numpy as np

BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
ALL = ["x", "v", "xx", "xv", "vv"]


def field_errors(a, b, mass_ref):
    """Per-field errors of a against b (dicts of arrays).

    m, x, v: max |a-b| / max(|b|, floor), floor = 1e-12 * max|b|.
    second moments: max |a-b| / sqrt(diag_i diag_j) of b, over cells whose
    mass is above 1e-9 of the largest cell mass (tiny slivers carry
    covariances dominated by rounding in any summation order)."""
    out = {}
    mmax = np.abs(mass_ref).max()
    keep = mass_ref > 1e-9 * mmax
    for f in b:
        if f not in a:
            continue
        d = np.abs(a[f] - b[f])
        if f in ("m", "x", "v"):
            scale = np.maximum(np.abs(b[f]), 1e-12 * np.abs(b[f]).max() + 1e-300)
            r = d / scale
            if f != "m":
                r = r[keep]
            out[f] = float(r.max()) if r.size else 0.0
        else:
            t = b[f]
            # scale of entry (i,j): sqrt(|t_ii t_jj|) of the cross-moment pair
            if f == "xx":
                di = np.abs(np.diagonal(t, axis1=-2, axis2=-1))
                dj = di
            elif f == "vv":
                di = np.abs(np.diagonal(t, axis1=-2, axis2=-1))
                dj = di
            else:
                di = np.abs(np.diagonal(b.get("xx", np.abs(t)), axis1=-2, axis2=-1)) if "xx" in b else np.abs(t).max(axis=-1)
                dj = np.abs(np.diagonal(b.get("vv", np.abs(t)), axis1=-2, axis2=-1)) if "vv" in b else np.abs(t).max(axis=-2)
            scale = np.sqrt(di[..., :, None] * dj[..., None, :]) + 1e-300
            r = (d / scale)[keep]
            out[f] = float(r.max()) if r.size else 0.0
    return out

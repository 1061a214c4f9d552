"""Shared helpers for the tests: comparison metrics.

Two error measures per field, both stated in the parity tests:
  rel  -- on "heavy" cells (mass >= 1e-3 of the peak cell mass):
            m:      |dm| / m
            x, v:   |dX| / (|X| + rms of the field over heavy cells)
            2nd:    |dT_ij| / sqrt(|T_ii| |T_jj|)  (xv: sqrt(|xx_ii| |vv_jj|) + |xv_ij|;
                    the reference's xv formula, src/geometry.c:139-140, adds
                    tet-sized terms, so xv is measured against its own size)
  wabs -- over all cells, mass-weighted and normalised by the field's peak:
            m: max |dm| / max m;  others: max |d(m T)| / max |m T|.
Cells holding a sliver of a tet carry moments whose relative error grows as
the sliver shrinks, whatever the summation order; wabs bounds what they
contribute.
"""
import numpy as np

BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
ALL = ["x", "v", "xx", "xv", "vv"]
HEAVY = 1e-3

# The FP32-clip column kernel against the FP64 oracle, as (rel, wabs):
# cell masses and means within 2e-5 relative on heavy cells (a heavy cell can
# hold a single sliver fragment of one tet, whose FP32 q-coordinates carry
# ~1e-8 of the tet's mass), mass-weighted errors within 1e-6 of the peak;
# second moments within 1e-3 and 3e-5.
TOL32 = {"m": (2e-5, 1e-6), "x": (2e-5, 1e-6), "v": (2e-5, 1e-6), "xx": (1e-3, 3e-5), "xv": (1e-3, 3e-5),
         "vv": (1e-3, 3e-5)}


def field_errors(a, b, mass_ref=None):
    """Errors of a against the reference b; returns {field: (rel, wabs)}."""
    m = b["m"] if mass_ref is None else mass_ref
    mmax = float(np.abs(m).max()) if m.size else 0.0
    heavy = m >= HEAVY * mmax if mmax > 0 else np.zeros(m.shape, bool)
    out = {}
    for f in b:
        if f not in a:
            continue
        A, B = np.asarray(a[f]), np.asarray(b[f])
        if f == "m":
            d = np.abs(A - B)
            rel = float((d[heavy] / B[heavy]).max()) if heavy.any() else 0.0
            wabs = float(d.max() / mmax) if mmax > 0 else float(d.max())
        elif f in ("x", "v"):
            d = np.abs(A - B)
            if heavy.any():
                rms = float(np.sqrt(np.mean(B[heavy] ** 2)))
                rel = float((d[heavy] / (np.abs(B[heavy]) + rms + 1e-300)).max())
            else:
                rel = 0.0
            W = m[..., None]
            den = float(np.abs(W * B).max()) or 1.0
            wabs = float(np.abs(W * (A - B)).max() / den)
        else:
            d = np.abs(A - B)
            if f == "xv":
                di = np.abs(np.diagonal(b["xx"], axis1=-2, axis2=-1)) if "xx" in b else np.abs(B).max(-1)
                dj = np.abs(np.diagonal(b["vv"], axis1=-2, axis2=-1)) if "vv" in b else np.abs(B).max(-2)
            else:
                di = dj = np.abs(np.diagonal(B, axis1=-2, axis2=-1))
            scale = np.sqrt(di[..., :, None] * dj[..., None, :]) + 1e-300
            if f == "xv":
                scale = scale + np.abs(B)
            rel = float((d / scale)[heavy].max()) if heavy.any() else 0.0
            W = m[..., None, None]
            den = float(np.abs(W * B).max()) or 1.0
            wabs = float(np.abs(W * (A - B)).max() / den)
        out[f] = (rel, wabs)
    return out


def assert_close(a, b, tol, what=""):
    """tol: {field: (rel_tol, wabs_tol)}"""
    errs = field_errors(a, b)
    for f, (rel, wabs) in errs.items():
        rt, wt = tol[f]
        assert rel <= rt and wabs <= wt, f"{what} field {f}: rel {rel:.3e} (tol {rt:.0e}), " \
                                         f"wabs {wabs:.3e} (tol {wt:.0e}); all: {errs}"
    return errs

# The FP64 kernel against the FP64 oracle (same measures): masses within
# 1e-12 relative, means within 1e-10, covariances within 1e-7 of their
# diagonal scale on heavy cells; mass-weighted errors within 1e-13 (1e-12 for
# the covariances) of the peak.
TOL64 = {"m": (1e-12, 1e-13), "x": (1e-10, 1e-13), "v": (1e-10, 1e-13), "xx": (1e-7, 1e-12), "xv": (1e-7, 1e-12),
         "vv": (1e-7, 1e-12)}

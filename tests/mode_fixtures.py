"""The reference's outputs for mode='annihilation', PSI.VDF and the R*-tree
query (tests/golden/{ann,vdf,tree}_*.npz, written by tests/golden/
make_golden.py from the reference's own C core), and the comparisons the
tests apply to them.

Bars: annihilation cells holding >= 1e-3 of the peak within 1e-8 relative,
every cell within 1e-9 of the peak, the total within 1e-10 (caustic cells
are ill-conditioned: rho = mass/volume of thin tets); VDF streams equal in
number per point, densities within 1e-11 relative, velocities within 1e-11
of the largest speed, in any order (the reference's follows its tree);
tree queries the same element sets.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def ann_cases():
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith("ann_") and f.endswith(".npz"))


def check_annihilation(got_m, exp_m, what):
    pk = np.abs(exp_m).max()
    heavy = exp_m >= 1e-3 * pk
    d = np.abs(got_m - exp_m)
    rel = (d[heavy] / exp_m[heavy]).max()
    assert rel <= 1e-8, f"{what}: heavy rel {rel:.2e}"
    assert d.max() <= 1e-9 * pk, f"{what}: abs {d.max() / pk:.2e} of peak"
    assert abs(got_m.sum() / exp_m.sum() - 1) <= 1e-10, what
    return rel, d.max() / pk


def vdf_expected(z):
    """per point: (rho, vel) of the reference"""
    off = np.concatenate([[0], np.cumsum(z["count"])])
    return [(z["rho"][off[i]:off[i + 1]], z["vel_out"][off[i]:off[i + 1]]) for i in range(len(z["count"]))]


def check_streams(got, exp, what):
    (r1, v1), (r2, v2) = got, exp
    assert len(r1) == len(r2), f"{what}: {len(r1)} vs {len(r2)} streams"
    if len(r1) == 0:
        return
    o1, o2 = np.lexsort((v1[:, 0], r1)), np.lexsort((v2[:, 0], r2))
    np.testing.assert_allclose(r1[o1], r2[o2], rtol=1e-11, atol=0, err_msg=what)
    assert np.abs(v1[o1] - v2[o2]).max() <= 1e-11 * max(np.abs(v2).max(), 1e-300), what


def tree_expected(z):
    off = np.concatenate([[0], np.cumsum(z["count"])])
    return [z["elems"][off[i]:off[i + 1]] for i in range(len(z["count"]))]

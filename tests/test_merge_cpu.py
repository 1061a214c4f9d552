"""The end-of-call merge of new deposits into a grid's present contents
(cell_merge_finalize in pypsi_amd/csrc/psivox_moments.h, built for the host
by tests/hostsim), used when large host fields travel to the device beside
the kernels.

Merging the raw moments of a second call's result into the first call's
fields equals the reference's in-place merge of the two calls (the online
update of src/geometry.c:164-186, one fragment at a time, computed by the
oracle) to the FP64 kernel's bar; cells the second call leaves without mass
keep their contents bit for bit, as the reference only updates cells a
fragment lands in.
"""
import numpy as np
import pytest

from helpers import ALL, BOX, assert_close

TOL = {"m": (1e-12, 1e-13), "x": (1e-10, 1e-13), "v": (1e-10, 1e-13), "xx": (1e-7, 1e-12),
       "xv": (1e-7, 1e-12), "vv": (1e-7, 1e-12)}


@pytest.mark.parametrize("fields", [ALL, ["v", "vv"], []])
def test_merge_at_end_equals_reference_merge(oracle_mod, fields):
    import hostsim_api as H
    from pypsi_amd.synth import zeldovich_block
    O = oracle_mod
    pos1, vel1 = zeldovich_block(8, 1.0, seed=1)
    pos2, vel2 = zeldovich_block(6, 0.5, seed=2)
    pos2 = np.ascontiguousarray(0.3 + 0.25 * pos2)
    P1 = O.Problem.block(pos1, vel1, periodic_box=BOX, n=(16,) * 3, bounds=BOX, fields=fields)
    P2 = O.Problem.block(pos2, vel2, periodic_box=None, n=(16,) * 3, bounds=BOX, fields=fields)
    first, _ = O.voxels(P1)
    second, _ = O.voxels(P2)
    ref, _ = O.voxels(P2, out={k: v.copy() for k, v in first.items()})
    got = H.cells_merge_finalize(P1, {k: v.copy() for k, v in first.items()}, H.cells_to_raw(P1, second))
    assert_close(got, ref, TOL)
    untouched = second["m"] == 0
    assert untouched.any() and (~untouched).any()
    for k in got:
        np.testing.assert_array_equal(got[k][untouched], first[k][untouched], err_msg=k)
    # into an empty grid the merge is the plain finalize of the new moments
    empty = {k: np.zeros_like(v) for k, v in first.items()}
    got = H.cells_merge_finalize(P1, empty, H.cells_to_raw(P1, second))
    assert_close(got, second, TOL)

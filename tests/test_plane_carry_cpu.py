"""The column path's plane-point carry (PlaneCarry in pypsi_amd/csrc/psivox_fast.h)
and its one-pass treatment of two-cell columns (slab_moments_2), built for the
host by tests/hostsim."""
import numpy as np

from helpers import ALL, BOX


def test_plane_carry_is_bit_identical(oracle_mod):
    """Each cell's lower-plane point is the previous cell's upper-plane point,
    carried instead of searched again: the host build of the column path
    gives bit-identical fields and counts either way (periodic, open with a
    window, refined)."""
    import ctypes as C
    import hostsim_api as H
    from pypsi_amd.synth import zeldovich_block
    O = oracle_mod
    cases = [(12, 2.5, (24, 20, 28), ALL, BOX, True, 0), (10, 1.0, (16, 12, 18), ["v", "vv"],
             ((0.2, 0.15, 0.3), (0.8, 0.7, 0.9)), False, 0), (8, 3.0, (32,) * 3, ["x", "xx"], BOX, True, 2)]
    try:
        for n, rms, gn, fields, bounds, periodic, lvl in cases:
            pos, vel = zeldovich_block(n, rms, seed=7)
            P = O.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=gn, bounds=bounds, fields=fields)
            H.lib().hs_set_plane_carry(C.c_int(1))
            a, sa = H.run(P, "fast", reftol=0.2, maxlvl=lvl)
            H.lib().hs_set_plane_carry(C.c_int(0))
            b, sb = H.run(P, "fast", reftol=0.2, maxlvl=lvl)
            np.testing.assert_array_equal(sa, sb)
            for k in a:
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    finally:
        H.lib().hs_set_plane_carry(C.c_int(1))


def test_two_cell_pass_matches_cell_by_cell(oracle_mod):
    """A column of two cells takes both in one pass over its fan triangles
    (slab_moments_2): each crossing triangle is cut once, into the same
    pieces, from the same crossing points, as the cell-by-cell passes cut it
    (tri_cross).  The two ways differ only where the compiler fuses a sum's
    multiply-adds differently (an ulp of a fragment's moments), which the
    bars here bound far below the FP32 parity bars; fragment counts agree,
    and both agree with the oracle within those bars."""
    import ctypes as C
    import hostsim_api as H
    from helpers import TOL32, assert_close, field_errors
    from pypsi_amd.synth import zeldovich_block
    O = oracle_mod
    cases = [(12, 2.5, (24, 20, 28), ALL, BOX, True, 0), (10, 1.0, (16, 12, 18), ["v", "vv"],
             ((0.2, 0.15, 0.3), (0.8, 0.7, 0.9)), False, 0), (8, 3.0, (32,) * 3, ["x", "xx"], BOX, True, 2)]
    try:
        for n, rms, gn, fields, bounds, periodic, lvl in cases:
            pos, vel = zeldovich_block(n, rms, seed=7)
            P = O.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=gn, bounds=bounds, fields=fields)
            ref, _ = O.voxels(P, reftol=0.2, maxlvl=lvl)
            H.lib().hs_set_pair_cells(C.c_int(1))
            a, sa = H.run(P, "fast", reftol=0.2, maxlvl=lvl)
            H.lib().hs_set_pair_cells(C.c_int(0))
            b, sb = H.run(P, "fast", reftol=0.2, maxlvl=lvl)
            assert abs(int(sa[0]) - int(sb[0])) <= 2 and sa[3] == 0
            for f, (rel, wabs) in field_errors(a, b).items():
                assert rel <= TOL32[f][0] / 10 and wabs <= TOL32[f][1] / 10, (f, rel, wabs)
            assert_close(a, ref, TOL32, f"two-cell pass {n}->{gn}")
    finally:
        H.lib().hs_set_pair_cells(C.c_int(1))

"""The column path's plane-point carry (PlaneCarry in pypsi_amd/csrc/psivox_fast.h),
built for the host by tests/hostsim."""
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

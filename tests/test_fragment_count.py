"""The fragment count (the metric's numerator) pinned by exact integer
arithmetic (tests/fragcount.py) on 2^3-4^3 lattices of random tets: the FP64
oracle's count of leaves with positive zeroth moment equals the exact count of
(tet copy, cell) pairs with positive-volume overlap.  On the GPU (the -m gpu
test at the end) the FP64 kernel is held to the exact count, the FP32 column
kernels to within two fragments (slivers whose FP32 zeroth moment rounds
across zero), with the mass still conserved to 1e-12."""
import time

import numpy as np
import pytest

from fragcount import count_fragments, kuhn_tets
from helpers import BOX

CASES = [(3, 0.05, 4, True), (4, 0.1, 8, True), (4, 0.15, 16, True), (2, 0.2, 8, False), (4, 0.25, 4, False),
         (3, 0.4, 16, False)]


def problem(O, side, rms, gn, periodic, seed):
    pos, conn = kuhn_tets(side, rms, seed, periodic)
    vel = np.zeros_like(pos)
    return pos, conn, O.Problem(pos, vel, np.ones(conn.shape[0]), conn, periodic_box=BOX if periodic else None,
                                n=(gn,) * 3, bounds=BOX)


def test_exact_count_small_cases(oracle_mod):
    """Hand-counted cases.  The corner tet {a, b, c >= 0, a + b + c <= 0.1}
    at (0.1, 0.1, 0.1) lies in one cell of 4^3 and of 2^3; on 16^3 its local
    cell boundaries fall at 0.025 and 0.0875, and the cells whose lower
    corner has a local coordinate sum below 0.1 number 1 + 3 + 3 + 1 + 3 =
    11.  A tet stretched to x = 0.4 crosses one plane; one with a face on a
    cell plane does not reach the next cell; a periodic tet across x = 1
    falls in both end cells.  The oracle agrees with each."""
    pos = np.array([[0.1, 0.1, 0.1], [0.2, 0.1, 0.1], [0.1, 0.2, 0.1], [0.1, 0.1, 0.2]])
    assert count_fragments(pos, np.array([[0, 1, 2, 3]]), (4, 4, 4), False) == 1
    assert count_fragments(pos, np.array([[0, 1, 2, 3]]), (2, 2, 2), False) == 1
    assert count_fragments(pos, np.array([[0, 1, 2, 3]]), (16, 16, 16), False) == 11
    # a tet spanning 2 cells in x only: x in [0.1, 0.4], cells of 0.25
    pos2 = pos.copy()
    pos2[1, 0] = 0.4
    assert count_fragments(pos2, np.array([[0, 1, 2, 3]]), (4, 4, 4), False) == 2
    # touching a plane is not overlapping: a tet with a face on x = 0.25
    pos3 = np.array([[0.25, 0.1, 0.1], [0.25, 0.2, 0.1], [0.25, 0.1, 0.2], [0.4, 0.1, 0.1]])
    assert count_fragments(pos3, np.array([[0, 1, 2, 3]]), (4, 4, 4), False) == 1
    # a periodic tet across the x = 1 face counts in both end cells
    pos4 = np.array([[0.95, 0.1, 0.1], [0.05, 0.1, 0.1], [0.95, 0.2, 0.1], [0.95, 0.1, 0.2]])
    assert count_fragments(pos4, np.array([[0, 1, 2, 3]]), (4, 4, 4), True) == 2
    for p, n, per, want in ((pos, 16, False, 11), (pos2, 4, False, 2), (pos3, 4, False, 1), (pos4, 4, True, 2)):
        P = oracle_mod.Problem(p, np.zeros_like(p), np.ones(1), np.array([[0, 1, 2, 3]], dtype=np.int32),
                               periodic_box=BOX if per else None, n=(n,) * 3, bounds=BOX)
        assert oracle_mod.voxels(P)[1]["nfrag"] == want


@pytest.mark.parametrize("side,rms,gn,periodic", CASES)
def test_oracle_count_is_exact(oracle_mod, side, rms, gn, periodic):
    pos, conn, P = problem(oracle_mod, side, rms, gn, periodic, seed=side * 10 + gn)
    t = time.time()
    exact = count_fragments(pos, conn, (gn,) * 3, periodic)
    out, st = oracle_mod.voxels(P)
    print(f"{side}^3 rms {rms} -> {gn}^3 periodic={periodic}: exact {exact}, oracle {st['nfrag']} "
          f"({time.time() - t:.1f} s)")
    assert st["nfrag"] == exact
    assert abs(out["m"].sum() / conn.shape[0] - 1.0) < 1e-12 if periodic else True


@pytest.mark.gpu
@pytest.mark.parametrize("side,rms,gn,periodic", CASES)
def test_gpu_count_is_exact(oracle_mod, side, rms, gn, periodic):
    import torch
    from pypsi_amd import device as D
    pos, conn, P = problem(oracle_mod, side, rms, gn, periodic, seed=side * 10 + gn)
    exact = count_fragments(pos, conn, (gn,) * 3, periodic)
    dev = torch.device("cuda:0")
    for kernel in ("plain", "fast", "auto"):
        print(f"case: fragment count {side}^3 -> {gn}^3 periodic={periodic} kernel={kernel}", flush=True)
        tf = {"m": torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)}
        mesh = D.make_mesh(torch.from_numpy(P.pos).to(dev), torch.from_numpy(P.vel).to(dev),
                           torch.from_numpy(P.mass).to(dev), torch.from_numpy(P.conn).to(dev),
                           BOX if periodic else None)
        st = D.voxels(mesh, D.make_grid((gn,) * 3, BOX, tf), D.make_opts(kernel=kernel))
        print(kernel, "exact", exact, "gpu", st["nfrag"], flush=True)
        if kernel == "plain":
            assert st["nfrag"] == exact
        else:
            # FP32 clips: a sliver whose zeroth moment is within rounding of
            # zero can land on either side of the m0 > 0 test
            assert abs(st["nfrag"] - exact) <= 2, kernel
        assert abs(float(tf["m"].sum()) / conn.shape[0] - 1.0) < 1e-12 if periodic else True

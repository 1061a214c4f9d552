"""The kernel algorithms of libpsivox, compiled for the host (tests/hostsim),
against the oracle on the CPU: the FP64 sweep (bar: the GPU plain kernel's)
and the FP32 cell-local path (bar: the GPU fast kernel's).  Catches logic
errors before any GPU run; the GPU tests check the same bars on the device."""
import numpy as np
import pytest

import hostsim_api as H
from helpers import ALL, BOX, assert_close, field_errors, TOL32

TOL64 = {"m": (1e-12, 1e-13), "x": (1e-10, 1e-13), "v": (1e-10, 1e-13), "xx": (1e-7, 1e-12), "xv": (1e-7, 1e-12),
         "vv": (1e-7, 1e-12)}


def zel(n, rms, seed):
    from pypsi_amd.synth import zeldovich_block
    return zeldovich_block(n, rms, seed)


@pytest.mark.parametrize("kernel", ["plain", "fast"])
@pytest.mark.parametrize("n,rms,gn,fields,lvl", [
    (8, 1.0, 16, ALL, 0),
    (8, 2.0, 16, ALL, 2),
    (12, 3.0, 20, ["v", "vv"], 1),
    (16, 0.3, 32, [], 0),
])
def test_block_meshes(oracle_mod, kernel, n, rms, gn, fields, lvl):
    pos, vel = zel(n, rms, 1)
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX, n=(gn,) * 3, bounds=BOX, fields=fields)
    ref, st = oracle_mod.voxels(P, reftol=0.1, maxlvl=lvl)
    got, hs = H.run(P, kernel, reftol=0.1, maxlvl=lvl)
    assert hs[3] == 0  # no failures / overflow
    assert_close(got, ref, TOL64 if kernel == "plain" else TOL32, f"{kernel} {n}->{gn}")
    assert abs(got["m"].sum() / P.mass.sum() - 1) < 1e-12
    if kernel == "plain":
        assert hs[0] == st["nfrag"]
    else:
        assert abs(hs[0] - st["nfrag"]) <= 1e-4 * st["nfrag"] + 2


@pytest.mark.parametrize("kernel", ["plain", "fast"])
def test_lattice_known_answer(oracle_mod, kernel):
    """Vertices exactly on grid planes (every tet vertex on three planes)."""
    for n in (4, 16):
        g = (np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1) + 0.5) / n
        P = oracle_mod.Problem.block(np.ascontiguousarray(g), None, periodic_box=BOX, n=(2 * n,) * 3)
        got, hs = H.run(P, kernel)
        np.testing.assert_allclose(got["m"], 0.125, rtol=1e-13 if kernel == "plain" else 1e-6, atol=0)


@pytest.mark.parametrize("kernel", ["plain", "fast"])
def test_windows(oracle_mod, kernel):
    pos, vel = zel(10, 1.5, 4)
    for per, bounds in ((False, ((0.2, 0.15, 0.3), (0.8, 0.7, 0.9))), (True, ((-0.1, 0.3, 0.55), (0.4, 0.8, 1.05)))):
        P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if per else None, n=(10, 9, 11), bounds=bounds,
                                     fields=["x", "v", "xx", "vv"])
        ref, _ = oracle_mod.voxels(P)
        got, hs = H.run(P, kernel)
        assert_close(got, ref, TOL64 if kernel == "plain" else TOL32, f"{kernel} window periodic={per}")
        assert abs(got["m"].sum() / ref["m"].sum() - 1) < 1e-12


@pytest.mark.parametrize("kernel", ["plain", "fast"])
def test_random_tets_and_quadratic(oracle_mod, kernel):
    rng = np.random.default_rng(42)
    pos = rng.uniform(0.05, 0.95, size=(40, 3))
    vel = rng.normal(size=(40, 3))
    conn = rng.integers(0, 40, size=(30, 4)).astype(np.int32)
    mass = rng.uniform(0.5, 2.0, size=30)
    P = oracle_mod.Problem(pos, vel, mass, conn, n=(12, 10, 14), bounds=BOX, fields=ALL)
    ref, _ = oracle_mod.voxels(P)
    got, _ = H.run(P, kernel)
    assert_close(got, ref, TOL64 if kernel == "plain" else TOL32, f"{kernel} random tets")
    g = np.stack(np.meshgrid(*(np.linspace(0, 1, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    qpos = 0.5 + (g - 0.5) * (0.2 + 0.4 * rng.uniform(size=g.shape))
    qconn = np.arange(27, dtype=np.int32).reshape(1, 27)
    P = oracle_mod.Problem(qpos, None, None, qconn, n=(32,) * 3, bounds=BOX, fields=["x", "v", "xx"])
    ref, _ = oracle_mod.voxels(P, reftol=0.05, maxlvl=3)
    got, _ = H.run(P, kernel, reftol=0.05, maxlvl=3)
    assert_close(got, ref, TOL64 if kernel == "plain" else TOL32, f"{kernel} quadratic")


def test_merge_existing_contents(oracle_mod):
    """Raw-moment conversion of existing fields reproduces the online merge."""
    pos1, vel1 = zel(6, 1.0, 1)
    pos2, vel2 = zel(6, 2.0, 2)
    P1 = oracle_mod.Problem.block(pos1, vel1, periodic_box=BOX, n=(12,) * 3, bounds=BOX, fields=ALL)
    P2 = oracle_mod.Problem.block(pos2, vel2, periodic_box=BOX, n=(12,) * 3, bounds=BOX, fields=ALL)
    ref, _ = oracle_mod.voxels(P1)
    ref, _ = oracle_mod.voxels(P2, out=ref)
    got, _ = H.run(P1, "plain")
    got, _ = H.run(P2, "plain", out=got)
    assert_close(got, ref, TOL64, "merge")


def test_wide_tets_routed_to_fp64_columns(oracle_mod):
    """Elements spanning 8 cells per edge (6^3 lattice onto 48^3): every tet
    is wider than the FP32 column path's 4-cell limit.  In FP32 their
    sliver fragments break the second-moment bar (about 3e-3 here); the
    automatic choice gives them to the FP64 column kernel, which meets the
    FP64 bars.  An open mesh of wrapped positions (elements spanning the
    whole grid) and a mixed mesh hold the FP32 bar the same way."""
    L = H.lib()
    try:
        pos, vel = zel(6, 1.5, 0)
        P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX, n=(48,) * 3, bounds=BOX, fields=["x", "v", "xx", "vv"])
        ref, st = oracle_mod.voxels(P)
        L.hs_set_route_wide(0)
        fp32, _ = H.run(P, "fast")
        e32 = field_errors(fp32, ref)
        assert e32["vv"][0] > TOL32["vv"][0]  # the regime the routing exists for
        L.hs_set_route_wide(1)
        got, hst = H.run(P, "fast")
        assert hst[0] == st["nfrag"]
        assert_close(got, ref, TOL64, "wide tets in FP64 columns")
        for n, gn, rms, per in ((8, 16, 2.0, False), (10, 30, 3.0, True)):
            pos, vel = zel(n, rms, n)
            P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if per else None, n=(gn,) * 3, bounds=BOX,
                                         fields=ALL)
            ref, _ = oracle_mod.voxels(P)
            got, _ = H.run(P, "fast")
            assert_close(got, ref, TOL32, f"routed {n}^3 -> {gn}^3 periodic={per}")
    finally:
        L.hs_set_route_wide(0)


@pytest.mark.parametrize("fp64", [False, True])
def test_y_table_matches_generic_clips(oracle_mod, fp64):
    """The column's two y clips through the y-slab table of the x-slab
    polyhedra (build_column) give the generic clips' polyhedron bit for bit
    (live slots, order, coordinates and link words) on every column task of
    a periodic block (levels 0 and 2) and of periodic and open meshes in a
    zoom window, in FP32 and FP64."""
    import hostsim_api as H
    from pypsi_amd.synth import zeldovich_block
    O = oracle_mod
    cases = []
    for n, rms, gn, lvl in ((16, 2.0, 32, 0), (12, 1.0, 24, 2), (8, 3.0, 40, 0)):
        pos, vel = zeldovich_block(n, rms, seed=n)
        cases.append((O.Problem.block(pos, vel, periodic_box=BOX, n=(gn,) * 3, bounds=BOX, fields=["v"]), lvl))
    win = ((0.13, 0.07, 0.21), (0.81, 0.93, 0.77))
    pos, vel = zeldovich_block(10, 2.0, seed=4)
    cases.append((O.Problem.block(pos, vel, periodic_box=BOX, n=(23, 17, 29), bounds=win, fields=["v"]), 1))
    g = (np.stack(np.meshgrid(*(np.arange(10),) * 3, indexing="ij"), -1) + 0.5) / 10
    cases.append((O.Problem.block(np.ascontiguousarray(g + vel), vel, n=(23, 17, 29), bounds=win, fields=["v"]), 1))
    for P, lvl in cases:
        ntask, ntable, bad = H.ytable_check(P, 0.2, lvl, fp64)
        print(P.n, lvl, "tasks", ntask, "y-clipped", ntable, "mismatches", bad)
        assert ntask > 0 and ntable > 0.5 * ntask
        assert bad == 0


def test_sort_key_cell_count(oracle_mod):
    """The ordering's key takes a column task's cell count from column_zext
    (the z-extremes of tet and column without building the column): on a
    periodic Zel'dovich mesh and on an open one in a window it is never
    below the built column's cell count and equal to it for at least 99% of
    the tasks (a vertex on a cell plane can add one)."""
    cases = [(16, 2.0, (32,) * 3, BOX, True), (12, 1.0, (24, 20, 28), ((0.1, 0.15, 0.2), (0.8, 0.9, 0.85)), False)]
    for n, rms, gn, bounds, periodic in cases:
        pos, vel = zel(n, rms, 3)
        P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=gn, bounds=bounds,
                                     fields=[])
        tasks, equal, below, above = H.zext_check(P)
        assert tasks > 10000
        assert below == 0, (tasks, equal, below, above)
        assert equal >= 0.99 * tasks, (tasks, equal, below, above)

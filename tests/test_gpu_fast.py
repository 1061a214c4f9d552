"""GPU parity of the fast kernel (FP32 cell-local clip, FP64 deposit) against
the FP64 oracle.

Tolerances (FP32 clipping; measures of tests/helpers.py): on cells holding
>= 1e-3 of the peak mass, mass within 2e-5 relative, means within 2e-5,
covariances within 1e-3 of their diagonal scale; mass-weighted absolute
errors within 1e-6 (m, x, v) and 3e-5 (second moments) of the field's peak;
total mass conserved to 1e-12 relative; fragment count within 1e-5 relative
of the oracle's (fragments of FP32 mass <= 0 are dropped as the reference
drops m0 <= 0, so slivers at the FP32 noise floor may differ).
"""
import numpy as np
import pytest

from helpers import ALL, BOX, assert_close, TOL32

pytestmark = pytest.mark.gpu



@pytest.fixture(scope="module")
def PSI():
    from pypsi_amd import PSI as mod
    return mod


def check(grid, ref, what):
    errs = assert_close(grid.fields, ref, TOL32, what)
    print(what, "errors (rel heavy / mass-weighted abs):", {k: "%.1e/%.1e" % v for k, v in errs.items()}, flush=True)


@pytest.mark.parametrize("n,rms,gn,fields,lvl", [
    (32, 0.3, 64, [], 0),
    (12, 1.0, 24, ALL, 0),
    (12, 2.5, 20, ["v", "vv"], 0),
    (8, 2.0, 16, ALL, 2),
])
def test_fast_vs_oracle(PSI, oracle_mod, n, rms, gn, fields, lvl):
    from pypsi_amd.synth import zeldovich_block
    print(f"case: fast {n}^3 rms{rms} -> {gn}^3 {fields} L{lvl}", flush=True)
    pos, vel = zeldovich_block(n, rms, seed=3)
    grid = PSI.Grid(n=(gn,) * 3, bounds=BOX, fields=fields)
    mesh = PSI.Mesh(loader="block", pos=pos, vel=vel, periodic_box=BOX)
    PSI.voxels(grid=grid, mesh=mesh, refine_tolerance=0.2, refine_max_lvl=lvl, kernel="fast")
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX, n=(gn,) * 3, bounds=BOX, fields=fields)
    ref, st = oracle_mod.voxels(P, reftol=0.2, maxlvl=lvl)
    check(grid, ref, f"{n}^3->{gn}^3")
    assert abs(1 - grid.fields["m"].sum() / mesh.mass.sum()) < 1e-12


def test_fast_window_open_mesh(PSI, oracle_mod):
    from pypsi_amd.synth import zeldovich_block
    print("case: fast window open mesh", flush=True)
    pos, vel = zeldovich_block(10, 1.5, seed=4)
    bounds = ((0.2, 0.15, 0.3), (0.8, 0.7, 0.9))
    grid = PSI.Grid(n=(16, 12, 18), bounds=bounds, fields=["v", "vv"])
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=pos, vel=vel), kernel="fast")
    P = oracle_mod.Problem.block(pos, vel, n=(16, 12, 18), bounds=bounds, fields=["v", "vv"])
    ref, _ = oracle_mod.voxels(P)
    check(grid, ref, "window")
    # mass inside the window: exact FP64 target per tet
    assert abs(grid.fields["m"].sum() / ref["m"].sum() - 1) < 1e-12


def test_fast_periodic_zoom(PSI, oracle_mod):
    """Periodic mesh onto a window smaller than the box (targets computed in FP64)."""
    from pypsi_amd.synth import zeldovich_block
    print("case: fast periodic zoom", flush=True)
    pos, vel = zeldovich_block(10, 2.0, seed=5)
    bounds = ((-0.1, 0.3, 0.55), (0.4, 0.8, 1.05))
    grid = PSI.Grid(n=(10, 10, 10), bounds=bounds, fields=["x", "xx"])
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=pos, vel=vel, periodic_box=BOX), kernel="fast")
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX, n=(10, 10, 10), bounds=bounds, fields=["x", "xx"])
    ref, _ = oracle_mod.voxels(P)
    check(grid, ref, "zoom")
    assert abs(grid.fields["m"].sum() / ref["m"].sum() - 1) < 1e-12


def test_fast_random_tets_and_quadratic(PSI, oracle_mod):
    rng = np.random.default_rng(42)
    print("case: fast random tets + quadratic", flush=True)
    pos = rng.uniform(0.05, 0.95, size=(40, 3))
    vel = rng.normal(size=(40, 3))
    conn = rng.integers(0, 40, size=(30, 4)).astype(np.int32)
    mass = rng.uniform(0.5, 2.0, size=30)
    grid = PSI.Grid(n=(12, 10, 14), bounds=BOX, fields=ALL)
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="array", pos=pos, vel=vel, mass=mass, connectivity=conn),
               kernel="fast")
    P = oracle_mod.Problem(pos, vel, mass, conn, n=(12, 10, 14), bounds=BOX, fields=ALL)
    ref, _ = oracle_mod.voxels(P)
    check(grid, ref, "random tets")
    assert abs(grid.fields["m"].sum() / mass.sum() - 1) < 1e-12
    g = np.stack(np.meshgrid(*(np.linspace(0, 1, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    qpos = 0.5 + (g - 0.5) * (0.2 + 0.4 * rng.uniform(size=g.shape))
    qconn = np.arange(27, dtype=np.int32).reshape(1, 27)
    grid = PSI.Grid(n=(32,) * 3, bounds=BOX, fields=["x", "v", "xx"])
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="array", pos=qpos, connectivity=qconn), refine_tolerance=0.05,
               refine_max_lvl=3, kernel="fast")
    P = oracle_mod.Problem(qpos, None, None, qconn, n=(32,) * 3, bounds=BOX, fields=["x", "v", "xx"])
    ref, _ = oracle_mod.voxels(P, reftol=0.05, maxlvl=3)
    check(grid, ref, "quadratic")
    assert abs(1 - grid.fields["m"].sum()) < 1e-12


def test_fast_uniform_lattice_kat(PSI):
    """Vertices exactly on grid planes: every cell exactly 0.125."""
    print("case: fast lattice KAT", flush=True)
    n = 16
    g = (np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1) + 0.5) / n
    grid = PSI.Grid(n=(2 * n,) * 3, bounds=BOX)
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=np.ascontiguousarray(g), periodic_box=BOX),
               kernel="fast")
    np.testing.assert_allclose(grid.fields["m"], 0.125, rtol=1e-6, atol=0)
    assert abs(grid.fields["m"].sum() / n ** 3 - 1) < 1e-12


def test_fast_c1_count_and_mass(PSI, oracle_mod):
    """The 32^3 -> 64^3 workload: fragment count and mass conservation."""
    import json
    import os
    from pypsi_amd import device as D
    import torch
    from oracle import block_connectivity
    from pypsi_amd.synth import zeldovich_block
    print("case: fast c1 count", flush=True)
    pos, vel = zeldovich_block(32, 0.3, seed=0)
    conn = block_connectivity((32,) * 3, True)
    dev = torch.device("cuda:0")
    tf = {"m": torch.zeros((64,) * 3, dtype=torch.float64, device=dev)}
    st = D.voxels(D.make_mesh(torch.from_numpy(pos.reshape(-1, 3)).to(dev), torch.from_numpy(vel.reshape(-1, 3)).to(dev),
                              torch.ones(conn.shape[0], dtype=torch.float64, device=dev), torch.from_numpy(conn).to(dev),
                              BOX), D.make_grid((64,) * 3, BOX, tf), D.make_opts(kernel="fast"))
    counts = json.load(open(os.path.join(os.path.dirname(__file__), "..", "profiles", "workload_counts.json")))
    ref = counts["zeldovich-32^3-rms0.3-to-64^3-m"]["nfrag"]
    assert abs(st["nfrag"] - ref) <= 1e-5 * ref, (st["nfrag"], ref)
    assert abs(tf["m"].sum().item() / conn.shape[0] - 1) < 1e-12


def test_refined_clumps_against_oracle(PSI, oracle_mod):
    """Shell crossing with level-2 refinement (the refined workload's
    construction at 32^3 -> 128^3): lattice rms 1.5 cells plus 8 seeded
    infall clumps, tolerance 0.1 cells, fields v, vv."""
    from oracle import block_connectivity
    from pypsi_amd.synth import add_infall_clumps, zeldovich_block
    print("case: refined clumps 32^3 L2 -> 128^3", flush=True)
    n = 32
    pos0, vel0 = zeldovich_block(n, 1.5, 0)
    pos = add_infall_clumps(pos0, n, nclumps=8, seed=100, radius_cells=4.0)
    d = pos - pos0
    d -= np.round(d)
    vel = vel0 + d
    grid = PSI.Grid(n=(128,) * 3, bounds=BOX, fields=["v", "vv"])
    mesh = PSI.Mesh(loader="block", pos=pos, vel=vel, periodic_box=BOX)
    PSI.voxels(grid=grid, mesh=mesh, refine_tolerance=0.1, refine_max_lvl=2, kernel="fast")
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX, n=(128,) * 3, bounds=BOX, fields=["v", "vv"])
    ref, st = oracle_mod.voxels(P, reftol=0.1, maxlvl=2, threads=-1)
    print("tets", st["ntets"], "fragments", st["nfrag"])
    errs = assert_close(grid.fields, ref, TOL32, "refined clumps")
    print({k: "%.1e/%.1e" % v for k, v in errs.items()})
    assert abs(grid.fields["m"].sum() / mesh.mass.sum() - 1) < 1e-12

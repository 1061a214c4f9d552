"""GPU parity: PSI.voxels on the MI355X (through the C ABI) against the FP64
oracle on the same seeded inputs.

Tolerances (FP64 plain kernel), with the measures of tests/helpers.py:
fragment count identical; on cells holding >= 1e-3 of the peak mass, mass
within 1e-12 relative, means within 1e-10, covariances within 1e-7 of their
diagonal scale; mass-weighted absolute errors over all cells within 1e-13 of
the field's peak; total mass conserved to 1e-12.  The FP32-clip fast kernel is held to the stated FP32 bar
in test_gpu_fast.py.
"""
import numpy as np
import pytest

from helpers import ALL, BOX, assert_close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def PSI():
    from pypsi_amd import PSI as mod
    return mod


def run_psi(PSI, pos, vel, n, fields, kernel="plain", bounds=BOX, periodic=True, reftol=1.0, maxlvl=0, mass=None):
    grid = PSI.Grid(n=n, bounds=bounds, fields=fields)
    kw = dict(loader="block", pos=pos, vel=vel)
    if periodic:
        kw["periodic_box"] = BOX
    if mass is not None:
        kw["mass"] = mass
    mesh = PSI.Mesh(**kw)
    PSI.voxels(grid=grid, mesh=mesh, refine_tolerance=reftol, refine_max_lvl=maxlvl, kernel=kernel)
    return grid, mesh


def oracle_run(O, pos, vel, n, fields, bounds=BOX, periodic=True, reftol=1.0, maxlvl=0, mass=None):
    P = O.Problem.block(pos, vel, mass=mass, periodic_box=BOX if periodic else None, n=n, bounds=bounds,
                        fields=fields)
    return O.voxels(P, reftol=reftol, maxlvl=maxlvl)


TOL = {"m": (1e-12, 1e-13), "x": (1e-10, 1e-13), "v": (1e-10, 1e-13), "xx": (1e-7, 1e-12),
       "xv": (1e-7, 1e-12), "vv": (1e-7, 1e-12)}


def check_close(grid, ref, tol=TOL):
    errs = assert_close(grid.fields, ref, tol)
    print("errors (rel on heavy cells, mass-weighted abs):", {k: "%.1e/%.1e" % v for k, v in errs.items()})
    return errs


def test_c1_workload_plain(PSI, oracle_mod):
    """32^3 lattice, rms 0.3 cells -> 64^3, mass only (the CPU-runnable workload)."""
    from pypsi_amd.synth import zeldovich_block
    print("case: c1 32^3 rms0.3 -> 64^3 plain", flush=True)
    pos, vel = zeldovich_block(32, 0.3, seed=0)
    grid, mesh = run_psi(PSI, pos, vel, (64,) * 3, [])
    ref, st = oracle_run(oracle_mod, pos, vel, (64,) * 3, [])
    check_close(grid, ref)
    err = abs(1 - grid.fields["m"].sum() / mesh.mass.sum())
    assert err < 1e-12


@pytest.mark.parametrize("rms,maxlvl", [(1.0, 0), (2.5, 0), (2.0, 2)])
def test_all_fields_plain(PSI, oracle_mod, rms, maxlvl):
    from pypsi_amd.synth import zeldovich_block
    print(f"case: all fields 12^3 rms{rms} L{maxlvl} plain", flush=True)
    pos, vel = zeldovich_block(12, rms, seed=7)
    grid, mesh = run_psi(PSI, pos, vel, (24, 20, 28), ALL, reftol=0.2, maxlvl=maxlvl)
    ref, st = oracle_run(oracle_mod, pos, vel, (24, 20, 28), ALL, reftol=0.2, maxlvl=maxlvl)
    check_close(grid, ref)
    assert abs(1 - grid.fields["m"].sum() / mesh.mass.sum()) < 1e-12


def test_window_nonperiodic_plain(PSI, oracle_mod):
    """Open mesh onto a zoom window cutting through it (window planes)."""
    from pypsi_amd.synth import zeldovich_block
    print("case: window nonperiodic", flush=True)
    pos, vel = zeldovich_block(10, 1.0, seed=11)
    bounds = ((0.2, 0.15, 0.3), (0.8, 0.7, 0.9))
    grid, _ = run_psi(PSI, pos, vel, (16, 12, 18), ["v", "vv"], bounds=bounds, periodic=False)
    ref, _ = oracle_run(oracle_mod, pos, vel, (16, 12, 18), ["v", "vv"], bounds=bounds, periodic=False)
    check_close(grid, ref)


def test_merge_two_calls(PSI, oracle_mod):
    """voxels() merges into the grid's existing contents like the reference."""
    from pypsi_amd.synth import zeldovich_block
    print("case: merge", flush=True)
    pos1, vel1 = zeldovich_block(8, 1.0, seed=1)
    pos2, vel2 = zeldovich_block(8, 1.5, seed=2)
    grid = PSI.Grid(n=(16,) * 3, bounds=BOX, fields=ALL)
    for p, v in ((pos1, vel1), (pos2, vel2)):
        PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=p, vel=v, periodic_box=BOX), kernel="plain")
    P1 = oracle_mod.Problem.block(pos1, vel1, periodic_box=BOX, n=(16,) * 3, bounds=BOX, fields=ALL)
    P2 = oracle_mod.Problem.block(pos2, vel2, periodic_box=BOX, n=(16,) * 3, bounds=BOX, fields=ALL)
    ref, _ = oracle_mod.voxels(P1)
    ref, _ = oracle_mod.voxels(P2, out=ref)
    check_close(grid, ref)


@pytest.mark.parametrize("kernel", ["plain", "fast"])
def test_merge_large_host_fields(PSI, oracle_mod, kernel):
    """Host fields of 16 MB and more travel to the device beside the kernels
    and are merged at the end: two calls onto a 48^3 grid with all six
    fields (29 MB) give the oracle's merged result, and cells the second
    mesh does not reach keep the first call's values bit for bit (as the
    reference, which only updates cells a fragment lands in)."""
    from helpers import TOL32
    from pypsi_amd.synth import zeldovich_block
    print(f"case: merge 48^3 all fields, host arrays, kernel={kernel}", flush=True)
    G = 48
    pos1, vel1 = zeldovich_block(24, 1.0, seed=1)
    pos2, vel2 = zeldovich_block(12, 0.5, seed=2)
    pos2 = np.ascontiguousarray(0.3 + 0.25 * pos2)  # an open mesh inside [0.3, 0.55]^3
    grid = PSI.Grid(n=(G,) * 3, bounds=BOX, fields=ALL)
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=pos1, vel=vel1, periodic_box=BOX), kernel=kernel)
    first = {k: np.array(v, copy=True) for k, v in grid.fields.items()}
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=pos2, vel=vel2), kernel=kernel)
    P1 = oracle_mod.Problem.block(pos1, vel1, periodic_box=BOX, n=(G,) * 3, bounds=BOX, fields=ALL)
    P2 = oracle_mod.Problem.block(pos2, vel2, periodic_box=None, n=(G,) * 3, bounds=BOX, fields=ALL)
    ref, _ = oracle_mod.voxels(P1)
    ref, _ = oracle_mod.voxels(P2, out=ref)
    check_close(grid, ref, TOL if kernel == "plain" else TOL32)
    # cells two cells or more away from the second mesh's box
    lo, hi = pos2.reshape(-1, 3).min(0) - 2 / G, pos2.reshape(-1, 3).max(0) + 2 / G
    c = (np.arange(G) + 0.5) / G
    near = [(c >= lo[a]) & (c <= hi[a]) for a in range(3)]
    far = ~(near[0][:, None, None] & near[1][None, :, None] & near[2][None, None, :])
    assert far.any() and (~far).any()
    for k, v in grid.fields.items():
        np.testing.assert_array_equal(v[far], first[k][far], err_msg=k)


def test_device_pointer_path(oracle_mod):
    """The C ABI with device-resident torch tensors (the benchmark path)."""
    import torch
    from pypsi_amd import device as D
    from pypsi_amd.synth import zeldovich_block
    from oracle import block_connectivity
    print("case: device pointers", flush=True)
    pos, vel = zeldovich_block(16, 1.0, seed=5)
    conn = block_connectivity(pos.shape[:3], True)
    mass = np.ones(conn.shape[0])
    dev = torch.device("cuda:0")
    tpos = torch.from_numpy(pos.reshape(-1, 3)).to(dev)
    tvel = torch.from_numpy(vel.reshape(-1, 3)).to(dev)
    tm = torch.from_numpy(mass).to(dev)
    tc = torch.from_numpy(conn).to(dev)
    tf = {"m": torch.zeros((32,) * 3, dtype=torch.float64, device=dev),
          "v": torch.zeros((32,) * 3 + (3,), dtype=torch.float64, device=dev),
          "vv": torch.zeros((32,) * 3 + (3, 3), dtype=torch.float64, device=dev)}
    st = D.voxels(D.make_mesh(tpos, tvel, tm, tc, BOX), D.make_grid((32,) * 3, BOX, tf),
                  D.make_opts(kernel="plain", stream=torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    ref, ost = oracle_run(oracle_mod, pos, vel, (32,) * 3, ["v", "vv"])
    assert st["nfrag"] == ost["nfrag"]
    got = {k: v.cpu().numpy() for k, v in tf.items()}
    assert_close(got, ref, TOL)


def test_simplex_and_quadratic_elements(PSI, oracle_mod):
    """loader='array' with tets (4) and triquadratic elements (27), refined."""
    rng = np.random.default_rng(42)
    print("case: array loader tets + quadratic", flush=True)
    # random tets inside the unit box
    npart = 40
    pos = rng.uniform(0.05, 0.95, size=(npart, 3))
    vel = rng.normal(size=(npart, 3))
    conn = rng.integers(0, npart, size=(30, 4)).astype(np.int32)
    mass = rng.uniform(0.5, 2.0, size=30)
    grid = PSI.Grid(n=(12, 10, 14), bounds=BOX, fields=ALL)
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="array", pos=pos, vel=vel, mass=mass, connectivity=conn),
               kernel="plain")
    P = oracle_mod.Problem(pos, vel, mass, conn, n=(12, 10, 14), bounds=BOX, fields=ALL)
    ref, _ = oracle_mod.voxels(P)
    check_close(grid, ref)
    # one perturbed triquadratic element, refined (tutorial Ex.1 shape, seeded)
    g = np.stack(np.meshgrid(*(np.linspace(0, 1, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    qpos = 0.5 + (g - 0.5) * (0.2 + 0.4 * rng.uniform(size=g.shape))
    qconn = np.arange(27, dtype=np.int32).reshape(1, 27)
    grid = PSI.Grid(n=(32,) * 3, bounds=BOX, fields=["x", "v", "xx"])
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="array", pos=qpos, connectivity=qconn), refine_tolerance=0.05,
               refine_max_lvl=3, kernel="plain")
    P = oracle_mod.Problem(qpos, None, None, qconn, n=(32,) * 3, bounds=BOX, fields=["x", "v", "xx"])
    ref, _ = oracle_mod.voxels(P, reftol=0.05, maxlvl=3)
    check_close(grid, ref)
    assert abs(1 - grid.fields["m"].sum()) < 1e-13


def test_empty_and_outside(PSI):
    """An empty mesh and a mesh entirely outside the window deposit nothing."""
    print("case: empty/outside", flush=True)
    grid = PSI.Grid(n=(8,) * 3, bounds=((2, 2, 2), (3, 3, 3)), fields=["v"])
    pos = np.random.default_rng(0).uniform(0, 1, size=(4, 4, 4, 3))
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=pos), kernel="plain")
    assert grid.fields["m"].sum() == 0.0
    conn = np.zeros((0, 4), dtype=np.int32)
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="array", pos=np.zeros((1, 3)), connectivity=conn), kernel="plain")
    assert grid.fields["m"].sum() == 0.0


def test_uniform_lattice_kat(PSI):
    """Zero displacement, N^3 periodic block onto (2N)^3: every cell holds
    exactly mass_total / (2N)^3 (vertices sit exactly on grid planes)."""
    print("case: uniform lattice KAT", flush=True)
    n = 16
    g = (np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1) + 0.5) / n
    grid = PSI.Grid(n=(2 * n,) * 3, bounds=BOX)
    PSI.voxels(grid=grid, mesh=PSI.Mesh(loader="block", pos=np.ascontiguousarray(g), periodic_box=BOX),
               kernel="plain")
    m = grid.fields["m"]
    np.testing.assert_allclose(m, n ** 3 / (2 * n) ** 3, rtol=1e-13, atol=0)

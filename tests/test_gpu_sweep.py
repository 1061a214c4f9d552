"""A seeded sweep of 24 random configurations on the MI355X against the
oracle: lattice sizes 6..12, displacements of 0.2..3 lattice cells, grids of
up to 3 cells per element edge with unequal axes, random field sets,
periodic and open meshes (open ones unwrapped), full grids and random
windows, refinement levels 0..2 (tolerance 0.2 cells).  The FP64 kernel is
held to the FP64 bar, the FP32 column kernel and the automatic choice to the
FP32 bar (both in tests/helpers.py and test_gpu_parity.py).

Then the regime that needs the automatic choice: elements spanning 8 cells
per edge, and an open mesh made of wrapped periodic positions (elements
spanning the whole grid).  kernel='auto' (PSI.voxels' default) sends those
tets to the FP64 column kernel and holds the FP32 bar there."""
import numpy as np
import pytest

from helpers import BOX, TOL32, assert_close

pytestmark = pytest.mark.gpu

TOL = {"m": (1e-12, 1e-13), "x": (1e-10, 1e-13), "v": (1e-10, 1e-13), "xx": (1e-7, 1e-12),
       "xv": (1e-7, 1e-12), "vv": (1e-7, 1e-12)}


def config(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(6, 13))
    gn = tuple(int(x) for x in rng.integers(n, 3 * n + 1, 3))
    rms = float(rng.uniform(0.2, 3.0))
    fields = [f for f in ("x", "v") if rng.uniform() < 0.7]
    if "x" in fields and rng.uniform() < 0.6:
        fields.append("xx")
    if "v" in fields and rng.uniform() < 0.6:
        fields.append("vv")
    if "x" in fields and "v" in fields and rng.uniform() < 0.5:
        fields.append("xv")
    periodic = bool(rng.uniform() < 0.6)
    bounds = BOX
    if rng.uniform() < 0.5:
        bounds = (tuple(rng.uniform(0.0, 0.3, 3)), tuple(rng.uniform(0.7, 1.0, 3)))
    lvl = int(rng.integers(0, 3))
    return n, gn, rms, fields, periodic, bounds, lvl


def mesh_arrays(n, rms, seed, periodic):
    from pypsi_amd.synth import zeldovich_block
    pos, vel = zeldovich_block(n, rms, seed=seed)
    if not periodic:  # an open mesh: the lattice displaced without the periodic wrap
        g = (np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1) + 0.5) / n
        pos = np.ascontiguousarray(g + vel)
    return pos, vel


@pytest.fixture(scope="module")
def PSI():
    from pypsi_amd import PSI as mod
    return mod


@pytest.mark.parametrize("seed", range(24))
def test_sweep(PSI, oracle_mod, seed):
    n, gn, rms, fields, periodic, bounds, lvl = config(seed)
    print(f"case: sweep seed {seed}: {n}^3 rms{rms:.2f} -> {gn} {fields} periodic={periodic} "
          f"window={bounds != BOX} L{lvl}", flush=True)
    pos, vel = mesh_arrays(n, rms, seed, periodic)
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=gn, bounds=bounds,
                                 fields=fields)
    ref, st = oracle_mod.voxels(P, reftol=0.2, maxlvl=lvl)
    kw = dict(loader="block", pos=pos, vel=vel)
    if periodic:
        kw["periodic_box"] = BOX
    for kernel, tol in (("plain", TOL), ("fast", TOL32), ("auto", TOL32)):
        grid = PSI.Grid(n=gn, bounds=bounds, fields=fields)
        PSI.voxels(grid=grid, mesh=PSI.Mesh(**kw), refine_tolerance=0.2, refine_max_lvl=lvl, kernel=kernel)
        errs = assert_close(grid.fields, ref, tol, kernel)
        print(kernel, {k: "%.1e/%.1e" % v for k, v in errs.items()}, flush=True)
        assert abs(1 - grid.fields["m"].sum() / ref["m"].sum()) < 1e-12


@pytest.mark.parametrize("n,gn,rms,periodic", [(6, 48, 1.5, True), (8, 64, 1.0, True), (8, 16, 2.0, False),
                                               (10, 40, 3.0, False)])
def test_auto_wide_tets(PSI, oracle_mod, n, gn, rms, periodic):
    import torch
    from pypsi_amd import device as D
    from pypsi_amd.synth import zeldovich_block
    fields = ["x", "v", "xx", "xv", "vv"]
    print(f"case: wide tets {n}^3 rms{rms} -> {gn}^3 periodic={periodic} (open: wrapped positions)", flush=True)
    pos, vel = zeldovich_block(n, rms, seed=n)
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=(gn,) * 3, bounds=BOX,
                                 fields=fields)
    ref, st = oracle_mod.voxels(P)
    # the drop-in call with its default kernel
    grid = PSI.Grid(n=(gn,) * 3, bounds=BOX, fields=fields)
    kw = dict(loader="block", pos=pos, vel=vel)
    if periodic:
        kw["periodic_box"] = BOX
    PSI.voxels(grid=grid, mesh=PSI.Mesh(**kw))
    errs = assert_close(grid.fields, ref, TOL32, "auto")
    print("auto", {k: "%.1e/%.1e" % v for k, v in errs.items()}, flush=True)
    assert abs(1 - grid.fields["m"].sum() / ref["m"].sum()) < 1e-12
    # the same through the C ABI: the wide tets went to the FP64 column kernel
    dev = torch.device("cuda:0")
    tf = {f: torch.zeros(v.shape, dtype=torch.float64, device=dev) for f, v in ref.items()}
    tp, tv = torch.from_numpy(P.pos).to(dev), torch.from_numpy(P.vel).to(dev)
    mesh = D.make_mesh(tp, tv, torch.from_numpy(P.mass).to(dev), torch.from_numpy(P.conn).to(dev),
                       BOX if periodic else None)
    stats = D.voxels(mesh, D.make_grid((gn,) * 3, BOX, tf), D.make_opts(kernel="auto"))
    assert stats["nwide"] > 0
    assert abs(stats["nfrag"] - st["nfrag"]) <= 1e-5 * st["nfrag"]
    assert_close({k: v.cpu().numpy() for k, v in tf.items()}, ref, TOL32, "auto (C ABI)")

"""A seeded sweep of 24 random configurations on the MI355X against the
oracle: lattice sizes 6..12, displacements of 0.2..3 lattice cells, grids of
up to 3 cells per element edge with unequal axes, random field sets,
periodic and open meshes (open ones unwrapped), full grids and random
windows, refinement levels 0..2 (tolerance 0.2 cells).  The FP64 kernel is
held to the FP64 bar and the FP32 column kernel to its FP32 bar (both in
tests/helpers.py and test_gpu_parity.py)."""
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
    for kernel, tol in (("plain", TOL), ("fast", TOL32)):
        grid = PSI.Grid(n=gn, bounds=bounds, fields=fields)
        PSI.voxels(grid=grid, mesh=PSI.Mesh(**kw), refine_tolerance=0.2, refine_max_lvl=lvl, kernel=kernel)
        errs = assert_close(grid.fields, ref, tol, kernel)
        print(kernel, {k: "%.1e/%.1e" % v for k, v in errs.items()}, flush=True)
        assert abs(1 - grid.fields["m"].sum() / ref["m"].sum()) < 1e-12

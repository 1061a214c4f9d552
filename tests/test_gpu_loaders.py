"""A Gadget-2 snapshot through PSI.Mesh(loader='gadget2') and PSI.voxels on
the MI355X, against the FP64 oracle on the loaded arrays: the notebook's
Example 2 settings (tutorial/Tutorial.ipynb:160-186: box from the snapshot,
fields v and vv, refinement level 2 at tolerance 0.1) on a written 16^3
Zel'dovich snapshot in a 40-unit box.  Bars: the FP32 column kernel to
helpers.TOL32, the FP64 kernel to the FP64 bar of test_gpu_parity; total mass
to 1e-12 (the notebook prints 6.99e-15 on its snapshot)."""
import numpy as np
import pytest

from pypsi_amd import snapshot as S
from helpers import TOL32, assert_close
from test_gpu_parity import TOL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def snapshot_case(oracle_mod, tmp_path_factory):
    from pypsi_amd import PSI
    fn = str(tmp_path_factory.mktemp("snap") / "snapshot")
    S.zeldovich_snapshot(fn, 16, 1.5, seed=2, box=40.0, mass1=14.96)
    mesh = PSI.Mesh(loader="gadget2", filename=fn)
    wb = np.array(mesh.wrapbox)
    P = oracle_mod.Problem(mesh.pos, mesh.vel, mesh.mass, mesh.connectivity, n=(64,) * 3, bounds=wb,
                           periodic_box=wb, fields=["v", "vv"])
    ref, _ = oracle_mod.voxels(P, reftol=0.1, maxlvl=2, threads=8)
    return PSI, mesh, ref


@pytest.mark.parametrize("kernel", ["fast", "plain"])
def test_gadget2_example2_settings(snapshot_case, kernel):
    PSI, mesh, ref = snapshot_case
    print(f"case: gadget2 16^3 box 40 -> 64^3 v,vv L2 tol0.1 kernel={kernel}", flush=True)
    grid = PSI.Grid(n=(64,) * 3, bounds=mesh.wrapbox, fields=["v", "vv"])
    PSI.voxels(grid=grid, mesh=mesh, refine_max_lvl=2, refine_tolerance=0.1, kernel=kernel)
    errs = assert_close(grid.fields, ref, TOL32 if kernel == "fast" else TOL, f"gadget2 {kernel}")
    print("errors:", {k: "%.1e/%.1e" % v for k, v in errs.items()})
    assert abs(1.0 - grid.fields["m"].sum() / mesh.mass.sum()) < 1e-12

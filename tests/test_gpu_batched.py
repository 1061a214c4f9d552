"""Batched voxelization (psivox_voxels_batched): many meshes of uneven size
in one device pass, against single-mesh calls on the same inputs (same
algorithm: fragment counts equal, fields within 1e-12 relative of their
peak, the atomics' summation order aside) and against the FP64 oracle at
the FP32 column kernel's bar (tests/test_gpu_fast.py)."""
import numpy as np
import pytest

from helpers import BOX, assert_close, TOL32

pytestmark = pytest.mark.gpu



def _fields(n, fields, dev):
    import torch
    out = {"m": torch.zeros(n, dtype=torch.float64, device=dev)}
    for f in fields:
        out[f] = torch.zeros(tuple(n) + ((3,) if len(f) == 1 else (3, 3)), dtype=torch.float64, device=dev)
    return out


def _case_list():
    from oracle import block_connectivity
    from pypsi_amd.synth import zeldovich_block
    specs = [(6, 0.5, (12, 12, 12), 1, True, BOX), (10, 2.0, (20, 16, 24), 2, True, BOX),
             (4, 1.0, (8, 8, 8), 3, True, BOX), (12, 1.5, (16, 16, 16), 4, False, ((0.1, 0.2, 0.15), (0.8, 0.9, 0.85))),
             (8, 3.0, (32, 32, 32), 5, True, BOX), (3, 0.2, (4, 6, 5), 6, True, BOX)]
    out = []
    for n, rms, gn, seed, per, bounds in specs:
        pos, vel = zeldovich_block(n, rms, seed=seed)
        conn = block_connectivity((n,) * 3, per)
        out.append((pos.reshape(-1, 3), vel.reshape(-1, 3), conn, gn, per, bounds))
    return out


def _descs(cases, fields):
    import torch
    from pypsi_amd import device as D
    dev = torch.device("cuda:0")
    meshes, grids, outs = [], [], []
    for pos, vel, conn, gn, per, bounds in cases:
        tp, tv = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
        tc = torch.from_numpy(conn).to(dev)
        tm = torch.ones(conn.shape[0], dtype=torch.float64, device=dev)
        meshes.append(D.make_mesh(tp, tv, tm, tc, BOX if per else None))
        tf = _fields(gn, fields, dev)
        grids.append(D.make_grid(gn, bounds, tf))
        outs.append(tf)
    return meshes, grids, outs


@pytest.mark.parametrize("fields", [[], ["v", "vv"], ["x", "v", "xx", "xv", "vv"]])
def test_batched_equals_single_calls(fields):
    import torch
    from pypsi_amd import device as D
    print("case: batched vs single", fields, flush=True)
    cases = _case_list()
    s = torch.cuda.current_stream().cuda_stream
    mb, gb, ob = _descs(cases, fields)
    stb = D.voxels_batched(mb, gb, D.make_opts(stream=s))
    ms, gs, os_ = _descs(cases, fields)
    sts = [D.voxels(m, g, D.make_opts(stream=s)) for m, g in zip(ms, gs)]
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(ob, os_)):
        assert stb[i]["nfrag"] == sts[i]["nfrag"], i
        for f in a:
            x, y = a[f].cpu().numpy(), b[f].cpu().numpy()
            assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), (i, f)


def test_batched_against_oracle(oracle_mod):
    import torch
    from pypsi_amd import device as D
    print("case: batched vs oracle", flush=True)
    cases = _case_list()
    fields = ["v", "vv"]
    mb, gb, ob = _descs(cases, fields)
    D.voxels_batched(mb, gb, D.make_opts(stream=torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    for i, (pos, vel, conn, gn, per, bounds) in enumerate(cases):
        P = oracle_mod.Problem(pos, vel, None, conn, periodic_box=BOX if per else None, n=gn, bounds=bounds,
                               fields=fields)
        ref, _ = oracle_mod.voxels(P)
        got = {k: v.cpu().numpy() for k, v in ob[i].items()}
        assert_close(got, ref, TOL32, what=f"mesh {i}")


def test_batched_mixed_fields_runs_mesh_by_mesh():
    """Grids asking for different fields cannot share a pass: the call runs
    them one by one with the same results."""
    import torch
    from pypsi_amd import device as D
    print("case: batched mixed fields", flush=True)
    cases = _case_list()[:3]
    dev = torch.device("cuda:0")
    s = torch.cuda.current_stream().cuda_stream
    mb, gb, ob = _descs(cases, [])
    ob[1] = _fields(cases[1][3], ["v"], dev)
    gb[1] = D.make_grid(cases[1][3], cases[1][5], ob[1])
    D.voxels_batched(mb, gb, D.make_opts(stream=s))
    ms, gs, os_ = _descs(cases, [])
    os_[1] = _fields(cases[1][3], ["v"], dev)
    gs[1] = D.make_grid(cases[1][3], cases[1][5], os_[1])
    for m, g in zip(ms, gs):
        D.voxels(m, g, D.make_opts(stream=s))
    torch.cuda.synchronize()
    for a, b in zip(ob, os_):
        for f in a:
            x, y = a[f].cpu().numpy(), b[f].cpu().numpy()
            assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300)


def test_periodic_simplex_mesh_fast(oracle_mod):
    """Array-loader tets in a periodic box on the FP32 path (mass targets of
    simplex records)."""
    from pypsi_amd import PSI
    print("case: periodic simplex fast", flush=True)
    rng = np.random.default_rng(9)
    pos = rng.uniform(0.0, 1.0, size=(60, 3))
    conn = np.array([rng.choice(60, 4, replace=False) for _ in range(80)], dtype=np.int32)
    # keep each tet compact so the periodic wrap keeps it whole
    pos[conn[:, 1:]] = (pos[conn[:, :1]] + rng.uniform(-0.1, 0.1, size=(80, 3, 3))) % 1.0
    grid = PSI.Grid(n=(16, 16, 16), bounds=BOX, fields=["v"])
    mesh = PSI.Mesh(loader="array", pos=pos, vel=pos.copy(), connectivity=conn, periodic_box=BOX)
    PSI.voxels(grid=grid, mesh=mesh, kernel="fast")
    P = oracle_mod.Problem(pos, pos.copy(), None, conn, periodic_box=BOX, n=(16,) * 3, bounds=BOX, fields=["v"])
    ref, _ = oracle_mod.voxels(P)
    assert_close(grid.fields, ref, TOL32)
    assert abs(grid.fields["m"].sum() / mesh.mass.sum() - 1.0) < 1e-12

"""The benchmark's own workloads on the MI355X against the FP64 oracle.

Headline (128^3 particles, rms 2.0 lattice cells, seed 0 -> 256^3 grid, fields
m, v, vv): the first 2^19 of the 2,097,152 elements (lattice x-slabs 0..31,
3.1M tets) onto the full grid, every field per voxel against the oracle, for
each kernel; and the whole mesh through its size-independent properties
(total mass to 1e-12, the fragment count within 1e-5 of the oracle's
210,589,023).

Batched uneven workload (64 meshes, pypsi_amd.synth.uneven_batch_specs): one
psivox_voxels_batched call over all 64; per mesh the total mass to 1e-12 and
the fragment count within 1e-5 of the oracle's (tests/golden/c3_counts.json);
per voxel against the oracle for the six smallest meshes.

The measured errors are printed (run with -s to see them).
"""
import json
import os

import numpy as np
import pytest

from helpers import BOX, TOL32, assert_close

pytestmark = pytest.mark.gpu

TOL64 = {"m": (1e-12, 1e-13), "x": (1e-10, 1e-13), "v": (1e-10, 1e-13), "xx": (1e-7, 1e-12),
         "xv": (1e-7, 1e-12), "vv": (1e-7, 1e-12)}
# the oracle's fragment count of the whole headline mesh
C2_NFRAG = 210589023
NSUB = 1 << 19
HERE = os.path.dirname(os.path.abspath(__file__))


def _block_conn(n):
    from oracle import block_connectivity
    return block_connectivity((n,) * 3, True)


@pytest.fixture(scope="module")
def c2(oracle_mod):
    from pypsi_amd.synth import zeldovich_block
    pos, vel = zeldovich_block(128, 2.0, 0)
    pos, vel = pos.reshape(-1, 3), vel.reshape(-1, 3)
    conn = _block_conn(128)
    P = oracle_mod.Problem(pos, vel, None, conn[:NSUB], periodic_box=BOX, n=(256,) * 3, bounds=BOX,
                           fields=["v", "vv"])
    ref, st = oracle_mod.voxels(P, threads=-1)
    return pos, vel, conn, ref, st


def _run_device(pos, vel, conn, gn, fields, kernel):
    import torch
    from pypsi_amd import device as D
    dev = torch.device("cuda:0")
    tf = {"m": torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)}
    for f in fields:
        tf[f] = torch.zeros((gn,) * 3 + ((3,) if len(f) == 1 else (3, 3)), dtype=torch.float64, device=dev)
    tp, tv = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
    tc = torch.from_numpy(np.ascontiguousarray(conn)).to(dev)
    tm = torch.ones(conn.shape[0], dtype=torch.float64, device=dev)
    mesh = D.make_mesh(tp, tv, tm, tc, BOX)
    grid = D.make_grid((gn,) * 3, BOX, tf)
    st = D.voxels(mesh, grid, D.make_opts(kernel=kernel, stream=torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in tf.items()}, st


@pytest.mark.parametrize("kernel", ["auto", "fast", "plain"])
def test_headline_subset_per_voxel(c2, kernel):
    pos, vel, conn, ref, ost = c2
    print(f"case: headline 128^3 rms2.0 -> 256^3 m,v,vv, first {NSUB} elements, kernel {kernel}", flush=True)
    out, st = _run_device(pos, vel, conn[:NSUB], 256, ("v", "vv"), kernel)
    tol = TOL64 if kernel == "plain" else TOL32
    errs = assert_close(out, ref, tol, kernel)
    mass_err = abs(1.0 - out["m"].sum() / ref["m"].sum())
    dn = abs(st["nfrag"] - ost["nfrag"]) / ost["nfrag"]
    print(json.dumps({"headline_subset": kernel, "errors": {k: {"rel": v[0], "wabs": v[1]} for k, v in errs.items()},
                      "mass_rel_err": mass_err, "nfrag": st["nfrag"], "oracle_nfrag": ost["nfrag"]}), flush=True)
    assert mass_err < 1e-12
    assert dn <= (0.0 if kernel == "plain" else 1e-5), (st["nfrag"], ost["nfrag"])


def test_headline_full_properties():
    from pypsi_amd.synth import zeldovich_block
    print("case: headline 128^3 rms2.0 -> 256^3 m,v,vv, whole mesh, kernel auto", flush=True)
    pos, vel = zeldovich_block(128, 2.0, 0)
    conn = _block_conn(128)
    out, st = _run_device(pos.reshape(-1, 3), vel.reshape(-1, 3), conn, 256, ("v", "vv"), "auto")
    mass_err = abs(1.0 - out["m"].sum() / conn.shape[0])
    dn = abs(st["nfrag"] - C2_NFRAG) / C2_NFRAG
    print(json.dumps({"headline_full": {"mass_rel_err": mass_err, "nfrag": st["nfrag"], "oracle_nfrag": C2_NFRAG}}),
          flush=True)
    assert mass_err < 1e-12
    assert dn <= 1e-5
    assert np.all(out["m"] >= 0.0)
    # v is a mass-weighted mean of vertex velocities: inside their range
    vmax = np.abs(vel).max()
    assert np.abs(out["v"]).max() <= vmax * (1 + 1e-6)
    # vv is a covariance: symmetric, non-negative diagonal
    vv = out["vv"]
    assert np.allclose(vv, np.swapaxes(vv, -1, -2), rtol=0, atol=1e-12 * np.abs(vv).max())
    assert np.diagonal(vv, axis1=-2, axis2=-1).min() >= -1e-9 * np.abs(vv).max()


@pytest.fixture(scope="module")
def c3():
    with open(os.path.join(HERE, "golden", "c3_counts.json")) as f:
        return json.load(f)


def test_batched_uneven_64(oracle_mod, c3):
    import torch
    from pypsi_amd import device as D
    from pypsi_amd.synth import uneven_batch_specs, zeldovich_block
    print("case: batched uneven 64 meshes, one psivox_voxels_batched call", flush=True)
    dev = torch.device("cuda:0")
    specs = uneven_batch_specs(64)
    meshes, grids, fields, host = [], [], [], []
    for (n, rms, gn, seed), rec in zip(specs, c3["meshes"]):
        assert (n, gn, seed) == (rec["n"], rec["grid"], rec["seed"])
        pos, _ = zeldovich_block(n, rms, seed)
        conn = _block_conn(n)
        tp = torch.from_numpy(pos.reshape(-1, 3)).to(dev)
        tc = torch.from_numpy(conn).to(dev)
        tm = torch.ones(conn.shape[0], dtype=torch.float64, device=dev)
        m = torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)
        meshes.append(D.make_mesh(tp, None, tm, tc, BOX))
        grids.append(D.make_grid((gn,) * 3, BOX, {"m": m}))
        fields.append(m)
        host.append((pos, conn, gn))
    sts = D.voxels_batched(meshes, grids, D.make_opts(stream=torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    worst_n, worst_m = 0.0, 0.0
    for q, (st, rec) in enumerate(zip(sts, c3["meshes"])):
        dn = abs(st["nfrag"] - rec["nfrag"]) / rec["nfrag"]
        dm = abs(float(fields[q].sum().item()) / meshes[q].nelem - 1.0)
        worst_n, worst_m = max(worst_n, dn), max(worst_m, dm)
        assert dn <= 1e-5, (q, st["nfrag"], rec["nfrag"])
        assert dm < 1e-12, (q, dm)
    # per voxel: the six smallest meshes
    small = sorted(range(64), key=lambda q: specs[q][0])[:6]
    errs_all = {}
    for q in small:
        pos, conn, gn = host[q]
        P = oracle_mod.Problem.block(pos, None, periodic_box=BOX, n=(gn,) * 3, bounds=BOX)
        ref, _ = oracle_mod.voxels(P, threads=-1)
        errs = assert_close({"m": fields[q].cpu().numpy()}, ref, TOL32, f"mesh {q}")
        errs_all[q] = errs["m"]
    print(json.dumps({"batched_uneven_64": {"max_nfrag_rel_diff": worst_n, "max_mass_rel_err": worst_m,
                                            "per_voxel_m_small": {str(k): v for k, v in errs_all.items()}}}),
          flush=True)

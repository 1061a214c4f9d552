"""psivox_voxels_noalloc through the C ABI: the allocation-free variant of
psivox_voxels (its buffers come from a fixed device scratch array) gives
what psivox_voxels gives on the same device mesh and grid -- the same
kernels, so the same fragment counts and fields equal up to the summation
order of the FP64 atomics -- on two calls in a
row, the second with a smaller input, and the FP32 bar against the oracle.
A mesh larger than one scratch batch runs in several batches; a grid whose
accumulator does not fit is refused with a message."""
import ctypes as C

import numpy as np
import pytest

from helpers import BOX, TOL32, assert_close

pytestmark = pytest.mark.gpu

# the same kernels in another FP64 atomic summation order: the covariance
# finalize (raw second moment minus the squared mean) amplifies that order's
# rounding on cells with a small spread, so covariances get 1e-9
SAME = {"m": (1e-12, 1e-12), "x": (1e-10, 1e-12), "v": (1e-10, 1e-12), "xx": (1e-9, 1e-12), "xv": (1e-9, 1e-12),
        "vv": (1e-9, 1e-12)}


def _case(n, rms, seed, gn, fields):
    import torch
    from oracle import block_connectivity
    from pypsi_amd import device as D
    from pypsi_amd.synth import zeldovich_block
    dev = torch.device("cuda:0")
    pos, vel = zeldovich_block(n, rms, seed)
    conn = block_connectivity((n,) * 3, True)
    tp = torch.from_numpy(pos.reshape(-1, 3)).to(dev)
    tv = torch.from_numpy(vel.reshape(-1, 3)).to(dev)
    tc = torch.from_numpy(conn).to(dev)
    tm = torch.ones(conn.shape[0], dtype=torch.float64, device=dev)
    mesh = D.make_mesh(tp, tv, tm, tc, BOX)

    def grid():
        tf = {"m": torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)}
        for f in fields:
            tf[f] = torch.zeros((gn,) * 3 + ((3,) if len(f) == 1 else (3, 3)), dtype=torch.float64, device=dev)
        return D.make_grid((gn,) * 3, BOX, tf), tf

    return pos, vel, mesh, grid


def _noalloc(mesh, grid, opts):
    from pypsi_amd import device as D
    st = D.Stats()
    rc = D.lib().psivox_voxels_noalloc(C.byref(mesh), C.byref(grid), C.byref(opts), C.byref(st))
    return rc, st.as_dict()


def _host(tf):
    return {k: v.cpu().numpy() for k, v in tf.items()}


def test_noalloc_matches_voxels_two_calls(oracle_mod):
    import torch
    from pypsi_amd import device as D
    s = torch.cuda.current_stream().cuda_stream
    for n, rms, gn, fields in ((16, 2.0, 40, ("x", "v", "xx", "xv", "vv")), (8, 1.0, 16, ("v", "vv"))):
        print(f"case: noalloc vs voxels {n}^3 rms{rms} -> {gn}^3 {fields}", flush=True)
        pos, vel, mesh, grid = _case(n, rms, n, gn, fields)
        ga, tfa = grid()
        gb, tfb = grid()
        sa = D.voxels(mesh, ga, D.make_opts(stream=s))
        rc, sb = _noalloc(mesh, gb, D.make_opts(stream=s))
        torch.cuda.synchronize()
        assert rc == 0, D.lib().psivox_last_error()
        assert sb["nfrag"] == sa["nfrag"] and sb["nwide"] == sa["nwide"] and sb["ntets"] == sa["ntets"]
        a, b = _host(tfa), _host(tfb)
        errs = assert_close(b, a, SAME, "noalloc vs voxels")
        print("noalloc vs voxels", {k: "%.1e/%.1e" % v for k, v in errs.items()}, sb["nfrag"], flush=True)
        P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX, n=(gn,) * 3, bounds=BOX, fields=list(fields))
        ref, _ = oracle_mod.voxels(P)
        assert_close(b, ref, TOL32, "noalloc vs oracle")


def test_noalloc_several_batches_and_limits():
    import torch
    from pypsi_amd import device as D
    s = torch.cuda.current_stream().cuda_stream
    # 64^3 lattice: 1.57M tets, more than one batch of the scratch
    print("case: noalloc 64^3 rms1 -> 128^3 m (several scratch batches)", flush=True)
    _, _, mesh, grid = _case(64, 1.0, 0, 128, ())
    ga, tfa = grid()
    gb, tfb = grid()
    sa = D.voxels(mesh, ga, D.make_opts(stream=s))
    rc, sb = _noalloc(mesh, gb, D.make_opts(stream=s))
    torch.cuda.synchronize()
    assert rc == 0, D.lib().psivox_last_error()
    assert sb["nfrag"] == sa["nfrag"]
    assert_close(_host(tfb), _host(tfa), SAME, "noalloc batches")
    # an accumulator larger than the scratch: refused, nothing written
    print("case: noalloc 400^3 grid (accumulator beyond the scratch)", flush=True)
    m = torch.zeros((400,) * 3, dtype=torch.float64, device="cuda:0")
    big = D.make_grid((400,) * 3, BOX, {"m": m})
    rc, _ = _noalloc(mesh, big, D.make_opts(stream=s))
    assert rc == -4 and b"scratch" in D.lib().psivox_last_error()
    assert float(m.abs().sum().item()) == 0.0
    # host arrays and refinement are refused
    print("case: noalloc argument errors", flush=True)
    rc, _ = _noalloc(mesh, gb, D.make_opts(stream=s, refine_max_lvl=1))
    assert rc == -6
    host_grid = D.make_grid((128,) * 3, BOX, {"m": tfb["m"]})
    host_grid.on_device = 0
    rc, _ = _noalloc(mesh, host_grid, D.make_opts(stream=s))
    assert rc == -1 and b"device" in D.lib().psivox_last_error()
    # the library's other entry points still work afterwards
    gc, tfc = grid()
    D.voxels(mesh, gc, D.make_opts(stream=s))
    torch.cuda.synchronize()
    assert_close(_host(tfc), _host(tfa), SAME, "voxels after noalloc")

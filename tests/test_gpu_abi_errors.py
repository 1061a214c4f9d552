"""Argument errors that the C ABI can only answer on a device: connectivity
entries outside [0, npart) in device meshes (every entry point checks them
before any kernel reads through them), and grids too long for the mass
adjoint's plane encoding.  Each call returns PSIVOX_E_ARG with a message and
leaves the GPU usable."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def _mesh(conn_rows):
    import torch
    from pypsi_amd import device as D
    dev = torch.device("cuda:0")
    pos = torch.rand((5, 3), dtype=torch.float64, device=dev)
    conn = torch.tensor(conn_rows, dtype=torch.int32, device=dev)
    mass = torch.ones(conn.shape[0], dtype=torch.float64, device=dev)
    return D.make_mesh(pos, None, mass, conn, BOX), dev


@pytest.mark.parametrize("rows", [[[0, 1, 2, 5]], [[-1, 1, 2, 3]], [[0, 1, 2, 3], [4, 3, 2, 2147483647]]])
def test_device_connectivity_out_of_range(rows):
    import torch
    from pypsi_amd import device as D
    print("case: device connectivity", rows, flush=True)
    mesh, dev = _mesh(rows)
    m = torch.zeros((8, 8, 8), dtype=torch.float64, device=dev)
    grid = D.make_grid((8, 8, 8), BOX, {"m": m})
    L = D.lib()
    st = D.Stats()
    for kernel in ("fast", "plain"):
        opts = D.make_opts(kernel=kernel)
        assert L.psivox_voxels(C.byref(mesh), C.byref(grid), C.byref(opts), C.byref(st)) == -1
        assert b"connectivity" in L.psivox_last_error()
    assert L.psivox_voxels_batched(1, C.byref(mesh), C.byref(grid), C.byref(D.make_opts()), C.byref(st)) == -1
    g = torch.zeros((5, 3), dtype=torch.float64, device=dev)
    assert L.psivox_mass_adjoint(C.byref(mesh), C.byref(grid), C.c_void_p(m.data_ptr()), C.c_void_p(g.data_ptr()),
                                 C.byref(D.make_opts())) == -1
    with pytest.raises(RuntimeError, match="connectivity"):
        D.ElemIndex(mesh, BOX)
    # the device is still usable: a valid mesh voxelizes
    ok, _ = _mesh([[0, 1, 2, 3]])
    D.voxels(ok, grid, D.make_opts())
    assert torch.isfinite(m).all()


def test_adjoint_plane_limit():
    import torch
    from pypsi_amd import device as D
    print("case: adjoint grid of 2^17 cells along x", flush=True)
    mesh, dev = _mesh([[0, 1, 2, 3]])
    grid = D.make_grid((1 << 17, 1, 1), BOX, {"m": torch.zeros(1, dtype=torch.float64, device=dev)})
    g = torch.zeros((5, 3), dtype=torch.float64, device=dev)
    L = D.lib()
    rc = L.psivox_mass_adjoint(C.byref(mesh), C.byref(grid), C.c_void_p(g.data_ptr()), C.c_void_p(g.data_ptr()),
                               C.byref(D.make_opts()))
    assert rc == -1 and b"cells per axis" in L.psivox_last_error()

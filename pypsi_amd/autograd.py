"""Differentiable mass deposit for PyTorch: the grid mass field of
PSI.voxels(mode='density') as a function of the particle positions, with
the backward pass computed by libpsivox's adjoint kernels
(psivox_mass_adjoint, include/psivox.h).

    m = voxel_mass(pos, conn, n, bounds, mass=None, periodic_box=None)

pos (npart, 3) is a CUDA tensor of any floating dtype (the kernels take
float64 positions and clip in FP32 cell-local coordinates); conn (nelem, 4|8|27)
int32; mass (nelem,) float64 or None (1 per element).  m has shape n and
float64 dtype.  Refinement level 0 only.  The clip polyhedra are recomputed
in the backward pass, nothing per fragment is stored.
"""
from __future__ import annotations

import torch

from pypsi_amd import device as D


def _descs(pos64, mass, conn, n, bounds, periodic_box, m=None):
    mesh = D.make_mesh(pos64, None, mass, conn, periodic_box)
    fields = {"m": m} if m is not None else {"m": torch.empty(0, dtype=torch.float64, device=pos64.device)}
    grid = D.make_grid(n, bounds, fields)
    return mesh, grid


class VoxelMass(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, conn, mass, n, bounds, periodic_box, kernel):
        if not pos.is_cuda:
            raise RuntimeError("voxel_mass needs CUDA (HIP) tensors")
        pos64 = pos.detach().to(torch.float64).contiguous()
        conn = conn.to(torch.int32).contiguous()
        if mass is None:
            mass = torch.ones(conn.shape[0], dtype=torch.float64, device=pos.device)
        mass = mass.to(torch.float64).contiguous()
        m = torch.zeros(tuple(n), dtype=torch.float64, device=pos.device)
        mesh, grid = _descs(pos64, mass, conn, n, bounds, periodic_box, m)
        stream = torch.cuda.current_stream(pos.device).cuda_stream
        D.voxels(mesh, grid, D.make_opts(kernel=kernel, stream=stream))
        ctx.save_for_backward(pos64, conn, mass)
        ctx.geom = (tuple(n), bounds, periodic_box, pos.dtype)
        return m

    @staticmethod
    def backward(ctx, grad_m):
        pos64, conn, mass = ctx.saved_tensors
        n, bounds, periodic_box, dtype = ctx.geom
        g = grad_m.to(torch.float64).contiguous()
        dpos = torch.zeros_like(pos64)
        mesh, grid = _descs(pos64, mass, conn, n, bounds, periodic_box)
        stream = torch.cuda.current_stream(pos64.device).cuda_stream
        D.mass_adjoint(mesh, grid, g, dpos, D.make_opts(stream=stream))
        return dpos.to(dtype), None, None, None, None, None, None


def voxel_mass(pos, conn, n, bounds, mass=None, periodic_box=None, kernel="auto"):
    """Grid mass field (float64, shape n) of PSI.voxels, differentiable in pos."""
    return VoxelMass.apply(pos, conn, mass, tuple(int(v) for v in n), bounds, periodic_box, kernel)


def voxel_density(pos, conn, n, bounds, mass=None, periodic_box=None, kernel="auto"):
    """Mass per unit volume (m / cell volume)."""
    import math
    vol = math.prod((float(bounds[1][a]) - float(bounds[0][a])) / int(n[a]) for a in range(3))
    return voxel_mass(pos, conn, n, bounds, mass, periodic_box, kernel) / vol

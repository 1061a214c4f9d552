"""Differentiable PSI.voxels for PyTorch: the grid mass field as a function
of the particle positions (voxel_mass, voxel_density; backward pass
psivox_mass_adjoint), and the fields m, v, vv as functions of positions and
velocities (voxel_fields; backward pass psivox_adjoint), both computed by
libpsivox's adjoint kernels (include/psivox.h).

    m = voxel_mass(pos, conn, n, bounds, mass=None, periodic_box=None)
    m, v, vv = voxel_fields(pos, vel, conn, n, bounds, mass=None, periodic_box=None)

pos (npart, 3) is a CUDA tensor of any floating dtype (the kernels take
float64 positions; voxel_mass's default forward clips in FP32 cell-local
coordinates, its gradient within 1e-4 of finite differences); conn (nelem, 4|8|27)
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
    def forward(ctx, pos, conn, mass, n, bounds, periodic_box, kernel, refine):
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
        D.voxels(mesh, grid, D.make_opts(kernel=kernel, stream=stream, refine_tolerance=refine[0],
                                         refine_max_lvl=refine[1]))
        ctx.save_for_backward(pos64, conn, mass)
        ctx.geom = (tuple(n), bounds, periodic_box, pos.dtype, refine)
        return m

    @staticmethod
    def backward(ctx, grad_m):
        pos64, conn, mass = ctx.saved_tensors
        n, bounds, periodic_box, dtype, refine = ctx.geom
        g = grad_m.to(torch.float64).contiguous()
        dpos = torch.zeros_like(pos64)
        mesh, grid = _descs(pos64, mass, conn, n, bounds, periodic_box)
        stream = torch.cuda.current_stream(pos64.device).cuda_stream
        D.mass_adjoint(mesh, grid, g, dpos, D.make_opts(stream=stream, refine_tolerance=refine[0],
                                                        refine_max_lvl=refine[1]))
        return dpos.to(dtype), None, None, None, None, None, None, None


def voxel_mass(pos, conn, n, bounds, mass=None, periodic_box=None, kernel="fast64", refine_tolerance=1.0,
               refine_max_lvl=0):
    """Grid mass field (float64, shape n) of PSI.voxels, differentiable in pos.
    The default forward clips in FP64 (kernel 'fast64'), so dL/dm carries no
    FP32 error: gradients agree with FP64 finite differences to 1e-4.  The
    FP32-clip forwards ('auto', 'fast') are faster; their cell masses' FP32
    error (about 1e-6 relative) reaches the gradient wherever the loss's
    dL/dm is small beside its neighbours'.  Refinement as PSI.voxels
    (refine_tolerance in cells, refine_max_lvl): a refined tet's vertices are
    the element's map at fixed parameter points, so their gradients reach
    the element's nodes through that map; the refinement decisions carry no
    gradient."""
    return VoxelMass.apply(pos, conn, mass, tuple(int(v) for v in n), bounds, periodic_box, kernel,
                           (float(refine_tolerance), int(refine_max_lvl)))


def voxel_density(pos, conn, n, bounds, mass=None, periodic_box=None, kernel="fast64", refine_tolerance=1.0,
                  refine_max_lvl=0):
    """Mass per unit volume (m / cell volume)."""
    import math
    vol = math.prod((float(bounds[1][a]) - float(bounds[0][a])) / int(n[a]) for a in range(3))
    return voxel_mass(pos, conn, n, bounds, mass, periodic_box, kernel, refine_tolerance, refine_max_lvl) / vol


FIELD_SHAPES = {"m": (), "v": (3,), "vv": (3, 3)}


class VoxelFields(torch.autograd.Function):
    """(m, v, vv) of PSI.voxels as functions of the particle positions and
    velocities; the backward pass is psivox_adjoint (positions: the
    plane-section integrals, velocities: the fragments' moments)."""

    @staticmethod
    def forward(ctx, pos, vel, conn, mass, n, bounds, periodic_box, kernel, refine):
        if not pos.is_cuda:
            raise RuntimeError("voxel_fields needs CUDA (HIP) tensors")
        pos64 = pos.detach().to(torch.float64).contiguous()
        vel64 = vel.detach().to(torch.float64).contiguous()
        conn = conn.to(torch.int32).contiguous()
        if mass is None:
            mass = torch.ones(conn.shape[0], dtype=torch.float64, device=pos.device)
        mass = mass.to(torch.float64).contiguous()
        f = {k: torch.zeros(tuple(n) + s, dtype=torch.float64, device=pos.device) for k, s in FIELD_SHAPES.items()}
        mesh = D.make_mesh(pos64, vel64, mass, conn, periodic_box)
        grid = D.make_grid(n, bounds, f)
        stream = torch.cuda.current_stream(pos.device).cuda_stream
        D.voxels(mesh, grid, D.make_opts(kernel=kernel, stream=stream, refine_tolerance=refine[0],
                                         refine_max_lvl=refine[1]))
        ctx.save_for_backward(pos64, vel64, conn, mass, f["m"], f["v"], f["vv"])
        ctx.geom = (tuple(n), bounds, periodic_box, pos.dtype, vel.dtype, refine)
        return f["m"], f["v"], f["vv"]

    @staticmethod
    def backward(ctx, gm, gv, gvv):
        pos64, vel64, conn, mass, m, v, vv = ctx.saved_tensors
        n, bounds, periodic_box, pdt, vdt, refine = ctx.geom
        mesh = D.make_mesh(pos64, vel64, mass, conn, periodic_box)
        grid = D.make_grid(n, bounds, {"m": m, "v": v, "vv": vv})
        grads = {k: (g.to(torch.float64).contiguous() if g is not None else None)
                 for k, g in (("m", gm), ("v", gv), ("vv", gvv))}
        if all(g is None for g in grads.values()):
            return (None,) * 9
        dpos = torch.zeros_like(pos64)
        dvel = torch.zeros_like(vel64)
        stream = torch.cuda.current_stream(pos64.device).cuda_stream
        D.adjoint(mesh, grid, grads, dpos, dvel, D.make_opts(stream=stream, refine_tolerance=refine[0],
                                                             refine_max_lvl=refine[1]))
        return dpos.to(pdt), dvel.to(vdt), None, None, None, None, None, None, None


def voxel_fields(pos, vel, conn, n, bounds, mass=None, periodic_box=None, kernel="fast64", refine_tolerance=1.0,
                 refine_max_lvl=0):
    """(m, v, vv) of PSI.voxels (mass, mass-weighted mean velocity, velocity
    dispersion; float64, shapes n, n + (3,), n + (3, 3)), differentiable in
    pos and vel, with refinement as voxel_mass.  The default forward is the column
    kernel in FP64: the weights of the velocity fields' adjoint are formed
    from the forward's cell means, and the FP32-clip forward ('auto') leaves
    up to 1e-4 of error in the velocity gradient."""
    return VoxelFields.apply(pos, vel, conn, mass, tuple(int(v) for v in n), bounds, periodic_box, kernel,
                             (float(refine_tolerance), int(refine_max_lvl)))

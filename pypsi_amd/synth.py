"""Seeded synthetic phase-space sheets (Zel'dovich-displaced Lagrangian lattices).

These build the inputs of the benchmark workloads: an N^3 particle lattice on
the unit periodic box, displaced by a Gaussian random field whose density power
spectrum goes as k^-2, with velocities equal to the displacement.  The lattice
matches the reference tutorial's block-mesh example (tutorial/Tutorial.ipynb,
cell 12: positions at ((i,j,k)+0.5)/N fed to Mesh(loader='block', ...,
periodic_box=unit box)).  Everything is float64 C-contiguous, shaped
(N, N, N, 3) as Mesh(loader='block') expects (src/PSI.c:530-545).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["zeldovich_block", "uneven_batch_specs", "add_infall_clumps"]


def _displacement(rng: np.random.Generator, n: int, rms_cells: float) -> np.ndarray:
    """Displacement field psi (n,n,n,3) with sqrt(<|psi|^2>) = rms_cells / n."""
    noise = rng.standard_normal((n, n, n))
    dk = np.fft.rfftn(noise)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    delta = dk / np.sqrt(k2)  # |k|^-1 amplitude -> P_delta ~ k^-2
    delta[0, 0, 0] = 0.0
    psi = np.empty((n, n, n, 3))
    for a, ka in enumerate((kx, ky, kz)):
        # psi_hat = i k delta_hat / k^2 (sign convention irrelevant for statistics)
        psi[..., a] = np.fft.irfftn(1j * ka * delta / k2, s=(n, n, n), axes=(0, 1, 2))
    rms = math.sqrt(float(np.mean(np.sum(psi * psi, axis=-1))))
    if rms > 0:
        psi *= (rms_cells / n) / rms
    return psi


def _wrap_unit(x: np.ndarray) -> np.ndarray:
    x = np.mod(x, 1.0)
    x[x >= 1.0] = 0.0  # np.mod of a tiny negative can round up to exactly 1
    return x


def zeldovich_block(n: int, rms_cells: float, seed: int = 0):
    """(pos, vel) of an n^3 lattice displaced with the given rms (in lattice
    cells), wrapped into [0, 1)."""
    rng = np.random.default_rng(seed)
    psi = _displacement(rng, n, rms_cells)
    g = (np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), axis=-1) + 0.5) / n
    pos = np.ascontiguousarray(_wrap_unit(g + psi))
    vel = np.ascontiguousarray(psi)
    return pos, vel


def add_infall_clumps(pos: np.ndarray, n: int, nclumps: int = 8, seed: int = 100, radius_cells: float = 10.0,
                      amp_cells: float = 1.0) -> np.ndarray:
    """Seeded spherical infall: inside radius r < radius_cells (lattice cells)
    particles move toward the clump centre by amp*(max(r,0.5))^-0.5 cells."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, size=(nclumps, 3))
    out = pos.copy()
    for c in centres:
        d = out - c
        d -= np.round(d)  # periodic minimum image
        r = np.sqrt(np.sum(d * d, axis=-1)) * n  # in cells
        inside = r < radius_cells
        mag = amp_cells * np.maximum(r, 0.5) ** -0.5 / n
        rhat = d / np.maximum(r[..., None] / n, 1e-30)
        step = np.minimum(mag, r / n)[..., None] * rhat
        out = np.where(inside[..., None], out - step, out)
    return np.ascontiguousarray(_wrap_unit(out))


def uneven_batch_specs(count: int = 64):
    """Per-mesh (n, rms_cells, grid_n, seed) of the uneven batched workload:
    n log-uniform in 16..96, rms uniform in 0.2..3.0, grid 2^round(log2(8n/3))
    clipped to 32..256."""
    specs = []
    for s in range(count):
        rng = np.random.default_rng(s)
        n = int(round(math.exp(rng.uniform(math.log(16), math.log(96)))))
        rms = float(rng.uniform(0.2, 3.0))
        gn = int(min(256, max(32, 2 ** round(math.log2(8 * n / 3)))))
        specs.append((n, rms, gn, s))
    return specs

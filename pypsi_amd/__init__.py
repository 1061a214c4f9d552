"""pypsi_amd -- MI355X-native PSI.voxels.

    from pypsi_amd import PSI          # drop-in for the reference module PSI
    grid = PSI.Grid(n=(256,)*3, bounds=((0,0,0),(1,1,1)), fields=['v', 'vv'])
    mesh = PSI.Mesh(loader='block', pos=pos, vel=vel, periodic_box=((0,0,0),(1,1,1)))
    PSI.voxels(grid=grid, mesh=mesh)

PSI is a C extension over libpsivox.so (HIP kernels for gfx950).  Nothing
here falls back to the CPU: if the native library is missing or no GPU is
visible, importing PSI or calling voxels() raises.
"""
from __future__ import annotations

import importlib
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def _load_psi():
    try:
        return importlib.import_module("pypsi_amd.PSI")
    except ImportError as e:  # pragma: no cover - exercised when unbuilt
        raise ImportError(
            "pypsi_amd.PSI is not built (run `python -m pypsi_amd.build` or __graft_entry__.build()): " + str(e)
        ) from e


def __getattr__(name):
    if name == "PSI":
        mod = _load_psi()
        globals()["PSI"] = mod
        return mod
    raise AttributeError(name)

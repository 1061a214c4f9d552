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

import ctypes
import importlib
import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def preload_hip_runtime() -> None:
    """Load the HIP runtime that PyTorch ships (when PyTorch is installed)
    before libpsivox.so, so a process that also uses torch holds one HIP
    runtime: libpsivox binds the runtime by soname (libamdhip64.so.7), torch
    by its own bundled path, and two copies in one process cannot both open
    the GPU."""
    if "torch" in sys.modules:
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if not spec or not spec.submodule_search_locations:
        return
    lib = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(lib):
        ctypes.CDLL(lib, mode=ctypes.RTLD_GLOBAL)


def _load_psi():
    preload_hip_runtime()
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

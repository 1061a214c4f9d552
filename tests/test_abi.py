"""The C ABI and the PSI module surface on the CPU: libpsivox.so exports every
function include/psivox.h declares, answers argument errors before touching a
GPU, and the PSI module validates arguments with the reference's exceptions
and messages (src/PSI.c:58-279, 367-722, 974-1018)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
HEADER = os.path.join(ROOT, "include", "psivox.h")
LIB = os.path.join(ROOT, "pypsi_amd", "libpsivox.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(psivox_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("psivox_voxels", "psivox_voxels_batched", "psivox_mass_adjoint", "psivox_last_error",
              "psivox_abi_version", "psivox_device_count", "psivox_elem_index_build", "psivox_elem_index_free",
              "psivox_elem_index_query", "psivox_sample_vdf"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from pypsi_amd import preload_hip_runtime
    preload_hip_runtime()
    lib = C.CDLL(LIB)
    for n in declared():
        assert hasattr(lib, n), n
    from pypsi_amd.device import EXPORTS
    assert sorted(EXPORTS) == declared()


def test_abi_version_and_argument_errors():
    from pypsi_amd import device as D
    L = D.lib()
    assert L.psivox_abi_version() == 2
    # NULL descriptors are rejected before any device work
    assert L.psivox_voxels(None, None, None, None) == -1
    assert b"required" in L.psivox_last_error()
    assert L.psivox_mass_adjoint(None, None, None, None, None) == -1
    assert L.psivox_voxels_batched(-1, None, None, None, None) == -1
    # the point-query entry points likewise
    assert L.psivox_elem_index_build(None, None, None, None) == -1
    assert L.psivox_elem_index_query(None, None, None, 0, None, None) == -1
    assert L.psivox_sample_vdf(None, None, 1.0, 0, None, None, 0, None, None) == -1
    L.psivox_elem_index_free(None)


def test_struct_layouts_match_header():
    """ctypes layouts used by the bindings (sizes of the C structs)."""
    from pypsi_amd import device as D
    assert C.sizeof(D.Mesh) == 4 + 4 + 8 + 8 + 4 * 8 + 48 + 4 + 4
    assert C.sizeof(D.Grid) == 12 + 4 + 48 + 24 + 48
    assert C.sizeof(D.Stats) == 4 * 8 + 2 * 8 + 4 + 4 + 8


def test_module_argument_errors():
    from pypsi_amd import PSI
    with pytest.raises(ValueError, match="Invalid grid type"):
        PSI.Grid(type="polar", n=(4, 4, 4))
    with pytest.raises(ValueError, match="requires you to also provide field 'x'"):
        PSI.Grid(n=(4, 4, 4), fields=["xx"])
    with pytest.raises(ValueError, match="Unrecognized field identifier"):
        PSI.Grid(n=(4, 4, 4), fields=["q"])
    g = PSI.Grid(n=(4, 5, 6), bounds=((0, 0, 0), (1, 2, 3)), fields=["x", "v", "xx", "vx", "vv"])
    assert set(g.fields) == {"m", "x", "v", "xx", "xv", "vv"}
    assert g.fields["xv"].shape == (4, 5, 6, 3, 3) and g.fields["m"].dtype == np.float64
    assert np.allclose(g.d, (0.25, 0.4, 0.5))
    with pytest.raises(ValueError, match="Invalid mesh loader"):
        PSI.Mesh(loader="nope")
    with pytest.raises(TypeError, match="pos must be"):
        PSI.Mesh(loader="block", pos=np.zeros((2, 2, 2, 3), dtype=np.float32))
    m = PSI.Mesh(loader="block", pos=np.zeros((3, 3, 3, 3)), periodic_box=((0, 0, 0), (1, 1, 1)))
    assert m.connectivity.shape == (27, 8) and m.mass.shape == (3, 3, 3)  # (ne, ne, ne) as src/PSI.c:530-545
    with pytest.raises(ValueError, match="Invalid mode for PSI.voxels"):
        PSI.voxels(grid=g, mesh=m, mode="nonsense")
    with pytest.raises(TypeError):
        PSI.VDF(m, (0.5, 0.5))
    with pytest.raises(TypeError):
        PSI.RStarTree(mesh=3)


def test_mesh_array_connectivity_range():
    """Connectivity entries outside [0, npart) are refused when the mesh is
    made (the reference reads pos[conn] unchecked, src/cpsi.c:85-98)."""
    from pypsi_amd import PSI
    pos = np.random.default_rng(0).uniform(size=(5, 3))
    PSI.Mesh(loader="array", pos=pos, connectivity=np.array([[0, 1, 2, 4]], dtype=np.int32))
    for bad in ([[0, 1, 2, 5]], [[-1, 1, 2, 3]], [[0, 1, 2, 3], [4, 3, 2, 7]]):
        with pytest.raises(ValueError, match="outside"):
            PSI.Mesh(loader="array", pos=pos, connectivity=np.array(bad, dtype=np.int32))

"""ctypes access to libpsivox's C ABI (include/psivox.h) with device-resident
torch tensors: used by the benchmark, the autograd wrapper and the GPU tests.

PyTorch only provides device memory and the current stream here; every
voxelization is done by libpsivox's HIP kernels.  Loading fails loudly when the
library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpsivox.so")

KERNELS = {"auto": 0, "plain": 1, "fast": 2, "fast64": 3}
FIELD_ORDER = ["m", "x", "v", "xx", "xv", "vv"]
FIELD_WIDTH = {"m": 1, "x": 3, "v": 3, "xx": 9, "xv": 9, "vv": 9}


class Mesh(C.Structure):
    _fields_ = [("elemtype", C.c_int32), ("periodic", C.c_int32), ("npart", C.c_int64), ("nelem", C.c_int64),
                ("pos", C.c_void_p), ("vel", C.c_void_p), ("mass", C.c_void_p), ("conn", C.c_void_p),
                ("box", (C.c_double * 3) * 2), ("on_device", C.c_int32), ("reserved", C.c_int32)]


class Grid(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("on_device", C.c_int32), ("window", (C.c_double * 3) * 2),
                ("d", C.c_double * 3), ("fields", C.c_void_p * 6)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_int64)


class Opts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("refine_max_lvl", C.c_int32), ("refine_tolerance", C.c_double),
                ("kernel", C.c_int32), ("xv_true_covariance", C.c_int32), ("stream", C.c_void_p),
                ("progress", C.c_void_p), ("progress_ctx", C.c_void_p), ("batch_elements", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("nelem_used", C.c_int64), ("ntets", C.c_int64), ("nghosts", C.c_int64), ("nfrag", C.c_int64),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("kernel", C.c_int32),
                ("reserved", C.c_int32), ("nwide", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built; run `python -m pypsi_amd.build`")
        from pypsi_amd import preload_hip_runtime
        preload_hip_runtime()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.psivox_voxels.argtypes = [P, P, P, P]
        L.psivox_voxels_batched.argtypes = [C.c_int32, P, P, P, P]
        L.psivox_voxels_noalloc.argtypes = [P, P, P, P]
        L.psivox_mass_adjoint.argtypes = [P, P, P, P, P]
        L.psivox_adjoint.argtypes = [P, P, P, P, P, P]
        L.psivox_last_error.restype = C.c_char_p
        L.psivox_device_count.restype = C.c_int
        L.psivox_abi_version.restype = C.c_int
        L.psivox_elem_index_build.argtypes = [P, P, P, P]
        L.psivox_elem_index_free.argtypes = [P]
        L.psivox_elem_index_free.restype = None
        L.psivox_elem_index_query.argtypes = [P, P, P, C.c_int64, P, P]
        L.psivox_sample_vdf.argtypes = [P, P, C.c_double, C.c_int32, P, P, C.c_int64, P, P]
        _lib = L
    return _lib


EXPORTS = ["psivox_voxels", "psivox_voxels_batched", "psivox_voxels_noalloc", "psivox_mass_adjoint", "psivox_adjoint", "psivox_last_error",
           "psivox_abi_version", "psivox_device_count", "psivox_elem_index_build", "psivox_elem_index_free",
           "psivox_elem_index_query", "psivox_sample_vdf"]


def check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(f"libpsivox error {rc}: {lib().psivox_last_error().decode()}")


def make_mesh(pos, vel, mass, conn, periodic_box=None) -> Mesh:
    """Descriptor for device tensors: pos/vel (npart,3) f64, mass (nelem,) f64,
    conn (nelem, vpe) int32, all contiguous on the GPU."""
    m = Mesh()
    m.elemtype = int(conn.shape[1])
    m.periodic = int(periodic_box is not None)
    m.npart = int(pos.shape[0])
    m.nelem = int(conn.shape[0])
    m.pos = pos.data_ptr()
    m.vel = vel.data_ptr() if vel is not None else None
    m.mass = mass.data_ptr()
    m.conn = conn.data_ptr()
    if periodic_box is not None:
        for s in range(2):
            for a in range(3):
                m.box[s][a] = float(periodic_box[s][a])
    m.on_device = 1
    # the descriptor carries raw device pointers: it keeps the tensors alive
    m._keep = (pos, vel, mass, conn)
    return m


def make_grid(n, bounds, fields: dict) -> Grid:
    g = Grid()
    for a in range(3):
        g.n[a] = int(n[a])
        g.window[0][a] = float(bounds[0][a])
        g.window[1][a] = float(bounds[1][a])
        g.d[a] = (float(bounds[1][a]) - float(bounds[0][a])) / int(n[a])
    for i, f in enumerate(FIELD_ORDER):
        g.fields[i] = fields[f].data_ptr() if f in fields else None
    g.on_device = 1
    g._keep = dict(fields)
    return g


def make_opts(kernel="auto", refine_tolerance=1.0, refine_max_lvl=0, stream=None, xv_true_covariance=False,
              mode="density") -> Opts:
    o = Opts()
    o.mode = {"density": 0, "annihilation": 1}[mode]
    o.refine_tolerance = float(refine_tolerance)
    o.refine_max_lvl = int(refine_max_lvl)
    o.kernel = KERNELS[kernel]
    o.xv_true_covariance = int(bool(xv_true_covariance))
    o.stream = stream
    return o


def voxels(mesh: Mesh, grid: Grid, opts: Opts) -> dict:
    st = Stats()
    check(lib().psivox_voxels(C.byref(mesh), C.byref(grid), C.byref(opts), C.byref(st)))
    return st.as_dict()


def voxels_batched(meshes, grids, opts: Opts) -> list:
    """Many (mesh, grid) pairs in one device pass (psivox_voxels_batched);
    returns one stats dict per pair."""
    n = len(meshes)
    if len(grids) != n:
        raise ValueError("meshes and grids must have the same length")
    ma = (Mesh * n)(*meshes)
    ga = (Grid * n)(*grids)
    sa = (Stats * n)()
    check(lib().psivox_voxels_batched(n, ma, ga, C.byref(opts), sa))
    return [st.as_dict() for st in sa]


def mass_adjoint(mesh: Mesh, grid: Grid, dl_dm, dl_dpos, opts: Opts) -> None:
    """dl_dpos += dL/dpos for dL/dm = dl_dm (device tensors: grid-shaped
    float64, and (npart, 3) float64)."""
    check(lib().psivox_mass_adjoint(C.byref(mesh), C.byref(grid), C.c_void_p(dl_dm.data_ptr()),
                                    C.c_void_p(dl_dpos.data_ptr()), C.byref(opts)))


def adjoint(mesh: Mesh, grid: Grid, dl_dfields: dict, dl_dpos=None, dl_dvel=None, opts: Opts = None) -> None:
    """dl_dpos += dL/dpos, dl_dvel += dL/dvel for the gradients dl_dfields
    {'m', 'v', 'vv': device tensors in the fields' layouts}; grid holds the
    forward's fields (psivox_adjoint)."""
    ptrs = (C.c_void_p * 6)(*[dl_dfields[f].data_ptr() if f in dl_dfields and dl_dfields[f] is not None else None
                              for f in FIELD_ORDER])
    check(lib().psivox_adjoint(C.byref(mesh), C.byref(grid), ptrs,
                               C.c_void_p(dl_dpos.data_ptr()) if dl_dpos is not None else None,
                               C.c_void_p(dl_dvel.data_ptr()) if dl_dvel is not None else None,
                               C.byref(opts if opts is not None else make_opts())))


class ElemIndex:
    """The element-box index of a mesh on the GPU (psivox_elem_index_*): the
    device form of PSI.RStarTree.  window: ((lo), (hi)) for the build's
    window test (the periodic box, or an unbounded window)."""

    def __init__(self, mesh: Mesh, window, stream=None):
        W = ((C.c_double * 3) * 2)(*[(C.c_double * 3)(*map(float, w)) for w in window])
        h = C.c_void_p()
        check(lib().psivox_elem_index_build(C.byref(mesh), C.byref(W), stream, C.byref(h)))
        self._h = h
        self._mesh = mesh  # a device mesh must outlive the index

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().psivox_elem_index_free(h)
            self._h = None

    def query(self, box, stream=None):
        """int32 element indices whose box meets the closed box, ascending."""
        import numpy as np
        Q = ((C.c_double * 3) * 2)(*[(C.c_double * 3)(*map(float, b)) for b in box])
        n = C.c_int64(0)
        cap = 1024
        while True:
            out = np.empty(cap, dtype=np.int32)
            check(lib().psivox_elem_index_query(self._h, C.byref(Q), out.ctypes.data, cap, C.byref(n), stream))
            if n.value <= cap:
                return out[:n.value].copy()
            cap = n.value

    def sample(self, x, refine_tolerance=1.0, refine_max_lvl=0, stream=None):
        """(rho, vel) streams at x, as PSI.VDF."""
        import numpy as np
        X = (C.c_double * 3)(*map(float, x))
        n = C.c_int64(0)
        cap = 256
        while True:
            rho, vel = np.empty(cap), np.empty((cap, 3))
            check(lib().psivox_sample_vdf(self._h, X, refine_tolerance, refine_max_lvl, rho.ctypes.data,
                                          vel.ctypes.data, cap, C.byref(n), stream))
            if n.value <= cap:
                return rho[:n.value].copy(), vel[:n.value].copy()
            cap = n.value

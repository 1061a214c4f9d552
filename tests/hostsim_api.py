"""ctypes wrapper of tests/hostsim (the kernel algorithm built for the host)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hostsim", "libhostsim.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostsim")], check=True)
        _lib = C.CDLL(LIB)
    return _lib


def run(P, kernel="fast", reftol=1.0, maxlvl=0, xvfix=False, out=None):
    """P: an oracle.Problem; returns (fields, stats[nfrag, ncopies, ntets, fails/status, column tasks])."""
    out = P.zero_fields() if out is None else out
    order = ["m", "x", "v", "xx", "xv", "vv"]
    fptr = (C.c_void_p * 6)(*[out[f].ctypes.data if f in out else None for f in order])
    st = np.zeros(5, np.int64)
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)
    fn = lib().hs_voxels_fast if kernel == "fast" else lib().hs_voxels_plain
    if kernel == "fast":
        fn(C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel), vp(P.mass), vp(P.conn),
           C.c_int(int(P.periodic)), box, n, win, d, fptr, C.c_int(int(xvfix)), C.c_double(reftol), C.c_int(maxlvl),
           vp(st))
    else:
        fn(C.c_int(P.elemtype), C.c_int64(P.pos.shape[0]), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel),
           vp(P.mass), vp(P.conn), C.c_int(int(P.periodic)), box, n, win, d, fptr, C.c_int(int(xvfix)),
           C.c_double(reftol), C.c_int(maxlvl), vp(st))
    return out, st


def mass_adjoint(P, w):
    """dL/dpos (npart, 3) of the column-kernel adjoint for w = dL/dm."""
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(P.n)
    out = np.zeros((P.pos.shape[0], 3))
    st = np.zeros(2, np.int64)
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)
    lib().hs_mass_adjoint(C.c_int(P.elemtype), C.c_int64(P.pos.shape[0]), C.c_int64(P.conn.shape[0]), vp(P.pos),
                          vp(P.mass), vp(P.conn), C.c_int(int(P.periodic)), box, n, win, d, vp(w), vp(out), vp(st))
    return out, st


def sample_vdf(P, x, reftol=1.0, maxlvl=0, window=None):
    """(rho, vel) of the device VDF path run on the host; window defaults to
    the mesh's box (as PSI.VDF without a tree)."""
    L = lib()
    L.hs_sample_vdf.restype = C.c_int64
    win = (C.c_double * 6)(*(P.box if window is None else np.asarray(window, dtype=np.float64)).reshape(-1))
    box = (C.c_double * 6)(*P.box.reshape(-1))
    X = (C.c_double * 3)(*[float(v) for v in x])
    vp = lambda a: C.c_void_p(a.ctypes.data)
    args = (C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel), vp(P.mass), vp(P.conn),
            C.c_int(int(P.periodic)), box, win, X, C.c_double(reftol), C.c_int(maxlvl))
    n = L.hs_sample_vdf(*args, None, None, C.c_int64(0))
    rho, vel = np.empty(n), np.empty((n, 3))
    L.hs_sample_vdf(*args, vp(rho), vp(vel), C.c_int64(n))
    return rho, vel


def annihilation(P, reftol=1.0, maxlvl=0, out=None):
    """mode='annihilation' with the device's pair code on the host: (fields,
    fragments, tet pairs)."""
    L = lib()
    L.hs_annihilation.restype = C.c_int64
    out = P.zero_fields() if out is None else out
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)
    pairs = C.c_int64(0)
    nf = L.hs_annihilation(C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.mass), vp(P.conn),
                           C.c_int(int(P.periodic)), box, n, win, d, vp(out["m"]), C.c_double(reftol),
                           C.c_int(maxlvl), C.byref(pairs))
    return out, nf, pairs.value


def _grid_args(P, out):
    order = ["m", "x", "v", "xx", "xv", "vv"]
    fptr = (C.c_void_p * 6)(*[out[f].ctypes.data if f in out else None for f in order])
    return (C.c_int * 3)(*P.n), (C.c_double * 6)(*P.bounds.reshape(-1)), (C.c_double * 3)(*P.d), fptr


def cells_to_raw(P, fields, xvfix=False):
    """the additive raw moments (nvox, nacc) of a grid's fields"""
    n, win, d, fptr = _grid_args(P, fields)
    nacc = 1 + 3 * ("x" in fields) + 3 * ("v" in fields) + 6 * ("xx" in fields) + 9 * ("xv" in fields) + \
        6 * ("vv" in fields)
    acc = np.zeros((int(np.prod(P.n)), nacc))
    lib().hs_cells_to_raw(n, win, d, fptr, C.c_int(int(xvfix)), C.c_void_p(acc.ctypes.data))
    return acc


def cells_merge_finalize(P, fields, acc, xvfix=False):
    """merge new raw moments acc into the fields in place (the end-of-call merge)"""
    n, win, d, fptr = _grid_args(P, fields)
    acc = np.ascontiguousarray(acc, dtype=np.float64).copy()
    lib().hs_cells_merge_finalize(n, win, d, fptr, C.c_int(int(xvfix)), C.c_void_p(acc.ctypes.data))
    return fields


def adjoint(P, W):
    """(dpos, dvel) of the device adjoint of m, v, vv run on the host, for
    the cells' weights W (nvox, 13)."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    dpos = np.zeros((P.pos.shape[0], 3))
    dvel = np.zeros((P.pos.shape[0], 3))
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)
    fails = lib().hs_adjoint(C.c_int(P.elemtype), C.c_int64(P.pos.shape[0]), C.c_int64(P.conn.shape[0]), vp(P.pos),
                             vp(P.vel), vp(P.mass), vp(P.conn), C.c_int(int(P.periodic)), box, n, win, d, vp(W),
                             vp(dpos), vp(dvel))
    return dpos, dvel, fails


def expand_params(P, reftol=1.0, maxlvl=0):
    """Check every refined tet's parameter record against its vertices (the
    element's map at the vertex parameters): (max position error, max
    velocity error, tet count)."""
    out = np.zeros(3)
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    rc = lib().hs_expand_params(C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel), vp(P.mass),
                                vp(P.conn), C.c_int(int(P.periodic)), box, n, win, d, C.c_double(reftol),
                                C.c_int(maxlvl), vp(out))
    if rc != 0:
        raise RuntimeError("parameter record names the wrong element")
    return out[0], out[1], int(out[2])


def expand_records(P, reftol=1.0, maxlvl=0, cap=1 << 20):
    """The refined expansion: (tets (nt, 28), params (nt, 3) uint64)."""
    tets = np.zeros((cap, 28))
    prm = np.zeros((cap, 3), np.uint64)
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    f = lib().hs_expand_records
    f.restype = C.c_int64
    nt = f(C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel), vp(P.mass), vp(P.conn),
           C.c_int(int(P.periodic)), box, n, win, d, C.c_double(reftol), C.c_int(maxlvl), C.c_int64(cap), vp(tets),
           vp(prm))
    if nt < 0:
        raise RuntimeError("expansion larger than cap")
    return tets[:nt], prm[:nt]


def ytable_check(P, reftol=1.0, maxlvl=0, fp64=False):
    """Every column task built with the y-slab table and with the generic
    clips alone: (tasks, tasks whose y planes cut the slab polyhedron,
    mismatches)."""
    out = np.zeros(3, np.int64)
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    lib().hs_ytable_check(C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel), vp(P.mass),
                          vp(P.conn), C.c_int(int(P.periodic)), box, n, win, d, C.c_double(reftol), C.c_int(maxlvl),
                          C.c_int(int(fp64)), vp(out))
    return tuple(int(x) for x in out)


def zext_check(P):
    """The ordering's sort key (column_cells) against the built columns'
    cell counts over every FP32 column task: (tasks, equal, key below, key
    above)."""
    out = np.zeros(4, np.int64)
    n = (C.c_int * 3)(*P.n)
    win = (C.c_double * 6)(*P.bounds.reshape(-1))
    d = (C.c_double * 3)(*P.d)
    box = (C.c_double * 6)(*P.box.reshape(-1))
    vp = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    lib().hs_zext_check(C.c_int(P.elemtype), C.c_int64(P.conn.shape[0]), vp(P.pos), vp(P.vel), vp(P.mass),
                        vp(P.conn), C.c_int(int(P.periodic)), box, n, win, d, vp(out))
    return tuple(int(x) for x in out)

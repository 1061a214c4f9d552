"""ctypes access to the oracle -- TEST INFRASTRUCTURE ONLY.

liboracle.so is the FP64 C restatement of PSI.voxels (psi_oracle.c);
_ref/libpsiref.so, when it was built, is the reference's own C core compiled
from its sources (Makefile target `ref`).  Only tests/, the smoke check and the
CPU-baseline leg of bench.py import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libpsiref.so")

FIELD_WIDTH = {"m": 1, "x": 3, "v": 3, "xx": 9, "xv": 9, "vv": 9}
FIELD_ORDER = ["m", "x", "v", "xx", "xv", "vv"]


def build(ref: bool = True) -> None:
    targets = ["liboracle.so"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class _OMesh(C.Structure):
    _fields_ = [("elemtype", C.c_int), ("npart", C.c_int64), ("nelem", C.c_int64),
                ("pos", C.c_void_p), ("vel", C.c_void_p), ("mass", C.c_void_p), ("conn", C.c_void_p),
                ("periodic", C.c_int), ("box", (C.c_double * 3) * 2)]


class _OGrid(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("win", (C.c_double * 3) * 2), ("d", C.c_double * 3),
                ("f", C.c_void_p * 6)]


class OStats(C.Structure):
    _fields_ = [("nfrag", C.c_int64), ("ntets", C.c_int64), ("nghosts", C.c_int64), ("nelem_used", C.c_int64),
                ("maxverts", C.c_int), ("maxstack", C.c_int), ("overflow", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.orc_voxels.argtypes = [P, P, C.c_double, C.c_int, C.c_int, P]
        L.orc_voxels_range.argtypes = [P, P, C.c_double, C.c_int, C.c_int, C.c_int64, C.c_int64, P]
        L.orc_voxels_omp.argtypes = [P, P, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, P]
        L.orc_list_fragments.argtypes = [P, P, C.c_double, C.c_int, C.c_int64, C.c_int64, P, P, C.c_int64]
        L.orc_list_fragments.restype = C.c_int64
        L.orc_count_tets.argtypes = [P, P, C.c_double, C.c_int, P]
        L.orc_voxelize_one_tet.argtypes = [P, P, P, P, C.c_double, C.c_int, P]
        L.orc_mass_adjoint.argtypes = [P, P, P, P]
        L.orc_fd_mass_loss.argtypes = [P, P, P, P, C.c_int, C.c_int64, C.c_int, C.c_double, P]
        L.orc_adjoint.argtypes = [P, P, P, P, P]
        L.orc_fd_fields_loss.argtypes = [P, P, P, P, P, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_double, P]
        L.orc_annihilation.argtypes = [P, P, C.c_double, C.c_int, P]
        L.orc_elem_query.argtypes = [P, P, P, P, C.c_int64]
        L.orc_elem_query.restype = C.c_int64
        L.orc_sample_vdf.argtypes = [P, P, P, C.c_double, C.c_int, P, P, C.c_int64]
        L.orc_sample_vdf.restype = C.c_int64
        _lib = L
    return _lib


def block_connectivity(shape, periodic: bool) -> np.ndarray:
    """Trilinear connectivity of a (nx,ny,nz) particle block, corner order
    4i+2j+k (src/PSI.c:636-656)."""
    nx, ny, nz = shape
    ex, ey, ez = (nx, ny, nz) if periodic else (nx - 1, ny - 1, nz - 1)
    i, j, k = np.meshgrid(np.arange(ex), np.arange(ey), np.arange(ez), indexing="ij")
    conn = np.empty((ex, ey, ez, 8), dtype=np.int32)
    for ii in range(2):
        for jj in range(2):
            for kk in range(2):
                conn[..., 4 * ii + 2 * jj + kk] = (ny * nz * ((i + ii) % nx) + nz * ((j + jj) % ny)
                                                   + ((k + kk) % nz))
    return conn.reshape(-1, 8)


class Problem:
    """A mesh + grid description shared by the oracle and the reference build."""

    def __init__(self, pos, vel, mass, conn, periodic_box=None, n=(32, 32, 32),
                 bounds=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), fields=()):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(vel if vel is not None else np.zeros_like(self.pos),
                                        dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.elemtype = self.conn.shape[1]
        nelem = self.conn.shape[0]
        self.mass = np.ascontiguousarray(mass if mass is not None else np.ones(nelem),
                                         dtype=np.float64).reshape(-1)
        self.periodic = periodic_box is not None
        self.box = np.array(periodic_box if periodic_box is not None else ((0, 0, 0), (0, 0, 0)), dtype=np.float64)
        self.n = tuple(int(v) for v in n)
        self.bounds = np.array(bounds, dtype=np.float64)
        self.d = tuple(float((self.bounds[1, a] - self.bounds[0, a]) / self.n[a]) for a in range(3))
        self.fields = ["m"] + [f for f in fields if f != "m"]

    @classmethod
    def block(cls, pos, vel=None, mass=None, periodic_box=None, **kw):
        shape = pos.shape[:3]
        conn = block_connectivity(shape, periodic_box is not None)
        if mass is not None:
            mass = np.asarray(mass).reshape(-1)
        return cls(pos.reshape(-1, 3), None if vel is None else vel.reshape(-1, 3), mass, conn,
                   periodic_box=periodic_box, **kw)

    def zero_fields(self):
        out = {}
        for f in self.fields:
            w = FIELD_WIDTH[f]
            shp = self.n if w == 1 else self.n + ((3,) if w == 3 else (3, 3))
            out[f] = np.zeros(shp, dtype=np.float64)
        return out

    def _omesh(self):
        m = _OMesh()
        m.elemtype = self.elemtype
        m.npart = self.pos.shape[0]
        m.nelem = self.conn.shape[0]
        m.pos = self.pos.ctypes.data
        m.vel = self.vel.ctypes.data
        m.mass = self.mass.ctypes.data
        m.conn = self.conn.ctypes.data
        m.periodic = int(self.periodic)
        for s in range(2):
            for a in range(3):
                m.box[s][a] = self.box[s, a]
        return m

    def _ogrid(self, out):
        g = _OGrid()
        for a in range(3):
            g.n[a] = self.n[a]
            g.win[0][a] = self.bounds[0, a]
            g.win[1][a] = self.bounds[1, a]
            g.d[a] = self.d[a]
        for i, f in enumerate(FIELD_ORDER):
            g.f[i] = out[f].ctypes.data if f in out else None
        return g


def voxels(prob: Problem, reftol=1.0, maxlvl=0, xvfix=False, out=None, threads=0, e0=0, e1=None):
    """Run the oracle.  threads=0: serial; threads>0 or -1: OpenMP."""
    L = lib()
    out = prob.zero_fields() if out is None else out
    m, g, st = prob._omesh(), prob._ogrid(out), OStats()
    e1 = prob.conn.shape[0] if e1 is None else e1
    if threads:
        rc = L.orc_voxels_omp(C.byref(m), C.byref(g), reftol, maxlvl, int(xvfix), int(threads), e0, e1,
                              C.byref(st))
    else:
        rc = L.orc_voxels_range(C.byref(m), C.byref(g), reftol, maxlvl, int(xvfix), e0, e1, C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle failed")
    return out, st.as_dict()


def list_fragments(prob: Problem, reftol=1.0, maxlvl=0, e0=0, e1=None, cap=1 << 22):
    L = lib()
    m, g = prob._omesh(), prob._ogrid(prob.zero_fields())
    e1 = prob.conn.shape[0] if e1 is None else e1
    gi = np.empty(cap, dtype=np.int64)
    mm = np.empty(cap, dtype=np.float64)
    n = L.orc_list_fragments(C.byref(m), C.byref(g), reftol, maxlvl, e0, e1, gi.ctypes.data, mm.ctypes.data, cap)
    if n > cap:
        return list_fragments(prob, reftol, maxlvl, e0, e1, cap=int(n))
    return gi[:n].copy(), mm[:n].copy()


def mass_adjoint(prob: Problem, w):
    """FP64 dL/dpos (npart, 3) for w = dL/dm on the grid (refine level 0)."""
    L = lib()
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(prob.n)
    m, g = prob._omesh(), prob._ogrid({})
    out = np.zeros((prob.pos.shape[0], 3))
    L.orc_mass_adjoint(C.byref(m), C.byref(g), w.ctypes.data, out.ctypes.data)
    return out


def particle_elements(prob: Problem):
    """CSR map particle -> elements that use it: (start, elems)."""
    conn = prob.conn.astype(np.int64)
    pid = conn.reshape(-1)
    eid = np.repeat(np.arange(conn.shape[0], dtype=np.int64), conn.shape[1])
    order = np.lexsort((eid, pid))
    pid, eid = pid[order], eid[order]
    keep = np.ones(pid.size, bool)
    keep[1:] = (pid[1:] != pid[:-1]) | (eid[1:] != eid[:-1])
    pid, eid = pid[keep], eid[keep]
    start = np.searchsorted(pid, np.arange(prob.pos.shape[0] + 1))
    return start, np.ascontiguousarray(eid)


def fd_mass_loss_grad(prob: Problem, r, particles, axes, h=1e-6, csr=None):
    """FP64 central differences of L = sum_C (m_C/V - rho_t)^2 with respect to
    pos[p, a] for each (p, a) pair; r = m/V - rho_t of the unperturbed mesh."""
    L = lib()
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(prob.n)
    start, elems = csr if csr is not None else particle_elements(prob)
    m, g = prob._omesh(), prob._ogrid({})
    out = np.zeros(len(particles))
    val = C.c_double()
    for i, (p, a) in enumerate(zip(particles, axes)):
        s0, s1 = int(start[p]), int(start[p + 1])
        rc = L.orc_fd_mass_loss(C.byref(m), C.byref(g), r.ctypes.data, elems[s0:].ctypes.data, s1 - s0, int(p),
                                int(a), float(h), C.byref(val))
        if rc != 0:
            raise RuntimeError("oracle finite difference failed")
        out[i] = val.value
    return out


def adjoint_weights(fields, grads):
    """Per-cell weights dL/d(M, Sv, Svv), (nvox, 13), from the forward's
    fields {'m', 'v', 'vv'} and the gradients {'m', 'v', 'vv'} of a loss
    with respect to them (m = M, v = Sv / M, vv = Svv / M - v v^T; a cell
    without mass keeps dL/dM = dL/dm only)."""
    m = fields["m"].reshape(-1)
    n = m.size
    W = np.zeros((n, 13))
    gm = grads.get("m")
    if gm is not None:
        W[:, 0] = gm.reshape(-1)
    ok = m != 0
    iM = np.where(ok, 1.0 / np.where(ok, m, 1.0), 0.0)
    v = fields["v"].reshape(n, 3) if "v" in fields else np.zeros((n, 3))
    gv = grads.get("v")
    if gv is not None:
        gv = gv.reshape(n, 3)
        W[:, 1:4] += gv * iM[:, None]
        W[:, 0] -= np.sum(gv * v, axis=1) * iM
    gvv = grads.get("vv")
    if gvv is not None:
        G = gvv.reshape(n, 3, 3)
        vv = fields["vv"].reshape(n, 3, 3)
        W[:, 4:] = (G * iM[:, None, None]).reshape(n, 9)
        W[:, 1:4] -= np.einsum("nab,nb->na", G + np.swapaxes(G, 1, 2), v) * iM[:, None]
        W[:, 0] += np.einsum("nab,nab->n", G, v[:, :, None] * v[:, None, :] - vv) * iM
    W[~ok, 1:] = 0.0
    return W


def adjoint(prob: Problem, W):
    """(dL/dpos, dL/dvel) of the fields m, v, vv from the cells' weights
    W (nvox, 13), FP64, refine level 0."""
    L = lib()
    W = np.ascontiguousarray(W, dtype=np.float64)
    dpos = np.zeros((prob.pos.shape[0], 3))
    dvel = np.zeros((prob.pos.shape[0], 3))
    m, g = prob._omesh(), prob._ogrid(prob.zero_fields())
    if L.orc_adjoint(C.byref(m), C.byref(g), W.ctypes.data, dpos.ctypes.data, dvel.ctypes.data) != 0:
        raise RuntimeError("oracle adjoint failed")
    return dpos, dvel


def raw_sums(fields):
    """(nvox, 10) additive sums M, Sv[3], Svv (xx xy xz yy yz zz) of the
    finalized fields m, v, vv."""
    m = fields["m"].reshape(-1)
    n = m.size
    v = fields["v"].reshape(n, 3)
    vv = fields["vv"].reshape(n, 3, 3)
    S2 = (vv + v[:, :, None] * v[:, None, :]) * m[:, None, None]
    out = np.empty((n, 10))
    out[:, 0] = m
    out[:, 1:4] = v * m[:, None]
    out[:, 4:] = S2[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    return out


def fd_fields_loss_grad(prob: Problem, base, G, particles, axes, which, h=1e-6, csr=None):
    """Central differences of sum_C G_m m + G_v . v + G_vv : vv (G: (nvox,
    13) coefficients; base: raw_sums of the unperturbed fields) for the
    position (which 0) or velocity (which 1) components (particles[i],
    axes[i])."""
    L = lib()
    start, elems = csr if csr is not None else particle_elements(prob)
    base = np.ascontiguousarray(base, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64)
    m, g = prob._omesh(), prob._ogrid(prob.zero_fields())
    out = np.empty(len(particles))
    val = C.c_double()
    for i, (p, a) in enumerate(zip(particles, axes)):
        el = np.ascontiguousarray(elems[start[p]:start[p + 1]])
        rc = L.orc_fd_fields_loss(C.byref(m), C.byref(g), base.ctypes.data, G.ctypes.data, el.ctypes.data, len(el),
                                  C.c_int64(int(p)), C.c_int(int(a)), C.c_int(int(which)), C.c_double(h),
                                  C.byref(val))
        if rc:
            raise RuntimeError("oracle finite differences failed")
        out[i] = val.value
    return out


def count_tets(prob: Problem, reftol=1.0, maxlvl=0):
    L = lib()
    m, g = prob._omesh(), prob._ogrid(prob.zero_fields())
    counts = np.zeros(prob.conn.shape[0], dtype=np.int32)
    L.orc_count_tets(C.byref(m), C.byref(g), reftol, maxlvl, counts.ctypes.data)
    return counts


# ---------------------------------------------------------------------------
# the reference's own C core (oracle/_ref), driven through its psi_voxels()
# ---------------------------------------------------------------------------

class _RGrid(C.Structure):  # include/grid.h:46-53
    _fields_ = [("type", C.c_int), ("dim", C.c_int), ("fields", C.c_void_p * 16),
                ("window", (C.c_double * 3) * 2), ("n", C.c_int * 3), ("d", C.c_double * 3)]


class _RMesh(C.Structure):  # include/mesh.h:42-52
    _fields_ = [("elemtype", C.c_int), ("dim", C.c_int), ("pos", C.c_void_p), ("vel", C.c_void_p),
                ("mass", C.c_void_p), ("npart", C.c_int), ("nelem", C.c_int), ("connectivity", C.c_void_p),
                ("box", (C.c_double * 3) * 2), ("periodic", C.c_int)]


_ref = None


def have_ref() -> bool:
    return os.path.exists(REFLIB)


def reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(REFLIB)
        L.psi_voxels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.psi_voxels.restype = None
        _ref = L
    return _ref


class _silence_stdout:
    """The reference prints a progress line to stdout; keep it off fd 1."""

    def __enter__(self):
        import sys
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        C.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def annihilation(prob: Problem, reftol=1.0, maxlvl=0, out=None):
    """mode='annihilation': the pairwise overlap rate added into out['m']."""
    L = lib()
    out = prob.zero_fields() if out is None else out
    m, g, st = prob._omesh(), prob._ogrid(out), OStats()
    if L.orc_annihilation(C.byref(m), C.byref(g), reftol, maxlvl, C.byref(st)) != 0:
        raise RuntimeError("oracle failed")
    return out, st.as_dict()


def ref_voxels(prob: Problem, reftol=1.0, maxlvl=0, out=None, e0=0, e1=None, mode=0):
    """Run the reference psi_voxels() (density mode, or mode=1 for
    annihilation) on the same problem, optionally on the element subset
    [e0, e1)."""
    L = reflib()
    out = prob.zero_fields() if out is None else out
    conn = prob.conn
    mass = prob.mass
    if e0 != 0 or e1 is not None:
        conn = np.ascontiguousarray(conn[e0:e1])
        mass = np.ascontiguousarray(mass[e0:e1])
    m = _RMesh()
    m.elemtype = prob.elemtype
    m.dim = 3
    m.pos = prob.pos.ctypes.data
    m.vel = prob.vel.ctypes.data
    m.mass = mass.ctypes.data
    m.npart = prob.pos.shape[0]
    m.nelem = conn.shape[0]
    m.connectivity = conn.ctypes.data
    m.periodic = int(prob.periodic)
    for s in range(2):
        for a in range(3):
            m.box[s][a] = prob.box[s, a]
    g = _RGrid()
    g.type = 0
    g.dim = 3
    for i, f in enumerate(FIELD_ORDER):
        g.fields[i] = out[f].ctypes.data if f in out else None
    for a in range(3):
        g.window[0][a] = prob.bounds[0, a]
        g.window[1][a] = prob.bounds[1, a]
        g.n[a] = prob.n[a]
        g.d[a] = prob.d[a]
    with _silence_stdout():
        L.psi_voxels(C.byref(g), C.byref(m), None, mode, reftol, maxlvl)
    return out


def ref_load_snapshot(filename: str, loader: str = "gadget2"):
    """The reference's own snapshot loader (src/mesh.c:35-302, peek then
    load, as src/PSI.c:665-702) on a file: (pos, vel, mass, connectivity,
    box) as it leaves them, or None when it reports failure."""
    L = reflib()
    for fn in (L.peek_gadget2, L.load_gadget2, L.load_gevolution):
        fn.argtypes = [C.c_void_p, C.c_char_p]
        fn.restype = C.c_int
    m = _RMesh()
    name = filename.encode()
    with _silence_stdout():
        if not L.peek_gadget2(C.byref(m), name):
            return None
    n = m.npart
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    conn = np.empty((m.nelem, m.elemtype), dtype=np.int32)
    m.pos, m.vel, m.mass, m.connectivity = (pos.ctypes.data, vel.ctypes.data, mass.ctypes.data,
                                            conn.ctypes.data)
    load = L.load_gevolution if loader == "gevolution" else L.load_gadget2
    with _silence_stdout():
        if not load(C.byref(m), name):
            return None
    box = np.array([[m.box[s][a] for a in range(3)] for s in range(2)])
    return pos, vel, mass, conn, box


# ---------------------------------------------------------------------------
# point queries: element boxes and the multistream sample (PSI.VDF)
# ---------------------------------------------------------------------------

def _box2(w):
    W = ((C.c_double * 3) * 2)()
    for s in range(2):
        for a in range(3):
            W[s][a] = float(w[s][a])
    return W


def tree_window(prob: Problem):
    """The window RStarTree(mesh) builds its tree with (src/PSI.c:875-886):
    the periodic box, or +-1e30 for a non-periodic mesh."""
    if prob.periodic:
        return prob.box.copy()
    return np.array([[-1.0e30] * 3, [1.0e30] * 3])


def vdf_window(prob: Problem):
    """The window psi_sample_vdf builds its internal tree with
    (src/cpsi.c:203-208): the mesh's box, all zeros for a non-periodic mesh."""
    return prob.box.copy()


def elem_query(prob: Problem, window, q):
    """Elements whose wrapped box meets the closed box q, ascending."""
    L = lib()
    m = prob._omesh()
    W, Q = _box2(window), _box2(q)
    n = L.orc_elem_query(C.byref(m), C.byref(W), C.byref(Q), None, 0)
    out = np.empty(n, dtype=np.int32)
    L.orc_elem_query(C.byref(m), C.byref(W), C.byref(Q), out.ctypes.data, n)
    return out


def sample_vdf(prob: Problem, x, reftol=1.0, maxlvl=0, window=None):
    """(rho, vel) streams at x, ascending element order."""
    L = lib()
    m = prob._omesh()
    W = _box2(vdf_window(prob) if window is None else window)
    X = (C.c_double * 3)(*[float(v) for v in x])
    n = L.orc_sample_vdf(C.byref(m), C.byref(W), X, reftol, maxlvl, None, None, 0)
    rho = np.empty(n)
    vel = np.empty((n, 3))
    L.orc_sample_vdf(C.byref(m), C.byref(W), X, reftol, maxlvl, rho.ctypes.data, vel.ctypes.data, n)
    return rho, vel


class _RVec(C.Structure):  # psi_rvec, include/psi.h:84-87
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


def _rmesh(prob: Problem):
    m = _RMesh()
    m.elemtype = prob.elemtype
    m.dim = 3
    m.pos = prob.pos.ctypes.data
    m.vel = prob.vel.ctypes.data
    m.mass = prob.mass.ctypes.data
    m.npart = prob.pos.shape[0]
    m.nelem = prob.conn.shape[0]
    m.connectivity = prob.conn.ctypes.data
    m.periodic = int(prob.periodic)
    for s in range(2):
        for a in range(3):
            m.box[s][a] = prob.box[s, a]
    return m


def ref_sample_vdf(prob: Problem, x, reftol=1.0, maxlvl=0):
    """The reference's psi_sample_vdf (src/cpsi.c:177-275) with its internal
    R*-tree: (rho, vel) in its tree's element order."""
    L = reflib()
    P = C.c_void_p
    L.psi_sample_vdf.argtypes = [_RVec, P, P, P, P, P, C.c_double, C.c_int]
    L.psi_sample_vdf.restype = None
    m = _rmesh(prob)
    rp, vp = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    n = C.c_int(0)
    with _silence_stdout():
        L.psi_sample_vdf(_RVec(*[float(v) for v in x]), C.byref(m), None, C.byref(rp), C.byref(vp), C.byref(n),
                         reftol, maxlvl)
    k = n.value
    rho = np.array([rp[i] for i in range(k)])
    vel = np.array([vp[i] for i in range(3 * k)]).reshape(k, 3)
    libc = C.CDLL(None)
    libc.free.argtypes = [P]
    libc.free(C.cast(rp, P))
    libc.free(C.cast(vp, P))
    return rho, vel


def ref_tree_query(prob: Problem, window, q):
    """The reference's R*-tree (src/rtree.c) built from the mesh with window
    and queried with the closed box q: element indices in its order."""
    L = reflib()
    P = C.c_void_p
    L.psi_rtree_from_mesh.argtypes = [P, P, P]
    L.psi_rtree_query_init.argtypes = [P, P, P]
    L.psi_rtree_query_next.argtypes = [P, P]
    L.psi_rtree_query_next.restype = C.c_int
    L.psi_rtree_destroy.argtypes = [P]
    m = _rmesh(prob)
    tree = (C.c_char * 256)()  # psi_rtree, include/rtree.h:53-60 (32 bytes)
    qry = (C.c_char * 16384)()  # psi_rtree_query: a 1024-entry stack and two pointers
    W = (_RVec * 2)(_RVec(*window[0]), _RVec(*window[1]))
    Q = (_RVec * 2)(_RVec(*q[0]), _RVec(*q[1]))
    with _silence_stdout():
        L.psi_rtree_from_mesh(tree, C.byref(m), W)
    out = []
    e = C.c_int(0)
    L.psi_rtree_query_init(qry, tree, Q)
    while L.psi_rtree_query_next(qry, C.byref(e)):
        out.append(e.value)
    L.psi_rtree_destroy(tree)
    return np.array(out, dtype=np.int32)


class RefTree:
    """The reference's R*-tree (src/rtree.c:48-73) built once over a mesh,
    then psi_sample_vdf (src/cpsi.c:177-275) answered through it, as the
    reference's PSI.VDF(mesh, x, tree=t) does (src/PSI.c:1022-1058): used
    to time the reference's own query path."""

    def __init__(self, prob: Problem, window):
        L = reflib()
        P = C.c_void_p
        L.psi_rtree_from_mesh.argtypes = [P, P, P]
        L.psi_rtree_destroy.argtypes = [P]
        L.psi_sample_vdf.argtypes = [_RVec, P, P, P, P, P, C.c_double, C.c_int]
        L.psi_sample_vdf.restype = None
        self._L = L
        self._prob = prob
        self._m = _rmesh(prob)
        self._tree = (C.c_char * 256)()  # psi_rtree, include/rtree.h:53-60 (32 bytes)
        W = (_RVec * 2)(_RVec(*window[0]), _RVec(*window[1]))
        with _silence_stdout():
            L.psi_rtree_from_mesh(self._tree, C.byref(self._m), W)
        libc = C.CDLL(None)
        libc.free.argtypes = [P]
        self._free = libc.free

    def sample_vdf(self, x, reftol=1.0, maxlvl=0):
        rp, vp = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        n = C.c_int(0)
        self._L.psi_sample_vdf(_RVec(*[float(v) for v in x]), C.byref(self._m), self._tree, C.byref(rp), C.byref(vp),
                               C.byref(n), reftol, maxlvl)
        k = n.value
        rho = np.array([rp[i] for i in range(k)])
        self._free(C.cast(rp, C.c_void_p))
        self._free(C.cast(vp, C.c_void_p))
        return rho

    def close(self):
        if self._tree is not None:
            self._L.psi_rtree_destroy(self._tree)
            self._tree = None

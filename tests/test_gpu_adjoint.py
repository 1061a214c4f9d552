"""GPU adjoints (psivox_mass_adjoint and psivox_adjoint through the C ABI,
and the torch.autograd wrappers) against the FP64 oracle adjoint and FP64
central finite differences.

Tolerances: the adjoint kernel (FP64 throughout), fed the same dL/dm as the
oracle, within 1e-9 of the largest gradient component; the 64^3 -> 128^3 workload (rms 1.0 cells, loss
sum (rho - rho_t)^2 with rho_t from seed 1) on 1000 seeded (vertex, axis)
pairs: |g - g_fd| / max(|g_fd|, 1e-3 rms(g_fd)) <= 1e-4 against FP64
central differences with h = 1e-6 box units, for the adjoint fed the
oracle's FP64 dL/dm and for the end-to-end autograd path with the FP64
forward, and the FP32-clip forward (its cell masses carry FP32 error into
dL/dm: 7.2e-5 measured).
"""
import numpy as np
import pytest

from adjoint_cases import BOX, cases, loss_weights, refined_fd, refined_problem, sample

pytestmark = pytest.mark.gpu


def gpu_adjoint(P, w):
    import torch
    from pypsi_amd import device as D
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(P.pos).to(dev)
    mass = torch.from_numpy(P.mass).to(dev)
    conn = torch.from_numpy(P.conn).to(dev)
    mesh = D.make_mesh(pos, None, mass, conn, P.box if P.periodic else None)
    grid = D.make_grid(P.n, P.bounds, {"m": torch.empty(0, dtype=torch.float64, device=dev)})
    tw = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    out = torch.zeros_like(pos)
    D.mass_adjoint(mesh, grid, tw, out, D.make_opts(stream=torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_adjoint_zero_volume_tets(oracle_mod):
    """Tets with a repeated vertex have zero volume and are dropped, as the
    oracle drops them.  With two equal edge columns the determinant's
    cofactor products cancel exactly only when rounded alike: fused
    multiply-adds there gave determinants of order 1e-20 and gradients of
    order 1e7 (found by the host model once it was built with the device's
    contraction); the determinant is now formed without them."""
    print("case: adjoint with zero-volume tets", flush=True)
    rng = np.random.default_rng(42)
    p = rng.uniform(0.05, 0.95, size=(40, 3))
    conn = rng.integers(0, 40, size=(30, 4)).astype(np.int32)
    deg = np.array([len(set(c)) < 4 for c in conn])
    assert deg.any()
    P = oracle_mod.Problem(p, None, rng.uniform(0.5, 2.0, 30), conn, n=(12, 10, 14), bounds=BOX)
    r, w = loss_weights(oracle_mod, P)
    gg = gpu_adjoint(P, w)
    go = oracle_mod.mass_adjoint(P, w)
    err = np.abs(gg - go).max() / np.abs(go).max()
    print("gpu vs oracle %.2e" % err, flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("idx", range(5))
def test_adjoint_vs_oracle(oracle_mod, idx):
    label, P = cases(oracle_mod)[idx]
    print("case: adjoint", label, flush=True)
    r, w = loss_weights(oracle_mod, P)
    go = oracle_mod.mass_adjoint(P, w)
    gg = gpu_adjoint(P, w)
    err = np.abs(gg - go).max() / np.abs(go).max()
    print(label, "gpu adjoint vs oracle: %.2e" % err)
    assert err < 1e-9


def test_adjoint_accumulates(oracle_mod):
    """dl_dpos is added to, not overwritten."""
    import torch
    from pypsi_amd import device as D
    label, P = cases(oracle_mod)[0]
    print("case: accumulate", flush=True)
    r, w = loss_weights(oracle_mod, P)
    dev = torch.device("cuda:0")
    # the descriptors hold raw device pointers: keep every tensor alive
    pos = torch.from_numpy(P.pos).to(dev)
    mass = torch.from_numpy(P.mass).to(dev)
    conn = torch.from_numpy(P.conn).to(dev)
    mesh = D.make_mesh(pos, None, mass, conn, P.box)
    grid = D.make_grid(P.n, P.bounds, {})
    tw = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    out = torch.ones_like(pos)
    opts = D.make_opts(stream=torch.cuda.current_stream().cuda_stream)
    D.mass_adjoint(mesh, grid, tw, out, opts)
    D.mass_adjoint(mesh, grid, tw, out, opts)
    torch.cuda.synchronize()
    go = oracle_mod.mass_adjoint(P, w)
    assert np.abs(out.cpu().numpy() - 1.0 - 2.0 * go).max() <= 1e-9 * np.abs(go).max()


@pytest.mark.parametrize("kind", ["trilinear", "triquadratic"])
def test_refined_adjoint_vs_fd(oracle_mod, kind):
    """The adjoint on refined meshes: every refined tet's vertex is the
    element's map (trilinear or triquadratic) at a dyadic parameter point,
    so a vertex's gradient reaches the element's nodes with the map's weights
    there; the refinement decisions are piecewise constant and carry none.
    The mass adjoint (loss sum_C w_C m_C) and the fields adjoint (loss
    sum_C G_m m + G_v . v + G_vv : vv, v and vv weighted on cells holding at
    least 1e-3 of the peak mass) through the C ABI, against FP64 central
    differences of the oracle's refined forward (h = 1e-6) on 60 position and
    60 velocity components: within 1e-6 (mass) and 1e-5 (fields) of the
    largest difference."""
    import torch
    from pypsi_amd import device as D
    O = oracle_mod
    print("case: refined adjoint vs fd,", kind, flush=True)
    P, reftol, lvl = refined_problem(O, kind)
    ref, st = O.voxels(P, reftol=reftol, maxlvl=lvl)
    per = 6 if P.conn.shape[1] != 4 else 1
    assert st["ntets"] > per * P.conn.shape[0], "the case must refine"
    rng = np.random.default_rng(11)
    npart = P.pos.shape[0]
    comps = [(int(p), int(a)) for p, a in zip(rng.integers(0, npart, 60), rng.integers(0, 3, 60))]
    dev = torch.device("cuda:0")
    tpos, tvel = torch.from_numpy(P.pos).to(dev), torch.from_numpy(P.vel).to(dev)
    mesh = D.make_mesh(tpos, tvel, torch.from_numpy(P.mass).to(dev), torch.from_numpy(P.conn).to(dev),
                       BOX if P.periodic else None)
    opts = D.make_opts(kernel="plain", refine_tolerance=reftol, refine_max_lvl=lvl,
                       stream=torch.cuda.current_stream().cuda_stream)
    # mass: L = sum w m
    w = rng.standard_normal(P.n)
    g = torch.zeros_like(tpos)
    D.mass_adjoint(mesh, D.make_grid(P.n, BOX, {}), torch.from_numpy(w).to(dev), g, opts)
    torch.cuda.synchronize()
    fd = refined_fd(O, P, lambda f: float((w * f["m"]).sum()), 0, comps, reftol, lvl)
    gm = g.cpu().numpy()[tuple(np.array(comps).T)]
    em = np.abs(gm - fd).max() / np.abs(fd).max()
    print("mass adjoint vs fd: %.2e (max |fd| %.3e)" % (em, np.abs(fd).max()), flush=True)
    assert em <= 1e-6
    # the same through torch.autograd (voxel_mass with the refinement settings)
    from pypsi_amd.autograd import voxel_mass
    pa = tpos.clone().requires_grad_(True)
    mm = voxel_mass(pa, torch.from_numpy(P.conn).to(dev), P.n, BOX, mass=torch.from_numpy(P.mass).to(dev),
                    periodic_box=BOX if P.periodic else None, kernel="plain", refine_tolerance=reftol,
                    refine_max_lvl=lvl)
    (mm * torch.from_numpy(w).to(dev)).sum().backward()
    assert torch.allclose(pa.grad, g, rtol=1e-12, atol=1e-12 * float(g.abs().max()))
    # fields: the forward on the GPU (FP64), then the adjoint of m, v, vv
    fl = {k: torch.zeros(v.shape, dtype=torch.float64, device=dev) for k, v in ref.items()}
    grid = D.make_grid(P.n, BOX, fl)
    D.voxels(mesh, grid, opts)
    heavy = ref["m"] >= 1e-3 * ref["m"].max()
    G = {k: rng.standard_normal(ref[k].shape) for k in ("m", "v", "vv")}
    G["v"][~heavy] = 0.0
    G["vv"][~heavy] = 0.0
    tg = {k: torch.from_numpy(v).to(dev) for k, v in G.items()}
    dpos, dvel = torch.zeros_like(tpos), torch.zeros_like(tvel)
    D.adjoint(mesh, grid, tg, dpos, dvel, opts)
    torch.cuda.synchronize()

    def loss(f):
        return float(sum((G[k] * f[k]).sum() for k in G))

    for which, gg, name in ((0, dpos, "positions"), (1, dvel, "velocities")):
        fd = refined_fd(O, P, loss, which, comps, reftol, lvl)
        gv = gg.cpu().numpy()[tuple(np.array(comps).T)]
        e = np.abs(gv - fd).max() / np.abs(fd).max()
        print("fields adjoint vs fd, %s: %.2e (max |fd| %.3e)" % (name, e, np.abs(fd).max()), flush=True)
        assert e <= 1e-5


def _c4(oracle_mod):
    from oracle import block_connectivity
    from pypsi_amd.synth import zeldovich_block
    pos, _ = zeldovich_block(64, 1.0, seed=0)
    post, _ = zeldovich_block(64, 1.0, seed=1)
    P = oracle_mod.Problem.block(pos, None, periodic_box=BOX, n=(128,) * 3, bounds=BOX)
    Pt = oracle_mod.Problem.block(post, None, periodic_box=BOX, n=(128,) * 3, bounds=BOX)
    V = float(np.prod(P.d))
    rho = oracle_mod.voxels(P, threads=-1)[0]["m"] / V
    rhot = oracle_mod.voxels(Pt, threads=-1)[0]["m"] / V
    rng = np.random.default_rng(1000)
    parts = rng.choice(P.pos.shape[0], size=1000, replace=False)
    axes = rng.integers(0, 3, size=1000)
    return P, rho, rhot, parts, axes


def _rel_err(g, fd):
    den = np.maximum(np.abs(fd), 1e-3 * np.sqrt(np.mean(fd ** 2)))
    return np.abs(g - fd) / den


def test_c4_adjoint_vs_fd_and_oracle(oracle_mod):
    """64^3 -> 128^3: adjoint kernel vs FP64 finite differences and vs the
    FP64 oracle adjoint on 1000 seeded (vertex, axis) pairs."""
    print("case: c4 adjoint vs fd", flush=True)
    P, rho, rhot, parts, axes = _c4(oracle_mod)
    V = float(np.prod(P.d))
    r = rho - rhot
    w = 2.0 * r / V
    gg = gpu_adjoint(P, w)[parts, axes]
    fd = oracle_mod.fd_mass_loss_grad(P, r, parts, axes, h=1e-6)
    e_fd = _rel_err(gg, fd)
    # the oracle adjoint on the elements that use the sampled particles
    start, elems = oracle_mod.particle_elements(P)
    sel = np.unique(np.concatenate([elems[start[p]:start[p + 1]] for p in parts]))
    Ps = oracle_mod.Problem(P.pos, None, P.mass[sel], P.conn[sel], periodic_box=BOX, n=P.n, bounds=BOX)
    go = oracle_mod.mass_adjoint(Ps, w)[parts, axes]
    e_or = _rel_err(gg, go)
    e_ofd = _rel_err(go, fd)
    print("max rel err: gpu-fd %.2e, gpu-oracle %.2e, oracle-fd %.2e" % (e_fd.max(), e_or.max(), e_ofd.max()))
    assert e_ofd.max() <= 1e-4
    assert e_fd.max() <= 1e-4
    assert e_or.max() <= 1e-6


@pytest.mark.parametrize("kernel,tol", [("fast64", 1e-4), ("plain", 1e-4), ("auto", 3e-4), ("fast", 3e-4)])
def test_c4_autograd_vs_fd(oracle_mod, kernel, tol):
    """The torch.autograd path end to end (GPU forward, loss in torch,
    backward through the adjoint kernel).  With an FP64 forward (the column
    kernel in FP64, voxel_density's default 'fast64', or the plain kernel)
    the error is the finite differences' own (4.0e-5 measured): bar 1e-4.
    The FP32-clip forwards ('auto', 'fast') carry their cell masses' FP32
    error (about 1e-6 relative) into dL/dm = 2 (rho - rho_t), which is small
    where the two densities nearly agree: 7e-5 to 1.5e-4 measured as the
    rounding pattern of the clips changes, bar 3e-4."""
    import torch
    from pypsi_amd.autograd import voxel_density
    print("case: c4 autograd", kernel, flush=True)
    P, rho, rhot, parts, axes = _c4(oracle_mod)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(P.pos).to(dev).requires_grad_(True)
    conn = torch.from_numpy(P.conn).to(dev)
    trho_t = torch.from_numpy(rhot).to(dev)
    rho_g = voxel_density(pos, conn, (128,) * 3, BOX, periodic_box=BOX, kernel=kernel)
    loss = ((rho_g - trho_t) ** 2).sum()
    loss.backward()
    g = pos.grad.cpu().numpy()[parts, axes]
    fd = oracle_mod.fd_mass_loss_grad(P, rho - rhot, parts, axes, h=1e-6)
    e = _rel_err(g, fd)
    print("autograd (%s forward) vs fd: max %.2e" % (kernel, e.max()))
    assert e.max() <= tol
    # FP32 positions: the gradient comes back in the input dtype
    p32 = torch.from_numpy(P.pos.astype(np.float32)).to(dev).requires_grad_(True)
    rho32 = voxel_density(p32, conn, (128,) * 3, BOX, periodic_box=BOX, kernel=kernel)
    ((rho32 - trho_t) ** 2).sum().backward()
    assert p32.grad.dtype == torch.float32 and torch.isfinite(p32.grad).all()


# ---------------------------------------------------------------------------
# the adjoint of the fields m, v, vv (positions and velocities)
# ---------------------------------------------------------------------------

def _fields_problem(oracle_mod, n, gn, rms, seed, periodic=True):
    from pypsi_amd.synth import zeldovich_block
    pos, vel = zeldovich_block(n, rms, seed)
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=(gn,) * 3, bounds=BOX,
                                 fields=["v", "vv"])
    ref, _ = oracle_mod.voxels(P, threads=-1)
    return P, ref


def _rand_grads(ref, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(ref[k].shape) for k in ("m", "v", "vv")}


@pytest.mark.parametrize("periodic", [True, False])
def test_fields_adjoint_vs_oracle(oracle_mod, periodic):
    """psivox_adjoint through the C ABI against the oracle's FP64 adjoint
    fed the same forward fields and gradients: within 1e-9 of the largest
    component (positions and velocities)."""
    import torch
    from pypsi_amd import device as D
    print("case: fields adjoint vs oracle, periodic", periodic, flush=True)
    P, ref = _fields_problem(oracle_mod, 7, 13, 1.5, 2, periodic)
    grads = _rand_grads(ref, 9)
    dpos_o, dvel_o = oracle_mod.adjoint(P, oracle_mod.adjoint_weights(ref, grads))
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in ref.items()}
    tg = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in grads.items()}
    pos, vel = torch.from_numpy(P.pos).to(dev), torch.from_numpy(P.vel).to(dev)
    mesh = D.make_mesh(pos, vel, torch.from_numpy(P.mass).to(dev), torch.from_numpy(P.conn).to(dev),
                       BOX if periodic else None)
    grid = D.make_grid(P.n, BOX, t)
    dpos, dvel = torch.zeros_like(pos), torch.zeros_like(vel)
    D.adjoint(mesh, grid, tg, dpos, dvel, D.make_opts(stream=torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    ep = np.abs(dpos.cpu().numpy() - dpos_o).max() / np.abs(dpos_o).max()
    ev = np.abs(dvel.cpu().numpy() - dvel_o).max() / np.abs(dvel_o).max()
    print("gpu fields adjoint vs oracle: positions %.2e, velocities %.2e" % (ep, ev), flush=True)
    assert ep <= 1e-9 and ev <= 1e-9


@pytest.mark.parametrize("kernel,tol", [("fast64", 1e-4), ("plain", 1e-4), ("auto", 3e-4)])
def test_fields_autograd_vs_fd(oracle_mod, kernel, tol):
    """voxel_fields end to end: GPU forward of m, v, vv (kernel; its default
    is the FP64 column kernel 'fast64', measured 1.6e-7; the FP32-clip
    forward's cell fields carry their FP32 error into the weights, measured
    8e-5 to 1.8e-4 on velocities as the clips' rounding pattern changes,
    bar 3e-4), the loss
    sum_C G_m m + G_v . v + G_vv : vv in torch with seeded random
    coefficients, backward through psivox_adjoint; against FP64 central
    differences of the same loss (oracle forward, h = 1e-6) on 300 seeded
    position and 300 velocity components, |g - fd| / max(|fd|, 1e-3 rms fd)."""
    import torch
    from pypsi_amd.autograd import voxel_fields
    from oracle import block_connectivity
    print("case: fields autograd vs fd, forward", kernel, flush=True)
    P, ref = _fields_problem(oracle_mod, 12, 24, 1.0, 3)
    grads = _rand_grads(ref, 4)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(P.pos).to(dev).requires_grad_(True)
    vel = torch.from_numpy(P.vel).to(dev).requires_grad_(True)
    conn = torch.from_numpy(P.conn).to(dev)
    m, v, vv = voxel_fields(pos, vel, conn, P.n, BOX, periodic_box=BOX, kernel=kernel)
    tg = {k: torch.from_numpy(g).to(dev) for k, g in grads.items()}
    loss = (tg["m"] * m).sum() + (tg["v"] * v).sum() + (tg["vv"] * vv).sum()
    loss.backward()
    rng = np.random.default_rng(77)
    parts = rng.choice(P.pos.shape[0], size=300, replace=False)
    axes = rng.integers(0, 3, size=300)
    base = oracle_mod.raw_sums(ref)
    coef = np.concatenate([grads["m"].reshape(-1, 1), grads["v"].reshape(-1, 3), grads["vv"].reshape(-1, 9)], axis=1)
    csr = oracle_mod.particle_elements(P)
    for which, t in ((0, pos), (1, vel)):
        g = t.grad.cpu().numpy()[parts, axes]
        fd = oracle_mod.fd_fields_loss_grad(P, base, coef, parts, axes, which, h=1e-6, csr=csr)
        e = _rel_err(g, fd)
        print("fields autograd (%s forward) vs fd, %s: max %.2e" % (kernel, ["positions", "velocities"][which], e.max()),
              flush=True)
        assert e.max() <= tol

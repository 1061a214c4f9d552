"""Shared meshes for the adjoint tests (seeded)."""
import numpy as np

BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def cases(O):
    """(label, oracle Problem) pairs covering periodic ghosts, windows,
    anisotropic grids, simplices and 27-node elements."""
    from pypsi_amd.synth import zeldovich_block
    out = []
    pos, vel = zeldovich_block(8, 1.0, seed=3)
    out.append(("periodic 8^3 rms1 -> 12x10x14", O.Problem.block(pos, vel, periodic_box=BOX, n=(12, 10, 14), bounds=BOX)))
    pos, vel = zeldovich_block(6, 2.5, seed=4)
    out.append(("periodic 6^3 rms2.5 -> 16^3", O.Problem.block(pos, vel, periodic_box=BOX, n=(16,) * 3, bounds=BOX)))
    pos, vel = zeldovich_block(10, 1.0, seed=11)
    out.append(("open mesh, window", O.Problem.block(pos, vel, n=(16, 12, 18), bounds=((0.2, 0.15, 0.3), (0.8, 0.7, 0.9)))))
    rng = np.random.default_rng(42)
    p = rng.uniform(0.05, 0.95, size=(40, 3))
    conn = np.array([rng.choice(40, 4, replace=False) for _ in range(30)], dtype=np.int32)
    out.append(("random tets", O.Problem(p, None, rng.uniform(0.5, 2.0, 30), conn, n=(12, 10, 14), bounds=BOX)))
    g = np.stack(np.meshgrid(*(np.linspace(0, 1, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    q = 0.5 + (g - 0.5) * (0.3 + 0.4 * rng.uniform(size=g.shape))
    out.append(("27-node element", O.Problem(q, None, None, np.arange(27, dtype=np.int32).reshape(1, 27), n=(10, 9, 11),
                                             bounds=BOX)))
    return out


def loss_weights(O, P, seed=0):
    """r = rho - rho_t and w = dL/dm for L = sum (rho - rho_t)^2, with a
    seeded random target."""
    out, _ = O.voxels(P)
    V = float(np.prod(P.d))
    rng = np.random.default_rng(seed)
    rho = out["m"] / V
    rhot = rng.uniform(0.0, 2.0 * max(rho.mean(), 1e-300), size=P.n)
    r = rho - rhot
    return r, 2.0 * r / V


def sample(P, k, seed=0):
    """k (particle, axis) pairs among the particles the mesh uses."""
    rng = np.random.default_rng(seed)
    used = np.unique(P.conn)
    parts = rng.choice(used, size=min(k, used.size), replace=False)
    axes = rng.integers(0, 3, size=parts.size)
    return parts, axes


def refined_problem(O, kind):
    """A refined mesh of each element type: a periodic 4^3 Zel'dovich block
    of trilinear elements (rms 1.5 cells) onto 8^3, tolerance 0.1, level 2;
    one perturbed triquadratic element onto 11^3, tolerance 0.05, level 3
    (its middle nodes keep the coordinate 0.5, which 11 cells per axis keep
    off the grid planes: a vertex on a plane is a kink of the mass, where the
    adjoint takes one side and central differences the mean of both).  Both
    with velocities and the fields v, vv."""
    from pypsi_amd.synth import zeldovich_block
    if kind == "trilinear":
        pos, vel = zeldovich_block(4, 1.5, seed=3)
        P = O.Problem.block(pos, vel, periodic_box=BOX, n=(8,) * 3, bounds=BOX, fields=["v", "vv"])
        return P, 0.1, 2
    rng = np.random.default_rng(5)
    g = np.stack(np.meshgrid(*(np.linspace(0.0, 1.0, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    p = 0.5 + (g - 0.5) * (0.3 + 0.4 * rng.uniform(size=g.shape))
    v = rng.normal(size=p.shape)
    P = O.Problem(p, v, np.array([1.0]), np.arange(27, dtype=np.int32).reshape(1, 27), n=(11,) * 3, bounds=BOX,
                  fields=["v", "vv"])
    return P, 0.05, 3


def refined_fd(O, P, loss, which, comps, reftol, lvl, h=1e-6):
    """Central differences of loss(forward fields) of the oracle's refined
    forward for (particle, axis) components of positions (which 0) or
    velocities (1)."""
    arr = P.pos if which == 0 else P.vel
    out = []
    for p, a in comps:
        x0 = arr[p, a]
        arr[p, a] = x0 + h
        lp = loss(O.voxels(P, reftol=reftol, maxlvl=lvl)[0])
        arr[p, a] = x0 - h
        lm = loss(O.voxels(P, reftol=reftol, maxlvl=lvl)[0])
        arr[p, a] = x0
        out.append((lp - lm) / (2 * h))
    return np.array(out)

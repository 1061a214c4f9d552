"""PSI.RStarTree and PSI.VDF on the MI355X (through the C ABI) against the
oracle on the same inputs (src/PSI.c:808-935, 1022-1058; src/cpsi.c:177-275).

Bars: RStarTree.query gives exactly the oracle's element list (ascending);
VDF gives the oracle's streams in the oracle's order (ascending element,
tets in refinement order), the same count, densities within 1e-11 relative
and velocities within 1e-11 of the largest speed (the oracle computes the
barycentric coordinates another way; 5e-13 differences are cancellation in
small refined tets).  Known answers as in test_vdf_cpu: one stream of
density N^3 on an undisplaced lattice; exact interpolation of a linear
velocity field; an odd number of streams inside a periodic sheet.
"""
import numpy as np
import pytest

from helpers import BOX

pytestmark = pytest.mark.gpu

RHO_TOL, VEL_TOL = 1e-11, 1e-11


@pytest.fixture(scope="module")
def PSI():
    from pypsi_amd import PSI as mod
    return mod


def zel(n, rms, seed):
    from pypsi_amd.synth import zeldovich_block
    return zeldovich_block(n, rms, seed)


def same_streams(got, ref, what):
    (r1, v1), (r2, v2) = got, ref
    assert r1.shape == r2.shape and v1.shape == (len(r1), 3), f"{what}: {len(r1)} vs {len(r2)} streams"
    if len(r1):
        np.testing.assert_allclose(r1, r2, rtol=RHO_TOL, atol=0, err_msg=what)
        assert np.abs(v1 - v2).max() <= VEL_TOL * max(np.abs(v2).max(), 1e-300), what


def test_tree_query(PSI, oracle_mod):
    O = oracle_mod
    pos, vel = zel(16, 2.0, 5)
    rng = np.random.default_rng(7)
    for periodic in (True, False):
        print(f"case: RStarTree.query 16^3 rms2 periodic={periodic}", flush=True)
        kw = dict(periodic_box=BOX) if periodic else {}
        mesh = PSI.Mesh(loader="block", pos=pos, vel=vel, **kw)
        P = O.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=(8,) * 3)
        tree = PSI.RStarTree(mesh)
        W = O.tree_window(P)
        for _ in range(30):
            c = rng.uniform(-0.1, 1.1, 3)
            h = rng.uniform(0, 0.15, 3)
            q = np.array([c - h, c + h])
            got = tree.query(box=(tuple(q[0]), tuple(q[1])))
            assert got.dtype == np.int32
            np.testing.assert_array_equal(got, O.elem_query(P, W, q))
        everything = tree.query(((-2.0,) * 3, (3.0,) * 3))
        np.testing.assert_array_equal(everything, O.elem_query(P, W, np.array([[-2.0] * 3, [3.0] * 3])))
        assert len(tree.query(((5.0,) * 3, (6.0,) * 3))) == 0
    with pytest.raises(TypeError, match="Query box must be a tuple"):
        tree.query((1, 2, 3))


@pytest.mark.parametrize("lvl", [0, 2])
def test_vdf_block(PSI, oracle_mod, lvl):
    O = oracle_mod
    pos, vel = zel(12, 2.5, 9)
    print(f"case: VDF 12^3 rms2.5 periodic lvl={lvl}", flush=True)
    mesh = PSI.Mesh(loader="block", pos=pos, vel=vel, periodic_box=BOX)
    P = O.Problem.block(pos, vel, periodic_box=BOX, n=(8,) * 3)
    tree = PSI.RStarTree(mesh)
    rng = np.random.default_rng(11)
    total = 0
    for i in range(40):
        x = rng.uniform(0.0, 1.0, 3)
        ref = O.sample_vdf(P, x, reftol=0.01, maxlvl=lvl, window=O.tree_window(P))
        same_streams(PSI.VDF(mesh, tuple(x), tree=tree, refine_tolerance=0.01, refine_max_lvl=lvl), ref, f"x={x}")
        if i % 4 == 0:  # without a tree: the mesh box is the window (the same for a periodic mesh)
            same_streams(PSI.VDF(mesh, tuple(x), refine_tolerance=0.01, refine_max_lvl=lvl), ref, f"x={x} no tree")
        total += len(ref[0])
    assert total > 20


def test_vdf_simplex_quadratic_and_open(PSI, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(42)
    pos = (rng.uniform(0.2, 0.8, size=(200, 1, 3)) + rng.uniform(-0.15, 0.15, size=(200, 4, 3))).reshape(-1, 3)
    vel = rng.normal(size=pos.shape)
    mass = rng.uniform(0.5, 2.0, size=200)
    conn = np.arange(800, dtype=np.int32).reshape(200, 4)
    print("case: VDF simplex array mesh", flush=True)
    mesh = PSI.Mesh(loader="array", pos=pos, vel=vel, mass=mass, connectivity=conn, periodic_box=BOX)
    P = O.Problem(pos, vel, mass, conn, periodic_box=BOX, n=(8,) * 3)
    g = np.stack(np.meshgrid(*(np.linspace(0, 1, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    qpos = np.ascontiguousarray(0.5 + (g - 0.5) * (0.2 + 0.4 * rng.uniform(size=g.shape)))
    qvel = rng.normal(size=qpos.shape)
    qconn = np.arange(27, dtype=np.int32).reshape(1, 27)
    print("case: VDF triquadratic element", flush=True)
    qmesh = PSI.Mesh(loader="array", pos=qpos, vel=qvel, connectivity=qconn, periodic_box=BOX)
    Q = O.Problem(qpos, qvel, None, qconn, periodic_box=BOX, n=(8,) * 3)
    hits = 0
    for _ in range(30):
        x = rng.uniform(0.3, 0.7, 3)
        a = PSI.VDF(mesh, tuple(x))
        same_streams(a, O.sample_vdf(P, x), f"simplex x={x}")
        b = PSI.VDF(qmesh, tuple(x), refine_tolerance=0.005, refine_max_lvl=3)
        same_streams(b, O.sample_vdf(Q, x, reftol=0.005, maxlvl=3), f"quadratic x={x}")
        hits += len(a[0]) + len(b[0])
    assert hits > 10
    # an open mesh: without a tree only elements whose box holds the origin take part
    # (src/cpsi.c:203-208); with one, every element does
    pos2, vel2 = zel(6, 1.0, 2)
    print("case: VDF open block", flush=True)
    omesh = PSI.Mesh(loader="block", pos=pos2, vel=vel2)
    OP = O.Problem.block(pos2, vel2, periodic_box=None, n=(8,) * 3)
    otree = PSI.RStarTree(omesh)
    for _ in range(10):
        x = rng.uniform(0.1, 0.9, 3)
        same_streams(PSI.VDF(omesh, tuple(x)), O.sample_vdf(OP, x), f"open x={x}")
        same_streams(PSI.VDF(omesh, tuple(x), tree=otree), O.sample_vdf(OP, x, window=O.tree_window(OP)),
                     f"open tree x={x}")
    with pytest.raises(ValueError, match="different mesh"):
        PSI.VDF(mesh, (0.5, 0.5, 0.5), tree=otree)


def test_vdf_known_answers(PSI):
    N = 6
    g = np.ascontiguousarray((np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1) + 0.5) / N)
    print("case: VDF undisplaced lattice", flush=True)
    mesh = PSI.Mesh(loader="block", pos=g, periodic_box=BOX)
    tree = PSI.RStarTree(mesh)
    rng = np.random.default_rng(0)
    for _ in range(20):
        rho, vel = PSI.VDF(mesh, tuple(rng.uniform(0.1, 0.9, 3)), tree=tree)
        assert len(rho) == 1
        np.testing.assert_allclose(rho, N ** 3, rtol=1e-12)
        assert np.abs(vel).max() == 0.0
    pos, _ = zel(12, 2.5, 9)
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [0.2, 0.1, -0.5]])
    b = np.array([0.01, -0.02, 0.03])
    print("case: VDF linear velocity field", flush=True)
    mesh = PSI.Mesh(loader="block", pos=pos, vel=np.ascontiguousarray(pos @ A.T + b), periodic_box=BOX)
    tree = PSI.RStarTree(mesh)
    odd = multi = 0
    for _ in range(30):
        x = rng.uniform(0.3, 0.7, 3)
        rho, vel = PSI.VDF(mesh, tuple(x), tree=tree, refine_tolerance=0.02, refine_max_lvl=1)
        odd += len(rho) % 2
        multi += len(rho) > 1
        if len(rho):
            np.testing.assert_allclose(vel, np.tile(A @ x + b, (len(rho), 1)), atol=1e-13)
    assert odd == 30 and multi > 0


def test_reference_golden(PSI):
    """VDF streams and tree queries against the reference's own outputs
    (tests/golden, written by its C core)."""
    import mode_fixtures as MF
    print("case: VDF / tree golden", flush=True)
    z = MF.load("vdf_block8.npz")
    mesh = PSI.Mesh(loader="block", pos=z["pos"], vel=z["vel"], periodic_box=BOX)
    tree = PSI.RStarTree(mesh)
    for x, e in zip(z["points"], MF.vdf_expected(z)):
        got = PSI.VDF(mesh, tuple(x), tree=tree, refine_tolerance=float(z["reftol"]), refine_max_lvl=int(z["maxlvl"]))
        MF.check_streams(got, e, f"x={x}")
        MF.check_streams(PSI.VDF(mesh, tuple(x), refine_tolerance=float(z["reftol"]),
                                 refine_max_lvl=int(z["maxlvl"])), e, f"x={x} no tree")
    t = MF.load("tree_block8.npz")
    tmesh = PSI.Mesh(loader="block", pos=t["pos"], vel=t["vel"], periodic_box=BOX)
    ttree = PSI.RStarTree(tmesh)
    for q, e in zip(t["boxes"], MF.tree_expected(t)):
        np.testing.assert_array_equal(ttree.query((tuple(q[0]), tuple(q[1]))), e)

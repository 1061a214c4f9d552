"""PSI.VDF and RStarTree's oracle on the CPU (src/cpsi.c:177-275,
src/rtree.c:48-123, src/geometry.c:335-389).

Pinned against the reference's own psi_sample_vdf and R*-tree (oracle/_ref)
on periodic and open meshes, refined and not, and by what the mathematics
fixes: an undisplaced lattice has one stream of density N^3 per unit element
mass; a velocity field linear in position is interpolated exactly; a
periodic sheet covers almost every point an odd number of times.
Tolerances: stream densities within 1e-11 relative and velocities within
1e-11 of the largest speed (the barycentric coordinates of small refined
tets lose a few digits to cancellation: 5e-13 measured); the same number of
streams and the same element sets.
"""
import numpy as np
import pytest

from helpers import BOX

RHO_TOL, VEL_TOL = 1e-11, 1e-11


def zel(n, rms, seed):
    from pypsi_amd.synth import zeldovich_block
    return zeldovich_block(n, rms, seed)


def match_streams(a, b, what):
    """the same streams in any order"""
    (r1, v1), (r2, v2) = a, b
    assert len(r1) == len(r2), f"{what}: {len(r1)} vs {len(r2)} streams"
    if len(r1) == 0:
        return
    o1, o2 = np.lexsort((v1[:, 0], r1)), np.lexsort((v2[:, 0], r2))
    np.testing.assert_allclose(r1[o1], r2[o2], rtol=RHO_TOL, atol=0, err_msg=what)
    scale = max(np.abs(v2).max(), 1e-300)
    assert np.abs(v1[o1] - v2[o2]).max() <= VEL_TOL * scale, what


needs_ref = pytest.mark.skipif("not __import__('oracle').have_ref()", reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("n,rms,lvl", [(8, 1.5, 0), (8, 1.5, 2), (12, 2.5, 1)])
def test_vdf_periodic_block_vs_reference(oracle_mod, n, rms, lvl):
    O = oracle_mod
    pos, vel = zel(n, rms, 3)
    P = O.Problem.block(pos, vel, periodic_box=BOX, n=(16,) * 3)
    rng = np.random.default_rng(1)
    total = 0
    for _ in range(25):
        x = rng.uniform(0, 1, 3)
        a = O.sample_vdf(P, x, reftol=0.01, maxlvl=lvl)
        match_streams(a, O.ref_sample_vdf(P, x, reftol=0.01, maxlvl=lvl), f"x={x} lvl={lvl}")
        total += len(a[0])
    # points in the layer of elements that straddle the periodic boundary get
    # no stream: the tets are sampled at their raw (unwrapped) positions
    # (src/cpsi.c:217-224), as the reference does
    assert total > 10


@needs_ref
def test_vdf_simplex_and_quadratic_vs_reference(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(42)
    # 200 overlapping tets of size <= 0.3 around points of [0.2, 0.8]^3
    pos = (rng.uniform(0.2, 0.8, size=(200, 1, 3)) + rng.uniform(-0.15, 0.15, size=(200, 4, 3))).reshape(-1, 3)
    vel = rng.normal(size=pos.shape)
    conn = np.arange(800, dtype=np.int32).reshape(200, 4)
    P = O.Problem(pos, vel, rng.uniform(0.5, 2.0, size=200), conn, periodic_box=BOX, n=(8,) * 3)
    g = np.stack(np.meshgrid(*(np.linspace(0, 1, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    qpos = 0.5 + (g - 0.5) * (0.2 + 0.4 * rng.uniform(size=g.shape))
    Q = O.Problem(qpos, rng.normal(size=qpos.shape), None, np.arange(27, dtype=np.int32).reshape(1, 27),
                  periodic_box=BOX, n=(8,) * 3)
    hits = 0
    for _ in range(40):
        x = rng.uniform(0.3, 0.7, 3)
        a = O.sample_vdf(P, x)
        match_streams(a, O.ref_sample_vdf(P, x), f"simplex x={x}")
        b = O.sample_vdf(Q, x, reftol=0.005, maxlvl=3)
        match_streams(b, O.ref_sample_vdf(Q, x, reftol=0.005, maxlvl=3), f"quadratic x={x}")
        hits += len(a[0]) + len(b[0])
    assert hits > 10


@needs_ref
def test_tree_query_vs_reference(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(7)
    pos, vel = zel(8, 2.0, 5)
    for periodic in (True, False):
        P = O.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=(8,) * 3)
        W = O.tree_window(P)
        for _ in range(20):
            c = rng.uniform(-0.1, 1.1, 3)
            h = rng.uniform(0, 0.2, 3)
            q = np.array([c - h, c + h])
            a = O.elem_query(P, W, q)
            b = O.ref_tree_query(P, W, q)
            assert np.all(np.diff(a) > 0)
            np.testing.assert_array_equal(a, np.sort(b))
        # a point query and the whole box
        a = O.elem_query(P, W, np.array([[0.5] * 3, [0.5] * 3]))
        np.testing.assert_array_equal(a, np.sort(O.ref_tree_query(P, W, np.array([[0.5] * 3, [0.5] * 3]))))


@needs_ref
def test_vdf_open_mesh_window_quirk(oracle_mod):
    """Without a tree, psi_sample_vdf indexes the elements with the mesh's box
    as the window (src/cpsi.c:203-208); for an open mesh that box is all zeros,
    so only elements whose box holds the origin take part."""
    O = oracle_mod
    pos, vel = zel(6, 1.0, 2)
    P = O.Problem.block(pos, vel, periodic_box=None, n=(8,) * 3)
    rng = np.random.default_rng(3)
    for _ in range(10):
        x = rng.uniform(0, 1, 3)
        match_streams(O.sample_vdf(P, x), O.ref_sample_vdf(P, x), f"open x={x}")
    match_streams(O.sample_vdf(P, [0.02, 0.02, 0.02]), O.ref_sample_vdf(P, [0.02, 0.02, 0.02]), "near origin")


def test_vdf_lattice_known_answer(oracle_mod):
    """Undisplaced periodic lattice of N^3 unit-mass elements: one stream,
    density N^3, zero velocity, at points off the tets' faces and inside the
    layer of elements that do not straddle the box edge."""
    O = oracle_mod
    N = 6
    g = (np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1) + 0.5) / N
    P = O.Problem.block(np.ascontiguousarray(g), None, periodic_box=BOX, n=(4,) * 3)
    rng = np.random.default_rng(0)
    for _ in range(20):
        x = rng.uniform(0.1, 0.9, 3)
        rho, vel = O.sample_vdf(P, x)
        assert len(rho) == 1
        np.testing.assert_allclose(rho, N ** 3, rtol=1e-12)
        assert np.abs(vel).max() == 0.0


def test_vdf_linear_velocity_and_odd_streams(oracle_mod):
    """vel = A pos + b on every particle: each stream's velocity is A x + b
    (barycentric interpolation of a linear field is exact), at any
    refinement.  A periodic sheet covers a point an odd number of times (away
    from the boundary layer of elements, which are sampled unwrapped)."""
    O = oracle_mod
    pos, _ = zel(12, 2.5, 9)
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [0.2, 0.1, -0.5]])
    b = np.array([0.01, -0.02, 0.03])
    P = O.Problem.block(pos, None, periodic_box=BOX, n=(8,) * 3)
    rng = np.random.default_rng(5)
    odd = 0
    multi = 0
    for _ in range(30):
        x = rng.uniform(0.3, 0.7, 3)
        rho, vel = O.sample_vdf(P, x, reftol=0.02, maxlvl=1)
        odd += len(rho) % 2
        multi += len(rho) > 1
    assert odd == 30 and multi > 0
    p = pos.reshape(-1, 3)
    Q = O.Problem(p, p @ A.T + b, None, P.conn, periodic_box=None, n=(8,) * 3)
    W = O.tree_window(Q)
    for _ in range(30):
        x = rng.uniform(0.2, 0.8, 3)
        for lvl in (0, 2):
            rho, vel = O.sample_vdf(Q, x, reftol=0.01, maxlvl=lvl, window=W)
            if len(rho):
                np.testing.assert_allclose(vel, np.tile(A @ x + b, (len(rho), 1)), atol=1e-13)


@pytest.mark.parametrize("lvl", [0, 2])
def test_device_path_on_host_vs_oracle(oracle_mod, lvl):
    """The device's query path compiled for the host (tests/hostsim: element
    boxes, box test, refinement, barycentric sample) gives the oracle's
    streams in the oracle's order (ascending element, tets in refinement
    order)."""
    import hostsim_api as H
    O = oracle_mod
    pos, vel = zel(12, 2.5, 9)
    P = O.Problem.block(pos, vel, periodic_box=BOX, n=(8,) * 3)
    rng = np.random.default_rng(11)
    total = 0
    for _ in range(30):
        x = rng.uniform(0.0, 1.0, 3)
        r1, v1 = H.sample_vdf(P, x, reftol=0.01, maxlvl=lvl)
        r2, v2 = O.sample_vdf(P, x, reftol=0.01, maxlvl=lvl)
        assert len(r1) == len(r2)
        np.testing.assert_allclose(r1, r2, rtol=RHO_TOL, atol=0)
        if len(r1):
            assert np.abs(v1 - v2).max() <= VEL_TOL * np.abs(v2).max()
        total += len(r1)
    assert total > 20

"""An exact count of fragments, the numerator of the fragments/s metric.

A fragment is one (periodic copy of a tet, grid cell) pair whose intersection
has positive volume: the reference keeps a leaf of its split DFS only when
its zeroth q-moment is positive (src/geometry.c:873-883), and that moment is
6 x the leaf's volume in the tet's barycentric frame.

This counts them with integers only.  Every float64 coordinate is a dyadic
rational, so scaling by one power of two makes the vertices and (for grids
of 2^k cells over the unit box) every cell plane integers.  A closed tet and
a closed cell have interiors that overlap exactly when no axis weakly
separates them, and the candidate axes are the three cell normals, the four
face normals of the tet and the eighteen products of a cell edge with a tet
edge (the facet normals of their Minkowski difference).  The test uses no
rounding at all, so the count is the one the reference would reach with
exact arithmetic."""
from fractions import Fraction
from itertools import product

import numpy as np

AXES0 = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _tet_axes(t):
    """Face normals and edge-cross-axis products of tet t (4 integer vertices)."""
    axes = []
    for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
        a, b, c = (t[i] for i in f)
        axes.append(_cross(tuple(b[k] - a[k] for k in range(3)), tuple(c[k] - a[k] for k in range(3))))
    for i, j in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
        e = tuple(t[j][k] - t[i][k] for k in range(3))
        for u in AXES0:
            axes.append(_cross(u, e))
    return [a for a in axes if a != (0, 0, 0)]


def _overlaps(t, axes, lo, hi):
    """Interiors of tet t and the box [lo, hi] overlap (no weak separation)."""
    for k in range(3):
        if max(v[k] for v in t) <= lo[k] or min(v[k] for v in t) >= hi[k]:
            return False
    for a in axes:
        p = [_dot(a, v) for v in t]
        q = [_dot(a, c) for c in product(*zip(lo, hi))]
        if max(p) <= min(q) or max(q) <= min(p):
            return False
    return True


def count_fragments(pos, conn, n, periodic):
    """Exact fragment count of the tets pos[conn] (conn: (ntet, 4)) on the
    unit box cut into n = (nx, ny, nz) cells, each a power of two.  With
    periodic=True the tets are unwrapped about their first vertex (minimum
    image, as the reference's element wrap), and every copy shifted by a box
    length in each axis is counted where it overlaps the grid."""
    for c in n:
        assert c & (c - 1) == 0, "cell counts must be powers of two"
    pos = np.asarray(pos, dtype=np.float64)
    den = max(Fraction(float(x)).denominator for x in pos.ravel())
    den = max(den, *n)
    # den is a power of two; every coordinate times den is an integer
    P = [tuple(int(Fraction(float(x)) * den) for x in p) for p in pos]
    total = 0
    for tet in np.asarray(conn):
        t = [P[i] for i in tet]
        if periodic:
            t = [t[0]] + [tuple(v[k] - den * ((2 * (v[k] - t[0][k]) + den) // (2 * den)) for k in range(3))
                           for v in t[1:]]
        shifts = product((-1, 0, 1), repeat=3) if periodic else [(0, 0, 0)]
        for s in shifts:
            ts = [tuple(v[k] + s[k] * den for k in range(3)) for v in t]
            axes = _tet_axes(ts)
            rng = []
            for k in range(3):
                step = den // n[k]
                a = max(0, min(v[k] for v in ts) // step)
                b = min(n[k] - 1, max(v[k] for v in ts) // step)
                rng.append(range(a, b + 1))
            for i, j, l in product(*rng):
                lo = (i * den // n[0], j * den // n[1], l * den // n[2])
                hi = ((i + 1) * den // n[0], (j + 1) * den // n[1], (l + 1) * den // n[2])
                total += _overlaps(ts, axes, lo, hi)
    return total


def kuhn_tets(side, rms, seed, periodic):
    """A displaced side^3 lattice of points, each cube cut into the six
    tets around its main diagonal (own decomposition); returns (pos, conn)
    with conn of shape (6 side^3, 4).  Periodic lattices wrap their indices
    and positions; open ones keep their points unwrapped.  A periodic tet
    must span less than half the box in every axis: the reference unwraps an
    element by moving its vertices in the upper half of the box down by one
    box length when its wrapped extent exceeds half the box (restated in the
    oracle's element wrap), which restores the element only when it is
    shorter than that."""
    rng = np.random.default_rng(seed)
    m = side if periodic else side + 1
    g = np.stack(np.meshgrid(*(np.arange(m),) * 3, indexing="ij"), -1).reshape(-1, 3)
    noise = rng.normal(scale=rms, size=g.shape)
    pos = (g + 0.5 + noise) / side
    if periodic:
        pos = pos % 1.0
    idx = lambda i, j, k: ((i % m) * m + (j % m)) * m + (k % m)  # noqa: E731
    conn = []
    for i, j, k in product(range(side), repeat=3):
        corners = list(product((0, 1), repeat=3))
        c = [idx(i + a, j + b, k + d) for a, b, d in corners]
        for perm in ((1, 2, 4), (1, 4, 2), (2, 1, 4), (2, 4, 1), (4, 1, 2), (4, 2, 1)):
            path = [0, perm[0], perm[0] + perm[1], 7]
            conn.append([c[p] for p in path])
            if not periodic:
                continue
            # the tet as the unwrapped lattice has it
            true = np.array([(np.array([i, j, k]) + corners[p] + 0.5 + noise[c[p]]) / side for p in path])
            assert (np.ptp(true, axis=0) < 0.45).all(), "a tet spans nearly half the box"
    return pos, np.array(conn, dtype=np.int32)

"""The parameter records the expansion writes for the adjoint on refined
meshes (pypsi_amd/csrc/psivox_refine.h): every refined tet's vertex, as the
expansion places it (lin27 midpoints, quad125 quarter-point weights, level
by level), equals the element's own map -- trilinear over its 8 corners or
triquadratic over its 27 nodes -- at the vertex's recorded dyadic parameter
point, for positions (the element's wrapped nodes) and velocities.  The
adjoint spreads a vertex's gradient over the nodes with exactly these
weights, so this pins its chain rule; the refined adjoint itself is held to
finite differences of the oracle on the GPU."""
import numpy as np
import pytest

import hostsim_api as H
from helpers import BOX


def cases(O):
    from pypsi_amd.synth import zeldovich_block
    out = []
    for lvl in (1, 2, 4):
        pos, vel = zeldovich_block(4, 1.5, seed=3)
        out.append((f"trilinear block L{lvl}", O.Problem.block(pos, vel, periodic_box=BOX, n=(8,) * 3, bounds=BOX),
                    0.1, lvl))
    rng = np.random.default_rng(5)
    g = np.stack(np.meshgrid(*(np.linspace(0.0, 1.0, 3),) * 3, indexing="ij"), -1).reshape(-1, 3)
    for lvl in (1, 3, 6):
        p = 0.5 + (g - 0.5) * (0.3 + 0.4 * rng.uniform(size=g.shape))
        v = rng.normal(size=p.shape)
        out.append((f"triquadratic L{lvl}", O.Problem(p, v, np.array([1.0]), np.arange(27, dtype=np.int32)[None],
                                                       n=(12,) * 3, bounds=BOX), 0.02, lvl))
    # a periodic block whose elements straddle the box faces (wrapped nodes)
    pos, vel = zeldovich_block(3, 0.8, seed=9)
    out.append(("trilinear wrapped L3", O.Problem.block(pos, vel, periodic_box=BOX, n=(6,) * 3, bounds=BOX), 0.05, 3))
    return out


@pytest.mark.parametrize("idx", range(7))
def test_refined_vertices_are_the_element_map(oracle_mod, idx):
    label, P, reftol, lvl = cases(oracle_mod)[idx]
    ep, eu, nt = H.expand_params(P, reftol=reftol, maxlvl=lvl)
    ne = P.conn.shape[0]
    print(label, "tets", nt, "max position error %.1e, velocity %.1e" % (ep, eu))
    assert nt > 6 * ne, "the case must refine"
    scale = max(1.0, float(np.abs(P.vel).max()))
    assert ep <= 1e-14 and eu <= 1e-13 * scale


def _weights(vpe, xi):
    """The element map's weights at xi (numpy restatement, test side)."""
    if vpe == 8:
        return np.array([(xi[0] if n >> 2 else 1 - xi[0]) * (xi[1] if (n >> 1) & 1 else 1 - xi[1]) *
                         (xi[2] if n & 1 else 1 - xi[2]) for n in range(8)])
    L = [[2 * (t - 0.5) * (t - 1), 4 * t * (1 - t), 2 * t * (t - 0.5)] for t in xi]
    return np.array([L[0][n // 9] * L[1][(n // 3) % 3] * L[2][n % 3] for n in range(27)])


@pytest.mark.parametrize("kind", ["trilinear", "triquadratic"])
def test_refined_chain_rule_vs_fd(oracle_mod, kind):
    """The chain rule the refined adjoint uses, on the CPU: the oracle's
    level-0 mass adjoint on the refined tets as an explicit tet mesh, spread
    over the element's nodes with the weights at each vertex's parameters,
    against FP64 central differences of the oracle's refined forward
    (h = 1e-6) on 30 node components: within 1e-7 of the largest."""
    from adjoint_cases import refined_fd, refined_problem
    O = oracle_mod
    P, reftol, lvl = refined_problem(O, kind)
    tets, prm = H.expand_records(P, reftol, lvl)
    nt = len(tets)
    Pt = O.Problem(tets[:, :12].reshape(-1, 3).copy(), None, tets[:, 24].copy(),
                   np.arange(4 * nt, dtype=np.int32).reshape(nt, 4), periodic_box=BOX if P.periodic else None,
                   n=P.n, bounds=BOX)
    rng = np.random.default_rng(11)
    w = rng.standard_normal(P.n)
    gt = O.mass_adjoint(Pt, w).reshape(nt, 4, 3)
    vpe = P.conn.shape[1]
    g = np.zeros_like(P.pos)
    for t in range(nt):
        e = int(prm[t, 0])
        for v in range(4):
            word = (int(prm[t, 1 + v // 2]) >> (32 * (v & 1))) & 0xFFFFFFFF
            xi = np.array([word & 255, (word >> 8) & 255, (word >> 16) & 255]) / 128.0
            np.add.at(g, P.conn[e], _weights(vpe, xi)[:, None] * gt[t, v][None, :])
    comps = [(int(p), int(a)) for p, a in zip(rng.integers(0, P.pos.shape[0], 30), rng.integers(0, 3, 30))]
    fd = refined_fd(O, P, lambda f: float((w * f["m"]).sum()), 0, comps, reftol, lvl)
    gc = g[tuple(np.array(comps).T)]
    err = np.abs(gc - fd).max() / np.abs(fd).max()
    print(kind, "chain rule vs fd: %.2e" % err)
    assert err <= 1e-7

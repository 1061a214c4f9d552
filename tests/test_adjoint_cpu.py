"""The mass adjoint on the CPU.

The oracle's FP64 adjoint (oracle/psi_adjoint.c: tet x grid-plane sections
ordered by angle and clipped per cell, Sutherland-Hodgman) is pinned by FP64
central finite differences of L = sum_C (m_C/V - rho_t)^2 (h = 1e-6 box
units) within 1e-7 of the largest gradient component.  The kernels' own
FP64 method (tests/hostsim running pypsi_amd/csrc/psivox_adjoint.h: one
task per tet copy, axis and plane, row/cell slab clips with an on-the-fly
shoelace) is held to the oracle within 1e-9 of the largest component.
Zero-volume tets have no x-space section; both drop them.
"""
import numpy as np
import pytest

import hostsim_api as H
from adjoint_cases import cases, loss_weights, sample


@pytest.fixture(scope="module")
def case_list(oracle_mod):
    return cases(oracle_mod)


@pytest.mark.parametrize("idx", range(5))
def test_oracle_adjoint_vs_fd(oracle_mod, case_list, idx):
    O = oracle_mod
    label, P = case_list[idx]
    r, w = loss_weights(O, P)
    g = O.mass_adjoint(P, w)
    parts, axes = sample(P, 40, seed=idx)
    fd = O.fd_mass_loss_grad(P, r, parts, axes, h=1e-6)
    err = np.abs(g[parts, axes] - fd).max() / np.abs(fd).max()
    print(label, "oracle adjoint vs FD: %.2e" % err)
    assert err < 1e-7


@pytest.mark.parametrize("idx", range(5))
def test_hostsim_adjoint_vs_oracle(oracle_mod, case_list, idx):
    O = oracle_mod
    label, P = case_list[idx]
    r, w = loss_weights(O, P)
    go = O.mass_adjoint(P, w)
    gh, st = H.mass_adjoint(P, w)
    assert st[1] == 0
    err = np.abs(gh - go).max() / np.abs(go).max()
    print(label, "kernel adjoint vs oracle: %.2e" % err)
    assert err < 1e-9


def test_adjoint_zero_volume_tets(oracle_mod):
    """Tets with a repeated vertex have zero volume: both adjoints drop them,
    and on particles used only by tets of nonzero volume they still match
    finite differences."""
    O = oracle_mod
    rng = np.random.default_rng(42)
    p = rng.uniform(0.05, 0.95, size=(40, 3))
    conn = rng.integers(0, 40, size=(30, 4)).astype(np.int32)
    deg = np.array([len(set(c)) < 4 for c in conn])
    assert deg.any()
    P = O.Problem(p, None, rng.uniform(0.5, 2.0, 30), conn, n=(12, 10, 14), bounds=((0, 0, 0), (1, 1, 1)))
    r, w = loss_weights(O, P)
    gh, _ = H.mass_adjoint(P, w)
    go = O.mass_adjoint(P, w)
    assert np.abs(gh - go).max() <= 1e-9 * np.abs(go).max()
    clean = np.setdiff1d(np.unique(conn[~deg]), np.unique(conn[deg]))
    parts, axes = np.repeat(clean, 3), np.tile(np.arange(3), clean.size)
    fd = O.fd_mass_loss_grad(P, r, parts, axes, h=1e-6)
    assert np.abs(gh[parts, axes] - fd).max() <= 1e-7 * np.abs(fd).max()


def test_fd_helper_matches_full_loss(oracle_mod):
    """The finite-difference helper (local re-voxelization) equals the
    difference of two full-grid losses."""
    O = oracle_mod
    from pypsi_amd.synth import zeldovich_block
    pos, vel = zeldovich_block(6, 1.0, seed=5)
    P = O.Problem.block(pos, vel, periodic_box=((0, 0, 0), (1, 1, 1)), n=(10,) * 3, bounds=((0, 0, 0), (1, 1, 1)))
    r, w = loss_weights(O, P)
    rhot = O.voxels(P)[0]["m"] / np.prod(P.d) - r
    h = 1e-4
    p, a = 17, 1

    def L(x):
        Q = O.Problem(P.pos.copy(), None, P.mass, P.conn, periodic_box=((0, 0, 0), (1, 1, 1)), n=P.n,
                      bounds=((0, 0, 0), (1, 1, 1)))
        Q.pos[p, a] = x
        return float((((O.voxels(Q)[0]["m"] / np.prod(P.d)) - rhot) ** 2).sum())

    x0 = P.pos[p, a]
    full = (L(x0 + h) - L(x0 - h)) / (2 * h)
    loc = O.fd_mass_loss_grad(P, r, [p], [a], h=h)[0]
    assert abs(full - loc) <= 1e-8 * abs(full)

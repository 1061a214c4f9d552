"""The adjoint of the fields m, v, vv (positions and velocities) on the
CPU: the oracle's FP64 adjoint (oracle/psi_adjoint.c: plane-section
integrals with a 7-point degree-5 rule, fragment q-moments for the
velocities) pinned by FP64 central finite differences and closed forms, and
the device's task functions built for the host (tests/hostsim: the 4-point
degree-3 rule, FP64 column moments) against it.

The loss is linear in the finalized fields with seeded random coefficients,
L = sum_C G_m m + G_v . v + G_vv : vv, so its gradient with respect to the
fields is G; dL/d(raw sums) follows from m = M, v = Sv/M, vv = Svv/M - v v^T
(oracle.adjoint_weights).  Bars: |g - fd| / max(|fd|, 1e-3 rms fd) within
2e-5 (finite differences with h = 1e-6 box units; their own error reaches
~5e-6 where a step crosses a grid plane), the host device algorithm within
1e-9 of the largest component of the oracle's (two FP64 decompositions of
the same fragments).
"""
import numpy as np
import pytest

import hostsim_api as H
from helpers import BOX


def _problem(oracle_mod, n=6, gn=12, rms=1.0, seed=0, periodic=True):
    from pypsi_amd.synth import zeldovich_block
    pos, vel = zeldovich_block(n, rms, seed)
    P = oracle_mod.Problem.block(pos, vel, periodic_box=BOX if periodic else None, n=(gn,) * 3, bounds=BOX,
                                 fields=["v", "vv"])
    ref, _ = oracle_mod.voxels(P)
    return P, ref


def _grads(ref, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(ref[k].shape) for k in ("m", "v", "vv")}


def _coef(g):
    return np.concatenate([g["m"].reshape(-1, 1), g["v"].reshape(-1, 3), g["vv"].reshape(-1, 9)], axis=1)


def _rel(g, fd):
    return np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.sqrt(np.mean(fd ** 2)))


def test_oracle_fields_adjoint_vs_fd(oracle_mod):
    P, ref = _problem(oracle_mod)
    grads = _grads(ref, 5)
    W = oracle_mod.adjoint_weights(ref, grads)
    dpos, dvel = oracle_mod.adjoint(P, W)
    base = oracle_mod.raw_sums(ref)
    npart = P.pos.shape[0]
    parts = np.repeat(np.arange(npart), 3)[:450]
    axes = np.tile(np.arange(3), npart)[:450]
    csr = oracle_mod.particle_elements(P)
    for which, g in ((0, dpos), (1, dvel)):
        fd = oracle_mod.fd_fields_loss_grad(P, base, _coef(grads), parts, axes, which, h=1e-6, csr=csr)
        e = _rel(g[parts, axes], fd)
        print(["positions", "velocities"][which], "oracle adjoint vs fd: max %.2e" % e.max())
        assert e.max() <= 2e-5


def test_fields_adjoint_closed_forms(oracle_mod):
    """A loss equal to the total x momentum (constant dL/dSv = e_x): its
    velocity gradients sum to the total mass, its position gradient
    vanishes (total momentum does not depend on where the sheet lies in a
    periodic box); a loss equal to the total mass: both vanish."""
    P, ref = _problem(oracle_mod)
    nvox = ref["m"].size
    W = np.zeros((nvox, 13))
    W[:, 1] = 1.0
    dpos, dvel = oracle_mod.adjoint(P, W)
    assert abs(dvel[:, 0].sum() - P.mass.sum()) <= 1e-12 * P.mass.sum()
    assert np.abs(dvel[:, 1:]).max() == 0.0
    assert np.abs(dpos).max() <= 1e-12 * np.abs(dvel).max()
    W = np.zeros((nvox, 13))
    W[:, 0] = 1.0
    dpos, dvel = oracle_mod.adjoint(P, W)
    assert np.abs(dpos).max() <= 1e-11 and np.abs(dvel).max() == 0.0


@pytest.mark.parametrize("periodic", [True, False])
def test_host_device_algorithm_vs_oracle(oracle_mod, periodic):
    P, ref = _problem(oracle_mod, n=7, gn=13, rms=1.5, seed=2, periodic=periodic)
    W = oracle_mod.adjoint_weights(ref, _grads(ref, 9))
    dpos, dvel = oracle_mod.adjoint(P, W)
    hp, hv, fails = H.adjoint(P, W)
    assert fails == 0
    ep = np.abs(hp - dpos).max() / np.abs(dpos).max()
    ev = np.abs(hv - dvel).max() / np.abs(dvel).max()
    print("host device algorithm vs oracle: positions %.2e, velocities %.2e" % (ep, ev))
    assert ep <= 1e-9 and ev <= 1e-9

"""PSI.voxels(mode='annihilation') on the CPU: the oracle against the
reference's own psi_voxels in that mode (oracle/_ref), the device's pair
code (tests/hostsim) against the oracle, and a known answer.

Tolerances, the same for every comparison here and on the GPU: cells
holding >= 1e-3 of the peak within 1e-8 relative, every cell within 1e-9 of
the peak, totals within 1e-10.  Cells on caustics, where rho = mass/volume of
thin tets is large, are ill-conditioned: the oracle and the reference build
the face planes with different rounding and add in a different order, and
differ by 2.4e-9 relative and 1.8e-10 of the peak on a 16^3 Zel'dovich mesh
(6.5e-11 on the 6^3 one below); the device code built for the host against
the oracle: 3.9e-12 and 1.9e-13 on 16^3.
"""
import numpy as np
import pytest

from helpers import BOX


def zel(n, rms, seed):
    from pypsi_amd.synth import zeldovich_block
    return zeldovich_block(n, rms, seed)


def close_m(got, ref, rel, absr, tot, what):
    a, b = got["m"], ref["m"]
    pk = np.abs(b).max()
    heavy = b >= 1e-3 * pk
    d = np.abs(a - b)
    assert (d[heavy] / b[heavy]).max() <= rel, f"{what}: heavy rel {(d[heavy] / b[heavy]).max():.2e}"
    assert d.max() <= absr * pk, f"{what}: abs {d.max() / pk:.2e} of peak"
    assert abs(a.sum() / b.sum() - 1) <= tot, what


def cases(O):
    pos, vel = zel(6, 1.5, 3)
    yield "zel 6^3 rms1.5 -> 12^3 L0", O.Problem.block(pos, vel, periodic_box=BOX, n=(12,) * 3), 0
    yield "zel 6^3 rms1.5 -> 12^3 L1", O.Problem.block(pos, vel, periodic_box=BOX, n=(12,) * 3), 1
    pos, vel = zel(5, 1.0, 4)
    yield "zel 5^3 window", O.Problem.block(pos, vel, periodic_box=BOX, n=(9, 8, 10),
                                            bounds=((0.2, 0.15, 0.3), (0.8, 0.7, 0.9))), 0
    rng = np.random.default_rng(42)
    p = (rng.uniform(0.2, 0.8, size=(40, 1, 3)) + rng.uniform(-0.12, 0.12, size=(40, 4, 3))).reshape(-1, 3)
    yield "40 tets", O.Problem(p, None, rng.uniform(0.5, 2.0, size=40), np.arange(160, dtype=np.int32).reshape(40, 4),
                               n=(10, 12, 11)), 0


@pytest.mark.skipif("not __import__('oracle').have_ref()", reason="oracle/_ref not built")
def test_oracle_vs_reference(oracle_mod):
    O = oracle_mod
    for what, P, lvl in cases(O):
        got, st = O.annihilation(P, reftol=0.2, maxlvl=lvl)
        assert st["overflow"] == 0
        close_m(got, O.ref_voxels(P, reftol=0.2, maxlvl=lvl, mode=1), 1e-8, 1e-9, 1e-10, what)


def test_device_code_vs_oracle(oracle_mod):
    import hostsim_api as H
    O = oracle_mod
    for what, P, lvl in cases(O):
        ref, st = O.annihilation(P, reftol=0.2, maxlvl=lvl)
        got, nf, npairs = H.annihilation(P, reftol=0.2, maxlvl=lvl)
        assert nf >= 0 and npairs == st["ntets"], what
        close_m(got, ref, 1e-8, 1e-9, 1e-10, what)


def test_uniform_lattice_known_answer(oracle_mod):
    """Undisplaced periodic lattice of N^3 unit-mass elements, density N^3:
    away from the elements that straddle the box edge and their neighbours
    (the reference pairs a straddling element at its raw positions,
    src/cpsi.c:124-131, so its tets span the box: cells 0-2 and G-1 differ
    there, in the reference too), each cell gets rho^2 x its volume: only a
    tet's overlap with itself has volume."""
    O = oracle_mod
    N, G = 6, 12
    g = (np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1) + 0.5) / N
    P = O.Problem.block(np.ascontiguousarray(g), None, periodic_box=BOX, n=(G,) * 3)
    out, _ = O.annihilation(P)
    inner = out["m"][3:G - 1, 3:G - 1, 3:G - 1]
    np.testing.assert_allclose(inner, float(N ** 3) ** 2 / G ** 3, rtol=1e-12)

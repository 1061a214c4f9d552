"""The oracle's annihilation mode, VDF sample and tree query against the
reference's own outputs (tests/golden, see tests/mode_fixtures.py for the
bars).  The GPU is held to the same fixtures in test_gpu_annihilation.py and
test_gpu_vdf.py."""
import numpy as np
import pytest

import mode_fixtures as MF


@pytest.mark.parametrize("name", MF.ann_cases())
def test_annihilation_golden(oracle_mod, name):
    O = oracle_mod
    z = MF.load(name)
    P = O.Problem.block(z["pos"], z["vel"], periodic_box=MF.BOX, n=tuple(int(v) for v in z["n"]))
    got, _ = O.annihilation(P, reftol=float(z["reftol"]), maxlvl=int(z["maxlvl"]))
    MF.check_annihilation(got["m"], z["out_m"], name)


def test_vdf_golden(oracle_mod):
    O = oracle_mod
    z = MF.load("vdf_block8.npz")
    P = O.Problem.block(z["pos"], z["vel"], periodic_box=MF.BOX, n=(8, 8, 8))
    exp = MF.vdf_expected(z)
    assert sum(len(r) for r, _ in exp) > 20
    for x, e in zip(z["points"], exp):
        MF.check_streams(O.sample_vdf(P, x, reftol=float(z["reftol"]), maxlvl=int(z["maxlvl"])), e, f"x={x}")


def test_tree_golden(oracle_mod):
    O = oracle_mod
    z = MF.load("tree_block8.npz")
    P = O.Problem.block(z["pos"], z["vel"], periodic_box=MF.BOX, n=(8, 8, 8))
    for q, e in zip(z["boxes"], MF.tree_expected(z)):
        np.testing.assert_array_equal(O.elem_query(P, O.tree_window(P), q), e)

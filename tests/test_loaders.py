"""The gadget2 and gevolution snapshot loaders of PSI.Mesh (src/mesh.c:35-302,
src/PSI.c:659-707) on the CPU: against the reference's own loaders
(oracle/_ref) on the tutorial snapshot and on written files, against known
rows of the gevolution tile order, and for the reference's errors.  The
notebook's Example 2 (tutorial/Tutorial.ipynb:186) runs through the oracle
as a known answer for the whole path on real data."""
import os

import numpy as np
import pytest

from pypsi_amd import snapshot as S

SNAP = "/root/reference/tutorial/snapshot_010"


def have_ref(oracle_mod):
    return oracle_mod.have_ref()


@pytest.fixture(scope="module")
def PSI():
    from pypsi_amd import PSI
    return PSI


def lattice_conn(side):
    i, j, k = np.meshgrid(*(np.arange(side),) * 3, indexing="ij")
    rows = []
    for c in range(8):
        a, b, d = (i + (c >> 2)) % side, (j + ((c >> 1) & 1)) % side, (k + (c & 1)) % side
        rows.append(((a * side + b) * side + d).reshape(-1))
    return np.stack(rows, -1).astype(np.int32)


def test_gadget2_written_file(PSI, oracle_mod, tmp_path):
    """Shuffled rows land at their ids; mass[1]/npart per element; periodic
    lattice connectivity; box from BoxSize."""
    side, box, m1 = 6, 25.0, 3.5
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, box, size=(side ** 3, 3)).astype(np.float32)
    vel = rng.normal(size=(side ** 3, 3)).astype(np.float32)
    order = rng.permutation(side ** 3)
    fn = str(tmp_path / "snap")
    S.write(fn, pos[order], vel[order], order, m1, box)
    m = PSI.Mesh(loader="gadget2", filename=fn)
    assert m.pos.shape == (side ** 3, 3) and m.pos.dtype == np.float64
    np.testing.assert_array_equal(m.pos, pos.astype(np.float64))
    np.testing.assert_array_equal(m.vel, vel.astype(np.float64))
    np.testing.assert_array_equal(m.mass, np.full(side ** 3, m1 / side ** 3))
    np.testing.assert_array_equal(m.connectivity, lattice_conn(side))
    assert m.wrapbox == ((0.0, 0.0, 0.0), (box, box, box))
    if have_ref(oracle_mod):
        r = oracle_mod.ref_load_snapshot(fn, "gadget2")
        for a, b in zip((m.pos, m.vel, m.mass, m.connectivity), r[:4]):
            np.testing.assert_array_equal(a, b)


def test_gevolution_known_rows(PSI, tmp_path):
    """Connectivity rows of an 8^3 gevolution mesh worked out by hand from
    the reference's tile tables (src/mesh.c:163-164, 168-209): particle 0 at
    tile (0,0,0) cell (0,0,0); particle 63 at cell (3,0,0), whose +i
    neighbours sit in tile (1,0,0) = ids 64.. ; particle 64 * 7 + 5 at tile
    (1,1,1), cell (1,1,1)."""
    side = 8
    n = side ** 3
    fn = str(tmp_path / "gev")
    S.write(fn, np.zeros((n, 3)), np.zeros((n, 3)), np.arange(n), 1.0, 1.0, id_bytes=8)
    m = PSI.Mesh(loader="gevolution", filename=fn)
    c = m.connectivity
    np.testing.assert_array_equal(c[0], [0, 1, 3, 2, 7, 6, 4, 5])
    # cell (3,0,0): (3,0,1)=62 (3,1,0)=60 (3,1,1)=61; +i wraps into tile 1 at cells (0,0,0)=0 (0,0,1)=1
    # (0,1,0)=3 (0,1,1)=2
    np.testing.assert_array_equal(c[63], [63, 62, 60, 61, 64, 65, 67, 66])
    # tile (1,1,1) = 7, cell (1,1,1) = flat 5; its corners stay in tile 7: (1,1,2)=12 (1,2,1)=30
    # (1,2,2)=23 (2,1,1)=58 (2,1,2)=51 (2,2,1)=33 (2,2,2)=40
    np.testing.assert_array_equal(c[7 * 64 + 5], 7 * 64 + np.array([5, 12, 30, 23, 58, 51, 33, 40]))


def test_gevolution_geometry(PSI, oracle_mod, tmp_path):
    """A lattice written in gevolution id order: every element is a unit
    lattice cube, corner 4ii+2jj+kk at offset (ii, jj, kk) across tiles and
    the periodic box; the reference's loader agrees row for row."""
    side, box = 8, 8.0
    cell = S.gevolution_order(side)
    rng = np.random.default_rng(5)
    order = rng.permutation(side ** 3)
    pos = cell.astype(np.float64) + 0.5
    fn = str(tmp_path / "gev")
    S.write(fn, pos[order], pos[order] * 0, order, 2.0, box, id_bytes=8)
    m = PSI.Mesh(loader="gevolution", filename=fn)
    np.testing.assert_array_equal(m.pos, pos)
    for corner in range(8):
        off = np.array([corner >> 2, (corner >> 1) & 1, corner & 1])
        np.testing.assert_array_equal(cell[m.connectivity[:, corner]], (cell + off) % side)
    assert len(np.unique(cell, axis=0)) == side ** 3
    if have_ref(oracle_mod):
        r = oracle_mod.ref_load_snapshot(fn, "gevolution")
        for a, b in zip((m.pos, m.vel, m.mass, m.connectivity), r[:4]):
            np.testing.assert_array_equal(a, b)


def test_loader_errors(PSI, tmp_path):
    with pytest.raises(ValueError, match="Bad file name: .*nowhere"):
        PSI.Mesh(loader="gadget2", filename=str(tmp_path / "nowhere"))
    with pytest.raises(ValueError, match="Bad file name"):
        PSI.Mesh(loader="gevolution")
    n = 27
    z = np.zeros((n, 3))
    for bad in ("pos", "vel", "ids"):
        fn = str(tmp_path / f"bad_{bad}")
        S.write(fn, z, z, np.arange(n), 1.0, 1.0, bad_block=bad)
        with pytest.raises(ValueError, match="Could not read"):
            PSI.Mesh(loader="gadget2", filename=fn)
    fn = str(tmp_path / "short")
    S.write(fn, z, z, np.arange(n), 1.0, 1.0)
    with open(fn, "rb") as f:
        data = f.read()
    with open(fn, "wb") as f:
        f.write(data[:-20])
    with pytest.raises(ValueError, match="Could not read"):
        PSI.Mesh(loader="gadget2", filename=fn)
    fn = str(tmp_path / "dup")
    S.write(fn, z, z, np.r_[np.arange(n - 1), 0], 1.0, 1.0)
    with pytest.raises(ValueError, match="Could not read"):
        PSI.Mesh(loader="gadget2", filename=fn)
    fn = str(tmp_path / "range")
    S.write(fn, z, z, np.arange(n) + 1, 1.0, 1.0)
    with pytest.raises(ValueError, match="Could not read"):
        PSI.Mesh(loader="gadget2", filename=fn)
    fn = str(tmp_path / "noncube")
    S.write(fn, np.zeros((10, 3)), np.zeros((10, 3)), np.arange(10), 1.0, 1.0)
    with pytest.raises(ValueError, match="Could not read"):
        PSI.Mesh(loader="gadget2", filename=fn)
    fn = str(tmp_path / "gev27")  # a cube, but not of a multiple of 4: no whole tiles
    S.write(fn, z, z, np.arange(n), 1.0, 1.0, id_bytes=8)
    with pytest.raises(ValueError, match="Could not read"):
        PSI.Mesh(loader="gevolution", filename=fn)
    with pytest.raises(ValueError, match="Invalid mesh loader"):
        PSI.Mesh(loader="gadget3", filename=fn)


needs_snapshot = pytest.mark.skipif(not os.path.exists(SNAP), reason="the reference's tutorial snapshot is absent")


@needs_snapshot
def test_snapshot_010_matches_reference_loader(PSI, oracle_mod):
    if not have_ref(oracle_mod):
        pytest.skip("oracle/_ref not built")
    m = PSI.Mesh(loader="gadget2", filename=SNAP)
    r = oracle_mod.ref_load_snapshot(SNAP, "gadget2")
    for a, b in zip((m.pos, m.vel, m.mass, m.connectivity), r[:4]):
        assert a.shape == b.shape and a.dtype == b.dtype
        np.testing.assert_array_equal(a, b)
    assert m.wrapbox == tuple(tuple(x) for x in r[4])
    assert m.pos.shape == (32768, 3) and m.wrapbox[1] == (40.0, 40.0, 40.0)


GOLDEN_SNAP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "snap010_example2.npz")


@needs_snapshot
def test_snapshot_010_fixture_inputs(PSI):
    """The fixture's arrays are what the loader makes of snapshot_010."""
    z = np.load(GOLDEN_SNAP)
    m = PSI.Mesh(loader="gadget2", filename=SNAP)
    np.testing.assert_array_equal(m.pos, z["pos"].astype(np.float64))
    np.testing.assert_array_equal(m.vel, z["vel"].astype(np.float64))
    np.testing.assert_array_equal(m.connectivity, z["conn"])
    assert np.all(m.mass == float(z["mass"]))
    assert m.wrapbox == tuple(tuple(x) for x in z["box"])


def test_snapshot_010_example2_oracle(oracle_mod):
    """Notebook Example 2: snapshot_010 onto 128^3 with v and vv, refinement
    level 2 at tolerance 0.1 (tutorial/Tutorial.ipynb:193-204); the notebook
    prints a global mass error of 6.99e-15 (tutorial/Tutorial.ipynb:186).
    The oracle, on the fixture's arrays, against the reference's sums and its
    central 16^3 block (OpenMP threads: the summation order differs, so the
    bar is 1e-12 rather than bit equality)."""
    from helpers import TOL64, assert_close
    O = oracle_mod
    z = np.load(GOLDEN_SNAP)
    box = np.array(z["box"])
    ne = z["conn"].shape[0]
    P = O.Problem(z["pos"].astype(np.float64), z["vel"].astype(np.float64), np.full(ne, float(z["mass"])),
                  z["conn"], n=(128,) * 3, bounds=box, periodic_box=box, fields=["v", "vv"])
    out, st = O.voxels(P, reftol=float(z["reftol"]), maxlvl=int(z["maxlvl"]), threads=min(8, os.cpu_count() or 1))
    assert st["overflow"] == 0 and st["nelem_used"] == 32768
    assert abs(1.0 - out["m"].sum() / (ne * float(z["mass"]))) < 1e-14
    assert abs(1.0 - out["m"].sum() / float(z["sum_m"])) < 1e-14
    mv = (out["m"][..., None] * out["v"]).sum(axis=(0, 1, 2))
    # the net momentum cancels: its scale is the total mass times the rms particle velocity
    pscale = ne * float(z["mass"]) * float(np.sqrt(np.mean(z["vel"].astype(np.float64) ** 2)))
    np.testing.assert_allclose(mv, z["sum_mv"], rtol=0, atol=1e-12 * pscale)
    np.testing.assert_allclose(out["vv"].sum(axis=(0, 1, 2)), z["sum_vv"], rtol=1e-12)
    S = tuple(slice(a, a + 16) for a in (int(v) for v in z["sub_lo"]))
    assert_close({"m": out["m"][S], "v": out["v"][S], "vv": out["vv"][S]},
                 {"m": z["sub_m"], "v": z["sub_v"], "vv": z["sub_vv"]}, TOL64, "oracle vs reference")

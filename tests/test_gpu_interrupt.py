"""The interrupt path: the reference checks PyErr_CheckSignals() after every
element and returns early with the exception pending, leaving the grid with
what was deposited so far (src/cpsi.c:160-163, src/PSI.c:1015-1017).  Here
the check runs between device batches of elements (the progress callback of
the C ABI), so a cancelled call leaves the fields holding whole batches.

* test_progress_cancel: the C ABI with 4096-element batches and a progress
  callback that asks to stop after two batches: PSIVOX_E_CANCELLED, and the
  grid holds exactly the mass of the first 8192 elements of a periodic mesh.
* test_sigint_mid_call: PSI.voxels in a child process that receives SIGINT
  while the call runs: KeyboardInterrupt, the grid holds the mass of the
  whole batches done (2^22 elements each), and the next call in the same
  process is correct.  The child keeps a stray signal away from the test
  runner."""
import ctypes as C
import os
import re
import signal
import subprocess
import sys
import time

import numpy as np
import pytest

from helpers import BOX

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BATCH = 1 << 22  # elements per device batch when the caller sets none
PROGRESS = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_int64)


def test_progress_cancel(oracle_mod):
    import torch
    from pypsi_amd import device as D
    from pypsi_amd.synth import zeldovich_block
    print("case: progress callback cancels after 8192 of 32768 elements", flush=True)
    pos, vel = (a.reshape(-1, 3) for a in zeldovich_block(32, 1.0, seed=3))
    dev = torch.device("cuda:0")
    conn = oracle_mod.block_connectivity((32, 32, 32), True)
    ne = conn.shape[0]
    mass = np.full(ne, 1.0 / ne)
    seen = []

    def cb(ctx, done, total):
        seen.append((done, total))
        return 1 if done >= 8192 else 0

    fn = PROGRESS(cb)
    m = torch.zeros((64,) * 3, dtype=torch.float64, device=dev)
    mesh = D.make_mesh(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev), torch.from_numpy(mass).to(dev),
                       torch.from_numpy(conn).to(dev), BOX)
    opts = D.make_opts(kernel="auto")
    opts.progress = C.cast(fn, C.c_void_p)
    opts.batch_elements = 4096
    with pytest.raises(RuntimeError, match="cancelled after 8192 of 32768"):
        D.voxels(mesh, D.make_grid((64,) * 3, BOX, {"m": m}), opts)
    assert seen == [(4096, ne), (8192, ne)]
    got = float(m.sum())
    print("deposited mass", got, "of", 1.0, flush=True)
    assert abs(got / (8192.0 / ne) - 1.0) < 1e-12
    # the same workspace then runs a whole call
    m.zero_()
    opts.progress = None
    D.voxels(mesh, D.make_grid((64,) * 3, BOX, {"m": m}), opts)
    assert abs(float(m.sum()) - 1.0) < 1e-12


CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
from pypsi_amd import PSI
from pypsi_amd.synth import zeldovich_block
BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
lat = (np.stack(np.meshgrid(*(np.arange(8),) * 3, indexing="ij"), -1) + 0.5) / 8
small = PSI.Mesh(loader="block", pos=lat, vel=np.zeros_like(lat), periodic_box=BOX)
g = PSI.Grid(n=(16,) * 3, bounds=BOX)
PSI.voxels(grid=g, mesh=small)  # device set-up before the timed call
n = 256
pos, vel = zeldovich_block(n, 1.0, seed=4)
mesh = PSI.Mesh(loader="block", pos=pos, vel=vel, periodic_box=BOX)
total = float(mesh.mass.sum())
grid = PSI.Grid(n=(256,) * 3, bounds=BOX)
print("start", flush=True)
try:
    PSI.voxels(grid=grid, mesh=mesh, kernel="plain")
    print("finished", flush=True)
except KeyboardInterrupt as e:
    print("interrupted", repr(float(grid.fields["m"].sum()) / total), mesh.connectivity.shape[0], flush=True)
g = PSI.Grid(n=(16,) * 3, bounds=BOX)
PSI.voxels(grid=g, mesh=small)
print("after", repr(float(g.fields["m"].sum() / small.mass.sum())), flush=True)
"""


def test_sigint_mid_call():
    print("case: SIGINT to PSI.voxels (plain kernel, 256^3 -> 256^3) in a child process", flush=True)
    p = subprocess.Popen([sys.executable, "-u", "-c", CHILD, ROOT], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    lines = []
    try:
        for line in p.stdout:
            lines.append(line.rstrip())
            if line.startswith("start"):
                time.sleep(0.5)
                p.send_signal(signal.SIGINT)
                break
        out, _ = p.communicate(timeout=120)
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    lines += out.splitlines()
    print("\n".join(lines), flush=True)
    assert p.returncode == 0, lines
    # the progress line ("\rElement k of n") precedes the child's own line without a newline
    got = [re.search(r"interrupted (\S+) (\d+)", ln) for ln in lines]
    got = [g for g in got if g]
    assert got, "the call finished before the signal arrived: " + str(lines)
    frac, ne = float(got[0].group(1)), int(got[0].group(2))
    # the grid holds whole batches of 2^22 elements (the default batch), each deposited in full
    done = round(frac * ne / BATCH) * BATCH
    assert 0 < done < ne, (frac, ne)
    assert abs(frac / (done / ne) - 1.0) < 1e-12, (frac, done, ne)
    after = [re.search(r"after (\S+)", ln) for ln in lines]
    after = [a for a in after if a]
    assert after and abs(float(after[0].group(1)) - 1.0) < 1e-12

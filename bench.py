#!/usr/bin/env python3
"""Benchmark: tet-voxel fragments/s of PSI.voxels on one MI355X.

Workload (default): 128^3 particles on the unit periodic box, each lattice
cube split into 6 tets (12,582,912 tets), Zel'dovich displacement with
P(k) ~ k^-2 and rms 2.0 lattice cells (seed 0), voxelized onto a 256^3 grid
with fields m, v, vv (mass, mean velocity, velocity dispersion tensor).

A step = one voxelization of the whole mesh with inputs resident in HBM:
zero the grid fields, convert to raw moments, expand elements to tets, the
voxelization kernel, finalize to the reference layout.  `value` =
N_frag / (mean step time), N_frag = fragments (copy, cell) with m0 > 0
counted by the FP64 oracle for this exact input (cached in
profiles/workload_counts.json; the GPU's own count is reported beside it).

Prints ONE JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP32-equivalent flops per fragment of the reference algorithm
# (canonical fragment: 8 vertices, one split, 12 fan triangles, transform and
# update), by field set
FLOPS_PER_FRAG = {"m": 337, "m,v": 550, "m,v,vv": 2287, "all": 3490}
PEAK_FP32_TFLOPS = 157.3

WORKLOADS = {
    # name: (lattice n, rms cells, seed, grid n, fields)
    "zeldovich-128^3-rms2.0-to-256^3-m,v,vv": (128, 2.0, 0, 256, ("v", "vv")),
    "zeldovich-32^3-rms0.3-to-64^3-m": (32, 0.3, 0, 64, ()),
    "zeldovich-64^3-rms1.0-to-128^3-m": (64, 1.0, 0, 128, ()),
}
DEFAULT = "zeldovich-128^3-rms2.0-to-256^3-m,v,vv"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def field_key(fields):
    s = ",".join(["m"] + list(fields))
    return s if s in FLOPS_PER_FRAG else "all"


def load_counts():
    p = os.path.join(ROOT, "profiles", "workload_counts.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def build_inputs(wname):
    n, rms, seed, gn, fields = WORKLOADS[wname]
    return build_inputs_seed(wname, seed)


def build_inputs_seed(wname, seed):
    import numpy as np
    from pypsi_amd.synth import zeldovich_block
    n, rms, _, gn, fields = WORKLOADS[wname]
    pos, vel = zeldovich_block(n, rms, seed)
    # trilinear periodic connectivity, corner order 4i+2j+k (src/PSI.c:636-656)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    conn = np.empty((n, n, n, 8), dtype=np.int32)
    for ii in range(2):
        for jj in range(2):
            for kk in range(2):
                conn[..., 4 * ii + 2 * jj + kk] = n * n * ((i + ii) % n) + n * ((j + jj) % n) + ((k + kk) % n)
    conn = conn.reshape(-1, 8)
    mass = np.ones(conn.shape[0])
    return pos.reshape(-1, 3), vel.reshape(-1, 3), mass, conn, gn, fields


def cpu_baseline(wname, pos, vel, mass, conn, gn, fields, budget_s=12.0):
    """The reference's own C core (oracle/_ref) single-threaded on the first
    elements of the same mesh, sized to about budget_s of CPU time; falls back
    to the oracle's C restatement if the reference build is absent."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    kind = "reference" if O.have_ref() else "port"

    def run(e1):
        P = O.Problem(pos, vel, mass[:e1], conn[:e1], periodic_box=box, n=(gn,) * 3, bounds=box, fields=fields)
        out = P.zero_fields()
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_voxels(P, out=out)
        else:
            O.voxels(P, out=out)
        dt = time.perf_counter() - t0
        return P, dt

    P, dt = run(512)
    rate_e = 512 / max(dt, 1e-6)
    e1 = int(min(conn.shape[0], max(1024, rate_e * budget_s)))
    P, dt = run(e1)
    # fragment count of the sample from the oracle (mass field only: the count
    # does not depend on the field set)
    Pm = O.Problem(pos, vel, mass[:e1], conn[:e1], periodic_box=box, n=(gn,) * 3, bounds=box)
    _, st = O.voxels(Pm, threads=-1)
    nfrag = st["nfrag"]
    return {"value": nfrag / dt, "unit": "fragments/s", "cores": 1, "kind": kind,
            "sample": f"first {e1} of {conn.shape[0]} elements ({6 * e1} tets, {nfrag} fragments) of the same "
                      f"mesh onto the full {gn}^3 grid, fields m{''.join(',' + f for f in fields)}; "
                      f"{dt:.1f} s single-threaded, -O3 x86-64-v3"}


def cpu_all_cores(pos, vel, mass, conn, gn, fields, budget_s=10.0):
    """The oracle's C restatement with OpenMP over elements on all host
    cores the process may use (per-thread private grids summed at the end),
    on the first elements of the same mesh, about budget_s of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    nth = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))

    def run(e1):
        P = O.Problem(pos, vel, mass[:e1], conn[:e1], periodic_box=box, n=(gn,) * 3, bounds=box, fields=fields)
        t0 = time.perf_counter()
        _, st = O.voxels(P, threads=nth)
        return st, time.perf_counter() - t0

    st, dt = run(min(conn.shape[0], 4096 * nth))
    e1 = int(min(conn.shape[0], max(4096, 4096 * nth * budget_s / max(dt, 1e-3))))
    st, dt = run(e1)
    return {"value": st["nfrag"] / dt, "unit": "fragments/s", "cores": nth, "kind": "port",
            "sample": f"first {e1} of {conn.shape[0]} elements ({st['nfrag']} fragments) onto the full {gn}^3 grid, "
                      f"fields m{''.join(',' + f for f in fields)}; {dt:.1f} s on {nth} threads, private grids summed"}


def end_to_end_leg(pos, vel, mass, conn, gn, fields, nfrag, steps=2, warmup=1):
    """PSI.voxels on NumPy arrays (the drop-in call): uploads of the mesh and
    the fields' contents, the device work, the download of the fields."""
    import numpy as np
    from pypsi_amd import PSI
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    n = round(conn.shape[0] ** (1.0 / 3.0))
    mesh = PSI.Mesh(loader="block", pos=pos.reshape(n, n, n, 3), vel=vel.reshape(n, n, n, 3), periodic_box=box)
    ts = []
    for i in range(warmup + steps):
        grid = PSI.Grid(type="cart", n=(gn,) * 3, bounds=box, fields=list(fields))
        t0 = time.perf_counter()
        PSI.voxels(grid=grid, mesh=mesh)
        if i >= warmup:
            ts.append(time.perf_counter() - t0)
    ms = 1e3 * sum(ts) / len(ts)
    return {"value": nfrag / (ms * 1e-3), "unit": "fragments/s", "ms_per_step": ms, "steps": steps,
            "what": "PSI.voxels(grid, mesh) on host NumPy arrays: PCIe in and out included"}


def snapshot_leg(steps=3, warmup=1):
    """The notebook's Example 2 call pattern (tutorial/Tutorial.ipynb:160-186)
    on a written Gadget-2 snapshot: PSI.Mesh(loader='gadget2') of a 32^3
    Zel'dovich lattice (rms 2.0 cells, seed 0) in a 40-unit box, then
    PSI.voxels onto 128^3 with v and vv at refinement level 2, tolerance 0.1,
    from host arrays.  Also the loader alone on a 128^3 file (2,097,152
    particles, 58.7 MB)."""
    import tempfile

    import numpy as np
    import torch
    from pypsi_amd import PSI, snapshot
    from pypsi_amd import device as D
    out = {"workload": "gadget2-32^3-rms2.0-box40-to-128^3-m,v,vv-L2-tol0.1"}
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "snap32")
        snapshot.zeldovich_snapshot(fn, 32, 2.0, seed=0, box=40.0, mass1=14.96)
        t0 = time.perf_counter()
        mesh = PSI.Mesh(loader="gadget2", filename=fn)
        out["load_ms"] = 1e3 * (time.perf_counter() - t0)
        # the GPU's own fragment count for this input (the FP32 kernel's is within 1e-6 of the FP64 count)
        dev = torch.device("cuda:0")
        tf = {"m": torch.zeros((128,) * 3, dtype=torch.float64, device=dev),
              "v": torch.zeros((128,) * 3 + (3,), dtype=torch.float64, device=dev),
              "vv": torch.zeros((128,) * 3 + (3, 3), dtype=torch.float64, device=dev)}
        dm = D.make_mesh(*(torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                           for a in (mesh.pos, mesh.vel, mesh.mass, mesh.connectivity)), periodic_box=mesh.wrapbox)
        st = D.voxels(dm, D.make_grid((128,) * 3, mesh.wrapbox, tf),
                      D.make_opts(refine_tolerance=0.1, refine_max_lvl=2,
                                   stream=torch.cuda.current_stream().cuda_stream))
        nfrag = int(st["nfrag"])
        ts = []
        for i in range(warmup + steps):
            grid = PSI.Grid(n=(128,) * 3, bounds=mesh.wrapbox, fields=["v", "vv"])
            t0 = time.perf_counter()
            PSI.voxels(grid=grid, mesh=mesh, refine_max_lvl=2, refine_tolerance=0.1)
            if i >= warmup:
                ts.append(time.perf_counter() - t0)
        ms = 1e3 * sum(ts) / len(ts)
        out.update({"value": nfrag / (ms * 1e-3), "unit": "fragments/s", "ms_per_step": ms, "steps": steps,
                    "gpu_nfrag": nfrag, "tets": int(st["ntets"]),
                    "mass_rel_err": abs(1.0 - float(grid.fields["m"].sum()) / float(mesh.mass.sum())),
                    "what": "PSI.voxels(grid, mesh) on the loaded host arrays: PCIe in and out included"})
        del mesh
        fn = os.path.join(d, "snap128")
        snapshot.zeldovich_snapshot(fn, 128, 2.0, seed=0, box=40.0, mass1=14.96)
        nbytes = os.path.getsize(fn)
        lt = []
        for _ in range(2):
            t0 = time.perf_counter()
            big = PSI.Mesh(loader="gadget2", filename=fn)
            lt.append(time.perf_counter() - t0)
            del big
        out["load_128^3"] = {"ms": 1e3 * min(lt), "file_bytes": nbytes, "MB_per_s": nbytes / min(lt) / 1e6,
                             "what": "PSI.Mesh(loader='gadget2'), particle ids in random order in the file"}
    return out


def vdf_leg(pos, vel, npts=400):
    """PSI.RStarTree over the 128^3 mesh (host arrays; the element boxes are
    built on the GPU), then PSI.VDF at random points of [0.1, 0.9]^3 with that
    tree, refinement level 0; the oracle's C restatement times a few of the
    same points (it scans every element's box per query)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pypsi_amd import PSI
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    n = round(pos.shape[0] ** (1.0 / 3.0))
    mesh = PSI.Mesh(loader="block", pos=pos.reshape(n, n, n, 3), vel=vel.reshape(n, n, n, 3), periodic_box=box)
    t0 = time.perf_counter()
    tree = PSI.RStarTree(mesh)
    build_ms = 1e3 * (time.perf_counter() - t0)
    pts = np.random.default_rng(0).uniform(0.1, 0.9, size=(npts, 3))
    for x in pts[:5]:
        PSI.VDF(mesh, tuple(x), tree=tree)
    streams = 0
    t0 = time.perf_counter()
    for x in pts:
        rho, _ = PSI.VDF(mesh, tuple(x), tree=tree)
        streams += len(rho)
    us = 1e6 * (time.perf_counter() - t0) / npts
    out = {"workload": f"zeldovich-{n}^3-rms2.0, VDF at {npts} random points, refine level 0",
           "tree_build_ms": build_ms, "us_per_point": us, "points_per_s": 1e6 / us,
           "streams_per_point": streams / npts,
           "what": "PSI.VDF(mesh, x, tree) from host: box scan, candidate refinement and barycentric tests on the GPU"}
    # against the reference's own path: its R*-tree built once, then
    # psi_sample_vdf through it (as its PSI.VDF(mesh, x, tree=t)); on the
    # 64^3 mesh of the same construction (the reference's tree over 128^3
    # takes some 40 s to build), the GPU on the same mesh and points
    try:
        import oracle as O
        from pypsi_amd.synth import zeldovich_block
        if O.have_ref():
            p64, v64 = zeldovich_block(64, 2.0, 0)
            P = O.Problem.block(p64, v64, periodic_box=box, n=(8,) * 3)
            m64 = PSI.Mesh(loader="block", pos=p64, vel=v64, periodic_box=box)
            t0 = time.perf_counter()
            t64 = PSI.RStarTree(m64)
            gbuild = 1e3 * (time.perf_counter() - t0)
            t0 = time.perf_counter()
            for x in pts[:200]:
                PSI.VDF(m64, tuple(x), tree=t64)
            gus = 1e6 * (time.perf_counter() - t0) / 200
            t0 = time.perf_counter()
            T = O.RefTree(P, O.tree_window(P))
            rbuild = 1e3 * (time.perf_counter() - t0)
            t0 = time.perf_counter()
            for x in pts[:200]:
                T.sample_vdf(x)
            rus = 1e6 * (time.perf_counter() - t0) / 200
            T.close()
            out["vs_reference_64^3"] = {"gpu_tree_build_ms": gbuild, "gpu_us_per_point": gus,
                                         "cpu_reference_tree_build_ms": rbuild, "cpu_reference_us_per_point": rus,
                                         "cpu_reference_cores": 1,
                                         "what": "200 points, refine level 0; the reference's R*-tree and "
                                                 "psi_sample_vdf single-threaded"}
    except Exception as e:  # report, never hide
        out["cpu_reference_error"] = repr(e)
    return out


def annihilation_leg(steps=3):
    """mode='annihilation' (the integral of rho^2 per cell from every pair of
    overlapping tets) from device-resident inputs: Zel'dovich 16^3 (rms 1.0
    cells) onto 32^3, with the reference's own C core (oracle/_ref) timed on
    the same input single-threaded, and 32^3 (rms 0.3) onto 64^3."""
    import numpy as np
    import torch
    from pypsi_amd import device as D
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dev = torch.device("cuda:0")
    out = {}
    for n, rms, gn in ((16, 1.0, 32), (32, 0.3, 64)):
        from pypsi_amd.synth import zeldovich_block
        pos, vel = zeldovich_block(n, rms, 0)
        i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        cn = np.empty((n, n, n, 8), dtype=np.int32)
        for c in range(8):
            cn[..., c] = n * n * ((i + (c >> 2)) % n) + n * ((j + ((c >> 1) & 1)) % n) + ((k + (c & 1)) % n)
        cn = cn.reshape(-1, 8)
        mass = np.ones(cn.shape[0])
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (pos.reshape(-1, 3), vel.reshape(-1, 3), mass, cn)]
        m = torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)
        mesh = D.make_mesh(*t, periodic_box=box)
        grid = D.make_grid((gn,) * 3, box, {"m": m})
        opts = D.make_opts(mode="annihilation", stream=torch.cuda.current_stream().cuda_stream)
        st = D.voxels(mesh, grid, opts)
        torch.cuda.synchronize()
        ts = []
        for _ in range(steps):
            m.zero_()
            t0 = time.perf_counter()
            st = D.voxels(mesh, grid, opts)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ms = 1e3 * sum(ts) / len(ts)
        leg = {"ms": ms, "tet_pairs": st["nghosts"], "tet_pairs_per_s": st["nghosts"] / (ms * 1e-3),
               "fragments": st["nfrag"], "pair_kernel_ms": st["kernel_ms"]}
        if n == 16:
            try:
                import oracle as O
                if O.have_ref():
                    P = O.Problem.block(pos, vel, periodic_box=box, n=(gn,) * 3)
                    t0 = time.perf_counter()
                    O.ref_voxels(P, mode=1)
                    dt = time.perf_counter() - t0
                    leg["cpu_reference_ms"] = 1e3 * dt
                    leg["cpu_reference_cores"] = 1
            except Exception as e:  # report, never hide
                leg["cpu_reference_error"] = repr(e)
        out[f"zeldovich-{n}^3-rms{rms}-to-{gn}^3"] = leg
    return out


def plain_leg(mesh, grid, tf, stream, nfrag, steps=2, warmup=1, kernel="plain"):
    """An FP64 kernel on the same inputs: 'plain', one thread per tet with the
    reference's cut sequence (the plain GPU method: its four polyhedra live
    in per-lane scratch memory), or 'fast64', the column kernel with every
    tet in FP64 (the column method at the reference's precision)."""
    import torch
    from pypsi_amd import device as D
    opts = D.make_opts(kernel=kernel, stream=stream)

    def step():
        for t in tf.values():
            t.zero_()
        return D.voxels(mesh, grid, opts)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = []
    for _ in range(steps):
        kms.append(step()["kernel_ms"])
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return {"value": nfrag / (ms * 1e-3), "unit": "fragments/s", "ms_per_step": ms, "kernel_ms_avg": sum(kms) / steps,
            "steps": steps, "dtype": "f64"}


UNEVEN = "zeldovich-uneven64-16^3..96^3-rms0.2..3.0-m"

# refined workload: lattice n, rms cells, seed, grid n, fields, refine level,
# refine tolerance (cells), infall clumps (seeded centres, radius 10 cells)
REFINED = {"clumps8-128^3-rms1.5-L2-tol0.1-to-512^3-m,v,vv": (128, 1.5, 0, 512, ("v", "vv"), 2, 0.1, 8)}


def build_refined(wname):
    """Zel'dovich lattice plus seeded spherical infall clumps; velocities are
    the total displacements (wrapped to the nearest image)."""
    import numpy as np
    from pypsi_amd.synth import add_infall_clumps, zeldovich_block
    n, rms, seed, gn, fields, lvl, tol, nclumps = REFINED[wname]
    pos0, vel0 = zeldovich_block(n, rms, seed)
    pos = add_infall_clumps(pos0, n, nclumps=nclumps, seed=100)
    d = pos - pos0
    d -= np.round(d)
    vel = np.ascontiguousarray(vel0 + d)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    conn = np.empty((n, n, n, 8), dtype=np.int32)
    for ii in range(2):
        for jj in range(2):
            for kk in range(2):
                conn[..., 4 * ii + 2 * jj + kk] = n * n * ((i + ii) % n) + n * ((j + jj) % n) + ((k + kk) % n)
    conn = conn.reshape(-1, 8)
    return pos.reshape(-1, 3), vel.reshape(-1, 3), np.ones(conn.shape[0]), conn, gn, fields, lvl, tol


def refined_leg(wname, steps=2, warmup=1, kernel="auto", cpu_sample_elements=0):
    """The refined workload on the GPU: on-device level-2 refinement then
    the column kernel, fields m, v, vv on a 512^3 FP64 grid."""
    import torch
    from pypsi_amd import device as D
    pos, vel, mass, conn, gn, fields, lvl, tol = build_refined(wname)
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dev = torch.device("cuda:0")
    tpos, tvel = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
    tmass, tconn = torch.from_numpy(mass).to(dev), torch.from_numpy(conn).to(dev)
    tf = {"m": torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)}
    for f in fields:
        tf[f] = torch.zeros((gn,) * 3 + ((3,) if len(f) == 1 else (3, 3)), dtype=torch.float64, device=dev)
    mesh = D.make_mesh(tpos, tvel, tmass, tconn, periodic_box=box)
    grid = D.make_grid((gn,) * 3, box, tf)
    opts = D.make_opts(kernel=kernel, refine_max_lvl=lvl, refine_tolerance=tol,
                       stream=torch.cuda.current_stream().cuda_stream)

    def step():
        for t in tf.values():
            t.zero_()
        return D.voxels(mesh, grid, opts)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = []
    for _ in range(steps):
        st = step()
        kms.append(st["kernel_ms"])
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    mass_err = abs(1.0 - float(tf["m"].sum().item()) / float(mass.sum()))
    rec = load_counts().get(wname, {})
    nfrag = rec.get("nfrag")
    src = rec.get("source", "oracle")
    cpu = None
    if cpu_sample_elements:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        e1 = cpu_sample_elements
        P = O.Problem(pos, vel, mass[:e1], conn[:e1], periodic_box=box, n=(gn,) * 3, bounds=box, fields=fields)
        cpu = cpu_sample([P], reftol=tol, maxlvl=lvl,
                         what=f"first {e1} of {conn.shape[0]} elements, refinement level {lvl} (tolerance {tol} cells), "
                              f"onto the full {gn}^3 grid, fields m,{','.join(fields)}; single-threaded")
    if nfrag is None:
        nfrag, src = int(st["nfrag"]), "gpu (count not cached)"
    return {"workload": wname, "value": nfrag / (ms * 1e-3), "unit": "fragments/s", "ms_per_step": ms,
            "kernel_ms_avg": sum(kms) / steps, "tets": int(st["ntets"]), "nfrag": nfrag, "nfrag_source": src,
            "gpu_nfrag": int(st["nfrag"]), "mass_rel_err": mass_err, "kernel": ["auto", "plain", "fast"][st["kernel"]],
            "wide_tets": int(st["nwide"]), "steps": steps, "cpu_baseline": cpu}


def count_fragments(wname):
    """Fragment count of a refined workload with the FP64 kernel (it keeps
    the reference's fragments one for one), stored beside the oracle counts."""
    t0 = time.perf_counter()
    r = refined_leg(wname, steps=1, warmup=0, kernel="plain")
    p = os.path.join(ROOT, "profiles", "workload_counts.json")
    data = load_counts()
    data[wname] = {"nfrag": r["gpu_nfrag"], "ntets": r["tets"], "source": "FP64 GPU kernel (k_vox_plain)",
                   "seconds": time.perf_counter() - t0}
    with open(p, "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps(data[wname]), flush=True)


def batched_leg(steps=3, warmup=1, cpu=True):
    """64 meshes (seeds 0..63): lattice n log-uniform in 16..96, rms uniform
    in 0.2..3.0 cells, each onto its own grid 2^round(log2(8n/3)) clipped to
    32..256, mass only.  One psivox_voxels_batched call per step, beside a
    loop of single-mesh calls on the same device-resident inputs."""
    import numpy as np
    import torch
    from pypsi_amd import device as D
    from pypsi_amd.synth import uneven_batch_specs, zeldovich_block
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dev = torch.device("cuda:0")
    meshes, grids, fields, host = [], [], [], []
    ntet = 0
    for n, rms, gn, seed in uneven_batch_specs(64):
        pos, _ = zeldovich_block(n, rms, seed)
        i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        conn = np.empty((n, n, n, 8), dtype=np.int32)
        for ii in range(2):
            for jj in range(2):
                for kk in range(2):
                    conn[..., 4 * ii + 2 * jj + kk] = n * n * ((i + ii) % n) + n * ((j + jj) % n) + ((k + kk) % n)
        conn = conn.reshape(-1, 8)
        host.append((pos, conn))
        tp = torch.from_numpy(pos.reshape(-1, 3)).to(dev)
        tc = torch.from_numpy(conn).to(dev)
        tm = torch.ones(conn.shape[0], dtype=torch.float64, device=dev)
        m = torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)
        meshes.append(D.make_mesh(tp, None, tm, tc, box))
        grids.append(D.make_grid((gn,) * 3, box, {"m": m}))
        fields.append(m)
        ntet += 6 * conn.shape[0]
    stream = torch.cuda.current_stream().cuda_stream
    opts = D.make_opts(kernel="fast", stream=stream)

    def run(batched):
        for m in fields:
            m.zero_()
        if batched:
            return D.voxels_batched(meshes, grids, opts)
        return [D.voxels(mm, g, opts) for mm, g in zip(meshes, grids)]

    res = {}
    for label, batched in (("batched", True), ("loop", False)):
        for _ in range(warmup):
            run(batched)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sts = run(batched)
        torch.cuda.synchronize()
        res[label] = 1e3 * (time.perf_counter() - t0) / steps
        res[label + "_gpu_nfrag"] = int(sum(st["nfrag"] for st in sts))
    mass_err = max(abs(float(m.sum().item()) / float(meshes[q].nelem) - 1.0) for q, m in enumerate(fields))
    # CPU: the reference's C core on the first 1% of every mesh's elements
    import oracle as O
    probs = []
    for (n, rms, gn, seed), (pos, conn) in zip(uneven_batch_specs(64), host):
        e1 = max(1, conn.shape[0] // 100)
        probs.append(O.Problem(pos.reshape(-1, 3), None, None, conn[:e1], periodic_box=box, n=(gn,) * 3, bounds=box))
    if cpu:
        cpu = cpu_sample(probs, what="the first 1% of the elements of each of the 64 meshes, each onto its own grid, "
                                     "mass; single-threaded, -O3 x86-64-v3")
    else:
        cpu = None
    nfrag = load_counts().get(UNEVEN, {}).get("nfrag")
    src = "oracle"
    if nfrag is None:
        nfrag, src = res["batched_gpu_nfrag"], "gpu (oracle count not cached)"
    return {"workload": UNEVEN, "meshes": 64, "tets": ntet, "nfrag": nfrag, "nfrag_source": src,
            "value": nfrag / (res["batched"] * 1e-3), "unit": "fragments/s", "ms_per_step": res["batched"],
            "loop_ms_per_step": res["loop"], "loop_value": nfrag / (res["loop"] * 1e-3),
            "gpu_nfrag": res["batched_gpu_nfrag"], "max_mass_rel_err": mass_err, "steps": steps,
            "cpu_baseline": cpu}


def parity_leg(pos, vel, mass, conn, gn, fields, nsub=1 << 19):
    """Per-voxel errors at the headline configuration: the first nsub
    elements of the mesh (lattice x-slabs 0..31 of 128: 3.1M tets) onto the
    full grid, every field, each GPU kernel against the FP64 oracle (the
    errors of tests/helpers.py: relative on cells holding >= 1e-3 of the
    peak mass, mass-weighted absolute over all cells)."""
    import torch
    from pypsi_amd import device as D
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from helpers import field_errors
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    P = O.Problem(pos, vel, mass[:nsub], conn[:nsub], periodic_box=box, n=(gn,) * 3, bounds=box, fields=fields)
    t0 = time.perf_counter()
    ref, ost = O.voxels(P, threads=-1)
    oracle_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    mesh = D.make_mesh(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev),
                       torch.from_numpy(mass[:nsub]).to(dev), torch.from_numpy(conn[:nsub]).to(dev), periodic_box=box)
    out = {"sample": f"first {nsub} of {conn.shape[0]} elements ({6 * nsub} tets) onto the full {gn}^3 grid",
           "oracle_nfrag": ost["nfrag"], "oracle_s": oracle_s}
    for kernel in ("auto", "plain"):
        tf = {"m": torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)}
        for f in fields:
            tf[f] = torch.zeros((gn,) * 3 + ((3,) if len(f) == 1 else (3, 3)), dtype=torch.float64, device=dev)
        st = D.voxels(mesh, D.make_grid((gn,) * 3, box, tf), D.make_opts(kernel=kernel))
        got = {k: v.cpu().numpy() for k, v in tf.items()}
        errs = field_errors(got, ref)
        out[kernel] = {"errors": {k: {"rel": v[0], "wabs": v[1]} for k, v in errs.items()},
                       "mass_rel_err": abs(1.0 - float(got["m"].sum()) / float(ref["m"].sum())),
                       "gpu_nfrag": int(st["nfrag"]), "wide_tets": int(st["nwide"])}
        del tf, got
    return out


C1 = "zeldovich-32^3-rms0.3-to-64^3-m"


def c1_leg(steps=20, warmup=3):
    """The reference's own CPU-runnable case: 32^3 particles (rms 0.3 cells)
    onto 64^3, mass.  The GPU kernels from device-resident inputs, beside the
    reference's C core on the whole mesh single-threaded and the oracle on
    all host cores."""
    import torch
    from pypsi_amd import device as D
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    pos, vel, mass, conn, gn, fields = build_inputs(C1)
    nfrag = load_counts().get(C1, {}).get("nfrag")
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(a).to(dev) for a in (pos, vel, mass, conn)]
    mesh = D.make_mesh(*t, periodic_box=box)
    m = torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)
    grid = D.make_grid((gn,) * 3, box, {"m": m})
    out = {"workload": C1, "nfrag": nfrag}
    for kernel in ("auto", "plain"):
        opts = D.make_opts(kernel=kernel, stream=torch.cuda.current_stream().cuda_stream)
        for _ in range(warmup):
            m.zero_()
            st = D.voxels(mesh, grid, opts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.zero_()
            st = D.voxels(mesh, grid, opts)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        n = nfrag or int(st["nfrag"])
        out[kernel] = {"value": n / (ms * 1e-3), "unit": "fragments/s", "ms_per_step": ms, "gpu_nfrag": int(st["nfrag"])}
    P = O.Problem(pos, vel, mass, conn, periodic_box=box, n=(gn,) * 3, bounds=box)
    if O.have_ref():
        t0 = time.perf_counter()
        O.ref_voxels(P)
        dt = time.perf_counter() - t0
        out["cpu_reference"] = {"value": (nfrag or 0) / dt, "unit": "fragments/s", "cores": 1, "kind": "reference",
                                "s": dt, "sample": "the whole mesh"}
    nth = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    O.voxels(P, threads=nth)
    dt = time.perf_counter() - t0
    out["cpu_oracle_all_cores"] = {"value": (nfrag or 0) / dt, "unit": "fragments/s", "cores": nth, "kind": "port",
                                   "s": dt, "sample": "the whole mesh"}
    return out


def cpu_sample(probs, reftol=1.0, maxlvl=0, what=""):
    """The reference's C core single-threaded over a list of oracle Problems
    (a bounded sample of a workload), fragments counted by the oracle on
    all host cores: {value fragments/s, ...}."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    kind = "reference" if O.have_ref() else "port"
    dt, nfrag = 0.0, 0
    for P in probs:
        out = P.zero_fields()
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_voxels(P, out=out, reftol=reftol, maxlvl=maxlvl)
        else:
            O.voxels(P, out=out, reftol=reftol, maxlvl=maxlvl)
        dt += time.perf_counter() - t0
        del out
        Pm = O.Problem(P.pos, P.vel, P.mass, P.conn, periodic_box=P.box if P.periodic else None, n=P.n,
                       bounds=P.bounds)
        _, st = O.voxels(Pm, reftol=reftol, maxlvl=maxlvl, threads=-1)
        nfrag += st["nfrag"]
    return {"value": nfrag / dt, "unit": "fragments/s", "cores": 1, "kind": kind, "sample": what,
            "sample_fragments": nfrag, "s": dt}


def fwd_adjoint_cpu(pos, mass, conn, gn, w, gpu_fwd_ms, gpu_adj_ms, frac=8):
    """The same forward and mass adjoint on the host: the reference's own C
    core single-threaded for the forward and the oracle's FP64 adjoint
    single-threaded, each on the first 1/frac of the elements (a bounded
    sample), scaled to the whole mesh by the element count."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    e1 = max(1, conn.shape[0] // frac)
    P = O.Problem(pos.reshape(-1, 3), None, np.ascontiguousarray(mass[:e1]), np.ascontiguousarray(conn[:e1]),
                  periodic_box=box, n=(gn,) * 3, bounds=box)
    scale = conn.shape[0] / e1
    t0 = time.perf_counter()
    kind = "reference"
    if O.have_ref():
        O.ref_voxels(P)
    else:
        O.voxels(P)
        kind = "port"
    t1 = time.perf_counter()
    O.mass_adjoint(P, w)
    t2 = time.perf_counter()
    fwd, adj = 1e3 * (t1 - t0) * scale, 1e3 * (t2 - t1) * scale
    return {"fwd_ms": fwd, "fwd_kind": kind, "adjoint_ms": adj, "adjoint_kind": "port (the oracle's FP64 adjoint)",
            "ratio_fwd_plus_adj_over_fwd": (fwd + adj) / fwd, "cores": 1,
            "sample": f"first {e1} of {conn.shape[0]} elements, scaled by the element count",
            "gpu_speedup_fwd_plus_adj": (fwd + adj) / (gpu_fwd_ms + gpu_adj_ms)}


def fwd_adjoint_leg(steps=5, warmup=2, cpu=True):
    """64^3 particles, rms 1.0 cells -> 128^3, mass: forward (FP64-clip
    column kernel, voxel_density's default) and the mass adjoint of
    L = sum (rho - rho_t)^2, rho_t from the same construction with seed 1.
    Reports t(fwd + adj) / t(fwd), and the same with the FP32-clip forward."""
    import torch
    from pypsi_amd import device as D
    pos, vel, mass, conn, gn, _ = build_inputs("zeldovich-64^3-rms1.0-to-128^3-m")
    post, _, _, _, _, _ = build_inputs_seed("zeldovich-64^3-rms1.0-to-128^3-m", 1)
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dev = torch.device("cuda:0")
    tconn = torch.from_numpy(conn).to(dev)
    tmass = torch.from_numpy(mass).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    opts = D.make_opts(kernel="fast64", stream=stream)
    o32 = D.make_opts(kernel="auto", stream=stream)
    V = (1.0 / gn) ** 3
    m = torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)
    grid = D.make_grid((gn,) * 3, box, {"m": m})
    # target density
    tpos_t = torch.from_numpy(post).to(dev)
    D.voxels(D.make_mesh(tpos_t, None, tmass, tconn, box), grid, opts)
    rho_t = m / V
    tpos = torch.from_numpy(pos).to(dev)
    mesh = D.make_mesh(tpos, None, tmass, tconn, box)
    g = torch.zeros_like(tpos)
    tf, ta = [], []
    for it in range(warmup + steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.zero_()
        D.voxels(mesh, grid, opts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        w = (2.0 / V) * (m / V - rho_t)
        g.zero_()
        D.mass_adjoint(mesh, grid, w, g, opts)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it >= warmup:
            tf.append(t1 - t0)
            ta.append(t2 - t1)
    f, a = 1e3 * sum(tf) / steps, 1e3 * sum(ta) / steps
    t32 = []
    for it in range(warmup + steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.zero_()
        D.voxels(mesh, grid, o32)
        torch.cuda.synchronize()
        if it >= warmup:
            t32.append(time.perf_counter() - t0)
    f32 = 1e3 * sum(t32) / steps
    out = {"workload": "zeldovich-64^3-rms1.0-to-128^3-m, loss sum (rho - rho_t)^2, rho_t seed 1",
           "forward_kernel": "fast64 (FP64 clip; the default of voxel_density, within 1e-4 of FP64 FD)",
           "fwd_ms": f, "adjoint_ms": a, "ratio_fwd_plus_adj_over_fwd": (f + a) / f, "steps": steps,
           "adjoint_dtype": "f64",
           "fp32_forward": {"kernel": "auto", "fwd_ms": f32, "ratio_fwd_plus_adj_over_fwd": (f32 + a) / f32,
                            "gradient_vs_fd": "within 3e-4 (FP32 cell masses)"}}
    if cpu:
        out["cpu"] = fwd_adjoint_cpu(pos, mass, conn, gn, w.cpu().numpy(), f, a)
    # the fields m, v, vv and their adjoint with respect to positions and
    # velocities (voxel_fields: FP64 column forward)
    vel = torch.from_numpy(build_inputs("zeldovich-64^3-rms1.0-to-128^3-m")[1]).to(dev)
    fl = {"m": m, "v": torch.zeros((gn,) * 3 + (3,), dtype=torch.float64, device=dev),
          "vv": torch.zeros((gn,) * 3 + (3, 3), dtype=torch.float64, device=dev)}
    fgrid = D.make_grid((gn,) * 3, box, fl)
    fmesh = D.make_mesh(tpos, vel, tmass, tconn, box)
    o64 = D.make_opts(kernel="fast64", stream=stream)
    gr = {k: torch.randn(t.shape, dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(7))
          for k, t in fl.items()}
    gv = torch.zeros_like(vel)
    tf, ta = [], []
    for it in range(warmup + steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in fl.values():
            t.zero_()
        D.voxels(fmesh, fgrid, o64)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g.zero_()
        gv.zero_()
        D.adjoint(fmesh, fgrid, gr, g, gv, opts)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it >= warmup:
            tf.append(t1 - t0)
            ta.append(t2 - t1)
    f, a = 1e3 * sum(tf) / steps, 1e3 * sum(ta) / steps
    out["fields_m_v_vv"] = {"forward_kernel": "fast64 (FP64 columns; the default of voxel_fields)", "fwd_ms": f,
                            "adjoint_ms": a, "ratio_fwd_plus_adj_over_fwd": (f + a) / f,
                            "what": "dL/dpos and dL/dvel of a loss on m, v, vv"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="auto", choices=["auto", "plain", "fast"])
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the per-voxel comparison with the oracle")
    ap.add_argument("--count-fragments", choices=sorted(REFINED),
                    help="count a refined workload's fragments with the FP64 kernel into profiles/ and exit")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the plain-kernel, forward+adjoint, batched, refined, end-to-end and snapshot legs")
    ap.add_argument("--only", choices=["snapshot", "vdf", "annihilation", "end_to_end"],
                    help="run one extra leg alone and print it")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported "
              "(the path shards by tets, one mesh slice and one grid copy per GPU, summed at the end)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pypsi_amd import device as D

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.count_fragments:
        count_fragments(args.count_fragments)
        return
    if args.only == "snapshot":
        print(json.dumps({"snapshot": snapshot_leg()}), flush=True)
        return
    if args.only == "annihilation":
        print(json.dumps({"annihilation": annihilation_leg()}), flush=True)
        return

    wname = args.workload
    log(f"[bench] building inputs for {wname}")
    pos, vel, mass, conn, gn, fields = build_inputs(wname)
    if args.only == "vdf":
        print(json.dumps({"vdf": vdf_leg(pos, vel)}), flush=True)
        return
    if args.only == "end_to_end":
        nfrag = load_counts().get(wname, {}).get("nfrag")
        print(json.dumps({"end_to_end": end_to_end_leg(pos, vel, mass, conn, gn, fields, nfrag)}), flush=True)
        return
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    dev = torch.device("cuda:0")
    tpos = torch.from_numpy(pos).to(dev)
    tvel = torch.from_numpy(vel).to(dev)
    tmass = torch.from_numpy(mass).to(dev)
    tconn = torch.from_numpy(conn).to(dev)
    tf = {"m": torch.zeros((gn,) * 3, dtype=torch.float64, device=dev)}
    for f in fields:
        tf[f] = torch.zeros((gn,) * 3 + ((3,) if len(f) == 1 else (3, 3)), dtype=torch.float64, device=dev)
    mesh = D.make_mesh(tpos, tvel, tmass, tconn, periodic_box=box)
    grid = D.make_grid((gn,) * 3, box, tf)
    stream = torch.cuda.current_stream().cuda_stream
    opts = D.make_opts(kernel=args.kernel, stream=stream)

    def step():
        for t in tf.values():
            t.zero_()
        return D.voxels(mesh, grid, opts)

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        st = step()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        kms.append(st["kernel_ms"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    gpu_nfrag = int(st["nfrag"])
    mass_err = abs(1.0 - float(tf["m"].sum().item()) / float(mass.sum()))

    counts = load_counts()
    nfrag = counts.get(wname, {}).get("nfrag")
    nfrag_src = "oracle"
    if nfrag is None:
        nfrag, nfrag_src = gpu_nfrag, "gpu (oracle count not cached)"
    value = nfrag / (ms_step * 1e-3)

    fk = field_key(fields)
    kavg = sum(kms) / len(kms)
    achieved = nfrag * FLOPS_PER_FRAG[fk] / (kavg * 1e-3) / 1e12
    # HBM bytes per launch of the voxelization kernel from rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel's reading
    # (pypsi_amd/profiling.py); None until measured for this workload
    tr = load_traffic().get(wname, {}).get("plain" if st["kernel"] == 1 else "fast") or {}
    traffic = tr.get("bytes")
    out = {
        "metric": "tet-voxel fragments/s, 128^3 mesh\u2192256^3 grid; plain vs batched, fwd+adjoint",
        "value": value,
        "unit": "fragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if st["kernel"] == 2 else "f64",
        "data": "synthetic (seeded Zel'dovich lattice)",
        "config": {"workload": wname, "kernel": args.kernel, "tets": conn.shape[0] * 6,
                   "grid": gn, "fields": ["m"] + list(fields), "nfrag": nfrag, "nfrag_source": nfrag_src,
                   "gpu_nfrag": gpu_nfrag, "mass_rel_err": mass_err, "kernel_ms_avg": kavg,
                   "clip": "f32" if st["kernel"] == 2 else "f64", "accumulate": "f64",
                   "tets_fp64_columns": int(st.get("nwide", 0))},
        # compute-bound: the column kernel is FP32 vector-ALU work, priced at
        # the FP32 peak (the bound field names compute as "mfma")
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "flops_per_fragment": FLOPS_PER_FRAG[fk],
                     "kernel": "k_vox_cols, FP32 pass plus the FP64 pass of the wide tets (durations summed)",
                     "note": "FP32 vector ALU (no matrix cores); traffic is the FP32 pass's FETCH_SIZE + WRITE_SIZE"},
    }
    if not args.no_parity:
        log("[bench] per-voxel errors against the oracle on a sample of the same workload")
        try:
            par = parity_leg(pos, vel, mass, conn, gn, fields)
            out["parity"] = par
            out["config"]["max_rel_err"] = {f: e["rel"] for f, e in par["auto"]["errors"].items()}
        except Exception as e:  # report, never hide
            out["parity"] = {"error": repr(e)}
    if not args.no_extra:
        log("[bench] the reference's CPU-runnable case, 32^3 -> 64^3")
        out["c1"] = c1_leg()
        log("[bench] plain FP64 kernel on the same workload")
        out["plain"] = plain_leg(mesh, grid, tf, stream, nfrag)
        log("[bench] FP64 column kernel on the same workload")
        out["fp64_columns"] = plain_leg(mesh, grid, tf, stream, nfrag, steps=3, kernel="fast64")
        out["fp64_columns"]["what"] = ("the column kernel with every tet clipped in FP64 (kernel 'fast64'): "
                                       "the FP32 path's speed-up at equal method, value / this value")
        log("[bench] forward + adjoint, 64^3 -> 128^3")
        out["fwd_adjoint"] = fwd_adjoint_leg(cpu=not args.no_cpu_baseline)
        log("[bench] batched uneven meshes")
        out["batched_uneven"] = batched_leg(cpu=not args.no_cpu_baseline)
        log("[bench] refined clumps, level 2, 512^3")
        out["refined"] = refined_leg(next(iter(REFINED)), cpu_sample_elements=0 if args.no_cpu_baseline else 9000)
        log("[bench] end to end from host arrays")
        out["end_to_end"] = end_to_end_leg(pos, vel, mass, conn, gn, fields, nfrag)
        log("[bench] gadget2 snapshot, Example 2 settings")
        out["snapshot"] = snapshot_leg()
        log("[bench] RStarTree + VDF point queries")
        out["vdf"] = vdf_leg(pos, vel)
        log("[bench] annihilation mode")
        out["annihilation"] = annihilation_leg()
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        try:
            out["cpu_baseline"] = cpu_baseline(wname, pos, vel, mass, conn, gn, fields)
        except Exception as e:  # report, never hide
            out["cpu_baseline"] = {"error": repr(e)}
        log("[bench] cpu, all host cores")
        try:
            out["cpu_baseline_all_cores"] = cpu_all_cores(pos, vel, mass, conn, gn, fields)
        except Exception as e:
            out["cpu_baseline_all_cores"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

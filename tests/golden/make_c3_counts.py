"""Writes tests/golden/c3_counts.json: the FP64 oracle's fragment count for
each of the 64 meshes of the uneven batched workload
(pypsi_amd.synth.uneven_batch_specs: lattice n log-uniform in 16..96, rms
0.2..3.0 lattice cells, periodic unit box, mass field on a grid of
2^round(log2(8n/3)) clipped to 32..256 cells per axis, refinement level 0).

The batched GPU test holds psivox_voxels_batched's per-mesh counts to these
(1e-5 relative).  About 5 minutes on 8 host threads:
    python tests/golden/make_c3_counts.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def main():
    import oracle as O
    from pypsi_amd.synth import uneven_batch_specs, zeldovich_block
    out = []
    t0 = time.perf_counter()
    for n, rms, gn, seed in uneven_batch_specs(64):
        pos, _ = zeldovich_block(n, rms, seed)
        P = O.Problem.block(pos, None, periodic_box=BOX, n=(gn,) * 3, bounds=BOX)
        res, st = O.voxels(P, threads=-1)
        out.append({"n": n, "rms": rms, "grid": gn, "seed": seed, "nfrag": int(st["nfrag"]),
                    "mass": float(res["m"].sum())})
        print(seed, n, gn, st["nfrag"], f"{time.perf_counter() - t0:.0f}s", flush=True)
    with open(os.path.join(HERE, "c3_counts.json"), "w") as f:
        json.dump({"workload": "zeldovich-uneven64-16^3..96^3-rms0.2..3.0-m", "meshes": out,
                   "nfrag_total": sum(m["nfrag"] for m in out)}, f, indent=0)


if __name__ == "__main__":
    main()

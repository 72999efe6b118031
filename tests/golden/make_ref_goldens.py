"""Writes tests/golden/ref_kernels.npz: CUP-CFD's own outputs, from its fvm
headers and mesh classes compiled from the reference sources
(``make -f oracle/ref.mk``, oracle/ref_harness.cpp), for

- every kernel case of tests/kernel_runs.py on the meshes of tests/ref_cases.py,
  FP64 and FP32, with the meshes, inputs and mesh state stored beside them;
- the reference's own StructGen meshes (MeshConfig + MeshSourceStructGenConfig
  over the naive partitioner) on the grids of ref_cases.STRUCTGEN.

Run in a tree that holds the reference sources:
    make -f oracle/ref.mk && python tests/golden/make_ref_goldens.py
The fixture is data only (arrays of inputs and outputs); nothing of the
reference's source travels with it.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("cup-cfd_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import refharness as rh  # noqa: E402
import ref_cases as rc  # noqa: E402
from kernel_runs import run, seed_state  # noqa: E402

OUT = os.path.join(HERE, "ref_kernels.npz")


def main():
    if not rh.available():
        rh.build()
    rec = {}
    for name in rc.MESHES:
        mesh = rc.golden_mesh(name)
        for k, v in rc.mesh_arrays(mesh).items():
            rec[f"{name}/mesh/{k}"] = v
        f = rc.fields(mesh)
        for k, v in f.__dict__.items():
            rec[f"{name}/in/{k}"] = v
        st0 = seed_state(mesh, rc.STATE_SEED)
        for k, v in st0.items():
            rec[f"{name}/state/{k}"] = v
        for prec in (64, 32):
            dt = np.float64 if prec == 64 else np.float32
            for kern in rc.kernels_for(name):
                rm = rh.RefMesh(mesh, prec)
                rm.set_state({k: v.astype(dt) for k, v in st0.items()})
                code, out = run(kern, rh, rm, f.cast(dt), rm.state)
                rec[f"{name}/{prec}/{kern}/rc"] = np.array([code], np.int32)
                for k, v in out.items():
                    rec[f"{name}/{prec}/{kern}/out/{k}"] = np.asarray(v)
                print(name, prec, kern, "rc", code, flush=True)
    for name, (nx, ny, nz, bounds) in rc.STRUCTGEN.items():
        r = rh.structgen(nx, ny, nz, bounds=bounds)
        for k in rc.SG_KEYS:
            rec[f"{name}/{k}"] = r[k]
        print(name, "structgen", r["counts"], flush=True)
    np.savez_compressed(OUT, **rec)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()

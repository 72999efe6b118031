"""Meshes and cases of the reference goldens (tests/golden/ref_kernels.npz).

The goldens hold CUP-CFD's own kernel outputs (its fvm headers compiled from
the reference sources, oracle/ref_harness.cpp) on small meshes, written by
tests/golden/make_ref_goldens.py. These helpers build the same meshes and
inputs, so the generator script, the CPU oracle test and the GPU test agree on
what each case is.

The reference's scalar and UVW face loops read carried locals (PhiFace,
uFace/vFace/wFace, visFace) that only OUTLET and INLET boundary faces assign
(ScalarKernels.ipp:262, 341; UVWKernels.ipp:270-276, 331-335); before the
first such face they are uninitialised. ``outlet_first`` swaps face 0 with the
first OUTLET face, so every carried value the reference reads is defined. On
meshes with only a DEFAULT region those loops read them on every interior
face, so the goldens leave the scalar and UVW cases out there.
"""
from __future__ import annotations

import os

import numpy as np

import cupcfd_hip as ch
from cupcfd_hip import workload as W

FACE_KEYS = ("faceCell1", "faceCell2", "faceBnd", "faceLambda", "faceRLencos", "faceArea", "faceNorm", "faceCenter",
             "faceXpac", "faceXnac")

# name -> HostMesh parameters and whether to put an OUTLET face first
MESHES = {
    # the reference's own StructGen grid (one DEFAULT region), uneven bounds
    "ref345": (dict(nx=3, ny=4, nz=5, bounds=(-1.0, 2.0, 0.0, 1.0, 0.5, 3.0)), False),
    # six typed regions, relabelled, non-zero xpac / xnac / rlencos
    "typed345": (dict(nx=3, ny=4, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                      seed=7), True),
    # jittered, typed, relabelled tetrahedra (2x2x2 hexes x 6)
    "tets222": (dict(nx=2, ny=2, nz=2, tets=True, jitter=0.2, typed=True, permute_cells=True, permute_faces=True,
                     face_extras=True, seed=5), True),
    # the reference's HDF5 test mesh (tests/geometry/mesh/data/MeshHDF5.hdf5) read by cupcfd_hdf5_mesh_create
    "hdf5": (dict(path=os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "MeshHDF5.hdf5")), False),
}

# kernels whose outputs read no uninitialised carried value on a DEFAULT-only mesh
DEFAULT_SAFE = ["gradient", "gradient3", "mass", "mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5",
                "mass_region", "uvw_region1", "uvw_bnds1", "visc_cell1", "visc_region", "visc_boundary", "visc_cell2"]

FIELD_SEED = 2024
STATE_SEED = 5

# StructGen grids whose geometry the goldens pin (nx, ny, nz, bounds)
STRUCTGEN = {
    "sg222": (2, 2, 2, (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)),
    "sg345": (3, 4, 5, (-1.0, 2.0, 0.0, 1.0, 0.5, 3.0)),
    "sg555": (5, 5, 5, (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)),
    "sg739": (7, 3, 9, (0.1, 0.7, -3.0, 5.0, 2.0, 2.3)),
}
SG_KEYS = ("faceCell1", "faceCell2", "faceBnd", "faceLambda", "faceNorm", "faceCenter", "faceXpac", "faceXnac",
           "faceRLencos", "faceArea", "cellCenter", "cellVolume", "bndFace", "bndRegion", "bndDistance", "regType")


def outlet_first(host):
    """ArrayMesh of ``host`` with face 0 and the first OUTLET face swapped (boundary -> face map kept)."""
    a = {k: np.array(v, copy=True) for k, v in host.a.items()}
    t = a["regType"][a["bndRegion"]]
    outs = np.flatnonzero(t == ch.RTYPE_OUTLET)
    assert outs.size, "mesh has no OUTLET boundary"
    f = int(a["bndFace"][outs].min())
    if f != 0:
        for k in FACE_KEYS:
            a[k][[0, f]] = a[k][[f, 0]]
        fb = a["faceBnd"]
        for face in (0, f):
            if fb[face] >= 0:
                a["bndFace"][fb[face]] = face
    return ch.ArrayMesh(a)


def golden_mesh(name):
    params, first = MESHES[name]
    host = ch.HostMesh.from_hdf5(params["path"]) if "path" in params else ch.HostMesh(**params)
    return outlet_first(host) if first else ch.ArrayMesh(host.a)


def kernels_for(name):
    from kernel_runs import KERNELS
    return KERNELS if MESHES[name][1] else DEFAULT_SAFE


def fields(mesh):
    return W.Fields(mesh, seed=FIELD_SEED)


def mesh_arrays(mesh):
    return {k: mesh.a[k] for k in ch.ARRAY_KEYS}

"""A mesh whose cells do not all have the same number of faces.

The generators only make all-hex or all-tet meshes, so every CSR row has the
same length there, and kernels that preload whole rows of uniform meshes always
take that path. Removing boundary faces from a generated mesh leaves their
cells with shorter rows, which sends those kernels down their row-by-row path.
The result is a valid MiniSoA view (every index in range, boundaries renumbered
densely), not a closed mesh, which the kernels do not require.
"""
import ctypes as ct

import numpy as np

import cupcfd_hip as ch


def drop_boundary_faces(host, count=3, seed=0):
    """(view, arrays) of `host` without `count` boundary faces; keep `arrays` alive as long as the view."""
    a = host.a
    fb = a["faceBnd"]
    rng = np.random.default_rng(seed)
    gone = rng.choice(np.flatnonzero(fb >= 0), size=count, replace=False)
    keepf = np.ones(host.nFaces, bool)
    keepf[gone] = False
    keepb = np.ones(host.nBoundaries, bool)
    keepb[fb[gone]] = False
    new_f = np.cumsum(keepf) - 1
    new_b = np.cumsum(keepb) - 1
    arr = {
        "faceCell1": a["faceCell1"][keepf], "faceCell2": a["faceCell2"][keepf],
        "faceBnd": np.where(fb >= 0, new_b[np.maximum(fb, 0)], -1)[keepf].astype(np.int32),
        "bndFace": new_f[a["bndFace"][keepb]].astype(np.int32), "bndRegion": a["bndRegion"][keepb],
    }
    for k in ("faceLambda", "faceRLencos", "faceArea", "faceNorm", "faceCenter", "faceXpac", "faceXnac"):
        arr[k] = a[k][keepf]
    for k in ("bndDistance", "bndYPlus", "bndUPlus", "bndQ", "bndH", "bndT", "bndShear"):
        arr[k] = a[k][keepb]
    arr = {k: np.ascontiguousarray(v) for k, v in arr.items()}
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    p = lambda k: arr[k].ctypes.data  # noqa: E731
    v.nFaces, v.nBoundaries = int(keepf.sum()), int(keepb.sum())
    v.faceCell1ID, v.faceCell2ID, v.faceBoundaryID = p("faceCell1"), p("faceCell2"), p("faceBnd")
    v.faceLambda, v.faceRLencos, v.faceArea = p("faceLambda"), p("faceRLencos"), p("faceArea")
    v.faceNorm, v.faceCenter = ch.V3F64(p("faceNorm"), 0), ch.V3F64(p("faceCenter"), 0)
    v.faceXpac, v.faceXnac = ch.V3F64(p("faceXpac"), 0), ch.V3F64(p("faceXnac"), 0)
    v.boundaryFaceID, v.boundaryRegionID, v.boundaryDistance = p("bndFace"), p("bndRegion"), p("bndDistance")
    v.boundaryYPlus, v.boundaryUPlus, v.boundaryQ = p("bndYPlus"), p("bndUPlus"), p("bndQ")
    v.boundaryH, v.boundaryT, v.boundaryShear = p("bndH"), p("bndT"), ch.V3F64(p("bndShear"), 0)
    return v, arr


def device_mesh(view, arrays, host):
    """A DeviceMesh (FP64, reordered) built from `view`; it keeps `arrays` and `host` (the region arrays) alive."""
    h = ct.c_void_p()
    ch.check("cupcfd_hip_mesh_create_f64", ch.lib().cupcfd_hip_mesh_create_f64(ct.byref(view), ch.MESH_REORDER,
                                                                                ct.byref(h)))
    dev = ch.DeviceMesh(host, 64, _handle=h)
    dev._keep = (view, arrays)
    return dev

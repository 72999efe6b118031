"""ctypes binding of libcupcfd_hip.so (the C ABI declared in include/cupcfd_hip.h).

The library is the product; this module only marshals numpy arrays and
torch device pointers into it. Loading fails loudly when the built library is
missing: there is no fallback implementation.

Host meshes come from ``HostMesh`` (StructGen numbering, see
``cupcfd_structgen_create``); ``DeviceMesh`` repacks one into HBM and exposes
the 16 reference kernels in drop-in form (numpy arrays in the mesh's
numbering, synchronous) and resident form (device pointers in internal order).
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CUPCFD_HIP_LIBRARY names another build of it (the sanitizer build of cup-cfd_amd/Makefile)
LIB_PATH = os.environ.get("CUPCFD_HIP_LIBRARY") or os.path.join(os.path.dirname(_HERE), "libcupcfd_hip.so")

E_SUCCESS = 95
E_INVALID_INDEX = 68
E_ARRAY_SIZE_MISMATCH = 25
E_ERROR = 98
E_NULL_PTR = 99
E_NOT_IMPLEMENTED = 109
E_FILE_MISSING = 10
E_INVALID_FILE_FORMAT = 105

RTYPE_DEFAULT, RTYPE_INLET, RTYPE_OUTLET, RTYPE_SYMP, RTYPE_WALL = range(5)
MESH_REORDER = 1


class CupCfdError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} returned eCodes {code}")
        self.code = code


_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        _lib = ct.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


# ---------------------------------------------------------------------------
# C structures
# ---------------------------------------------------------------------------
class V3F64(ct.Structure):
    _fields_ = [("base", ct.c_void_p), ("stride", ct.c_int64)]


class V3F32(ct.Structure):
    _fields_ = [("base", ct.c_void_p), ("stride", ct.c_int64)]


def _view_fields(v3):
    P = ct.c_void_p
    return [
        ("nCells", ct.c_int32), ("nFaces", ct.c_int32), ("nBoundaries", ct.c_int32), ("nRegions", ct.c_int32),
        ("faceCell1ID", P), ("faceCell2ID", P), ("faceBoundaryID", P),
        ("faceLambda", P), ("faceRLencos", P), ("faceArea", P),
        ("faceNorm", v3), ("faceCenter", v3), ("faceXpac", v3), ("faceXnac", v3),
        ("cellCenter", v3), ("cellVolume", P),
        ("boundaryFaceID", P), ("boundaryRegionID", P), ("boundaryDistance", P),
        ("boundaryYPlus", P), ("boundaryUPlus", P), ("boundaryQ", P), ("boundaryH", P), ("boundaryT", P),
        ("boundaryShear", v3),
        ("regionType", P), ("regionStd", P), ("regionFlux", P), ("regionAdiab", P),
        ("regionYLog", P), ("regionELog", P), ("regionDensity", P), ("regionTurbKE", P), ("regionTurbDiss", P),
        ("regionSplvl", P), ("regionDen", P), ("regionR", P), ("regionT", P),
        ("regionForceTangent", v3), ("regionUVW", v3),
    ]


class MeshViewF64(ct.Structure):
    _fields_ = _view_fields(V3F64)


class MeshViewF32(ct.Structure):
    _fields_ = _view_fields(V3F32)


class StructGenParams(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("xmin", ct.c_double), ("xmax", ct.c_double), ("ymin", ct.c_double), ("ymax", ct.c_double),
        ("zmin", ct.c_double), ("zmax", ct.c_double),
        ("seed", ct.c_uint64),
        ("permute_cells", ct.c_int32), ("permute_faces", ct.c_int32),
        ("regions", ct.c_int32), ("face_extras", ct.c_int32), ("tets", ct.c_int32),
        ("jitter", ct.c_double),
    ]


KERNELS = [
    "gradient_phi_gauss_dolfyn", "flux_mass_face_loop", "flux_mass_boundary_loop1", "flux_mass_boundary_loop2",
    "flux_mass_boundary_loop3", "flux_mass_boundary_loop4", "flux_mass_boundary_loop5", "flux_mass_region_loop",
    "flux_scalar_face_loop", "flux_uvw_face_loop1", "flux_uvw_region_loop1", "flux_uvw_bnds_loop1",
    "viscosity_cell_loop1", "viscosity_region_loop", "viscosity_boundary_loop", "viscosity_cell_loop2",
]

HOST_SYMBOLS = [
    "cupcfd_structgen_create", "cupcfd_host_mesh_view_f64", "cupcfd_host_mesh_view_f32",
    "cupcfd_host_mesh_destroy", "cupcfd_hdf5_mesh_create", "cupcfd_hip_mesh_create_f64", "cupcfd_hip_mesh_create_f32",
    "cupcfd_hip_mesh_destroy", "cupcfd_hip_mesh_counts", "cupcfd_hip_mesh_cell_order",
    "cupcfd_hip_mesh_face_order", "cupcfd_hip_mesh_set_carry_init", "cupcfd_hip_mesh_set_ordered_sums",
    "cupcfd_hip_mesh_tile_layout", "cupcfd_hip_mesh_set_schedule",
    "cupcfd_hip_mesh_autotune", "cupcfd_hip_mesh_create_device_f64", "cupcfd_hip_mesh_create_device_f32",
    "cupcfd_hip_mesh_digest", "cupcfd_hip_mesh_digest_field", "cupcfd_hip_mesh_generate_f64",
    "cupcfd_hip_mesh_generate_f32", "cupcfd_hip_mesh_face_colours", "cupcfd_hip_autotune_forget",
    "cupcfd_hip_mesh_push_state",
    "cupcfd_hip_mesh_pull_state", "cupcfd_dev_cells_to_internal", "cupcfd_dev_cells_from_internal",
    "cupcfd_dev_faces_to_internal", "cupcfd_dev_faces_from_internal", "cupcfd_hip_version",
]


# batched resident entry points (cupcfd_dev_<name>_<f64|f32>)
BATCHED = ["gradient_batched", "flux_scalar_batched"]
# adjoint entry points (FP64 only)
ADJOINT = ["cupcfd_dev_scalar_adjoint_f64", "cupcfd_dev_scalar_adjoint_batched_f64",
           "cupcfd_dev_scalar_adjoint_noalloc_f64", "cupcfd_dev_scalar_adjoint_full_f64"]
# FP32-storage entry points on 64-bit meshes
F32S = ["cupcfd_dev_gradient_batched_f32s", "cupcfd_dev_flux_scalar_batched_f32s",
        "cupcfd_dev_scalar_adjoint_batched_f32s"]
# residual norms of cell fields
NORMS = ["cupcfd_dev_residual_norms_f64", "cupcfd_dev_residual_norms_f32"]
# mesh sets: persistent launches over several meshes
MESHSET = ["cupcfd_hip_meshset_create", "cupcfd_hip_meshset_destroy", "cupcfd_dev_gradient_batched_set_f64",
           "cupcfd_dev_gradient_batched_set_f32"]


def all_symbols():
    """Every symbol include/cupcfd_hip.h declares."""
    out = list(HOST_SYMBOLS)
    for s in ("f64", "f32"):
        for k in KERNELS:
            out.append(f"cupcfd_fvm_{k}_{s}")
            out.append(f"cupcfd_dev_{k}_{s}")
        for k in BATCHED:
            out.append(f"cupcfd_dev_{k}_{s}")
    out += ADJOINT + F32S + MESHSET + NORMS
    return out


def _declare(L):
    L.cupcfd_structgen_create.argtypes = [ct.POINTER(StructGenParams), ct.POINTER(ct.c_void_p)]
    L.cupcfd_hdf5_mesh_create.argtypes = [ct.c_char_p, ct.POINTER(ct.c_void_p)]
    L.cupcfd_host_mesh_view_f64.argtypes = [ct.c_void_p, ct.POINTER(MeshViewF64)]
    L.cupcfd_host_mesh_view_f32.argtypes = [ct.c_void_p, ct.POINTER(MeshViewF32)]
    L.cupcfd_host_mesh_destroy.argtypes = [ct.c_void_p]
    L.cupcfd_host_mesh_destroy.restype = None
    L.cupcfd_hip_mesh_create_f64.argtypes = [ct.POINTER(MeshViewF64), ct.c_int32, ct.POINTER(ct.c_void_p)]
    L.cupcfd_hip_mesh_create_f32.argtypes = [ct.POINTER(MeshViewF32), ct.c_int32, ct.POINTER(ct.c_void_p)]
    L.cupcfd_hip_mesh_create_device_f64.argtypes = [ct.POINTER(MeshViewF64), ct.c_int32, ct.POINTER(ct.c_void_p)]
    L.cupcfd_hip_mesh_create_device_f32.argtypes = [ct.POINTER(MeshViewF32), ct.c_int32, ct.POINTER(ct.c_void_p)]
    for k in ("f64", "f32"):
        getattr(L, f"cupcfd_hip_mesh_generate_{k}").argtypes = [ct.POINTER(StructGenParams), ct.c_int32,
                                                               ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p)]
    L.cupcfd_hip_mesh_face_colours.argtypes = [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int32)]
    L.cupcfd_hip_mesh_digest.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_int32)]
    L.cupcfd_hip_mesh_digest_field.argtypes = [ct.c_int32]
    L.cupcfd_hip_mesh_digest_field.restype = ct.c_char_p
    L.cupcfd_hip_mesh_destroy.argtypes = [ct.c_void_p]
    L.cupcfd_hip_mesh_counts.argtypes = [ct.c_void_p] + [ct.POINTER(ct.c_int32)] * 4
    L.cupcfd_hip_mesh_cell_order.argtypes = [ct.c_void_p, ct.c_void_p]
    L.cupcfd_hip_mesh_face_order.argtypes = [ct.c_void_p, ct.c_void_p]
    L.cupcfd_hip_mesh_set_carry_init.argtypes = [ct.c_void_p, ct.c_double]
    L.cupcfd_hip_mesh_set_schedule.argtypes = [ct.c_void_p, ct.c_int32]
    L.cupcfd_hip_mesh_set_ordered_sums.argtypes = [ct.c_void_p, ct.c_int32]
    L.cupcfd_hip_mesh_tile_layout.argtypes = [ct.c_void_p, ct.c_void_p]
    L.cupcfd_hip_mesh_autotune.argtypes = [ct.c_void_p, ct.c_void_p]
    L.cupcfd_hip_mesh_push_state.argtypes = [ct.c_void_p]
    L.cupcfd_hip_mesh_pull_state.argtypes = [ct.c_void_p]
    for n in ("cells_to_internal", "cells_from_internal", "faces_to_internal", "faces_from_internal"):
        getattr(L, f"cupcfd_dev_{n}").argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_int32,
                                                   ct.c_void_p]
    L.cupcfd_hip_version.restype = ct.c_char_p
    # kernel entry points take pointers, int32 and T scalars; arguments are
    # converted explicitly by the wrappers below
    for s in ("f64", "f32"):
        for k in KERNELS:
            for mode in ("fvm", "dev"):
                getattr(L, f"cupcfd_{mode}_{k}_{s}").restype = ct.c_int
        for k in BATCHED:
            getattr(L, f"cupcfd_dev_{k}_{s}").restype = ct.c_int
    for k in ADJOINT + F32S + MESHSET + NORMS:
        getattr(L, k).restype = ct.c_int
    L.cupcfd_hip_meshset_create.argtypes = [ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_void_p)]
    L.cupcfd_hip_meshset_destroy.argtypes = [ct.c_void_p]


def check(fn, code):
    if code != E_SUCCESS:
        raise CupCfdError(fn, code)


# ---------------------------------------------------------------------------
# Host mesh
# ---------------------------------------------------------------------------
def _np(ptr, n, dtype, shape=None, owner=None):
    """numpy view of n elements at ptr; the view keeps ``owner`` (the object whose memory it is) alive."""
    if n == 0:
        return np.zeros(shape if shape else (0,), dtype=dtype)
    ctype = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
    addr = ptr if isinstance(ptr, int) else ct.cast(ptr, ct.c_void_p).value
    buf = (ctype * n).from_address(addr)
    buf._owner = owner
    a = np.ndarray((n,), dtype=dtype, buffer=buf)
    return a.reshape(shape) if shape else a


def structgen_params(nx, ny=None, nz=None, bounds=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), seed=0, permute_cells=False,
                     permute_faces=False, typed=False, face_extras=False, tets=False, jitter=0.0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return StructGenParams(nx, ny, nz, *[float(b) for b in bounds], seed, int(permute_cells), int(permute_faces),
                           int(typed), int(face_extras), int(tets), float(jitter))


class HostMesh:
    """A StructGen-numbered mesh held by the library (cupcfd_structgen_create)."""

    def __init__(self, nx, ny=None, nz=None, bounds=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), seed=0,
                 permute_cells=False, permute_faces=False, typed=False, face_extras=False, tets=False,
                 jitter=0.0, _handle=None):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        L = lib()
        if _handle is None:
            p = structgen_params(nx, ny, nz, bounds, seed, permute_cells, permute_faces, typed, face_extras, tets,
                                 jitter)
            self._h = ct.c_void_p()
            check("cupcfd_structgen_create", L.cupcfd_structgen_create(ct.byref(p), ct.byref(self._h)))
        else:  # a host mesh the library made (cupcfd_hip_mesh_generate_*)
            self._h = _handle
        self.params = dict(nx=nx, ny=ny, nz=nz, bounds=bounds, seed=seed, permute_cells=permute_cells,
                           permute_faces=permute_faces, typed=typed, face_extras=face_extras, tets=tets,
                           jitter=jitter)
        self.view64 = MeshViewF64()
        check("cupcfd_host_mesh_view_f64", L.cupcfd_host_mesh_view_f64(self._h, ct.byref(self.view64)))
        self._view32 = None
        v = self.view64
        self.nCells, self.nFaces, self.nBoundaries, self.nRegions = v.nCells, v.nFaces, v.nBoundaries, v.nRegions
        C, F, B, R = self.nCells, self.nFaces, self.nBoundaries, self.nRegions
        i32, f64, u8 = np.int32, np.float64, np.uint8
        self.a = {
            "faceCell1": _np(v.faceCell1ID, F, i32, owner=self), "faceCell2": _np(v.faceCell2ID, F, i32, owner=self),
            "faceBnd": _np(v.faceBoundaryID, F, i32, owner=self), "faceLambda": _np(v.faceLambda, F, f64, owner=self),
            "faceRLencos": _np(v.faceRLencos, F, f64, owner=self), "faceArea": _np(v.faceArea, F, f64, owner=self),
            "faceNorm": _np(v.faceNorm.base, 3 * F, f64, (F, 3), owner=self),
            "faceCenter": _np(v.faceCenter.base, 3 * F, f64, (F, 3), owner=self),
            "faceXpac": _np(v.faceXpac.base, 3 * F, f64, (F, 3), owner=self),
            "faceXnac": _np(v.faceXnac.base, 3 * F, f64, (F, 3), owner=self),
            "cellCenter": _np(v.cellCenter.base, 3 * C, f64, (C, 3), owner=self), "cellVolume": _np(v.cellVolume, C, f64, owner=self),
            "bndFace": _np(v.boundaryFaceID, B, i32, owner=self), "bndRegion": _np(v.boundaryRegionID, B, i32, owner=self),
            "bndDistance": _np(v.boundaryDistance, B, f64, owner=self), "bndYPlus": _np(v.boundaryYPlus, B, f64, owner=self),
            "bndUPlus": _np(v.boundaryUPlus, B, f64, owner=self), "bndQ": _np(v.boundaryQ, B, f64, owner=self),
            "bndH": _np(v.boundaryH, B, f64, owner=self), "bndT": _np(v.boundaryT, B, f64, owner=self),
            "bndShear": _np(v.boundaryShear.base, 3 * B, f64, (B, 3), owner=self),
            "regType": _np(v.regionType, R, i32, owner=self), "regStd": _np(v.regionStd, R, u8, owner=self),
            "regFlux": _np(v.regionFlux, R, u8, owner=self), "regAdiab": _np(v.regionAdiab, R, u8, owner=self),
            "regYLog": _np(v.regionYLog, R, f64, owner=self), "regELog": _np(v.regionELog, R, f64, owner=self),
            "regDensity": _np(v.regionDensity, R, f64, owner=self), "regTurbKE": _np(v.regionTurbKE, R, f64, owner=self),
            "regTurbDiss": _np(v.regionTurbDiss, R, f64, owner=self), "regSplvl": _np(v.regionSplvl, R, f64, owner=self),
            "regDen": _np(v.regionDen, R, f64, owner=self), "regR": _np(v.regionR, R, f64, owner=self), "regT": _np(v.regionT, R, f64, owner=self),
            "regForceTangent": _np(v.regionForceTangent.base, 3 * R, f64, (R, 3), owner=self),
            "regUVW": _np(v.regionUVW.base, 3 * R, f64, (R, 3), owner=self),
        }

    STATE = ("bndYPlus", "bndUPlus", "bndQ", "bndH", "bndT", "bndShear", "regYLog", "regForceTangent")

    @classmethod
    def from_hdf5(cls, path):
        """A mesh read from one of the reference's HDF5 mesh files (cupcfd_hdf5_mesh_create;
        MeshHDF5Source, src/geometry/mesh/implementation/source/MeshHDF5Source.cpp)."""
        h = ct.c_void_p()
        check("cupcfd_hdf5_mesh_create", lib().cupcfd_hdf5_mesh_create(os.fsencode(path), ct.byref(h)))
        m = cls(0, _handle=h)
        m.params = dict(path=os.fspath(path))
        return m

    def view32(self):
        if self._view32 is None:
            self._view32 = MeshViewF32()
            check("cupcfd_host_mesh_view_f32", lib().cupcfd_host_mesh_view_f32(self._h, ct.byref(self._view32)))
        return self._view32

    def state(self, precision=64):
        """Copies of the mutable mesh state arrays (as the view of that precision holds them)."""
        if precision == 64:
            return {k: self.a[k].copy() for k in self.STATE}
        v = self.view32()
        B, R = self.nBoundaries, self.nRegions
        f32 = np.float32
        return {
            "bndYPlus": _np(v.boundaryYPlus, B, f32, owner=self).copy(), "bndUPlus": _np(v.boundaryUPlus, B, f32, owner=self).copy(),
            "bndQ": _np(v.boundaryQ, B, f32, owner=self).copy(), "bndH": _np(v.boundaryH, B, f32, owner=self).copy(),
            "bndT": _np(v.boundaryT, B, f32, owner=self).copy(),
            "bndShear": _np(v.boundaryShear.base, 3 * B, f32, (B, 3), owner=self).copy(),
            "regYLog": _np(v.regionYLog, R, f32, owner=self).copy(),
            "regForceTangent": _np(v.regionForceTangent.base, 3 * R, f32, (R, 3), owner=self).copy(),
        }

    def set_state(self, st, precision=64):
        if precision == 64:
            for k, v in st.items():
                self.a[k][...] = v
            return
        v = self.view32()
        B, R = self.nBoundaries, self.nRegions
        f32 = np.float32
        tgt = {
            "bndYPlus": _np(v.boundaryYPlus, B, f32, owner=self), "bndUPlus": _np(v.boundaryUPlus, B, f32, owner=self),
            "bndQ": _np(v.boundaryQ, B, f32, owner=self), "bndH": _np(v.boundaryH, B, f32, owner=self), "bndT": _np(v.boundaryT, B, f32, owner=self),
            "bndShear": _np(v.boundaryShear.base, 3 * B, f32, (B, 3), owner=self), "regYLog": _np(v.regionYLog, R, f32, owner=self),
            "regForceTangent": _np(v.regionForceTangent.base, 3 * R, f32, (R, 3), owner=self),
        }
        for k, val in st.items():
            tgt[k][...] = val

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.cupcfd_host_mesh_destroy(h)
            self._h = None


ARRAY_KEYS = ("faceCell1", "faceCell2", "faceBnd", "faceLambda", "faceRLencos", "faceArea", "faceNorm",
              "faceCenter", "faceXpac", "faceXnac", "cellCenter", "cellVolume", "bndFace", "bndRegion",
              "bndDistance", "bndYPlus", "bndUPlus", "bndQ", "bndH", "bndT", "bndShear", "regType", "regStd",
              "regFlux", "regAdiab", "regYLog", "regELog", "regDensity", "regTurbKE", "regTurbDiss", "regSplvl",
              "regDen", "regR", "regT", "regForceTangent", "regUVW")
_INT_KEYS = {"faceCell1", "faceCell2", "faceBnd", "bndFace", "bndRegion", "regType"}
_U8_KEYS = {"regStd", "regFlux", "regAdiab"}


class ArrayMesh:
    """A MiniSoA mesh given as numpy arrays (the keys of ``HostMesh.a``; 3-vectors as (n, 3)), in the mesh's
    own numbering: a mesh read from a file, or one a test edits. Offers the same ``a``, counts, views and state
    accessors as HostMesh, so DeviceMesh, the oracle and the tests take either. The arrays are copied; the FP32
    view holds the FP64 values rounded once."""

    STATE = HostMesh.STATE

    def __init__(self, arrays):
        a = {}
        for k in ARRAY_KEYS:
            v = np.asarray(arrays[k])
            dt = np.int32 if k in _INT_KEYS else (np.uint8 if k in _U8_KEYS else np.float64)
            v = np.ascontiguousarray(v, dtype=dt).copy()
            if k in ("faceNorm", "faceCenter", "faceXpac", "faceXnac", "cellCenter", "bndShear", "regForceTangent",
                     "regUVW"):
                v = v.reshape(-1, 3)
            a[k] = v
        self.a = a
        self.nCells = a["cellVolume"].size
        self.nFaces = a["faceCell1"].size
        self.nBoundaries = a["bndFace"].size
        self.nRegions = a["regType"].size
        self.params = {}
        self.view64 = self._view(MeshViewF64, V3F64, a)
        self._a32 = None
        self._view32 = None

    def _view(self, View, V3, a):
        v = View()
        v.nCells, v.nFaces, v.nBoundaries, v.nRegions = self.nCells, self.nFaces, self.nBoundaries, self.nRegions
        p = lambda k: a[k].ctypes.data if a[k].size else None  # noqa: E731
        q = lambda k: V3(a[k].ctypes.data if a[k].size else None, 0)  # noqa: E731
        v.faceCell1ID, v.faceCell2ID, v.faceBoundaryID = p("faceCell1"), p("faceCell2"), p("faceBnd")
        v.faceLambda, v.faceRLencos, v.faceArea = p("faceLambda"), p("faceRLencos"), p("faceArea")
        v.faceNorm, v.faceCenter, v.faceXpac, v.faceXnac = q("faceNorm"), q("faceCenter"), q("faceXpac"), q("faceXnac")
        v.cellCenter, v.cellVolume = q("cellCenter"), p("cellVolume")
        v.boundaryFaceID, v.boundaryRegionID, v.boundaryDistance = p("bndFace"), p("bndRegion"), p("bndDistance")
        v.boundaryYPlus, v.boundaryUPlus, v.boundaryQ = p("bndYPlus"), p("bndUPlus"), p("bndQ")
        v.boundaryH, v.boundaryT, v.boundaryShear = p("bndH"), p("bndT"), q("bndShear")
        v.regionType, v.regionStd, v.regionFlux, v.regionAdiab = p("regType"), p("regStd"), p("regFlux"), p("regAdiab")
        v.regionYLog, v.regionELog, v.regionDensity = p("regYLog"), p("regELog"), p("regDensity")
        v.regionTurbKE, v.regionTurbDiss, v.regionSplvl = p("regTurbKE"), p("regTurbDiss"), p("regSplvl")
        v.regionDen, v.regionR, v.regionT = p("regDen"), p("regR"), p("regT")
        v.regionForceTangent, v.regionUVW = q("regForceTangent"), q("regUVW")
        return v

    def view32(self):
        if self._view32 is None:
            self._a32 = {k: (v.astype(np.float32) if v.dtype == np.float64 else v) for k, v in self.a.items()}
            self._view32 = self._view(MeshViewF32, V3F32, self._a32)
        return self._view32

    def state(self, precision=64):
        src = self.a if precision == 64 else (self.view32() and self._a32)
        return {k: src[k].copy() for k in self.STATE}

    def set_state(self, st, precision=64):
        if precision != 64:
            self.view32()
        dst = self.a if precision == 64 else self._a32
        for k, v in st.items():
            dst[k][...] = v


# ---------------------------------------------------------------------------
# Device mesh and kernels
# ---------------------------------------------------------------------------
def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
        return a.ctypes.data_as(ct.c_void_p)
    if hasattr(a, "data_ptr"):  # torch tensor on the device
        return ct.c_void_p(a.data_ptr())
    if isinstance(a, int):
        return ct.c_void_p(a)
    raise TypeError(type(a))


def _n(a, per=1):
    return ct.c_int32(0 if a is None else (a.size if isinstance(a, np.ndarray) else a.numel()) // per)


class DeviceMesh:
    """A mesh repacked into HBM (cupcfd_hip_mesh_create_f64 / _f32, or with ``on_device`` the GPU repack
    cupcfd_hip_mesh_create_device_f64 / _f32, which builds the same arrays)."""

    def __init__(self, host: HostMesh, precision=64, reorder=True, on_device=False, _handle=None):
        self.host = host
        self.precision = precision
        self.T = ct.c_double if precision == 64 else ct.c_float
        self.np_t = np.float64 if precision == 64 else np.float32
        self.sfx = "f64" if precision == 64 else "f32"
        self.V3 = V3F64 if precision == 64 else V3F32
        self._h = ct.c_void_p()
        L = lib()
        flags = MESH_REORDER if reorder else 0
        if _handle is not None:  # a device mesh the library made (cupcfd_hip_mesh_generate_*)
            self._h = _handle
        else:
            fn = f"cupcfd_hip_mesh_create{'_device' if on_device else ''}_{self.sfx}"
            view = host.view64 if precision == 64 else host.view32()
            check(fn, getattr(L, fn)(ct.byref(view), flags, ct.byref(self._h)))
        c = [ct.c_int32() for _ in range(4)]
        check("cupcfd_hip_mesh_counts", L.cupcfd_hip_mesh_counts(self._h, *[ct.byref(x) for x in c]))
        self.nCells, self.nFaces, self.nBoundaries, self.nRegions = [x.value for x in c]
        self.cell_order = np.zeros(self.nCells, np.int32)
        self.face_order = np.zeros(self.nFaces, np.int32)
        L.cupcfd_hip_mesh_cell_order(self._h, _ptr(self.cell_order))
        L.cupcfd_hip_mesh_face_order(self._h, _ptr(self.face_order))

    @property
    def handle(self):
        return self._h

    def digest(self):
        """{field: 64-bit hash} of every device array and host-side order of this mesh (cupcfd_hip_mesh_digest)."""
        L = lib()
        out = np.zeros(128, np.uint64)
        n = ct.c_int32()
        check("cupcfd_hip_mesh_digest", L.cupcfd_hip_mesh_digest(self._h, _ptr(out), out.size, ct.byref(n)))
        return {L.cupcfd_hip_mesh_digest_field(i).decode(): int(out[i]) for i in range(n.value)}

    def face_colours(self):
        """(colour per face in mesh numbering, number of colours) of the order-preserving face colouring."""
        out = np.zeros(self.nFaces, np.int32)
        n = ct.c_int32()
        check("cupcfd_hip_mesh_face_colours", lib().cupcfd_hip_mesh_face_colours(self._h, _ptr(out), ct.byref(n)))
        return out, n.value

    def set_carry_init(self, v):
        check("set_carry_init", lib().cupcfd_hip_mesh_set_carry_init(self._h, float(v)))

    def set_ordered_sums(self, sequential=True):
        """Boundary-loop sums in the reference's sequential order (True, the default: its bits) or as a
        fixed-shape parallel tree (False: faster, the same bits every call, not the sequential sum's)."""
        check("set_ordered_sums", lib().cupcfd_hip_mesh_set_ordered_sums(self._h, 1 if sequential else 0))

    def tile_layout(self):
        """The cell-tile layout: dict of tiles, slots (most of any tile), entries (most CSR entries of any
        tile), outside (most slots with a cell outside their tile or a boundary face), inside (most slots
        between two cells of one tile) and uvw_split (the UVW tile runs its split LDS layout)."""
        info = (ct.c_int32 * 6)()
        check("tile_layout", lib().cupcfd_hip_mesh_tile_layout(self._h, info))
        return dict(zip(("tiles", "slots", "entries", "outside", "inside", "uvw_split"), list(info)))

    def autotune(self):
        """Time the gather schedules of gradient, scalar and UVW on this mesh and keep the fastest of each;
        returns the three choices (64 tiles, 4/6/8 lanes per cell, 1 thread per cell)."""
        out = np.zeros(3, np.int32)
        check("autotune", lib().cupcfd_hip_mesh_autotune(self._h, _ptr(out)))
        return [int(x) for x in out]

    def set_schedule(self, lanes_per_cell):
        """0 = automatic, 1 = one thread per cell, 2 = colour groups (gradient), 4, 6 or 8 lanes per cell,
        64 = 64-cell tiles."""
        check("set_schedule", lib().cupcfd_hip_mesh_set_schedule(self._h, int(lanes_per_cell)))

    def push_state(self):
        check("push_state", lib().cupcfd_hip_mesh_push_state(self._h))

    def pull_state(self):
        check("pull_state", lib().cupcfd_hip_mesh_pull_state(self._h))

    def v3(self, a):
        """cupcfd_v3 descriptor of a packed (n,3) numpy array."""
        assert a.dtype == self.np_t and a.flags["C_CONTIGUOUS"]
        return self.V3(a.ctypes.data, 0)

    def _s(self, x):
        return self.T(x)

    def call(self, mode, name, *args):
        """Call cupcfd_<mode>_<name>_<sfx> with already-converted ctypes args."""
        fn = f"cupcfd_{mode}_{name}_{self.sfx}"
        code = getattr(lib(), fn)(*args)
        check(fn, code)
        return code

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.cupcfd_hip_mesh_destroy(h)
            self._h = None


def version():
    return lib().cupcfd_hip_version().decode()


def generate_mesh(nx, ny=None, nz=None, precision=64, reorder=True, host_copy=True, **params):
    """Generate the mesh HostMesh(nx, ny, nz, **params) describes on the GPU and repack it there
    (cupcfd_hip_mesh_generate_f64 / _f32). Returns (HostMesh or None, DeviceMesh); the host copy,
    when asked for, equals HostMesh(nx, ny, nz, **params) and the device mesh equals DeviceMesh of it."""
    L = lib()
    p = structgen_params(nx, ny, nz, **params)
    hh, dh = ct.c_void_p(), ct.c_void_p()
    fn = f"cupcfd_hip_mesh_generate_{'f64' if precision == 64 else 'f32'}"
    check(fn, getattr(L, fn)(ct.byref(p), MESH_REORDER if reorder else 0, ct.byref(hh) if host_copy else None,
                             ct.byref(dh)))
    host = HostMesh(nx, ny, nz, _handle=hh, **params) if host_copy else None
    dev = DeviceMesh(host, precision, reorder=reorder, _handle=dh)
    if host is not None:
        dev._keep_host = host  # the device mesh mirrors state through the host copy's view
    return host, dev


from .fvm import *  # noqa: E402,F401,F403  (kernel wrappers)


class DeviceMeshSet:
    """A set of DeviceMesh objects of one precision for persistent multi-mesh launches
    (cupcfd_hip_meshset_create). Keeps its members alive; destroyed before them."""

    def __init__(self, meshes):
        self.meshes = list(meshes)
        assert self.meshes and len({m.precision for m in self.meshes}) == 1
        self.precision = self.meshes[0].precision
        self.sfx = self.meshes[0].sfx
        arr = (ct.c_void_p * len(self.meshes))(*[m._h.value for m in self.meshes])
        self._h = ct.c_void_p()
        check("cupcfd_hip_meshset_create", lib().cupcfd_hip_meshset_create(arr, len(self.meshes), ct.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cupcfd_hip_meshset_destroy(self._h)
            self._h = None

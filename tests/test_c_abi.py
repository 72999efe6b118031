"""The C-ABI library on the CPU: it loads, exports every symbol the public
header declares, and its host-side logic (mesh sources, view validation)
behaves; no kernel is launched here."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import cupcfd_hip as ch

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cupcfd_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    names = set(re.findall(r"^(?:int|void|const char \*)\s*(cupcfd_\w+)\(", text, re.M))
    macro = re.findall(r"(cupcfd_(?:fvm|dev)_\w+)_##S\(", text)
    for s in ("f64", "f32"):
        names.update(f"{m}_{s}" for m in macro)
    return names


def test_header_symbols_exported():
    L = ch.lib()
    declared = declared_symbols()
    assert len(declared) == len(ch.all_symbols()) == 31 + 2 * 2 * 16 + 2 * len(ch.BATCHED) + len(ch.ADJOINT) + len(ch.F32S) + len(ch.MESHSET) + len(ch.NORMS)
    assert declared == set(ch.all_symbols())
    for name in declared:
        assert hasattr(L, name), name


def test_ecodes_match_reference_enum():
    # cupcfd::error::eCodes ordinals (include/error/Error.h, 0-based enum)
    text = open(HEADER).read()
    vals = dict(re.findall(r"#define CUPCFD_(E_\w+) (\d+)", text))
    assert vals == {"E_FILE_MISSING": "10", "E_ARRAY_SIZE_MISMATCH": "25", "E_INVALID_INDEX": "68",
                    "E_CONFIG_INVALID_VALUE": "90", "E_SUCCESS": "95", "E_ERROR": "98", "E_NULL_PTR": "99",
                    "E_INVALID_FILE_FORMAT": "105", "E_NOT_IMPLEMENTED": "109"}


def test_version():
    assert "gfx950" in ch.version()


def _create(view):
    h = ct.c_void_p()
    rc = ch.lib().cupcfd_hip_mesh_create_f64(ct.byref(view), 0, ct.byref(h))
    return rc, h


def test_view_validation_rejects_bad_indices():
    """Repacking validates the view before touching the device."""
    host = ch.HostMesh(3)
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    bad = host.a["faceCell1"].copy()
    bad[5] = host.nCells  # out of range
    v.faceCell1ID = bad.ctypes.data
    assert _create(v)[0] == ch.E_INVALID_INDEX
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    badb = host.a["faceBnd"].copy()
    badb[0] = host.nBoundaries
    v.faceBoundaryID = badb.ctypes.data
    assert _create(v)[0] == ch.E_INVALID_INDEX
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    badr = host.a["regType"].copy()
    badr[0] = 7
    v.regionType = badr.ctypes.data
    assert _create(v)[0] == 90
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    v.cellVolume = None
    assert _create(v)[0] == ch.E_NULL_PTR


def test_null_handles():
    L = ch.lib()
    assert L.cupcfd_hip_mesh_destroy(None) == ch.E_NULL_PTR
    assert L.cupcfd_fvm_flux_uvw_region_loop1_f64(None) == ch.E_NULL_PTR
    assert L.cupcfd_dev_gradient_phi_gauss_dolfyn_f64(None, 1, None, None, None, None, None) == ch.E_NULL_PTR
    assert L.cupcfd_structgen_create(None, None) == ch.E_NULL_PTR


def test_float_view_matches_double():
    host = ch.HostMesh(4, typed=True, face_extras=True, seed=3)
    v32 = host.view32()
    F = host.nFaces
    S32 = np.ctypeslib.as_array(ct.cast(v32.faceNorm.base, ct.POINTER(ct.c_float)), shape=(3 * F,))
    assert np.array_equal(S32, host.a["faceNorm"].ravel().astype(np.float32))


@pytest.mark.parametrize("n", [1, 2])
def test_tiny_meshes(n):
    m = ch.HostMesh(n, 1, 1)
    assert m.nCells == n and m.nBoundaries == 2 * (1 + 2 * n)


def test_fields_copies_never_alias():
    """Parity runs feed the oracle and the GPU separate copies of every array."""
    from cupcfd_hip import workload as W
    host = ch.HostMesh(3)
    f = W.Fields(host)
    for g in (f.copy(), f.cast(np.float64), f.cast(np.float32)):
        for k, v in f.__dict__.items():
            assert not np.shares_memory(v, g.__dict__[k]), k

"""The GPU mesh repack (cupcfd_hip_mesh_create_device_*) against the host repack.

The host repack is what every kernel parity test runs on. The GPU repack must
produce the same device mesh: every array, every count and every host-side
order, compared field by field through cupcfd_hip_mesh_digest (integer and
byte work: bit-exact). Kernel runs on both builds must give bit-identical
outputs, and bad views must be refused with the same eCodes.
"""
import ctypes as ct

import numpy as np
import pytest

import cupcfd_hip as ch
from cupcfd_hip import fvm
from cupcfd_hip import workload as W
from kernel_runs import run, seed_state

MESHES = {
    # the reference's fvm unit-test mesh: 5^3 StructGen over [-1,1]^3, one DEFAULT region
    "ref5": dict(nx=5, bounds=(-1, 1, -1, 1, -1, 1)),
    # six typed regions, relabelled cells and faces, non-zero xpac/xnac/rlencos
    "typed_perm": dict(nx=7, ny=6, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=11),
    "typed32": dict(nx=32, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=3),
    "tets_typed": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, typed=True, permute_cells=True,
                       permute_faces=True, face_extras=True, seed=5),
    # 24576 jittered tets: many halo faces per tile
    "tets16": dict(nx=16, tets=True, jitter=0.25, typed=True, permute_cells=True, permute_faces=True, seed=8),
    # a slab one cell thick: every cell owns boundary faces
    "slab": dict(nx=40, ny=3, nz=1, permute_faces=True, seed=2),
    # one cell: no interior faces, no Morton sort
    "one_cell": dict(nx=1, ny=1, nz=1, typed=True),
    # a line of cells whose tiles have no halo faces on two axes
    "line": dict(nx=200, ny=1, nz=1, typed=True, permute_cells=True, seed=4),
}

N_FIELDS = 49


def _differ(a, b):
    return [k for k in a if a[k] != b[k]]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("reorder", [True, False])
@pytest.mark.parametrize("meshname", list(MESHES))
def test_device_build_matches_host(meshname, reorder, precision):
    host = ch.HostMesh(**MESHES[meshname])
    a = ch.DeviceMesh(host, precision, reorder=reorder)
    b = ch.DeviceMesh(host, precision, reorder=reorder, on_device=True)
    da, db = a.digest(), b.digest()
    assert len(da) == N_FIELDS and list(da) == list(db)
    assert _differ(da, db) == []
    assert np.array_equal(a.cell_order, b.cell_order) and np.array_equal(a.face_order, b.face_order)


@pytest.mark.gpu
def test_device_build_matches_host_64cubed():
    host = ch.HostMesh(64, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=4)
    da = ch.DeviceMesh(host, 64).digest()
    db = ch.DeviceMesh(host, 64, on_device=True).digest()
    assert _differ(da, db) == []


@pytest.mark.gpu
def test_digest_sees_a_different_mesh():
    """The digest is not blind: another relabelling changes the orders and the CSR."""
    a = ch.DeviceMesh(ch.HostMesh(6, permute_cells=True, permute_faces=True, seed=1), 64, on_device=True).digest()
    b = ch.DeviceMesh(ch.HostMesh(6, permute_cells=True, permute_faces=True, seed=2), 64, on_device=True).digest()
    d = set(_differ(a, b))
    assert {"cell_order", "face_order", "ent", "fconn"} <= d
    assert "counts" not in d and "rtype" not in d


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["gradient", "mass", "scalar", "uvw"])
def test_kernels_agree_on_both_builds(kernel):
    host = ch.HostMesh(**MESHES["typed_perm"])
    outs = []
    for on_device in (False, True):
        dev = ch.DeviceMesh(host, 64, on_device=on_device)
        host.set_state(seed_state(host))
        rc, out = run(kernel, fvm, dev, W.Fields(host, seed=21).copy(), lambda: host.state(64))
        assert rc == ch.E_SUCCESS
        outs.append(out)
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def _create(view, on_device, flags=ch.MESH_REORDER):
    L = ch.lib()
    h = ct.c_void_p()
    fn = L.cupcfd_hip_mesh_create_device_f64 if on_device else L.cupcfd_hip_mesh_create_f64
    rc = fn(ct.byref(view), flags, ct.byref(h))
    return rc, h


def _digest(h):
    L = ch.lib()
    out = np.zeros(64, np.uint64)
    n = ct.c_int32()
    assert L.cupcfd_hip_mesh_digest(h, out.ctypes.data_as(ct.c_void_p), out.size, ct.byref(n)) == ch.E_SUCCESS
    return {L.cupcfd_hip_mesh_digest_field(i).decode(): int(out[i]) for i in range(n.value)}


@pytest.mark.gpu
def test_device_build_strided_view():
    """3-vectors laid out as the reference's EuclideanVector<double,3> (32-byte stride)."""
    host = ch.HostMesh(**MESHES["typed_perm"])
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    keep = []
    for name in ("cellCenter", "faceNorm", "faceCenter", "faceXpac", "faceXnac"):
        a = host.a[name]
        s = np.full((a.shape[0], 4), np.nan)
        s[:, 1:] = a
        keep.append(s)
        setattr(v, name, ch.V3F64(s.ctypes.data + 8, 32))
    L = ch.lib()
    rc_p, hp = _create(host.view64, False)
    rc_s, hs = _create(v, True)
    try:
        assert rc_p == ch.E_SUCCESS and rc_s == ch.E_SUCCESS
        assert _differ(_digest(hp), _digest(hs)) == []
    finally:
        L.cupcfd_hip_mesh_destroy(hp)
        L.cupcfd_hip_mesh_destroy(hs)


@pytest.mark.gpu
def test_device_build_rejects_bad_views():
    """The same eCodes as the host repack, and no device mesh."""
    host = ch.HostMesh(3, typed=True)
    L = ch.lib()
    cases = []
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    bad = host.a["faceCell1"].copy()
    bad[5] = host.nCells
    v.faceCell1ID = bad.ctypes.data
    cases.append((v, ch.E_INVALID_INDEX, bad))
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    bad2 = host.a["faceCell2"].copy()
    bad2[np.flatnonzero(host.a["faceBnd"] < 0)[3]] = -2
    v.faceCell2ID = bad2.ctypes.data
    cases.append((v, ch.E_INVALID_INDEX, bad2))
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    badb = host.a["bndRegion"].copy()
    badb[-1] = host.nRegions
    v.boundaryRegionID = badb.ctypes.data
    cases.append((v, ch.E_INVALID_INDEX, badb))
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    badr = host.a["regType"].copy()
    badr[0] = 7
    v.regionType = badr.ctypes.data
    cases.append((v, 90, badr))
    v = ch.MeshViewF64.from_buffer_copy(host.view64)
    v.cellVolume = None
    cases.append((v, ch.E_NULL_PTR, None))
    for v, code, _ in cases:
        for on_device in (False, True):
            rc, h = _create(v, on_device)
            assert rc == code, (on_device, rc, code)
            assert not h.value
    assert L.cupcfd_hip_mesh_create_device_f64(None, 0, None) == ch.E_NULL_PTR


# ---- meshes generated on the GPU (cupcfd_hip_mesh_generate_*) -------------------
GEN = dict(MESHES)
GEN.update({
    # uneven spacing and bounds on every axis, typed, extras, faces relabelled
    "aniso": dict(nx=9, ny=4, nz=6, bounds=(0.0, 2.5, -1.0, 0.3, 0.1, 7.0), typed=True, face_extras=True,
                  permute_faces=True, seed=99),
    # strongly jittered tets on an uneven box, cells relabelled
    "tets_aniso": dict(nx=3, ny=5, nz=2, tets=True, jitter=0.45, bounds=(-2.0, 1.0, 0.0, 0.5, 3.0, 4.0), typed=True,
                       face_extras=True, permute_cells=True, seed=17),
})


def _same_bytes(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("meshname", list(GEN))
def test_generated_mesh_matches_host_generator(meshname, precision):
    params = GEN[meshname]
    ref = ch.HostMesh(**params)
    host, dev = ch.generate_mesh(precision=precision, **params)
    bad = [k for k in ref.a if not _same_bytes(ref.a[k], host.a[k])]
    assert bad == []
    dref = ch.DeviceMesh(ref, precision).digest()
    assert _differ(dev.digest(), dref) == []
    _, dev2 = ch.generate_mesh(precision=precision, host_copy=False, **params)
    assert _differ(dev2.digest(), dref) == []


@pytest.mark.gpu
def test_generated_mesh_matches_host_generator_96cubed():
    params = dict(nx=96, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=23)
    ref = ch.HostMesh(**params)
    host, dev = ch.generate_mesh(**params)
    assert [k for k in ref.a if not _same_bytes(ref.a[k], host.a[k])] == []
    assert _differ(dev.digest(), ch.DeviceMesh(ref, 64).digest()) == []


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["gradient", "scalar", "uvw"])
def test_kernels_on_generated_mesh_without_host_copy(kernel):
    """A generated mesh with no host copy starts from the generator's state and runs the drop-in
    kernels bit-identically to a mesh built from the host generator's mesh."""
    params = GEN["tets_typed"]
    ref = ch.HostMesh(**params)
    outs = []
    for dev in (ch.DeviceMesh(ref, 64), ch.generate_mesh(host_copy=False, **params)[1]):
        rc, out = run(kernel, fvm, dev, W.Fields(ref, seed=8).copy(), lambda: {})
        assert rc == ch.E_SUCCESS
        outs.append(out)
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.gpu
def test_generate_rejects_bad_params():
    L = ch.lib()
    for kw, code in ((dict(nx=0), 90), (dict(nx=2, ny=-1), 90), (dict(nx=2, tets=True, jitter=0.5), 90),
                     (dict(nx=2, tets=True, jitter=-0.1), 90)):
        p = ch.structgen_params(**kw)
        hh, dh = ct.c_void_p(), ct.c_void_p()
        assert L.cupcfd_hip_mesh_generate_f64(ct.byref(p), 1, ct.byref(hh), ct.byref(dh)) == code, kw
        assert not hh.value and not dh.value
    assert L.cupcfd_hip_mesh_generate_f64(None, 1, None, None) == ch.E_NULL_PTR


@pytest.mark.gpu
@pytest.mark.parametrize("on_device", [False, True])
@pytest.mark.parametrize("meshname", ["typed_perm", "typed32", "tets16", "line"])
def test_face_order_groups_faces_by_neighbour_tile(meshname, on_device):
    """The internal face order, from either repack: interior faces before boundary faces; interior faces by
    owner tile (64 consecutive internal cells), and within a tile first the faces whose cell2 is in the same
    tile, then those shared with each other tile, by that tile, in owner then mesh order within each group.
    So the faces a tile reads from one neighbour (its halo faces) are one contiguous run of records."""
    host = ch.HostMesh(**MESHES[meshname])
    dev = ch.DeviceMesh(host, 64, on_device=on_device)
    a = host.a
    fo = dev.face_order
    rank = np.empty(host.nCells, np.int64)
    rank[dev.cell_order] = np.arange(host.nCells)
    bnd = a["faceBnd"][fo] > -1
    Fi = int((~bnd).sum())
    assert not bnd[:Fi].any() and bnd[Fi:].all()
    own = rank[a["faceCell1"][fo[:Fi]]]
    nb = rank[a["faceCell2"][fo[:Fi]]]
    t1, t2 = own // 64, nb // 64
    grp = np.where(t2 == t1, 0, t2 + 1)
    key = np.stack([fo[:Fi], own, grp, t1])  # lexsort: last row is the primary key
    assert np.array_equal(np.lexsort(key), np.arange(Fi))
    # each (owner tile, neighbour tile) pair's faces are one contiguous run
    pair = t1 * (host.nCells // 64 + 2) + grp
    starts = np.flatnonzero(np.r_[True, pair[1:] != pair[:-1]])
    assert len(np.unique(pair[starts])) == len(starts)

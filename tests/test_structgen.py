"""Host mesh source: StructGen numbering against the reference's own known
answers (tests/golden/structgen_kat.json, extracted from
tests/geometry/mesh/implementation/source/MeshStructGenSourceTests.cpp), plus
closed forms of the axis-aligned geometry and relabelling invariants."""
import json
import os

import numpy as np
import pytest

import cupcfd_hip as ch

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "structgen_kat.json")))["cases"]


def mesh_of(case, **kw):
    nx, ny, nz, x0, x1, y0, y1, z0, z1 = case["source"]
    return ch.HostMesh(nx, ny, nz, bounds=(x0, x1, y0, y1, z0, z1), **kw)


def test_counts():
    assert mesh_of(KAT["getCellCount_test1"]).nCells == KAT["getCellCount_test1"]["expected"]
    assert mesh_of(KAT["getFaceCount_test1"]).nFaces == KAT["getFaceCount_test1"]["expected"]
    assert mesh_of(KAT["getBoundaryCount_test1"]).nBoundaries == KAT["getBoundaryCount_test1"]["expected"]
    assert mesh_of(KAT["getRegionCount_test1"]).nRegions == KAT["getRegionCount_test1"]["expected"]
    for n in (1, 2, 7, 32):  # F = 3 n^2 (n+1), B = 6 n^2
        m = ch.HostMesh(n)
        assert (m.nCells, m.nFaces, m.nBoundaries) == (n ** 3, 3 * n * n * (n + 1), 6 * n * n)


def test_cell_centres_exact():
    c = KAT["getCellCenter_test1"]
    m = mesh_of(c)
    assert np.array_equal(m.a["cellCenter"], np.array(c["cmp"], dtype=np.float64))


def test_cell_face_labels():
    c = KAT["getCellFaceLabels_test1"]
    m = mesh_of(c)
    idx, data = c["cmpIndices"], c["cmpData"]
    c1, c2 = m.a["faceCell1"], m.a["faceCell2"]
    for cell in range(m.nCells):
        expect = sorted(data[idx[cell]:idx[cell + 1]])
        got = sorted(np.nonzero((c1 == cell) | (c2 == cell))[0].tolist())
        assert got == expect, cell


def test_face_is_boundary():
    c = KAT["getFaceIsBoundary_test1"]
    m = mesh_of(c)
    got = (m.a["faceBnd"] > -1)[np.array(c["faceLabels"])]
    assert got.tolist() == c["cmp"]


def test_face_boundary_labels():
    c = KAT["getFaceBoundaryLabels_test1"]
    m = mesh_of(c)
    assert m.a["faceBnd"][np.array(c["faceLabels"])].tolist() == c["cmp"]
    # boundary -> face is the inverse map
    assert np.array_equal(m.a["faceBnd"][m.a["bndFace"]], np.arange(m.nBoundaries))


def test_face_cells():
    c1k, c2k = KAT["getFaceCell1Labels_test1"], KAT["getFaceCell2Labels_test1"]
    m = mesh_of(c1k)
    assert m.a["faceCell1"][np.array(c1k["faceLabels"])].tolist() == c1k["cmp"]
    assert m.a["faceCell2"][np.array(c2k["faceLabels"])].tolist() == c2k["cmp"]
    # lambda: 0.5 interior, -1 boundary (MeshStructGenSource.cpp:824-849)
    bnd = m.a["faceBnd"] > -1
    assert (m.a["faceLambda"][bnd] == -1.0).all() and (m.a["faceLambda"][~bnd] == 0.5).all()


@pytest.mark.parametrize("dims", [(3, 4, 5), (8, 8, 8)])
def test_geometry_closed_forms(dims):
    nx, ny, nz = dims
    b = (-1.0, 2.0, 0.0, 1.0, 0.5, 3.0)
    m = ch.HostMesh(nx, ny, nz, bounds=b)
    L = np.array([b[1] - b[0], b[3] - b[2], b[5] - b[4]])
    h = L / np.array(dims)
    a = m.a
    # volumes and areas
    assert np.allclose(a["cellVolume"], h.prod(), rtol=1e-14)
    assert np.isclose(a["cellVolume"].sum(), L.prod(), rtol=1e-12)
    bnd = a["faceBnd"] > -1
    assert np.isclose(a["faceArea"][bnd].sum(), 2 * (L[0] * L[1] + L[1] * L[2] + L[0] * L[2]), rtol=1e-12)
    # |S| = area, S axis-aligned
    S = a["faceNorm"]
    assert np.allclose(np.linalg.norm(S, axis=1), a["faceArea"], rtol=1e-14)
    assert ((S != 0).sum(axis=1) == 1).all()
    # S points out of cell1: (Xf - C1) . S > 0
    d = a["faceCenter"] - a["cellCenter"][a["faceCell1"]]
    assert ((d * S).sum(axis=1) > 0).all()
    # every cell is closed: sum of outward S over its faces vanishes
    out = np.zeros((m.nCells, 3))
    np.add.at(out, a["faceCell1"], S)
    inner = ~bnd
    np.add.at(out, a["faceCell2"][inner], -S[inner])
    assert np.abs(out).max() <= 1e-14 * np.abs(S).max()
    # interior face centre is the midpoint of its two cell centres
    mid = 0.5 * (a["cellCenter"][a["faceCell1"][inner]] + a["cellCenter"][a["faceCell2"][inner]])
    assert np.allclose(a["faceCenter"][inner], mid, rtol=0, atol=1e-14 * np.abs(L).max())
    # reference meshes: one DEFAULT region, boundary distance 0, zero xpac/xnac/rlencos
    assert a["regType"].tolist() == [ch.RTYPE_DEFAULT]
    assert (a["bndDistance"] == 0).all() and (a["faceXpac"] == 0).all() and (a["faceRLencos"] == 0).all()


def test_typed_regions():
    n = 6
    m = ch.HostMesh(n, typed=True)
    a = m.a
    assert a["regType"].tolist() == [ch.RTYPE_INLET, ch.RTYPE_OUTLET, ch.RTYPE_WALL, ch.RTYPE_WALL, ch.RTYPE_SYMP,
                                      ch.RTYPE_SYMP]
    t = a["regType"][a["bndRegion"]]
    for rt in (ch.RTYPE_INLET, ch.RTYPE_OUTLET):
        assert (t == rt).sum() == n * n
    assert (t == ch.RTYPE_WALL).sum() == 2 * n * n and (t == ch.RTYPE_SYMP).sum() == 2 * n * n
    # INLET faces sit on x = xmin with outward normal -x
    f = a["bndFace"][t == ch.RTYPE_INLET]
    assert (a["faceCenter"][f, 0] == 0.0).all() and (a["faceNorm"][f, 0] < 0).all()
    assert np.allclose(a["bndDistance"], 0.5 / n)


def test_relabel_is_a_permutation():
    base = ch.HostMesh(5, 4, 3, typed=True)
    perm = ch.HostMesh(5, 4, 3, typed=True, permute_cells=True, permute_faces=True, seed=99)
    a, b = base.a, perm.a
    # same multiset of cell centres / volumes
    key = lambda c: np.lexsort(c.T[::-1])  # noqa: E731
    assert np.array_equal(a["cellCenter"][key(a["cellCenter"])], b["cellCenter"][key(b["cellCenter"])])
    # each face keeps its geometry and its cells' centres
    fa = np.lexsort(a["faceCenter"].T[::-1])
    fb = np.lexsort(b["faceCenter"].T[::-1])
    assert np.array_equal(a["faceNorm"][fa], b["faceNorm"][fb])
    assert np.array_equal(a["cellCenter"][a["faceCell1"][fa]], b["cellCenter"][b["faceCell1"][fb]])
    # boundary ids keep StructGen labels; the boundary -> face map follows the faces
    assert np.array_equal(a["faceCenter"][a["bndFace"]], b["faceCenter"][b["bndFace"]])
    assert np.array_equal(b["faceBnd"][b["bndFace"]], np.arange(perm.nBoundaries))
    # deterministic for a seed
    again = ch.HostMesh(5, 4, 3, typed=True, permute_cells=True, permute_faces=True, seed=99)
    assert np.array_equal(again.a["faceCell1"], b["faceCell1"])
    other = ch.HostMesh(5, 4, 3, typed=True, permute_cells=True, permute_faces=True, seed=100)
    assert not np.array_equal(other.a["faceCell1"], b["faceCell1"])


def test_invalid_params():
    with pytest.raises(ch.CupCfdError) as e:
        ch.HostMesh(0, 3, 3)
    assert e.value.code == 90  # E_CONFIG_INVALID_VALUE

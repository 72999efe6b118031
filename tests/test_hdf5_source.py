"""HDF5 mesh source (cupcfd_hdf5_mesh_create) on the reference's own HDF5 test mesh.

tests/golden/MeshHDF5.hdf5 is tests/geometry/mesh/data/MeshHDF5.hdf5 of
CUP-CFD, unchanged. The expected values are the known answers of the
reference's MeshHDF5Source tests
(tests/geometry/mesh/implementation/source/MeshHDF5SourceTests.cpp:216-600),
mapped through the assembly UnstructuredMeshInterface::addData performs on one
rank (cells, faces and boundaries in ascending label order, 0-based ids), and,
for what those tests leave open (face areas, boundary distances, cell centres),
the file's own datasets as h5dump prints them. The file stores reals as
float32: the mesh holds them converted exactly to double, so each expected
double is the float32 nearest the test's decimal.
"""
import os

import numpy as np
import pytest

import cupcfd_hip as ch

HERE = os.path.dirname(os.path.abspath(__file__))
MESH = os.path.join(HERE, "golden", "MeshHDF5.hdf5")
NOT_A_MESH = os.path.join(HERE, "golden", "TestReadSource.h5")  # tests/io/data/TestReadSource.h5


def f32(x):
    return np.asarray(np.asarray(x, np.float32), np.float64)


@pytest.fixture(scope="module")
def mesh():
    return ch.HostMesh.from_hdf5(MESH)


def test_counts(mesh):
    # getCellCount / getFaceCount / getBoundaryCount / getRegionCount (MeshHDF5SourceTests.cpp:216-278)
    assert (mesh.nCells, mesh.nFaces, mesh.nBoundaries, mesh.nRegions) == (4, 20, 17, 1)


def test_face_cells(mesh):
    # getFaceCell1Labels / getFaceCell2Labels (MeshHDF5SourceTests.cpp:450-484), labels 1-based, -1 = boundary
    c1 = np.array([1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4, 1, 2, 2])
    c2 = np.array([-1] * 17 + [2, 3, 4])
    assert np.array_equal(mesh.a["faceCell1"], c1 - 1)
    assert np.array_equal(mesh.a["faceCell2"], np.where(c2 > 0, c2 - 1, -1))


def test_boundaries(mesh):
    # getFaceIsBoundary / getFaceBoundaryLabels (MeshHDF5SourceTests.cpp:395-432): faces 1..17 are boundary
    # faces with boundary labels below; boundary ids are the labels' ranks, so bndFace inverts faceBnd.
    blab = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 12, 14, 17, 16, 15])
    fb = mesh.a["faceBnd"]
    assert np.array_equal(fb[:17], blab - 1)
    assert (fb[17:] == -1).all()
    assert np.array_equal(mesh.a["bndFace"][blab - 1], np.arange(17))
    assert (mesh.a["bndRegion"] == 0).all()
    # /bnd/distance by boundary label (h5dump of the file)
    dist = [0.67, 0.89, 2.4, 2.5, 2.1, 5.3, 5.1, 4.2, 7.0, 3.5, 1.5, 3.7, 23.6, 2.51, 5.0, 643.5, 513.3]
    assert np.array_equal(mesh.a["bndDistance"], f32(dist))


def test_cell_faces_and_volumes(mesh):
    # getCellFaceLabels (MeshHDF5SourceTests.cpp:372-393) for cells 1, 3, 4 and getCellNFaces {6, 5, 6}
    c1, c2 = mesh.a["faceCell1"], mesh.a["faceCell2"]
    for cell, faces in ((1, [1, 2, 3, 4, 5, 18]), (3, [9, 10, 11, 12, 19]), (4, [13, 14, 15, 16, 17, 20])):
        mine = np.flatnonzero((c1 == cell - 1) | (c2 == cell - 1)) + 1
        assert mine.tolist() == faces, cell
    # getCellVolume for labels {1, 3, 4} (MeshHDF5SourceTests.cpp:346-364); cell 2 from the file
    assert np.array_equal(mesh.a["cellVolume"], [1.0, 2.0, 3.0, 4.0])
    assert np.array_equal(mesh.a["cellCenter"], [[5, 15, 25], [10, 110, 210], [15, 115, 215], [25, 125, 225]])


def test_face_geometry(mesh):
    # getFaceLambda (MeshHDF5SourceTests.cpp:486-509) and getFaceNormal (512-543)
    lam = [12.0, 3.2, 0.53, 634.0, -431.3, 23.351, 54.6352, 4246.2, 421.4, 2.53, 351.4, 23.0, 41.0, 0.0, 426.2,
           42111.5, 2462.2, 14.3, 15.0, -252.2]
    assert np.array_equal(mesh.a["faceLambda"], f32(lam))
    nx = np.arange(1.0, 21.0)
    ny = [100.0, 101.0, 102.0, 103.0, 104.0, 105.5, 106.0, 107.0, 108.0, 109.0, 110.0, 111.0, 112.3, 113.0, 114.0,
          115.0, 116.0, 117.0, 118.0, 119.0]
    nz = [200.8, 201.0, 202.0, 203.0, 204.5, 205.0, 206.0, 207.0, 208.0, 209.0, 210.0, 211.1, 212.0, 213.0, 214.0,
          215.0, 216.0, 217.0, 218.9, 219.0]
    assert np.array_equal(mesh.a["faceNorm"], f32(np.stack([nx, ny, nz], axis=1)))
    # /face/area and /face/x (h5dump): 0.1 .. 1.9 then 0; every face centre at the origin
    assert np.array_equal(mesh.a["faceArea"], f32(np.r_[np.arange(1, 20) / 10.0, 0.0]))
    assert (mesh.a["faceCenter"] == 0).all()
    # addData passes zero xpac, xnac and rlencos
    for k in ("faceXpac", "faceXnac", "faceRLencos"):
        assert (mesh.a[k] == 0).all(), k


def test_regions_default(mesh):
    assert mesh.a["regType"].tolist() == [ch.RTYPE_DEFAULT]
    for k in ("regStd", "regFlux", "regAdiab", "regYLog", "regELog", "regDensity", "regTurbKE", "regTurbDiss",
              "regSplvl", "regDen", "regR", "regT", "regForceTangent", "regUVW"):
        assert (mesh.a[k] == 0).all(), k


def test_float_view(mesh):
    v = mesh.view32()
    assert (v.nCells, v.nFaces, v.nBoundaries, v.nRegions) == (4, 20, 17, 1)


def test_errors(tmp_path):
    L = ch.lib()
    h = ch.ct.c_void_p(1)
    assert L.cupcfd_hdf5_mesh_create(os.fsencode(str(tmp_path / "absent.hdf5")), ch.ct.byref(h)) == ch.E_FILE_MISSING
    assert not h.value
    text = tmp_path / "mesh.hdf5"
    text.write_text("not an HDF5 file\n")
    assert L.cupcfd_hdf5_mesh_create(os.fsencode(str(text)), ch.ct.byref(h)) == ch.E_INVALID_FILE_FORMAT
    # an HDF5 file without the mesh attributes and datasets
    assert L.cupcfd_hdf5_mesh_create(os.fsencode(NOT_A_MESH), ch.ct.byref(h)) == ch.E_INVALID_FILE_FORMAT
    assert L.cupcfd_hdf5_mesh_create(None, ch.ct.byref(h)) == ch.E_NULL_PTR
    assert L.cupcfd_hdf5_mesh_create(os.fsencode(MESH), None) == ch.E_NULL_PTR
    with pytest.raises(ch.CupCfdError):
        ch.HostMesh.from_hdf5(str(tmp_path / "absent.hdf5"))

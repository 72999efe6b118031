"""The tetrahedral stress-mesh generator (cupcfd_structgen_create with tets = 1).

Each hex of an n_x x n_y x n_z grid becomes six tetrahedra around its main
diagonal; interior vertices are jittered. The reference has no tet
generator, so these checks pin the generator by what the geometry fixes:
  * counts: C = 6 n^3, F = 12 n^3 + 6 n^2, B = 12 n^2 for a cube, and the
    general-box forms;
  * every cell is closed (sum of outward S = 0) and the divergence identity
    sum_f S_f . Xf = 3 vol holds per cell, which ties normals, face centres
    and volumes together; volumes sum to the box volume; no cell is inverted;
  * Heron's area equals |S|; lambda lies in (0, 1) on interior faces, -1 on
    boundary faces; boundary faces lie on the box planes, with the typed
    region of their side;
  * unjittered cubes give the closed-form Kuhn geometry (vol = h^3 / 6, all
    lambdas 1/2 by symmetry of the centroids);
  * the same seed gives the same mesh.
"""
import numpy as np
import pytest

import cupcfd_hip as ch


def _cellsum(C, c1, c2, v):
    acc = np.zeros((C,) + v.shape[1:])
    m = c2 >= 0
    np.add.at(acc, c1, v)
    np.add.at(acc, c2[m], -v[m])
    return acc


@pytest.mark.parametrize("dims,jitter", [((3, 3, 3), 0.0), ((5, 4, 3), 0.2), ((6, 6, 6), 0.2)])
def test_counts_and_geometry(dims, jitter):
    nx, ny, nz = dims
    h = ch.HostMesh(nx, ny, nz, tets=True, jitter=jitter, seed=3)
    a = h.a
    H = nx * ny * nz
    assert h.nCells == 6 * H
    assert h.nBoundaries == 4 * (nx * ny + ny * nz + nx * nz)
    assert h.nFaces == (4 * 6 * H + h.nBoundaries) // 2
    if nx == ny == nz:
        n = nx
        assert (h.nCells, h.nFaces, h.nBoundaries) == (6 * n**3, 12 * n**3 + 6 * n**2, 12 * n**2)
    S, Xf, vol = a["faceNorm"], a["faceCenter"], a["cellVolume"]
    c1, c2 = a["faceCell1"], a["faceCell2"]
    assert vol.min() > 0
    assert abs(vol.sum() - 1.0) < 1e-12
    assert np.abs(_cellsum(h.nCells, c1, c2, S)).max() < 1e-15
    div = _cellsum(h.nCells, c1, c2, (S * Xf).sum(1))
    assert np.abs(div - 3 * vol).max() < 1e-13 * vol.mean()
    assert np.allclose(a["faceArea"], np.linalg.norm(S, axis=1), rtol=1e-12, atol=0)
    inter = c2 >= 0
    lam = a["faceLambda"]
    assert ((lam[inter] > 0) & (lam[inter] < 1)).all() and (lam[~inter] == -1).all()
    # boundary faces: on a box plane, outward
    bf = a["bndFace"]
    xb = Xf[bf]
    on = (np.isclose(xb, 0, atol=1e-15) | np.isclose(xb, 1, atol=1e-15)).any(1)
    assert on.all()
    out = (S[bf] * (xb - a["cellCenter"][c1[bf]])).sum(1)
    assert (out > 0).all()


def test_unjittered_kuhn_closed_forms():
    n = 4
    h = ch.HostMesh(n, tets=True, jitter=0.0)
    hh = 1.0 / n
    assert np.allclose(h.a["cellVolume"], hh**3 / 6, rtol=1e-13)
    lam = h.a["faceLambda"][h.a["faceCell2"] >= 0]
    assert np.allclose(lam, 0.5, rtol=0, atol=1e-14)


def test_typed_regions_by_side():
    h = ch.HostMesh(3, tets=True, jitter=0.2, typed=True, seed=4)
    a = h.a
    xb = a["faceCenter"][a["bndFace"]]
    reg = a["bndRegion"]
    # INLET x=0, OUTLET x=1, WALL y=0 / y=1, SYMP z=0 / z=1
    want = np.select([xb[:, 0] < 1e-15, xb[:, 0] > 1 - 1e-15, xb[:, 1] < 1e-15, xb[:, 1] > 1 - 1e-15,
                      xb[:, 2] < 1e-15], [0, 1, 2, 3, 4], 5)
    assert np.array_equal(reg, want)
    # boundary distance: normal distance from the cell centre to the face plane
    S = a["faceNorm"][a["bndFace"]]
    d = np.abs(((xb - a["cellCenter"][a["faceCell1"][a["bndFace"]]]) * S).sum(1)) / np.linalg.norm(S, axis=1)
    assert np.allclose(a["bndDistance"], d, rtol=1e-14)


def test_seeded_and_jitter_bounded():
    a1 = ch.HostMesh(4, tets=True, jitter=0.2, seed=9).a
    a2 = ch.HostMesh(4, tets=True, jitter=0.2, seed=9).a
    a3 = ch.HostMesh(4, tets=True, jitter=0.2, seed=10).a
    assert np.array_equal(a1["cellCenter"], a2["cellCenter"])
    assert not np.array_equal(a1["cellCenter"], a3["cellCenter"])
    a0 = ch.HostMesh(4, tets=True, jitter=0.0).a
    # centroids move by at most the jitter bound (0.2 h per component)
    assert np.abs(a1["cellCenter"] - a0["cellCenter"]).max() <= 0.2 * 0.25 + 1e-15


def test_bad_jitter_rejected():
    with pytest.raises(ch.CupCfdError):
        ch.HostMesh(3, tets=True, jitter=0.6)

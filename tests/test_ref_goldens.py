"""The oracle against CUP-CFD's own kernels (CPU).

tests/golden/ref_kernels.npz holds the reference's outputs, from its fvm
headers and mesh classes compiled from the reference sources
(oracle/ref_harness.cpp, written by tests/golden/make_ref_goldens.py), for
every kernel case on three meshes: the reference's own StructGen grid (one
DEFAULT region), a typed relabelled hex mesh with non-zero xpac/xnac/rlencos
and a jittered typed tet mesh, in FP64 and FP32. The oracle must give the
same bits on the same mesh arrays, inputs and mesh state: every branch the
goldens reach (typed INLET / OUTLET / SYMP / WALL boundaries, the carried
PhiFace and uFace / vFace / wFace, GammaBlend upwind blending, rlencos cross
diffusion, the wall functions) is pinned by the reference itself.

The goldens also hold the reference's StructGen meshes, which the host
generator must reproduce bit for bit (Heron face areas, area-weighted face
centroids, five-tetrahedron cell volumes, clockwise-tested normals).

Where the compiled reference (oracle/_ref) is present, the same comparison
also runs live on cases the fixture does not hold.
"""
import os

import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
import ref_cases as rc
import refharness as rh
from cupcfd_hip import workload as W
from kernel_runs import run, seed_state

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz"))


def bits(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.view(np.uint64)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    return a


def fixture_case(name):
    mesh = ch.ArrayMesh({k: GOLD[f"{name}/mesh/{k}"] for k in ch.ARRAY_KEYS})
    f = W.Fields.__new__(W.Fields)
    prefix = f"{name}/in/"
    for k in GOLD.files:
        if k.startswith(prefix):
            f.__dict__[k[len(prefix):]] = GOLD[k]
    st = {k[len(f"{name}/state/"):]: GOLD[k] for k in GOLD.files if k.startswith(f"{name}/state/")}
    return mesh, f, st


def cases():
    out = []
    for name in rc.MESHES:
        for prec in (64, 32):
            for kern in rc.kernels_for(name):
                out.append((name, prec, kern))
    return out


@pytest.mark.parametrize("name,prec,kern", cases())
def test_oracle_matches_reference(name, prec, kern):
    mesh, f, st0 = fixture_case(name)
    dt = np.float64 if prec == 64 else np.float32
    om = po.OracleMesh(mesh, prec)
    om.set_state({k: v.astype(dt) for k, v in st0.items()})
    code, out = run(kern, po, om, f.cast(dt), om.state)
    key = f"{name}/{prec}/{kern}"
    assert code == int(GOLD[key + "/rc"][0]) == ch.E_SUCCESS
    expect = {k[len(key) + 5:]: GOLD[k] for k in GOLD.files if k.startswith(key + "/out/")}
    assert expect.keys() == out.keys()
    for k, v in expect.items():
        assert np.array_equal(bits(out[k]), bits(v)), f"{key}: {k} differs from the reference"


def test_goldens_exercise_every_boundary_type():
    """The typed meshes reach each region type, and the reference's outputs there are not trivial."""
    mesh, _, _ = fixture_case("typed345")
    t = mesh.a["regType"][mesh.a["bndRegion"]]
    for rt in (ch.RTYPE_INLET, ch.RTYPE_OUTLET, ch.RTYPE_SYMP, ch.RTYPE_WALL):
        assert (t == rt).any()
    assert mesh.a["regType"][mesh.a["bndRegion"][mesh.a["faceBnd"][0]]] == ch.RTYPE_OUTLET
    assert (mesh.a["faceRLencos"] != 0).all() and (mesh.a["faceXpac"] != 0).any()
    cnt = GOLD["typed345/64/mass/out/counters"]
    assert (cnt > 0).all()  # INLET, OUTLET, SYMP and WALL faces counted by the reference
    # the scalar WALL-enthalpy branch writes boundary heat fluxes
    assert np.abs(GOLD["typed345/64/scalar_wall/out/mesh.bndQ"]).max() > 0


@pytest.mark.parametrize("name", list(rc.STRUCTGEN))
def test_structgen_matches_reference(name):
    nx, ny, nz, bounds = rc.STRUCTGEN[name]
    h = ch.HostMesh(nx, ny, nz, bounds=bounds)
    for k in rc.SG_KEYS:
        assert np.array_equal(bits(h.a[k]), bits(GOLD[f"{name}/{k}"])), f"{name}: {k}"


@pytest.mark.parametrize("name", list(rc.MESHES))
def test_fixture_meshes_are_the_generators(name):
    """The fixture meshes are what the generator builds today (so GPU runs on them test today's generator too)."""
    mesh = rc.golden_mesh(name)
    for k in ch.ARRAY_KEYS:
        assert np.array_equal(bits(mesh.a[k]), bits(GOLD[f"{name}/mesh/{k}"])), f"{name}: {k}"


# ---------------------------------------------------------------------------
# live comparison with the compiled reference, where it is present
# ---------------------------------------------------------------------------
LIVE = [
    ("typed654", dict(nx=6, ny=5, nz=4, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=31)),
    ("tets322", dict(nx=3, ny=2, nz=2, tets=True, jitter=0.3, typed=True, permute_cells=True, permute_faces=True,
                     face_extras=True, seed=8)),
]


@pytest.mark.skipif(not rh.available(), reason="compiled reference (oracle/_ref) not built in this tree")
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("mesh_name,params", LIVE)
def test_oracle_matches_live_reference(mesh_name, params, prec):
    from kernel_runs import KERNELS
    mesh = rc.outlet_first(ch.HostMesh(**params))
    dt = np.float64 if prec == 64 else np.float32
    f = W.Fields(mesh, seed=77)
    st0 = {k: v.astype(dt) for k, v in seed_state(mesh, 9).items()}
    for kern in KERNELS:
        om = po.OracleMesh(mesh, prec)
        om.set_state(st0)
        rm = rh.RefMesh(mesh, prec)
        rm.set_state(st0)
        c1, out_o = run(kern, po, om, f.cast(dt), om.state)
        c2, out_r = run(kern, rh, rm, f.cast(dt), rm.state)
        assert c1 == c2 == ch.E_SUCCESS
        for k in out_r:
            assert np.array_equal(bits(out_o[k]), bits(out_r[k])), f"{mesh_name}/{prec}/{kern}: {k}"


@pytest.mark.skipif(not rh.available(), reason="compiled reference (oracle/_ref) not built in this tree")
def test_live_structgen_relabel_free_grid():
    """An odd grid the fixture does not hold: the generator equals the reference StructGen."""
    r = rh.structgen(9, 4, 6, bounds=(-2.0, 0.3, 1.0, 1.7, -0.25, 4.0))
    h = ch.HostMesh(9, 4, 6, bounds=(-2.0, 0.3, 1.0, 1.7, -0.25, 4.0))
    for k in rc.SG_KEYS:
        assert np.array_equal(bits(h.a[k]), bits(r[k])), k

"""Parity of every HIP kernel with the CPU oracle, through the drop-in C ABI.

Each case feeds identical seeded inputs (and identical mesh state) to the
oracle (pyoracle, the sequential restatement of include/fvm/implementation/component/*.ipp)
and to libcupcfd_hip, then compares every output array and every piece of
mesh state the kernel sets.

Tolerances (max-norm relative, ||gpu - oracle||inf / ||oracle||inf):
  FP64: 1e-12; FP32: 2e-6 against the float oracle. Kernels with no
  transcendental calls are additionally required to be bit-identical in both
  precisions: the per-cell CSR gather keeps the reference's face order, the
  library is built with -ffp-contract=off, and the float kernels keep the
  reference's promotions (a float meeting a double literal is computed in
  double, e.g. fmin(MassFlux, 0.0)). Only the log/pow/sqrt-based wall
  functions may differ by an ulp between libm and the GPU math library.
"""
import zlib

import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import fvm
from cupcfd_hip import workload as W
from kernel_runs import KERNELS, compare, run, seed_state

MESHES = {
    # the reference's fvm unit-test mesh: 5^3 StructGen over [-1,1]^3, one DEFAULT region
    "ref5": dict(nx=5, bounds=(-1, 1, -1, 1, -1, 1)),
    # six typed regions, relabelled cells and faces, non-zero xpac/xnac/rlencos
    "typed_perm": dict(nx=7, ny=6, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=11),
    # 32^3 over [0,1]^3 (the CPU benchmark mesh), typed, relabelled
    "typed32": dict(nx=32, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=3),
    # 4x3x3 hexes split into 6 jittered tetrahedra each (216 cells, 4-face rows), typed, relabelled
    "tets_typed": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, typed=True, permute_cells=True,
                       permute_faces=True, face_extras=True, seed=5),
}
# meshes at the edges of the tile schedule: one cell (a tile that owns no interior face), and a slab one cell
# thick (every cell owns boundary faces; the last tile is partial)
ODD_MESHES = {
    "one_cell": dict(nx=1, ny=1, nz=1, typed=True),
    "slab": dict(nx=40, ny=3, nz=1, typed=True, permute_cells=True, permute_faces=True, seed=2),
}
# kernels whose FP64 and FP32 results must be bit-identical to the oracle
BITEXACT = {"gradient", "gradient3", "mass", "mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5",
            "mass_region", "scalar", "scalar_t", "uvw", "uvw_region1", "uvw_bnds1", "visc_cell1", "visc_cell2"}

_cache = {}


def _setup(meshname, precision):
    key = (meshname, precision)
    if key not in _cache:
        host = ch.HostMesh(**(MESHES[meshname] if meshname in MESHES else ODD_MESHES[meshname]))
        dev = ch.DeviceMesh(host, precision=precision, reorder=True)
        _cache[key] = (host, dev)
    return _cache[key]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("meshname", list(MESHES) + list(ODD_MESHES))
@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_parity(kernel, meshname, precision):
    host, dev = _setup(meshname, precision)
    dt = np.float64 if precision == 64 else np.float32
    fields = W.Fields(host, seed=zlib.crc32(f"{kernel}/{meshname}".encode()) % 1000)
    st0 = seed_state(host)
    st0 = {k: v.astype(dt) for k, v in st0.items()}

    om = po.OracleMesh(host, precision)
    om.set_state(st0)
    rc_o, out_o = run(kernel, po, om, fields.cast(dt), om.state)

    host.set_state(st0, precision)
    rc_g, out_g = run(kernel, fvm, dev, fields.cast(dt), lambda: host.state(precision))

    assert rc_o == ch.E_SUCCESS and rc_g == ch.E_SUCCESS, (rc_o, rc_g)
    assert out_o.keys() == out_g.keys()
    tol = 1e-12 if precision == 64 else 2e-6
    for k in out_o:
        same, rel = compare(out_g[k], out_o[k])
        if kernel in BITEXACT:
            assert same, f"{kernel}/{meshname}/{precision}: {k} not bit-identical (max rel {rel:.3e})"
        assert rel <= tol, f"{kernel}/{meshname}/{precision}: {k} max rel {rel:.3e} > {tol}"


@pytest.mark.gpu
def test_reorder_is_invisible():
    """Internal renumbering must not change any result bit."""
    host = ch.HostMesh(**MESHES["typed_perm"])
    a = ch.DeviceMesh(host, 64, reorder=True)
    b = ch.DeviceMesh(host, 64, reorder=False)
    assert not np.array_equal(a.cell_order, b.cell_order)
    for kernel in ("gradient", "mass", "scalar", "uvw"):
        f = W.Fields(host, seed=5)
        outs = []
        for dm in (a, b):
            host.set_state(seed_state(host))
            outs.append(run(kernel, fvm, dm, f.copy(), lambda: host.state(64))[1])
        for k in outs[0]:
            assert compare(outs[0][k], outs[1][k])[0], (kernel, k)


@pytest.mark.gpu
def test_deterministic_repeat():
    host = ch.HostMesh(**MESHES["typed32"])
    dev = ch.DeviceMesh(host, 64)
    f = W.Fields(host, seed=9)
    r1 = run("uvw", fvm, dev, f.copy(), lambda: host.state(64))[1]
    r2 = run("uvw", fvm, dev, f.copy(), lambda: host.state(64))[1]
    for k in r1:
        assert np.array_equal(r1[k], r2[k]), k


@pytest.mark.gpu
def test_strided_vectors_match_packed():
    """3-vectors laid out as the reference's EuclideanVector<double,3> (an 8-byte
    vptr slot, then cmp[3]: 32-byte stride) give the packed-layout results."""
    host = ch.HostMesh(**MESHES["typed_perm"])
    dev = ch.DeviceMesh(host, 64)
    f = W.Fields(host, seed=31)
    C = host.nCells

    def strided(a):
        s = np.full((a.shape[0], 4), np.nan)
        s[:, 1:] = a
        return s

    g_p, go_p = np.zeros((C, 3)), np.zeros((C, 3))
    assert fvm.GradientPhiGaussDolfyn(dev, 1, f.phi, f.phiB, g_p, go_p) == ch.E_SUCCESS
    g_s, go_s = strided(np.zeros((C, 3))), strided(np.zeros((C, 3)))
    L = ch.lib()
    V = ch.V3F64
    rc = L.cupcfd_fvm_gradient_phi_gauss_dolfyn_f64(
        dev.handle, 1, f.phi.ctypes.data_as(ch.ct.c_void_p), C, f.phiB.ctypes.data_as(ch.ct.c_void_p),
        host.nBoundaries, V(g_s.ctypes.data + 8, 32), C, V(go_s.ctypes.data + 8, 32), C)
    assert rc == ch.E_SUCCESS
    assert np.array_equal(g_s[:, 1:], g_p) and np.array_equal(go_s[:, 1:], go_p)
    assert np.isnan(g_s[:, 0]).all()  # the vptr slot is never written


@pytest.mark.gpu
def test_size_errors_are_reported():
    """Arrays shorter than the mesh needs: E_INVALID_INDEX, as the reference's
    bounds checks return (ScalarKernels.ipp:102-124)."""
    host, dev = _setup("ref5", 64)
    f = W.Fields(host, seed=1)
    g, go = np.zeros((host.nCells, 3)), np.zeros((host.nCells, 3))
    assert fvm.GradientPhiGaussDolfyn(dev, 1, f.phi[:-1].copy(), f.phiB, g, go) == ch.E_INVALID_INDEX
    assert fvm.GradientPhiGaussDolfyn(dev, 1, f.phi, f.phiB, g[:-1].copy(), go) == ch.E_INVALID_INDEX
    # longer arrays are fine; entries past the mesh's cells are zeroed like the reference's reset loop
    g2, go2 = np.full((host.nCells + 4, 3), 5.0), np.full((host.nCells + 4, 3), 5.0)
    assert fvm.GradientPhiGaussDolfyn(dev, 1, f.phi, f.phiB, g2, go2) == ch.E_SUCCESS
    assert (g2[host.nCells:] == 0).all() and (go2[host.nCells:] == 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("meshname", ["ref5", "typed_perm", "typed32", "tets_typed"])
def test_schedules_bit_identical(meshname):
    """One thread per cell, colour groups, 6 or 8 lanes per cell and 64-cell tiles give the same bits."""
    host, _ = _setup(meshname, 64)
    outs = {}
    for lpc in (1, 2, 6, 8, 64):
        dev = ch.DeviceMesh(host, 64)
        dev.set_schedule(lpc)
        for kernel in ("gradient", "gradient3", "scalar", "scalar_t", "scalar_wall", "uvw"):
            f = W.Fields(host, seed=77)
            host.set_state(seed_state(host))
            outs[(lpc, kernel)] = run(kernel, fvm, dev, f, lambda: host.state(64))[1]
    for kernel in ("gradient", "gradient3", "scalar", "scalar_t", "scalar_wall", "uvw"):
        for k in outs[(1, kernel)]:
            for lpc in (2, 6, 8, 64):
                assert compare(outs[(1, kernel)][k], outs[(lpc, kernel)][k])[0], (kernel, k, lpc)


def _greedy_colours(host):
    """The order-preserving colouring, sequentially: faces in mesh order, colour = 1 + the last colour given to
    either of its cells (0 if none)."""
    c1, c2 = host.a["faceCell1"], host.a["faceCell2"]
    last = np.full(host.nCells, -1, np.int64)
    col = np.empty(host.nFaces, np.int32)
    for f in range(host.nFaces):
        a, b = c1[f], c2[f]
        q = (last[a] if b < 0 else max(last[a], last[b])) + 1
        col[f] = q
        last[a] = q
        if b >= 0:
            last[b] = q
    return col


@pytest.mark.gpu
@pytest.mark.parametrize("meshname", ["ref5", "typed_perm", "typed32", "tets_typed", "line"])
def test_face_colouring_is_the_sequential_greedy(meshname):
    """The device colouring equals its sequential twin, and no two faces of one colour share a cell."""
    spec = {"line": dict(nx=150, ny=1, nz=1, permute_faces=True, seed=6)}
    host = ch.HostMesh(**spec[meshname]) if meshname in spec else _setup(meshname, 64)[0]
    col, n = ch.DeviceMesh(host, 64).face_colours()
    ref = _greedy_colours(host)
    assert n == ref.max() + 1
    assert np.array_equal(col, ref)
    c1, c2 = host.a["faceCell1"], host.a["faceCell2"]
    for q in range(n):
        cells = np.concatenate([c1[col == q], c2[(col == q) & (c2 >= 0)]])
        assert np.unique(cells).size == cells.size, q


# unordered, these tiles have 342 (typed_perm), 384 (typed32), 228 (tets_typed) and 247 (tets_scattered) face
# slots: the scalar and UVW tiles take their run-time row width, the 304-column and the 384-column layouts
SCATTERED = {"tets_scattered": dict(nx=6, ny=5, nz=4, tets=True, jitter=0.3, typed=True, permute_cells=True,
                                    permute_faces=True, face_extras=True, seed=9)}


@pytest.mark.gpu
@pytest.mark.parametrize("meshname", ["typed_perm", "typed32", "tets_typed", "tets_scattered"])
def test_tiles_of_scattered_cells(meshname):
    """Without the Morton renumbering a tile's 64 cells are scattered, so a tile can have more than 256 face
    slots and threads take a second slot, whose outside cell is staged after the first barrier: still the bits
    of one thread per cell, in every LDS row width. The mass flux on tiles (a tile may own more faces than its
    workgroup has threads) gives the bits of one thread per face."""
    host = ch.HostMesh(**SCATTERED[meshname]) if meshname in SCATTERED else _setup(meshname, 64)[0]
    outs = {}
    for lpc in (1, 64):
        dev = ch.DeviceMesh(host, 64, reorder=False)
        dev.set_schedule(lpc)
        for kernel in ("gradient", "mass", "scalar", "scalar_wall", "uvw"):
            f = W.Fields(host, seed=13)
            host.set_state(seed_state(host))
            outs[(lpc, kernel)] = run(kernel, fvm, dev, f, lambda: host.state(64))[1]
    for kernel in ("gradient", "mass", "scalar", "scalar_wall", "uvw"):
        for k in outs[(1, kernel)]:
            assert compare(outs[(1, kernel)][k], outs[(64, kernel)][k])[0], (kernel, k)


@pytest.mark.gpu
def test_autotune_cache_per_mesh_class(tmp_path, monkeypatch):
    """Autotune results are kept per (architecture, precision, mesh class) and, with CUPCFD_HIP_TUNE_CACHE set,
    in that file: a mesh of the same class takes the cached choice, and a fresh cache reads the file."""
    L = ch.lib()
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("CUPCFD_HIP_TUNE_CACHE", str(path))
    L.cupcfd_hip_autotune_forget()
    try:
        a = ch.DeviceMesh(ch.HostMesh(12, permute_cells=True, seed=1), 64)
        ca = a.autotune()
        lines = [ln.split() for ln in path.read_text().splitlines() if ln.strip()]
        # "key gradient scalar uvw mass": the mass flux's choice is cell tiles (64) or one thread per face (1)
        assert len(lines) == 1 and [int(x) for x in lines[0][1:4]] == ca and int(lines[0][4]) in (1, 64)
        b = ch.DeviceMesh(ch.HostMesh(12, permute_cells=True, permute_faces=True, seed=2), 64)
        assert b.autotune() == ca
        assert len([ln for ln in path.read_text().splitlines() if ln.strip()]) == 1
        # a cache entry written by another process is what the next call takes (a line without the mass
        # flux's entry is read too)
        path.write_text(f"{lines[0][0]} 1 1 1\n")
        L.cupcfd_hip_autotune_forget()
        c = ch.DeviceMesh(ch.HostMesh(12, seed=3), 64)
        assert c.autotune() == [1, 1, 1]
    finally:
        monkeypatch.delenv("CUPCFD_HIP_TUNE_CACHE")
        L.cupcfd_hip_autotune_forget()


@pytest.mark.gpu
@pytest.mark.parametrize("meshname", ["typed_perm", "typed32", "tets_typed"])
def test_autotune_keeps_bits_and_state(meshname):
    """cupcfd_hip_mesh_autotune picks one valid schedule per gather kernel, leaves the mesh state as it was,
    and the kernels keep giving the bits of one thread per cell."""
    host, _ = _setup(meshname, 64)
    dev = ch.DeviceMesh(host, 64)
    host.set_state(seed_state(host))
    dev.push_state()
    before = host.state(64)
    chosen = dev.autotune()
    assert len(chosen) == 3 and all(c in (64, 4, 6, 8, 1, 2) for c in chosen), chosen
    assert chosen[1] != 2 and chosen[2] != 2  # colour groups are a gradient schedule
    dev.pull_state()
    after = host.state(64)
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    ref = ch.DeviceMesh(host, 64)
    ref.set_schedule(1)
    for kernel in ("gradient", "scalar", "scalar_wall", "uvw"):
        outs = []
        for d in (dev, ref):
            f = W.Fields(host, seed=5)
            host.set_state(seed_state(host))
            outs.append(run(kernel, fvm, d, f, lambda: host.state(64))[1])
        for k in outs[1]:
            assert compare(outs[0][k], outs[1][k])[0], (kernel, k)


@pytest.mark.gpu
def test_uvw_split_layout_in_use_and_bit_identical():
    """FP64 meshes of 4x4x4-cell hex tiles run the UVW tile's split LDS layout (96 slots with a cell outside the
    tile or a boundary face, 144 between two cells of the tile); its results equal the one-thread-per-cell
    schedule's bit for bit, and FP32 meshes and tet tiles keep the plain layout."""
    host, dev = _setup("typed32", 64)
    lay = dev.tile_layout()
    assert lay["uvw_split"] == 1 and lay["outside"] == 96 and lay["inside"] == 144, lay
    assert _setup("typed32", 32)[1].tile_layout()["uvw_split"] == 0
    assert _setup("tets_typed", 64)[1].tile_layout()["uvw_split"] == 0
    outs = []
    for lpc in (64, 1):
        d = ch.DeviceMesh(host, 64)
        d.set_schedule(lpc)
        f = W.Fields(host, seed=78)
        host.set_state(seed_state(host))
        outs.append(run("uvw", fvm, d, f, lambda: host.state(64))[1])
    for k in outs[0]:
        assert compare(outs[0][k], outs[1][k])[0], k

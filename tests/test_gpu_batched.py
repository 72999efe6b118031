"""Batched K-scalar kernels against K calls of the single-scalar kernels.

cupcfd_dev_gradient_batched_* and cupcfd_dev_flux_scalar_batched_* promise,
for every scalar k, the bits of the single-scalar resident entry point on
scalar k's arrays (which test_gpu_parity.py pins to the oracle). Per-scalar
arrays are scalar-innermost: (C, K), (B, K), (C, 3, K). Cases cover the
three lane-group widths (K = 1, 3 -> 4 lanes; 16; 64) and a K above 64
(several chunks of 64), on the reference test mesh, a typed relabelled mesh
(carried PhiFace, INLET/OUTLET/SYMP/WALL branches) and a 32^3 typed mesh, in
FP64 and FP32, with nGradient 1 and 3.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R
from kernel_runs import seed_state

torch = pytest.importorskip("torch")

MESHES = {
    "ref5": dict(nx=5, bounds=(-1, 1, -1, 1, -1, 1)),
    "typed_perm": dict(nx=7, ny=6, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=11),
    "typed32": dict(nx=32, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=3),
    "tets_typed": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, typed=True, permute_cells=True,
                       permute_faces=True, face_extras=True, seed=5),
    # one DEFAULT region: the scalar flux evaluates interior entries only (4-face rows; ref5 has 6)
    "tets_plain": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, permute_cells=True, permute_faces=True,
                       face_extras=True, seed=6),
}
_cache = {}


def _mesh(name, precision):
    if (name, precision) not in _cache:
        host = ch.HostMesh(**MESHES[name])
        _cache[(name, precision)] = (host, ch.DeviceMesh(host, precision=precision, reorder=True))
    return _cache[(name, precision)]


def _dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).cuda()


def _same(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("meshname", list(MESHES))
@pytest.mark.parametrize("K", [1, 3, 16, 64, 70])
@pytest.mark.parametrize("nG", [1, 3])
def test_gradient_batched_matches_single(K, meshname, precision, nG):
    if meshname == "typed32" and K > 16 and nG == 3:
        pytest.skip("covered by the smaller meshes")
    host, dev = _mesh(meshname, precision)
    dt = np.float64 if precision == 64 else np.float32
    C, B = host.nCells, host.nBoundaries
    rng = np.random.default_rng(100 + K)
    phi = rng.uniform(-1, 1, (C, K))
    phiB = rng.uniform(-1, 1, (B, K))
    g = torch.zeros((C, 3, K), dtype=torch.float64 if precision == 64 else torch.float32, device="cuda")
    go = torch.full_like(g, 7.0)
    R.gradient_batched(dev, nG, K, _dev(phi, dt), _dev(phiB, dt), g, go)
    for k in range(K):
        g1 = torch.zeros((C, 3), dtype=g.dtype, device="cuda")
        go1 = torch.full_like(g1, 7.0)
        R.gradient(dev, nG, _dev(phi[:, k], dt), _dev(phiB[:, k], dt), g1, go1)
        torch.cuda.synchronize()
        assert _same(g[:, :, k].contiguous(), g1), k
        assert _same(go[:, :, k].contiguous(), go1), k


def _scalar_case(host, dev, precision, K, p, seed):
    dt = np.float64 if precision == 64 else np.float32
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    rng = np.random.default_rng(seed)
    f = W.Fields(host, seed=seed)
    per = dict(phi=rng.uniform(-1, 1, (C, K)), phiB=rng.uniform(-1, 1, (B, K)), au=rng.uniform(0.5, 1.5, (C, K)),
               su=rng.uniform(-1, 1, (C, K)), dphidx=rng.uniform(-1, 1, (C, 3, K)))
    shared = dict(visEff=f.visEff, den=f.den, mf=f.massFlux, te=f.te, cpB=f.cpB, visEffB=f.visEffB)
    st0 = seed_state(host)

    def reset():
        host.set_state(st0, precision)
        dev.push_state()

    # batched
    reset()
    bt = {k: _dev(v, dt) for k, v in per.items()}
    sh = {k: _dev(v, dt) for k, v in shared.items()}
    rface_b = torch.zeros(2 * F, dtype=bt["phi"].dtype, device="cuda")
    R.scalar_batched(dev, K, bt["phi"], bt["phiB"], sh["visEff"], bt["au"], bt["su"], sh["den"], sh["mf"], sh["te"],
                     sh["cpB"], sh["visEffB"], rface_b, bt["dphidx"], p)
    torch.cuda.synchronize()
    dev.pull_state()
    state_b = host.state(precision)
    # K single-scalar calls
    reset()
    outs = []
    for k in range(K):
        one = {n: _dev(v[..., k], dt) for n, v in per.items()}
        rface = torch.zeros(2 * F, dtype=rface_b.dtype, device="cuda")
        R.scalar(dev, one["phi"], one["phiB"], sh["visEff"], one["au"], one["su"], sh["den"], sh["mf"], sh["te"],
                 sh["cpB"], sh["visEffB"], rface, one["dphidx"], p)
        outs.append((one, rface))
    torch.cuda.synchronize()
    dev.pull_state()
    state_1 = host.state(precision)
    for k, (one, rface) in enumerate(outs):
        for n in ("su", "au", "phiB"):
            assert _same(bt[n][..., k].contiguous(), one[n]), (n, k)
        assert _same(rface_b, rface), ("RFace", k)
    for n in state_b:
        assert np.array_equal(state_b[n], state_1[n]), n


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("meshname", list(MESHES))
@pytest.mark.parametrize("K", [1, 3, 16, 64, 70])
@pytest.mark.parametrize("variant", ["scalar", "scalar_wall"])
def test_scalar_batched_matches_single(K, meshname, precision, variant):
    if meshname == "typed32" and K > 16:
        pytest.skip("covered by the smaller meshes")
    host, dev = _mesh(meshname, precision)
    p = dict(W.SCALAR)
    if variant == "scalar_wall":
        p["ivar"] = p["VarT"]
        p["SolveEnthalpy"] = 1
    _scalar_case(host, dev, precision, K, p, seed=7 + K)


@pytest.mark.gpu
def test_batched_rejects_bad_k():
    host, dev = _mesh("ref5", 64)
    z = torch.zeros(8, dtype=torch.float64, device="cuda")
    with pytest.raises(ch.CupCfdError):
        R.gradient_batched(dev, 1, 0, z, z, z, z)


@pytest.mark.gpu
@pytest.mark.parametrize("meshname", ["ref5", "typed_perm", "tets_typed"])
@pytest.mark.parametrize("K", [3, 16, 64])
@pytest.mark.parametrize("nG", [1, 3])
def test_gradient_batched_f32_storage_is_rounded_f64(K, meshname, nG):
    """FP32 storage, FP64 arithmetic: the float result is the FP64 result on the widened inputs, rounded."""
    host, dev = _mesh(meshname, 64)
    C, B = host.nCells, host.nBoundaries
    rng = np.random.default_rng(300 + K)
    phi = torch.from_numpy(rng.uniform(-1, 1, (C, K)).astype(np.float32)).cuda()
    phiB = torch.from_numpy(rng.uniform(-1, 1, (B, K)).astype(np.float32)).cuda()
    g32 = torch.zeros((C, 3, K), dtype=torch.float32, device="cuda")
    go32 = torch.zeros_like(g32)
    R.gradient_batched_f32s(dev, nG, K, phi, phiB, g32, go32)
    g64 = torch.zeros((C, 3, K), dtype=torch.float64, device="cuda")
    go64 = torch.zeros_like(g64)
    R.gradient_batched(dev, nG, K, phi.double(), phiB.double(), g64, go64)
    torch.cuda.synchronize()
    assert _same(g32, g64.to(torch.float32))
    assert _same(go32, go64.to(torch.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("meshname", ["ref5", "typed_perm", "tets_typed", "tets_plain", "typed32"])
@pytest.mark.parametrize("K", [3, 16, 64, 70])
@pytest.mark.parametrize("variant", ["scalar", "scalar_wall"])
def test_scalar_batched_f32_storage_is_rounded_f64(K, meshname, variant):
    """FP32 storage of the per-scalar arrays, FP64 everything else: Su, Au and PhiB are the FP64 batched
    results on the widened inputs, each rounded to float once; RFace and the mesh state are the FP64 call's."""
    if meshname == "typed32" and K > 16:
        pytest.skip("covered by the smaller meshes")
    host, dev = _mesh(meshname, 64)
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    p = dict(W.SCALAR)
    if variant == "scalar_wall":
        p["ivar"] = p["VarT"]
        p["SolveEnthalpy"] = 1
    rng = np.random.default_rng(500 + K)
    f = W.Fields(host, seed=9)
    per = dict(phi=rng.uniform(-1, 1, (C, K)), phiB=rng.uniform(-1, 1, (B, K)), au=rng.uniform(0.5, 1.5, (C, K)),
               su=rng.uniform(-1, 1, (C, K)), dphidx=rng.uniform(-1, 1, (C, 3, K)))
    sh = {k: _dev(v, np.float64) for k, v in
          dict(visEff=f.visEff, den=f.den, mf=f.massFlux, te=f.te, cpB=f.cpB, visEffB=f.visEffB).items()}
    st0 = seed_state(host)
    outs = []
    for storage in (np.float32, np.float64):
        host.set_state(st0, 64)
        dev.push_state()
        # the FP64 call reads the float inputs widened
        a = {k: _dev(v.astype(np.float32).astype(storage), storage) for k, v in per.items()}
        rface = torch.zeros(2 * F, dtype=torch.float64, device="cuda")
        call = R.scalar_batched_f32s if storage == np.float32 else R.scalar_batched
        call(dev, K, a["phi"], a["phiB"], sh["visEff"], a["au"], a["su"], sh["den"], sh["mf"], sh["te"], sh["cpB"],
             sh["visEffB"], rface, a["dphidx"], p)
        torch.cuda.synchronize()
        dev.pull_state()
        outs.append((a, rface, host.state(64)))
    (a32, r32, st32), (a64, r64, st64) = outs
    for n in ("su", "au", "phiB"):
        assert _same(a32[n], a64[n].to(torch.float32)), n
    assert _same(r32, r64)
    for n in st32:
        assert np.array_equal(st32[n], st64[n]), n


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("K", [3, 16, 64, 70])
def test_gradient_batched_meshset_matches_per_mesh(K, precision):
    """One persistent launch over a set of uneven meshes (hex and tet rows mixed) gives, per member, the bits of
    the single-mesh batched gradient."""
    specs = [dict(nx=9, ny=8, nz=8, seed=1, permute_cells=True, permute_faces=True),
             dict(nx=3, ny=3, nz=3),
             dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, seed=2, permute_cells=True, permute_faces=True),
             dict(nx=17, ny=9, nz=8, seed=3, permute_cells=True),
             dict(nx=8, ny=8, nz=8, typed=True, seed=4)]
    hosts = [ch.HostMesh(**sp) for sp in specs]
    devs = [ch.DeviceMesh(h, precision=precision, reorder=True) for h in hosts]
    dt = torch.float64 if precision == 64 else torch.float32
    rng = np.random.default_rng(K)
    phi = [torch.from_numpy(rng.uniform(-1, 1, (h.nCells, K))).to(dt).cuda() for h in hosts]
    phiB = [torch.from_numpy(rng.uniform(-1, 1, (h.nBoundaries, K))).to(dt).cuda() for h in hosts]
    g = [torch.zeros((h.nCells, 3, K), dtype=dt, device="cuda") for h in hosts]
    go = [torch.zeros_like(x) for x in g]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device="cuda")  # noqa: E731
    mset = ch.DeviceMeshSet(devs)
    R.gradient_batched_set(mset, K, ptr(phi), ptr(phiB), ptr(g), ptr(go))
    for i, d in enumerate(devs):
        g1 = torch.zeros_like(g[i])
        go1 = torch.zeros_like(g[i])
        R.gradient_batched(d, 1, K, phi[i], phiB[i], g1, go1)
        torch.cuda.synchronize()
        assert _same(g[i], g1) and _same(go[i], go1), i
    del mset


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("meshname", ["typed_perm", "tets_typed", "tets_plain"])
@pytest.mark.parametrize("variant", ["scalar", "scalar_wall"])
def test_batched_every_scalar_matches_oracle(meshname, precision, variant):
    """Every scalar k of the batched gradient (nGradient 1 and 3) and batched scalar flux against a direct
    oracle call on scalar k's arrays (mesh numbering, mesh state reset to the same seed each call).
    Bit-identical, except the WALL-enthalpy variant's log-based wall function (1e-12 / 2e-6)."""
    import pyoracle as po
    from kernel_runs import compare

    K = 5
    host, dev = _mesh(meshname, precision)
    dt = np.float64 if precision == 64 else np.float32
    co, fo = dev.cell_order, dev.face_order
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    rng = np.random.default_rng(5000 + precision)
    per = dict(phi=rng.uniform(-1, 1, (C, K)), phiB=rng.uniform(-1, 1, (B, K)), au=rng.uniform(0.5, 1.5, (C, K)),
               su=rng.uniform(-1, 1, (C, K)), dphidx=rng.uniform(-1, 1, (C, 3, K)))
    per = {k: v.astype(dt) for k, v in per.items()}
    f = W.Fields(host, seed=77).cast(dt)
    p = dict(W.SCALAR)
    if variant == "scalar_wall":
        p["ivar"] = p["VarT"]
        p["SolveEnthalpy"] = 1
    st0 = {k: v.astype(dt) for k, v in seed_state(host).items()}
    om = po.OracleMesh(host, precision)
    tol = 1e-12 if precision == 64 else 2e-6

    def check(got, want, what):
        same, rel = compare(got, want)
        if variant == "scalar":
            assert np.array_equal(np.ascontiguousarray(got).view(np.uint8),
                                  np.ascontiguousarray(want).view(np.uint8)), f"{what} (max rel {rel:.3e})"
        assert rel <= tol, f"{what}: max rel {rel:.3e}"

    # batched gradient against the oracle, every scalar
    for nG in (1, 3):
        g = torch.zeros((C, 3, K), dtype=torch.float64 if precision == 64 else torch.float32, device="cuda")
        go = torch.zeros_like(g)
        R.gradient_batched(dev, nG, K, _dev(per["phi"][co], dt), _dev(per["phiB"], dt), g, go)
        g_h, go_h = g.cpu().numpy(), go.cpu().numpy()
        for k in range(K):
            g1, go1 = np.zeros((C, 3), dt), np.zeros((C, 3), dt)
            assert po.GradientPhiGaussDolfyn(om, nG, per["phi"][:, k].copy(), per["phiB"][:, k].copy(), g1,
                                             go1) == ch.E_SUCCESS
            assert np.array_equal(g_h[:, :, k], g1[co]), (nG, k)
            assert np.array_equal(go_h[:, :, k], go1[co]), (nG, k)
    # batched scalar flux against the oracle, every scalar
    host.set_state(st0, precision)
    dev.push_state()
    bt = dict(phi=_dev(per["phi"][co], dt), phiB=_dev(per["phiB"], dt), au=_dev(per["au"][co], dt),
              su=_dev(per["su"][co], dt), dphidx=_dev(per["dphidx"][co], dt))
    sh = dict(visEff=_dev(f.visEff[co], dt), den=_dev(f.den[co], dt), mf=_dev(f.massFlux[fo], dt),
              te=_dev(f.te[co], dt), cpB=_dev(f.cpB, dt), visEffB=_dev(f.visEffB, dt))
    rface = torch.zeros(2 * F, dtype=bt["phi"].dtype, device="cuda")
    R.scalar_batched(dev, K, bt["phi"], bt["phiB"], sh["visEff"], bt["au"], bt["su"], sh["den"], sh["mf"], sh["te"],
                     sh["cpB"], sh["visEffB"], rface, bt["dphidx"], p)
    torch.cuda.synchronize()
    su_h, au_h, pb_h = bt["su"].cpu().numpy(), bt["au"].cpu().numpy(), bt["phiB"].cpu().numpy()
    rface_h = rface.cpu().numpy().reshape(F, 2)
    for k in range(K):
        om.set_state(st0)
        s1, a1, pb1 = per["su"][:, k].copy(), per["au"][:, k].copy(), per["phiB"][:, k].copy()
        rf1 = np.zeros(2 * F, dt)
        rc = po.FluxScalarDolfynFaceLoop(om, per["phi"][:, k].copy(), pb1, f.visEff.copy(), a1, s1, f.den.copy(),
                                         f.massFlux.copy(), f.te.copy(), f.cpB.copy(), f.visEffB.copy(), rf1,
                                         np.ascontiguousarray(per["dphidx"][:, :, k]), p["SolveTurb"],
                                         p["SolveEnthalpy"], p["vislam"], p["ivar"], p["VarT"], p["Sigma_T"],
                                         p["Prandtl"], p["VarTE"], p["Sigma_k"], p["VarED"], p["Sigma_e"],
                                         p["Sigma_s"], p["Schmidt"], p["GammaBlend"], p["Small"], p["Large"],
                                         p["TMCmu"])
        assert rc == ch.E_SUCCESS
        check(su_h[:, k], s1[co], f"Su scalar {k}")
        check(au_h[:, k], a1[co], f"Au scalar {k}")
        check(pb_h[:, k], pb1, f"PhiB scalar {k}")
        check(rface_h, rf1.reshape(F, 2)[fo], f"RFace scalar {k}")

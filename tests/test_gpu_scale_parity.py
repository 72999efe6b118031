"""Parity with the oracle at the workloads' full sizes (GPU).

- StructGen 128^3, cell and face ids relabelled (the sweep bench.py times):
  gradient, mass flux, scalar flux and UVW through the drop-in C ABI against
  the oracle on identical inputs, every output bit-identical (FP64, no
  transcendental branch on this one-DEFAULT-region mesh).
- 100^3 hexes x 6 jittered tets, relabelled: the same four kernels, bit-identical.
- 96^3 relabelled, K = 64 scalars: the batched gradient and batched scalar flux
  (resident, internal numbering), scalars 0, 31 and 63 against direct oracle
  calls and every scalar against its single-scalar call, bit-identical.
- the persistent mesh-set launch captured in a hipGraph: replays equal the
  eager launch and the per-mesh batched calls, bit for bit.

The oracle and the HIP path start the carried face values (PhiFace,
uFace/vFace/wFace) at the same defined value on meshes without INLET/OUTLET
faces, where the reference reads them uninitialised.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import fvm
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R
from kernel_runs import compare, run

torch = pytest.importorskip("torch")


def _bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _sweep_vs_oracle(host, seed):
    dev = ch.DeviceMesh(host, 64, reorder=True)
    om = po.OracleMesh(host, 64)
    f = W.Fields(host, seed=seed)
    for kern in ("gradient", "mass", "scalar", "uvw"):
        print(f"[scale parity] {kern} on {host.nCells} cells", flush=True)
        rc_o, out_o = run(kern, po, om, f.copy(), lambda: {})
        rc_g, out_g = run(kern, fvm, dev, f.copy(), lambda: {})
        assert rc_o == rc_g == ch.E_SUCCESS
        for k in out_o:
            same, rel = compare(out_g[k], out_o[k])
            assert _bits_equal(out_g[k], out_o[k]), f"{kern}: {k} not bit-identical (max rel {rel:.3e})"


@pytest.mark.gpu
def test_sweep_128_relabelled_bit_identical_to_oracle():
    host = ch.HostMesh(128, seed=12345, permute_cells=True, permute_faces=True)
    _sweep_vs_oracle(host, seed=777)


@pytest.mark.gpu
def test_tets_100_bit_identical_to_oracle():
    host = ch.HostMesh(100, seed=606, tets=True, jitter=0.2, permute_cells=True, permute_faces=True)
    _sweep_vs_oracle(host, seed=778)


def _cells(a, co, dt):
    return torch.from_numpy(np.ascontiguousarray(a[co], dtype=dt)).cuda()


@pytest.mark.gpu
def test_batched_96_k64_matches_oracle_and_single_calls():
    K = 64
    host = ch.HostMesh(96, seed=4242, permute_cells=True, permute_faces=True)
    dev = ch.DeviceMesh(host, 64, reorder=True)
    co, fo = dev.cell_order, dev.face_order
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    rng = np.random.default_rng(96)
    phi = rng.uniform(-1, 1, (C, K))
    phiB = rng.uniform(-1, 1, (B, K))
    f = W.Fields(host, seed=961)
    au = rng.uniform(0.5, 1.5, (C, K))
    su = rng.uniform(-1, 1, (C, K))
    dphidx = rng.uniform(-1, 1, (C, 3, K))
    dt = np.float64
    p = W.SCALAR
    # batched gradient and scalar flux, internal numbering
    g = torch.zeros((C, 3, K), dtype=torch.float64, device="cuda")
    go = torch.zeros_like(g)
    R.gradient_batched(dev, 1, K, _cells(phi, co, dt), torch.from_numpy(phiB).cuda(), g, go)
    sh = dict(visEff=_cells(f.visEff, co, dt), den=_cells(f.den, co, dt), te=_cells(f.te, co, dt),
              mf=torch.from_numpy(np.ascontiguousarray(f.massFlux[fo])).cuda(),
              cpB=torch.from_numpy(f.cpB).cuda(), visEffB=torch.from_numpy(f.visEffB).cuda())
    bt = dict(phi=_cells(phi, co, dt), phiB=torch.from_numpy(phiB.copy()).cuda(), au=_cells(au, co, dt),
              su=_cells(su, co, dt), dphidx=_cells(dphidx, co, dt))
    rface = torch.zeros(2 * F, dtype=torch.float64, device="cuda")
    R.scalar_batched(dev, K, bt["phi"], bt["phiB"], sh["visEff"], bt["au"], bt["su"], sh["den"], sh["mf"], sh["te"],
                     sh["cpB"], sh["visEffB"], rface, bt["dphidx"], p)
    torch.cuda.synchronize()
    g_h, go_h = g.cpu().numpy(), go.cpu().numpy()
    su_h, au_h = bt["su"].cpu().numpy(), bt["au"].cpu().numpy()
    rface_h = rface.cpu().numpy().reshape(F, 2)
    # scalars 0, 31, 63 against the oracle (mesh numbering)
    om = po.OracleMesh(host, 64)
    for k in (0, 31, 63):
        print(f"[scale parity] batched scalar {k} vs oracle", flush=True)
        g1, go1 = np.zeros((C, 3)), np.zeros((C, 3))
        assert po.GradientPhiGaussDolfyn(om, 1, phi[:, k].copy(), phiB[:, k].copy(), g1, go1) == ch.E_SUCCESS
        assert _bits_equal(g_h[:, :, k], g1[co]), f"gradient scalar {k}"
        assert _bits_equal(go_h[:, :, k], go1[co]), f"gradient (dPhidxo) scalar {k}"
        s1, a1, pb1 = su[:, k].copy(), au[:, k].copy(), phiB[:, k].copy()
        rf1 = np.zeros(2 * F)
        rc_o = po.FluxScalarDolfynFaceLoop(om, phi[:, k].copy(), pb1, f.visEff.copy(), a1, s1, f.den.copy(),
                                           f.massFlux.copy(), f.te.copy(), f.cpB.copy(), f.visEffB.copy(), rf1,
                                           np.ascontiguousarray(dphidx[:, :, k]), p["SolveTurb"], p["SolveEnthalpy"],
                                           p["vislam"], p["ivar"], p["VarT"], p["Sigma_T"], p["Prandtl"], p["VarTE"],
                                           p["Sigma_k"], p["VarED"], p["Sigma_e"], p["Sigma_s"], p["Schmidt"],
                                           p["GammaBlend"], p["Small"], p["Large"], p["TMCmu"])
        assert rc_o == ch.E_SUCCESS
        assert _bits_equal(su_h[:, k], s1[co]), f"Su scalar {k}"
        assert _bits_equal(au_h[:, k], a1[co]), f"Au scalar {k}"
        assert _bits_equal(bt["phiB"].cpu().numpy()[:, k], pb1), f"PhiB scalar {k}"
        assert _bits_equal(rface_h, rf1.reshape(F, 2)[fo]), "RFace"
    # every scalar against its single-scalar resident call
    print("[scale parity] batched vs 64 single-scalar calls", flush=True)
    for k in range(K):
        g1 = torch.zeros((C, 3), dtype=torch.float64, device="cuda")
        go1 = torch.zeros_like(g1)
        R.gradient(dev, 1, bt["phi"][:, k].contiguous(), torch.from_numpy(phiB[:, k].copy()).cuda(), g1, go1)
        one_su, one_au = _cells(su[:, k], co, dt), _cells(au[:, k], co, dt)
        one_pb = torch.from_numpy(phiB[:, k].copy()).cuda()
        rf = torch.zeros(2 * F, dtype=torch.float64, device="cuda")
        R.scalar(dev, bt["phi"][:, k].contiguous(), one_pb, sh["visEff"], one_au, one_su, sh["den"], sh["mf"],
                 sh["te"], sh["cpB"], sh["visEffB"], rf, _cells(dphidx[:, :, k], co, dt), p)
        torch.cuda.synchronize()
        assert _bits_equal(g1.cpu().numpy(), g_h[:, :, k]), k
        assert _bits_equal(go1.cpu().numpy(), go_h[:, :, k]), k
        assert _bits_equal(one_su.cpu().numpy(), su_h[:, k]), k
        assert _bits_equal(one_au.cpu().numpy(), au_h[:, k]), k


@pytest.mark.gpu
def test_meshset_persistent_launch_in_hipgraph():
    """The persistent mesh-set gradient captured once and replayed: each replay (with eager work and
    synchronisations between replays) equals the eager launch and the per-mesh batched calls."""
    K = 16
    specs = [dict(nx=12, ny=10, nz=9, seed=1, permute_cells=True, permute_faces=True),
             dict(nx=5, ny=4, nz=3),
             dict(nx=5, ny=4, nz=4, tets=True, jitter=0.2, seed=2, permute_cells=True, permute_faces=True),
             dict(nx=20, ny=9, nz=8, seed=3, permute_cells=True)]
    hosts = [ch.HostMesh(**sp) for sp in specs]
    devs = [ch.DeviceMesh(h, 64, reorder=True) for h in hosts]
    rng = np.random.default_rng(5)
    phi = [torch.from_numpy(rng.uniform(-1, 1, (h.nCells, K))).cuda() for h in hosts]
    phiB = [torch.from_numpy(rng.uniform(-1, 1, (h.nBoundaries, K))).cuda() for h in hosts]
    g = [torch.zeros((h.nCells, 3, K), dtype=torch.float64, device="cuda") for h in hosts]
    go = [torch.zeros_like(x) for x in g]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device="cuda")  # noqa: E731
    pp, pb, pg, pgo = ptr(phi), ptr(phiB), ptr(g), ptr(go)
    mset = ch.DeviceMeshSet(devs)
    R.gradient_batched_set(mset, K, pp, pb, pg, pgo)  # eager: creates any lazily allocated state
    torch.cuda.synchronize()
    eager = [(x.clone(), y.clone()) for x, y in zip(g, go)]
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            R.gradient_batched_set(mset, K, pp, pb, pg, pgo, stream=stream)
    torch.cuda.synchronize()
    for rep in range(3):
        for x, y in zip(g, go):
            x.fill_(float(rep + 1))
            y.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i, ((e, eo), x, y) in enumerate(zip(eager, g, go)):
            assert torch.equal(e, x) and torch.equal(eo, y), (rep, i)
    for i, d in enumerate(devs):
        g1, go1 = torch.zeros_like(g[i]), torch.zeros_like(g[i])
        R.gradient_batched(d, 1, K, phi[i], phiB[i], g1, go1)
        torch.cuda.synchronize()
        assert torch.equal(g1, g[i]) and torch.equal(go1, go[i]), i
    del graph, mset


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_dropin_large_arrays_bit_identical_to_oracle(precision):
    """Drop-in calls whose host arrays are large enough to go through the pinned staging (2 MB and more, in
    8 MB chunks with a partial last chunk): a typed relabelled 61 x 63 x 65 mesh, FP64 and FP32, the four
    face loops against the oracle in the same precision, every output and every piece of mesh state
    bit-identical."""
    from kernel_runs import seed_state

    host = ch.HostMesh(61, ny=63, nz=65, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=21)
    dt = np.float64 if precision == 64 else np.float32
    dev = ch.DeviceMesh(host, precision, reorder=True)
    om = po.OracleMesh(host, precision)
    f = W.Fields(host, seed=33).cast(dt)
    st0 = {k: v.astype(dt) for k, v in seed_state(host).items()}
    for kern in ("gradient", "mass", "scalar", "uvw"):
        print(f"[drop-in large] {kern} fp{precision}", flush=True)
        om.set_state(st0)
        rc_o, out_o = run(kern, po, om, f.copy(), om.state)
        host.set_state(st0, precision)
        rc_g, out_g = run(kern, fvm, dev, f.copy(), lambda: host.state(precision))
        assert rc_o == rc_g == ch.E_SUCCESS
        for k in out_o:
            assert _bits_equal(out_g[k], out_o[k]), f"{kern}/{precision}: {k} (max rel {compare(out_g[k], out_o[k])[1]:.3e})"

"""Batched kernels on a mesh whose rows have different lengths.

Uniform meshes (all-hex, all-tet) take the row-preloading paths of the batched
gradient, scalar flux and adjoint; this mesh (the reference's 5^3 unit mesh
with three boundary faces removed, so three cells have 5 faces) takes their
row-by-row paths. The contract is the same: scalar k's results are, bit for
bit, those of the single-scalar resident call on scalar k's arrays.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R
from mixed_mesh import device_mesh, drop_boundary_faces

torch = pytest.importorskip("torch")

_cache = {}


def _mesh():
    if "m" not in _cache:
        host = ch.HostMesh(5, bounds=(-1, 1, -1, 1, -1, 1))
        view, arrays = drop_boundary_faces(host, count=3, seed=1)
        _cache["m"] = device_mesh(view, arrays, host)
    return _cache["m"]


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("K", [3, 16, 64, 70])
@pytest.mark.parametrize("nG", [1, 3])
def test_gradient_batched_mixed_rows(K, nG):
    dev = _mesh()
    C, B = dev.nCells, dev.nBoundaries
    rng = np.random.default_rng(K + 10 * nG)
    phi, phiB = rng.uniform(-1, 1, (C, K)), rng.uniform(-1, 1, (B, K))
    g = torch.zeros((C, 3, K), dtype=torch.float64, device="cuda")
    go = torch.zeros_like(g)
    R.gradient_batched(dev, nG, K, _cuda(phi), _cuda(phiB), g, go)
    for k in range(K):
        g1 = torch.zeros((C, 3), dtype=torch.float64, device="cuda")
        go1 = torch.zeros_like(g1)
        R.gradient(dev, nG, _cuda(phi[:, k]), _cuda(phiB[:, k]), g1, go1)
        torch.cuda.synchronize()
        assert torch.equal(g[:, :, k], g1) and torch.equal(go[:, :, k], go1), k


@pytest.mark.gpu
@pytest.mark.parametrize("K", [3, 16, 64, 70])
def test_scalar_batched_mixed_rows(K):
    dev = _mesh()
    C, F, B = dev.nCells, dev.nFaces, dev.nBoundaries
    rng = np.random.default_rng(100 + K)
    per = dict(phi=rng.uniform(-1, 1, (C, K)), phiB=rng.uniform(-1, 1, (B, K)), au=rng.uniform(0.5, 1.5, (C, K)),
               su=rng.uniform(-1, 1, (C, K)), dphidx=rng.uniform(-1, 1, (C, 3, K)))
    sh = {k: _cuda(v) for k, v in dict(visEff=rng.uniform(1e-3, 2e-3, C), den=rng.uniform(0.9, 1.1, C),
                                       mf=rng.uniform(-1, 1, F), te=rng.uniform(0.1, 1, C),
                                       cpB=rng.uniform(0.9, 1.1, B), visEffB=rng.uniform(1e-3, 2e-3, B)).items()}
    p = W.SCALAR
    bt = {k: _cuda(v) for k, v in per.items()}
    rface_b = torch.zeros(2 * F, dtype=torch.float64, device="cuda")
    R.scalar_batched(dev, K, bt["phi"], bt["phiB"], sh["visEff"], bt["au"], bt["su"], sh["den"], sh["mf"], sh["te"],
                     sh["cpB"], sh["visEffB"], rface_b, bt["dphidx"], p)
    for k in range(K):
        one = {n: _cuda(v[..., k]) for n, v in per.items()}
        rface = torch.zeros(2 * F, dtype=torch.float64, device="cuda")
        R.scalar(dev, one["phi"], one["phiB"], sh["visEff"], one["au"], one["su"], sh["den"], sh["mf"], sh["te"],
                 sh["cpB"], sh["visEffB"], rface, one["dphidx"], p)
        torch.cuda.synchronize()
        for n in ("su", "au", "phiB"):
            assert torch.equal(bt[n][..., k], one[n]), (n, k)
        assert torch.equal(rface_b, rface), k


@pytest.mark.gpu
@pytest.mark.parametrize("K", [3, 16, 70])
def test_adjoint_batched_mixed_rows(K):
    dev = _mesh()
    C, F = dev.nCells, dev.nFaces
    rng = np.random.default_rng(200 + K)
    phi, sb, rb = rng.uniform(-1, 1, (C, K)), rng.normal(size=(C, K)), rng.normal(size=2 * F)
    mf, vis = _cuda(rng.uniform(-1, 1, F)), _cuda(rng.uniform(1e-3, 2e-3, C))
    P = W.SCALAR
    pb = torch.empty((C, K), dtype=torch.float64, device="cuda")
    mb = torch.empty(F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint_batched(dev, K, _cuda(phi), vis, mf, _cuda(sb), _cuda(rb), P, pb, mb)
    zero_c = torch.zeros(C, dtype=torch.float64, device="cuda")
    zero_r = torch.zeros(2 * F, dtype=torch.float64, device="cuda")
    p1 = torch.empty(C, dtype=torch.float64, device="cuda")
    m1 = torch.empty(F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint(dev, _cuda(phi[:, 0]), vis, mf, zero_c, _cuda(rb), P, p1, m1)
    want = m1.cpu().numpy().copy()
    for k in range(K):
        R.scalar_adjoint(dev, _cuda(phi[:, k]), vis, mf, _cuda(sb[:, k]), zero_r, P, p1, m1)
        torch.cuda.synchronize()
        assert torch.equal(pb[:, k], p1), k
        want = want + m1.cpu().numpy()
    assert np.array_equal(mb.cpu().numpy(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("K", [3, 70])
def test_adjoint_batched_f32_storage_mixed_rows(K):
    dev = _mesh()
    C, F = dev.nCells, dev.nFaces
    rng = np.random.default_rng(300 + K)
    phi = rng.uniform(-1, 1, (C, K)).astype(np.float32)
    sb = rng.normal(size=(C, K)).astype(np.float32)
    rb, mf, vis = _cuda(rng.normal(size=2 * F)), _cuda(rng.uniform(-1, 1, F)), _cuda(rng.uniform(1e-3, 2e-3, C))
    pb32 = torch.empty((C, K), dtype=torch.float32, device="cuda")
    pb64 = torch.empty((C, K), dtype=torch.float64, device="cuda")
    mb32 = torch.empty(F, dtype=torch.float64, device="cuda")
    mb64 = torch.empty_like(mb32)
    R.scalar_adjoint_batched_f32s(dev, K, _cuda(phi), vis, mf, _cuda(sb), rb, W.SCALAR, pb32, mb32)
    R.scalar_adjoint_batched(dev, K, _cuda(phi.astype(np.float64)), vis, mf, _cuda(sb.astype(np.float64)), rb,
                             W.SCALAR, pb64, mb64)
    torch.cuda.synchronize()
    assert torch.equal(pb32, pb64.to(torch.float32)) and torch.equal(mb32, mb64)

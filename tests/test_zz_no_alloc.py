"""The allocation-free adjoint entry point (cupcfd_dev_scalar_adjoint_noalloc_f64).

Through the C ABI, its PhiBar and MassFluxBar equal cupcfd_dev_scalar_adjoint_f64's
bit for bit on two calls in a row, the second on a smaller mesh (the fixed
scratch keeps the first call's contents, which the second must not read);
a mesh above the scratch's 2,097,152-cell limit is refused with
E_CONFIG_INVALID_VALUE before any launch, leaving the outputs untouched.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R

torch = pytest.importorskip("torch")
P = dict(W.SCALAR)


def _inputs(host, dev, seed):
    rng = np.random.default_rng(seed)
    co, fo = dev.cell_order, dev.face_order
    C, F = host.nCells, host.nFaces
    f = W.Fields(host, seed=seed)
    cells = lambda a: torch.from_numpy(np.ascontiguousarray(a[co])).cuda()  # noqa: E731
    faces = lambda a, k=1: torch.from_numpy(np.ascontiguousarray(a.reshape(F, k)[fo].ravel())).cuda()  # noqa: E731
    return dict(phi=cells(f.phi), visEff=cells(f.visEff), mf=faces(f.massFlux), subar=cells(rng.normal(size=C)),
                rbar=faces(rng.normal(size=(F, 2)), 2))


def _adjoint(dev, x, noalloc):
    C = dev.nCells
    pb = torch.full((C,), np.nan, dtype=torch.float64, device="cuda")
    mb = torch.full((dev.nFaces,), np.nan, dtype=torch.float64, device="cuda")
    R.scalar_adjoint(dev, x["phi"], x["visEff"], x["mf"], x["subar"], x["rbar"], P, pb, mb, noalloc=noalloc)
    torch.cuda.synchronize()
    return pb.cpu().numpy(), mb.cpu().numpy()


def _same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.gpu
def test_noalloc_adjoint_matches_allocating_entry_point():
    for n, seed in ((64, 1), (20, 2)):  # second call on a smaller mesh
        host = ch.HostMesh(n, n - 1, n - 2, seed=seed, permute_cells=True, permute_faces=True)
        dev = ch.DeviceMesh(host, 64, reorder=True)
        x = _inputs(host, dev, seed)
        pa, ma = _adjoint(dev, x, noalloc=False)
        pn, mn = _adjoint(dev, x, noalloc=True)
        assert np.isfinite(pa).all() and np.isfinite(ma).all()
        assert _same(pn, pa), f"PhiBar differs on {host.nCells} cells"
        assert _same(mn, ma), f"MassFluxBar differs on {host.nCells} cells"


@pytest.mark.gpu
def test_noalloc_adjoint_refuses_meshes_above_its_scratch():
    host = ch.HostMesh(129, 128, 128)  # 2,113,536 cells: 32 B each exceeds 64 MiB
    dev = ch.DeviceMesh(host, 64, reorder=True)
    x = _inputs(host, dev, 3)
    pb = torch.full((dev.nCells,), 7.0, dtype=torch.float64, device="cuda")
    mb = torch.full((dev.nFaces,), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(ch.CupCfdError) as e:
        R.scalar_adjoint(dev, x["phi"], x["visEff"], x["mf"], x["subar"], x["rbar"], P, pb, mb, noalloc=True)
    assert e.value.code == 90  # E_CONFIG_INVALID_VALUE
    torch.cuda.synchronize()
    assert bool((pb == 7.0).all()) and bool((mb == 7.0).all())

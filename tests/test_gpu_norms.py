"""Residual norms of cell fields (cupcfd_dev_residual_norms_*), against numpy on oracle outputs.

The fields are the oracle's Su (scalar loop) and su / sv / sw (UVW loop) on a
typed relabelled 32^3 mesh, in the device mesh's internal order. max |x| must
equal numpy's exactly; sqrt(sum x^2) must agree with an exactly rounded sum
(math.fsum) to 1e-14 relative. The same call twice gives the same bits, a
captured hipGraph replays it exactly, NaN propagates, FP32 storage is
accumulated in double.
"""
import math

import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R
from kernel_runs import run

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def setup():
    host = ch.HostMesh(32, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=3)
    dev = ch.DeviceMesh(host, 64, reorder=True)
    om = po.OracleMesh(host, 64)
    f = W.Fields(host, seed=41)
    _, s_out = run("scalar", po, om, f.copy(), lambda: {})
    _, u_out = run("uvw", po, om, f.copy(), lambda: {})
    fields = [s_out["Su"], u_out["su"], u_out["sv"], u_out["sw"]]
    return host, dev, fields


def _expect(x):
    x = np.asarray(x, dtype=np.float64)
    return float(np.abs(x).max()), math.sqrt(math.fsum(float(v) * float(v) for v in x))


def _check(got, fields):
    for i, x in enumerate(fields):
        mx, l2 = _expect(x)
        assert got[i, 0] == mx, (i, got[i, 0], mx)
        assert abs(got[i, 1] - l2) <= 1e-14 * l2, (i, got[i, 1], l2)


@pytest.mark.gpu
def test_norms_of_oracle_residuals(setup):
    host, dev, fields = setup
    co = dev.cell_order
    t = [torch.from_numpy(np.ascontiguousarray(x[co])).cuda() for x in fields]
    n1 = R.residual_norms(dev, t).cpu().numpy()
    _check(n1, fields)
    n2 = R.residual_norms(dev, t).cpu().numpy()
    assert np.array_equal(n1.view(np.uint64), n2.view(np.uint64))
    # order of the fields does not change a field's norms
    n3 = R.residual_norms(dev, t[::-1]).cpu().numpy()
    assert np.array_equal(n3[::-1].view(np.uint64), n1.view(np.uint64))


@pytest.mark.gpu
def test_norms_fp32_storage_accumulate_in_double(setup):
    host, dev, fields = setup
    co = dev.cell_order
    f32 = [np.ascontiguousarray(x[co]).astype(np.float32) for x in fields]
    got = R.residual_norms(dev, [torch.from_numpy(x).cuda() for x in f32]).cpu().numpy()
    _check(got, f32)


@pytest.mark.gpu
def test_norms_nan_propagates_and_bad_counts_refused(setup):
    host, dev, fields = setup
    x = torch.from_numpy(np.ascontiguousarray(fields[0])).cuda()
    x[5] = float("nan")
    got = R.residual_norms(dev, [x]).cpu().numpy()
    assert np.isnan(got).all()
    L = ch.lib()
    out = torch.zeros(18, dtype=torch.float64, device="cuda")
    arr = (ch.ct.c_void_p * 9)(*([x.data_ptr()] * 9))
    for n in (0, 9):
        rc = L.cupcfd_dev_residual_norms_f64(dev.handle, n, arr, ch.ct.c_void_p(out.data_ptr()), None)
        assert rc == 90, n


@pytest.mark.gpu
def test_norms_in_hipgraph(setup):
    host, dev, fields = setup
    co = dev.cell_order
    t = [torch.from_numpy(np.ascontiguousarray(x[co])).cuda() for x in fields]
    out = torch.zeros((4, 2), dtype=torch.float64, device="cuda")
    R.residual_norms(dev, t, out)  # sizes the workspace
    torch.cuda.synchronize()
    eager = out.clone()
    stream = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        with torch.cuda.graph(g, stream=stream):
            R.residual_norms(dev, t, out, stream=stream)
    for _ in range(3):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
    del g

"""The boundary, region and viscosity loops through the resident C ABI (cupcfd_dev_*, torch device arrays in
the device mesh's internal numbering) against the same kernels through the drop-in ABI, whose parity with the
oracle tests/test_gpu_parity.py checks. The mesh state is pushed before the resident call and pulled after
it. Both paths must agree bit for bit, on every output and every piece of mesh state the kernel sets."""
import zlib

import numpy as np
import pytest

import cupcfd_hip as ch
from cupcfd_hip import fvm
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R
from kernel_runs import compare, run, seed_state

torch = pytest.importorskip("torch")

SMALL = ["mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5", "mass_region", "uvw_region1", "uvw_bnds1",
         "visc_cell1", "visc_region", "visc_boundary", "visc_cell2"]
MESHES = {
    "typed_perm": dict(nx=7, ny=6, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=11),
    "tets_typed": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, typed=True, permute_cells=True,
                       permute_faces=True, face_extras=True, seed=5),
}
_cache = {}


def _setup(meshname, precision):
    key = (meshname, precision)
    if key not in _cache:
        host = ch.HostMesh(**MESHES[meshname])
        _cache[key] = (host, ch.DeviceMesh(host, precision=precision, reorder=True))
    return _cache[key]


def _resident(kernel, dev, f, tdt):
    """Run `kernel` through the resident ABI on inputs f (mesh numbering); its outputs in mesh numbering."""
    co, fo = dev.cell_order, dev.face_order

    def cell(a):
        return torch.from_numpy(np.ascontiguousarray(a[co])).cuda()

    def face(a):
        return torch.from_numpy(np.ascontiguousarray(a[fo])).cuda()

    def plain(a):  # boundary and region arrays keep the mesh's numbering
        return torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def one(v):
        return torch.full((1,), v, dtype=tdt, device="cuda")

    def back_cell(t):
        out = np.empty(t.shape, dtype=t.cpu().numpy().dtype)
        out[co] = t.cpu().numpy()
        return out

    def back_face(t):
        out = np.empty(t.shape, dtype=t.cpu().numpy().dtype)
        out[fo] = t.cpu().numpy()
        return out

    out = {}
    if kernel == "mass_bl1":
        fl = one(0.0)
        R.mass_bl1(dev, face(f.massFlux), fl)
        out = dict(flowin=fl)
    elif kernel == "mass_bl2":
        fl, fr = one(0.0), plain(f.flowRegion)
        R.mass_bl2(dev, face(f.massFlux), fr, fl)
        out = dict(flowout=fl, flowRegion=fr)
    elif kernel == "mass_bl3":
        r = one(0.0)
        R.mass_bl3(dev, W.MASS_LOOPS["flowin"], r)
        out = dict(ratearea=r)
    elif kernel == "mass_bl4":
        fl, mf = one(0.0), face(f.massFlux)
        uB, vB, wB = plain(f.uB), plain(f.vB), plain(f.wB)
        R.mass_bl4(dev, mf, uB, vB, wB, plain(f.denB), W.MASS_LOOPS["ratearea"], fl)
        out = dict(massFlux=back_face(mf), uB=uB, vB=vB, wB=wB, flowout=fl)
    elif kernel == "mass_bl5":
        q = W.MASS_LOOPS
        fl, mf, su = one(0.25), face(f.massFlux), cell(f.su)
        uB, vB, wB = plain(f.uB), plain(f.vB), plain(f.wB)
        R.mass_bl5(dev, mf, su, uB, vB, wB, q["fact"], q["solveU"], q["solveV"], q["solveW"], plain(f.flowFact), fl)
        out = dict(massFlux=back_face(mf), su=back_cell(su), uB=uB, vB=vB, wB=wB, flowout2=fl)
    elif kernel == "mass_region":
        ff = plain(f.flowFact)
        R.mass_region(dev, ff, plain(f.flowRegion), 0.7)
        out = dict(flowFact=ff)
    elif kernel == "uvw_region1":
        R.uvw_region1(dev)
    elif kernel == "uvw_bnds1":
        R.uvw_bnds1(dev)
    elif kernel == "visc_cell1":
        q = W.VISC["cell1"]
        ve = cell(f.visEff)
        R.visc_cell1(dev, q["small"], q["cmu"], q["visURF"], q["visLam"], cell(f.te), cell(f.ed), cell(f.den), ve)
        out = dict(visEff=back_cell(ve))
    elif kernel == "visc_region":
        R.visc_region(dev, W.VISC["region"]["kappa"])
    elif kernel == "visc_boundary":
        q = W.VISC["boundary"]
        vb = plain(f.visEffB)
        R.visc_boundary(dev, q["tmcmu"], q["large"], q["small"], q["kappa"], q["visLam"], cell(f.te), cell(f.den),
                        cell(f.visEff), vb)
        out = dict(visEffB=vb)
    elif kernel == "visc_cell2":
        ve, vb = cell(f.visEff), plain(f.visEffB)
        R.visc_cell2(dev, W.VISC["cell2"]["visLam"], ve, vb)
        out = dict(visEff=back_cell(ve), visEffB=vb)
    else:
        raise KeyError(kernel)
    torch.cuda.synchronize()
    return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in out.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("meshname", list(MESHES))
@pytest.mark.parametrize("kernel", SMALL)
def test_resident_loop_matches_dropin(kernel, meshname, precision):
    host, dev = _setup(meshname, precision)
    dt = np.float64 if precision == 64 else np.float32
    tdt = torch.float64 if precision == 64 else torch.float32
    fields = W.Fields(host, seed=zlib.crc32(f"resident/{kernel}/{meshname}".encode()) % 1000)
    st0 = {k: v.astype(dt) for k, v in seed_state(host).items()}

    host.set_state(st0, precision)
    rc, want = run(kernel, fvm, dev, fields.cast(dt), lambda: host.state(precision))
    assert rc == ch.E_SUCCESS

    host.set_state(st0, precision)
    dev.push_state()
    got = _resident(kernel, dev, fields.cast(dt), tdt)
    dev.pull_state()
    for k, v in host.state(precision).items():
        got["mesh." + k] = v

    assert got.keys() == want.keys()
    for k in want:
        same, rel = compare(got[k], want[k])
        assert same, f"{kernel}/{meshname}/{precision}: {k} differs from the drop-in call (max rel {rel:.3e})"


_big = {}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("kernel", ["mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5", "uvw_bnds1"])
def test_ordered_sums_over_several_chunks(kernel, precision):
    """The boundary loops' ordered sums against the oracle on a typed 64^3 mesh, where every side has 4096
    boundaries: the flow sums run over two LDS chunks of 3072 values, and each WALL region's force tangent over
    four chunks of 1024. Bit for bit in FP64 and FP32."""
    import pyoracle as po

    if precision not in _big:
        host = ch.HostMesh(64, typed=True, permute_cells=True, permute_faces=True, seed=17)
        _big[precision] = (host, ch.DeviceMesh(host, precision=precision, reorder=True))
    host, dev = _big[precision]
    dt = np.float64 if precision == 64 else np.float32
    fields = W.Fields(host, seed=zlib.crc32(f"chunks/{kernel}".encode()) % 1000)
    st0 = {k: v.astype(dt) for k, v in seed_state(host).items()}
    om = po.OracleMesh(host, precision)
    om.set_state(st0)
    rc_o, want = run(kernel, po, om, fields.cast(dt), om.state)
    host.set_state(st0, precision)
    rc_g, got = run(kernel, fvm, dev, fields.cast(dt), lambda: host.state(precision))
    assert rc_o == ch.E_SUCCESS and rc_g == ch.E_SUCCESS
    for k in want:
        same, rel = compare(got[k], want[k])
        assert same, f"{kernel}/{precision}: {k} not bit-identical (max rel {rel:.3e})"

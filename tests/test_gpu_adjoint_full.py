"""cupcfd_dev_scalar_adjoint_full_f64 (GPU): the adjoint of gradient (nGradient sweeps) + scalar face loop
on any mesh, against the oracle's (or_scalar_adjoint_full, pinned on the CPU by central differences and the
dot-product identity in tests/test_oracle_adjoint_full.py), bit for bit.

Cases: typed relabelled hexes with non-zero xpac/xnac/rlencos (INLET, OUTLET, WALL and SYMP faces), typed
jittered tets, a DEFAULT relabelled hex mesh and a typed 24^3 mesh; the scalar loop with INLET assigning
PhiFace, with INLET carrying it (ivar = VarT) and with the enthalpy wall on flux and non-flux regions, the
mesh state seeded so both wall-function branches run; nGradient 1, 2 and 3; every lane schedule. With
nGradient = 1 on a DEFAULT mesh the result equals cupcfd_dev_scalar_adjoint_f64's.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R
from kernel_runs import seed_state

torch = pytest.importorskip("torch")

MESHES = {
    "typed_perm": dict(nx=7, ny=6, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=11),
    "tets_typed": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, typed=True, permute_cells=True, permute_faces=True,
                       face_extras=True, seed=5),
    "plain": dict(nx=9, ny=8, nz=7, seed=5, permute_cells=True, permute_faces=True),
    "typed24": dict(nx=24, typed=True, permute_cells=True, permute_faces=True, face_extras=True, seed=3),
}


def params(variant):
    p = dict(W.SCALAR)
    if variant in ("scalar_T", "scalar_wall"):
        p["ivar"] = p["VarT"]
    if variant == "scalar_wall":
        p["SolveEnthalpy"] = 1
    return p


_cases = {}


def case(name):
    if name not in _cases:
        host = ch.HostMesh(**MESHES[name])
        st = seed_state(host)
        host.set_state(st)
        dev = ch.DeviceMesh(host, 64, reorder=True)
        om = po.OracleMesh(host, 64)
        om.set_state(st)
        _cases[name] = (host, dev, om, W.Fields(host, seed=23))
    return _cases[name]


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def run(name, variant, nG, seed=0):
    host, dev, om, f = case(name)
    p = params(variant)
    C, F = host.nCells, host.nFaces
    co, fo = dev.cell_order, dev.face_order
    rng = np.random.default_rng(1000 + seed)
    sb = rng.normal(size=C)
    rb = rng.normal(size=2 * F)
    g = np.zeros((C, 3))
    go = np.zeros((C, 3))
    assert po.GradientPhiGaussDolfyn(om, nG, f.phi.copy(), f.phiB.copy(), g, go) == 95
    rc, pb_o, mb_o = po.scalar_adjoint_full(om, nG, f.phi, f.visEff, f.massFlux, g, f.te, f.den, f.cpB, f.visEffB,
                                            sb, rb, p)
    assert rc == 95
    pb = torch.full((C,), np.nan, dtype=torch.float64, device="cuda")
    mb = torch.full((F,), np.nan, dtype=torch.float64, device="cuda")
    R.scalar_adjoint_full(dev, nG, _cuda(f.phi[co]), _cuda(f.visEff[co]), _cuda(f.massFlux[fo]), _cuda(g[co]),
                          _cuda(f.te[co]), _cuda(f.den[co]), _cuda(f.cpB), _cuda(f.visEffB), _cuda(sb[co]),
                          _cuda(rb.reshape(F, 2)[fo].ravel()), p, pb, mb)
    torch.cuda.synchronize()
    return pb.cpu().numpy(), mb.cpu().numpy(), pb_o[co], mb_o[fo]


@pytest.mark.gpu
@pytest.mark.parametrize("nG", [1, 2, 3])
@pytest.mark.parametrize("variant", ["scalar", "scalar_T", "scalar_wall"])
@pytest.mark.parametrize("name", list(MESHES))
def test_full_adjoint_matches_oracle_bitwise(name, variant, nG):
    print(f"[adjoint full] {name} {variant} nG={nG}", flush=True)
    pb, mb, pb_o, mb_o = run(name, variant, nG)
    assert np.array_equal(pb.view(np.uint64), pb_o.view(np.uint64)), (name, variant, nG,
                                                                         np.abs(pb - pb_o).max())
    assert np.array_equal(mb.view(np.uint64), mb_o.view(np.uint64)), (name, variant, nG,
                                                                         np.abs(mb - mb_o).max())


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [1, 4, 8])
def test_full_adjoint_every_schedule(schedule):
    host, dev, om, f = case("tets_typed")
    dev.set_schedule(schedule)
    try:
        pb, mb, pb_o, mb_o = run("tets_typed", "scalar_wall", 3, seed=schedule)
    finally:
        dev.set_schedule(0)
    assert np.array_equal(pb.view(np.uint64), pb_o.view(np.uint64))
    assert np.array_equal(mb.view(np.uint64), mb_o.view(np.uint64))


@pytest.mark.gpu
def test_full_adjoint_default_one_sweep_is_the_plain_adjoint():
    host, dev, om, f = case("plain")
    p = params("scalar")
    C, F = host.nCells, host.nFaces
    co, fo = dev.cell_order, dev.face_order
    rng = np.random.default_rng(5)
    sb, rb = _cuda(rng.normal(size=C)), _cuda(rng.normal(size=2 * F))
    phi, vis, mf = _cuda(f.phi[co]), _cuda(f.visEff[co]), _cuda(f.massFlux[fo])
    out = []
    for full in (False, True):
        pb = torch.empty(C, dtype=torch.float64, device="cuda")
        mb = torch.empty(F, dtype=torch.float64, device="cuda")
        if full:
            R.scalar_adjoint_full(dev, 1, phi, vis, mf, None, None, None, None, None, sb, rb, p, pb, mb)
        else:
            R.scalar_adjoint(dev, phi, vis, mf, sb, rb, p, pb, mb)
        torch.cuda.synchronize()
        out.append((pb.cpu().numpy(), mb.cpu().numpy()))
    assert np.array_equal(out[0][0].view(np.uint64), out[1][0].view(np.uint64))
    assert np.array_equal(out[0][1].view(np.uint64), out[1][1].view(np.uint64))


@pytest.mark.gpu
def test_full_adjoint_refuses_bad_arguments():
    host, dev, om, f = case("typed_perm")
    p = params("scalar")
    z = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")  # noqa: E731
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    with pytest.raises(ch.CupCfdError) as e:  # typed mesh without the forward gradient
        R.scalar_adjoint_full(dev, 1, z(C), z(C), z(F), None, z(C), z(C), z(B), z(B), z(C), z(2 * F), p, z(C), z(F))
    assert "99" in str(e.value)
    with pytest.raises(ch.CupCfdError) as e:  # nGradient < 1
        R.scalar_adjoint_full(dev, 0, z(C), z(C), z(F), z(3 * C), z(C), z(C), z(B), z(B), z(C), z(2 * F), p, z(C),
                              z(F))
    assert "90" in str(e.value)


@pytest.mark.gpu
def test_autograd_typed_two_sweeps_matches_central_differences():
    """torch.autograd through Chain(nGradient = 2) on the typed mesh with the enthalpy wall: the gradient of
    <sb, Su> + <rb, RFace> along a direction equals central differences of the GPU forward (the chain is at
    most bilinear, so the differences are exact up to rounding)."""
    from cupcfd_hip import adjoint as A

    host, dev, om, f = case("typed_perm")
    p = params("scalar_wall")
    co, fo = dev.cell_order, dev.face_order
    C, F = host.nCells, host.nFaces
    chain = A.Chain(dev, _cuda(f.phiB), _cuda(f.visEff[co]), p, den=_cuda(f.den[co]), te=_cuda(f.te[co]),
                    cpB=_cuda(f.cpB), visEffB=_cuda(f.visEffB), nGradient=2)
    rng = np.random.default_rng(12)
    sb, rb = _cuda(rng.normal(size=C)), _cuda(rng.normal(size=2 * F))
    phi0, mf0 = _cuda(f.phi[co]), _cuda(f.massFlux[fo])
    phi = phi0.clone().requires_grad_(True)
    mf = mf0.clone().requires_grad_(True)
    su, rface = A.scalar_flux(phi, mf, chain)
    ((su * sb).sum() + (rface * rb).sum()).backward()
    pb, mb = chain.adjoint(phi0, mf0, sb, rb)
    assert torch.equal(phi.grad, pb) and torch.equal(mf.grad, mb)
    pd = rng.normal(size=C)
    md = rng.normal(size=F)
    md[np.abs(f.massFlux[fo]) < 1e-5] = 0.0
    pd, md = _cuda(pd), _cuda(md)
    eps = 1e-6

    def loss(s):
        a, b = chain.forward(phi0 + s * eps * pd, mf0 + s * eps * md)
        return ((a * sb).sum() + (b * rb).sum()).item()

    fd = (loss(1) - loss(-1)) / (2 * eps)
    ad = (torch.dot(pb, pd) + torch.dot(mb, md)).item()
    assert abs(fd - ad) <= 1e-6 * max(1.0, abs(ad)), (fd, ad)

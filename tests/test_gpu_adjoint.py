"""GPU adjoint of gradient (nGradient = 1) + scalar face loop w.r.t. (phi, MassFlux).

* cupcfd_dev_scalar_adjoint_f64 equals the oracle's adjoint bit for bit
  (both sum each cell's faces in ascending face id), on the reference test
  mesh and on relabelled meshes, across gather schedules;
* central differences of the GPU forward chain along 256 seeded unit
  directions at 64^3 relabelled (faces with |mf| < 1e-5 left out of the
  MassFlux directions) agree with the adjoint's directional derivatives;
* the dot-product identity holds between the GPU adjoint and the FP64 CPU
  tangent;
* torch.autograd through ScalarFlux gives the adjoint;
* 50 explicit pseudo-time steps reversed with checkpoints every 10 give the
  same bits as the reverse sweep that keeps every state, and match
  central differences of the marched state;
* typed meshes are refused with E_NOT_IMPLEMENTED.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import adjoint as A
from cupcfd_hip import workload as W
from cupcfd_hip.fvm import resident as R

torch = pytest.importorskip("torch")
P = dict(W.SCALAR)


class Case:
    def __init__(self, **mesh):
        self.host = ch.HostMesh(**mesh)
        self.dev = ch.DeviceMesh(self.host, precision=64, reorder=True)
        self.co, self.fo = self.dev.cell_order, self.dev.face_order
        self.f = W.Fields(self.host, seed=31)
        self.C, self.F = self.host.nCells, self.host.nFaces

    def cells(self, a):
        return torch.from_numpy(np.ascontiguousarray(a[self.co])).cuda()

    def faces(self, a, k=1):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(self.F, k)[self.fo].ravel())).cuda()

    def to_mesh_cells(self, t):
        out = np.empty(self.C)
        out[self.co] = t.cpu().numpy()
        return out

    def to_mesh_faces(self, t):
        out = np.empty(self.F)
        out[self.fo] = t.cpu().numpy()
        return out

    def chain(self):
        f = self.f
        return A.Chain(self.dev, torch.from_numpy(f.phiB).cuda(), self.cells(f.visEff), P, den=self.cells(f.den),
                       te=self.cells(f.te), cpB=torch.from_numpy(f.cpB).cuda(),
                       visEffB=torch.from_numpy(f.visEffB).cuda())


_cases = {}


def case(name):
    spec = {"ref5": dict(nx=5, bounds=(-1, 1, -1, 1, -1, 1)),
            "perm12": dict(nx=12, ny=11, nz=10, seed=8, permute_cells=True, permute_faces=True),
            "perm64": dict(nx=64, seed=12345, permute_cells=True, permute_faces=True),
            "tets": dict(nx=5, ny=4, nz=4, tets=True, jitter=0.2, seed=6, permute_cells=True, permute_faces=True)}
    if name not in _cases:
        _cases[name] = Case(**spec[name])
    return _cases[name]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [0, 1, 8])
@pytest.mark.parametrize("name", ["ref5", "perm12", "tets"])
def test_adjoint_matches_oracle_bitwise(name, schedule):
    c = case(name)
    c.dev.set_schedule(schedule)
    rng = np.random.default_rng(4)
    sb = rng.normal(size=c.C)
    rb = rng.normal(size=2 * c.F)
    m = po.OracleMesh(c.host, 64)
    rc, pb_ref, mb_ref = po.scalar_adjoint(m, c.f.phi, c.f.visEff, c.f.massFlux, sb, rb, P)
    assert rc == 95
    pb = torch.empty(c.C, dtype=torch.float64, device="cuda")
    mb = torch.empty(c.F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint(c.dev, c.cells(c.f.phi), c.cells(c.f.visEff), c.faces(c.f.massFlux), c.cells(sb),
                     c.faces(rb, 2), P, pb, mb)
    torch.cuda.synchronize()
    c.dev.set_schedule(0)
    assert np.array_equal(c.to_mesh_cells(pb).view(np.uint64), pb_ref.view(np.uint64))
    assert np.array_equal(c.to_mesh_faces(mb).view(np.uint64), mb_ref.view(np.uint64))


@pytest.mark.gpu
def test_typed_mesh_refused():
    host = ch.HostMesh(4, typed=True)
    dev = ch.DeviceMesh(host, precision=64)
    z = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")  # noqa: E731
    with pytest.raises(ch.CupCfdError) as e:
        R.scalar_adjoint(dev, z(host.nCells), z(host.nCells), z(host.nFaces), z(host.nCells), z(2 * host.nFaces), P,
                         z(host.nCells), z(host.nFaces))
    assert "109" in str(e.value)


def _unit_direction(rng, c):
    pd = rng.normal(size=c.C)
    md = rng.normal(size=c.F)
    md[np.abs(c.f.massFlux) < 1e-5] = 0.0
    md[c.host.a["faceBnd"] > -1] = 0.0
    n = np.sqrt(pd @ pd + md @ md)
    return pd / n, md / n


@pytest.mark.gpu
def test_central_differences_256_directions_64cubed():
    c = case("perm64")
    ch_ = c.chain()
    rng = np.random.default_rng(2026)
    phi, mf = c.cells(c.f.phi), c.faces(c.f.massFlux)
    sb = torch.from_numpy(rng.normal(size=c.C)).cuda()
    rb = torch.from_numpy(rng.normal(size=2 * c.F)).cuda()
    pb, mb = ch_.adjoint(phi, mf, sb, rb)
    eps = 1e-6
    worst = 0.0
    for _ in range(256):
        pd, md = _unit_direction(rng, c)
        pd_t, md_t = c.cells(pd), c.faces(md)
        sp, rp = ch_.forward(phi + eps * pd_t, mf + eps * md_t)
        sm, rm = ch_.forward(phi - eps * pd_t, mf - eps * md_t)
        fd = (torch.dot(sb, sp - sm) + torch.dot(rb, rp - rm)).item() / (2 * eps)
        ad = (torch.dot(pb, pd_t) + torch.dot(mb, md_t)).item()
        worst = max(worst, abs(fd - ad) / max(1.0, abs(ad)))
    assert worst < 1e-6, worst


@pytest.mark.gpu
def test_dot_product_against_cpu_tangent():
    c = case("perm12")
    rng = np.random.default_rng(7)
    m = po.OracleMesh(c.host, 64)
    ch_ = c.chain()
    phi, mf = c.cells(c.f.phi), c.faces(c.f.massFlux)
    for _ in range(4):
        pd, md = _unit_direction(rng, c)
        sb = rng.normal(size=c.C)
        rb = rng.normal(size=2 * c.F)
        _, sd, rd = po.scalar_tangent(m, c.f.phi, c.f.visEff, c.f.massFlux, pd, md, P)
        pb, mb = ch_.adjoint(phi, mf, c.cells(sb), c.faces(rb, 2))
        lhs = sb @ sd + rb @ rd
        rhs = c.to_mesh_cells(pb) @ pd + c.to_mesh_faces(mb) @ md
        assert abs(lhs - rhs) <= 1e-12 * (np.abs(sb) @ np.abs(sd) + np.abs(rb) @ np.abs(rd))


@pytest.mark.gpu
def test_torch_autograd():
    c = case("perm12")
    ch_ = c.chain()
    rng = np.random.default_rng(9)
    phi = c.cells(c.f.phi).requires_grad_(True)
    mf = c.faces(c.f.massFlux).requires_grad_(True)
    sb = torch.from_numpy(rng.normal(size=c.C)).cuda()
    rb = torch.from_numpy(rng.normal(size=2 * c.F)).cuda()
    su, rface = A.scalar_flux(phi, mf, ch_)
    loss = (su * sb).sum() + (rface * rb).sum()
    loss.backward()
    pb, mb = ch_.adjoint(phi.detach(), mf.detach(), sb, rb)
    assert torch.equal(phi.grad, pb) and torch.equal(mf.grad, mb)


@pytest.mark.gpu
def test_march_checkpointed_reverse():
    c = case("perm12")
    ch_ = c.chain()
    rng = np.random.default_rng(11)
    vol = c.cells(c.host.a["cellVolume"])
    phi0, mf = c.cells(c.f.phi), c.faces(c.f.massFlux)
    dtau = 2e-6
    wT = torch.from_numpy(rng.normal(size=c.C)).cuda()
    pb10, mb10 = A.march_adjoint(ch_, phi0, mf, 50, dtau, vol, wT, every=10)
    pball, mball = A.march_adjoint(ch_, phi0, mf, 50, dtau, vol, wT, every=50)
    assert torch.equal(pb10, pball) and torch.equal(mb10, mball)
    eps = 1e-6
    for _ in range(4):
        pd, md = _unit_direction(rng, c)
        pd_t, md_t = c.cells(pd), c.faces(md)
        jp = torch.dot(wT, A.march(ch_, phi0 + eps * pd_t, mf + eps * md_t, 50, dtau, vol)).item()
        jm = torch.dot(wT, A.march(ch_, phi0 - eps * pd_t, mf - eps * md_t, 50, dtau, vol)).item()
        fd = (jp - jm) / (2 * eps)
        ad = (torch.dot(pb10, pd_t) + torch.dot(mb10, md_t)).item()
        assert abs(fd - ad) <= 1e-6 * max(1.0, abs(ad)), (fd, ad)


@pytest.mark.gpu
@pytest.mark.parametrize("K", [1, 3, 16, 40, 64, 70])
@pytest.mark.parametrize("name", ["ref5", "perm12", "tets"])
def test_batched_adjoint_matches_single(name, K):
    """PhiBar_k equals the single-scalar adjoint's and the oracle adjoint's bit for bit; MassFluxBar is the
    RFace term followed by the K scalars' terms added in order (each term taken from a single-scalar call, and
    from an oracle call, with a zero RFace cotangent)."""
    c = case(name)
    rng = np.random.default_rng(40 + K)
    phi = rng.uniform(-1, 1, (c.C, K))
    sb = rng.normal(size=(c.C, K))
    rb = rng.normal(size=2 * c.F)
    mf, vis = c.faces(c.f.massFlux), c.cells(c.f.visEff)
    pb = torch.empty((c.C, K), dtype=torch.float64, device="cuda")
    mb = torch.empty(c.F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint_batched(c.dev, K, c.cells(phi), vis, mf, c.cells(sb), c.faces(rb, 2), P, pb, mb)
    zero_c = torch.zeros(c.C, dtype=torch.float64, device="cuda")
    zero_r = torch.zeros(2 * c.F, dtype=torch.float64, device="cuda")
    p1 = torch.empty(c.C, dtype=torch.float64, device="cuda")
    m1 = torch.empty(c.F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint(c.dev, c.cells(phi[:, 0]), vis, mf, zero_c, c.faces(rb, 2), P, p1, m1)
    want = m1.cpu().numpy().copy()  # the RFace term alone
    for k in range(K):
        R.scalar_adjoint(c.dev, c.cells(phi[:, k]), vis, mf, c.cells(sb[:, k]), zero_r, P, p1, m1)
        torch.cuda.synchronize()
        assert torch.equal(pb[:, k].contiguous(), p1), k
        want = want + m1.cpu().numpy()
    assert np.array_equal(mb.cpu().numpy(), want)
    # every scalar against the oracle's adjoint directly (mesh numbering), MassFluxBar folded the same way
    m = po.OracleMesh(c.host, 64)
    pb_h, mb_h = pb.cpu().numpy(), mb.cpu().numpy()
    _, _, mb_o = po.scalar_adjoint(m, phi[:, 0].copy(), c.f.visEff, c.f.massFlux, np.zeros(c.C), rb, P)
    for k in range(K):
        rc, pk, mk = po.scalar_adjoint(m, phi[:, k].copy(), c.f.visEff, c.f.massFlux, sb[:, k].copy(),
                                       np.zeros(2 * c.F), P)
        assert rc == 95
        assert np.array_equal(pb_h[:, k].view(np.uint64), pk[c.co].view(np.uint64)), k
        mb_o = mb_o + mk
    assert np.array_equal(mb_h.view(np.uint64), mb_o[c.fo].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("K", [3, 16, 64, 70])
@pytest.mark.parametrize("name", ["perm12", "tets"])
def test_batched_adjoint_f32_storage_is_rounded_f64(name, K):
    """FP32 storage of phi, SuBar and PhiBar: PhiBar is the FP64 batched adjoint on the widened inputs, rounded
    once; MassFluxBar is that call's exactly."""
    c = case(name)
    rng = np.random.default_rng(70 + K)
    phi = rng.uniform(-1, 1, (c.C, K)).astype(np.float32)
    sb = rng.normal(size=(c.C, K)).astype(np.float32)
    rb = rng.normal(size=2 * c.F)
    mf, vis = c.faces(c.f.massFlux), c.cells(c.f.visEff)
    out = []
    for dt in (torch.float32, torch.float64):
        pb = torch.empty((c.C, K), dtype=dt, device="cuda")
        mb = torch.empty(c.F, dtype=torch.float64, device="cuda")
        ph, s_ = c.cells(phi.astype(np.float64)).to(dt), c.cells(sb.astype(np.float64)).to(dt)
        call = R.scalar_adjoint_batched_f32s if dt == torch.float32 else R.scalar_adjoint_batched
        call(c.dev, K, ph.contiguous(), vis, mf, s_.contiguous(), c.faces(rb, 2), P, pb, mb)
        torch.cuda.synchronize()
        out.append((pb, mb))
    (pb32, mb32), (pb64, mb64) = out
    assert torch.equal(pb32, pb64.to(torch.float32))
    assert torch.equal(mb32, mb64)


@pytest.mark.gpu
def test_graph_survives_workspace_growth():
    """A batched adjoint captured in a hipGraph at K = 3 keeps replaying correctly after an eager call at
    K = 64 has grown the mesh's scratch: the outgrown buffers stay allocated until the mesh is destroyed."""
    c = Case(nx=9, ny=8, nz=7, seed=21, permute_cells=True, permute_faces=True)
    rng = np.random.default_rng(123)
    mf, vis = c.faces(c.f.massFlux), c.cells(c.f.visEff)
    rb = c.faces(rng.normal(size=2 * c.F), 2)

    def inputs(K):
        return c.cells(rng.uniform(-1, 1, (c.C, K))), c.cells(rng.normal(size=(c.C, K)))

    phi3, sb3 = inputs(3)
    pb3 = torch.empty((c.C, 3), dtype=torch.float64, device="cuda")
    mb3 = torch.empty(c.F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint_batched(c.dev, 3, phi3, vis, mf, sb3, rb, P, pb3, mb3)  # sizes the K = 3 scratch
    torch.cuda.synchronize()
    want_p, want_m = pb3.clone(), mb3.clone()
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            R.scalar_adjoint_batched(c.dev, 3, phi3, vis, mf, sb3, rb, P, pb3, mb3, stream=stream)
    torch.cuda.synchronize()
    phi64, sb64 = inputs(64)
    pb64 = torch.empty((c.C, 64), dtype=torch.float64, device="cuda")
    mb64 = torch.empty(c.F, dtype=torch.float64, device="cuda")
    R.scalar_adjoint_batched(c.dev, 64, phi64, vis, mf, sb64, rb, P, pb64, mb64)  # grows the scratch
    torch.cuda.synchronize()
    eager64 = (pb64.clone(), mb64.clone())
    for _ in range(3):
        pb3.fill_(0.0)
        mb3.fill_(0.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(pb3, want_p) and torch.equal(mb3, want_m)
    # the K = 64 results were not disturbed by the replays, and a fresh K = 64 call still agrees
    R.scalar_adjoint_batched(c.dev, 64, phi64, vis, mf, sb64, rb, P, pb64, mb64)
    torch.cuda.synchronize()
    assert torch.equal(pb64, eager64[0]) and torch.equal(mb64, eager64[1])
    del graph

"""Pins of the CPU oracle (oracle/oracle_kernels.inc) against what the
mathematics fixes, independently of its own formulas.

The reference's tests hold no kernel-value fixtures (its fvm tests only check
E_SUCCESS on unpopulated inputs), so each oracle function is checked here by
closed forms on axis-aligned StructGen meshes, conservation invariants, and
special cases chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand, the wrong cell of a face) changes the result.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import workload as W
from kernel_runs import compare, run

N = 6
H = 1.0 / N


@pytest.fixture(scope="module")
def cube():
    return ch.HostMesh(N)  # [0,1]^3, one DEFAULT region


@pytest.fixture(scope="module")
def typed():
    return ch.HostMesh(N, typed=True)


def interior_cells(host):
    cc = host.a["cellCenter"]
    return np.all((cc > H) & (cc < 1 - H), axis=1)


def test_gradient_constant_field(cube):
    """phi = c everywhere: the reference adds phiF*S to BOTH cells of an interior
    face (GradientKernels.ipp:101-102), so an interior cell collects +S from its
    three high faces and +S from its three low faces: 2c/h per axis."""
    m = po.OracleMesh(cube)
    c = 0.75
    phi = np.full(cube.nCells, c)
    phiB = np.full(cube.nBoundaries, c)
    g = np.zeros((cube.nCells, 3))
    go = np.zeros((cube.nCells, 3))
    assert po.GradientPhiGaussDolfyn(m, 1, phi, phiB, g, go) == ch.E_SUCCESS
    inside = interior_cells(cube)
    assert np.allclose(g[inside], 2 * c / H, rtol=1e-13)
    assert np.array_equal(g, go)
    # a cell on the x-low boundary: low face is a boundary face of cell1 with
    # S = -A x, high face adds +A x: x-component cancels
    cc = cube.a["cellCenter"]
    xlow = (cc[:, 0] < H) & np.all((cc[:, 1:] > H) & (cc[:, 1:] < 1 - H), axis=1)
    assert np.allclose(g[xlow, 0], 0.0, atol=1e-12)
    xhigh = (cc[:, 0] > 1 - H) & np.all((cc[:, 1:] > H) & (cc[:, 1:] < 1 - H), axis=1)
    assert np.allclose(g[xhigh, 0], 2 * c / H, rtol=1e-13)


def test_gradient_linear_field(cube):
    """phi = a.x: the interior face value is a.Xf, so an interior cell's x
    component is A a.(Xf_hi + Xf_lo) / vol = 2 (a.Cc) / h."""
    m = po.OracleMesh(cube)
    av = np.array([0.3, -1.1, 2.0])
    phi = cube.a["cellCenter"] @ av
    phiB = cube.a["faceCenter"][cube.a["bndFace"]] @ av
    g = np.zeros((cube.nCells, 3))
    go = np.zeros((cube.nCells, 3))
    po.GradientPhiGaussDolfyn(m, 1, phi, phiB, g, go)
    inside = interior_cells(cube)
    expect = 2.0 * phi[inside, None] / H * np.ones(3)
    assert np.allclose(g[inside], expect, rtol=1e-12, atol=1e-12)
    # second sweep: on this grid Xf equals lerp(Cc) on interior faces, so the
    # dPhidxo correction term vanishes and the result repeats
    g2 = np.zeros_like(g)
    go2 = np.zeros_like(g)
    po.GradientPhiGaussDolfyn(m, 2, phi, phiB, g2, go2)
    assert np.allclose(g2, g, rtol=1e-13, atol=1e-13)


def test_gradient_sweep_uses_previous_gradient():
    """On a mesh whose face centres are off the centre line (seeded relabelled
    StructGen with shifted Xf is not available, so shift phiB and use the
    lambda weighting): the second sweep differs from the first exactly by
    S * dot(lerp(g1), Xf - lerp(Cc)) summed per cell."""
    host = ch.HostMesh(4)
    m = po.OracleMesh(host)
    m.arr["faceLambda"][m.arr["faceLambda"] > 0] = 0.3  # lerp(Cc) no longer equals Xf
    rng = np.random.default_rng(0)
    phi = rng.uniform(-1, 1, host.nCells)
    phiB = rng.uniform(-1, 1, host.nBoundaries)
    g1 = np.zeros((host.nCells, 3))
    o1 = np.zeros_like(g1)
    po.GradientPhiGaussDolfyn(m, 1, phi, phiB, g1, o1)
    g2 = np.zeros_like(g1)
    o2 = np.zeros_like(g1)
    po.GradientPhiGaussDolfyn(m, 2, phi, phiB, g2, o2)
    a = m.arr
    inner = a["faceBnd"] < 0
    ip, inn = a["faceCell1"][inner], a["faceCell2"][inner]
    lam = a["faceLambda"][inner][:, None]
    xac = a["cellCenter"][inn] * lam + a["cellCenter"][ip] * (1 - lam)
    corr = ((g1[inn] * lam + g1[ip] * (1 - lam)) * (a["faceCenter"][inner] - xac)).sum(axis=1)
    d = np.zeros_like(g1)
    np.add.at(d, ip, corr[:, None] * a["faceNorm"][inner])
    np.add.at(d, inn, corr[:, None] * a["faceNorm"][inner])
    d /= a["cellVolume"][:, None]
    assert np.allclose(g2 - g1, d, rtol=1e-9, atol=1e-12)


def test_mass_constant_velocity(cube):
    """Constant u, rho=1, zero gradients, constant p: mf = rho u.S and
    rface = -apv with apv = rho ar A vol / h (MassKernels.ipp:151-195)."""
    m = po.OracleMesh(cube)
    C, F, B = cube.nCells, cube.nFaces, cube.nBoundaries
    f = W.Fields(cube, seed=1)
    U = np.array([0.4, -0.7, 1.3])
    f.u[:], f.v[:], f.w[:] = U
    for g in (f.dudx, f.dvdx, f.dwdx, f.dpdx):
        g[:] = 0
    f.p[:] = 2.5
    f.ar[:] = 1.7
    mf0 = f.massFlux.copy()
    rc, out = run("mass", po, m, f, m.state)
    inner = cube.a["faceBnd"] < 0
    S = cube.a["faceNorm"]
    assert np.allclose(out["massFlux"][inner], S[inner] @ U, rtol=1e-13, atol=1e-15)
    assert np.array_equal(out["massFlux"][~inner], mf0[~inner])  # DEFAULT boundaries untouched
    apv = 1.7 * H ** 2 * H ** 3 / H
    rf = out["rface"].reshape(F, 2)
    assert np.allclose(rf[inner], -apv, rtol=1e-12)
    assert out["counters"].tolist() == [0, 0, 0, 0]


def test_mass_pressure_quirk():
    """Linear p = g.x with dpdx[c] = g + delta_c: the reference evaluates
    pin with dpdx[ip] (MassKernels.ipp:169), giving
    pin - pip = -delta_ip.(Cin - Cip) + terms in xpac that cancel."""
    host = ch.HostMesh(5, face_extras=True, seed=4)
    m = po.OracleMesh(host)
    f = W.Fields(host, seed=2)
    gv = np.array([0.5, -0.25, 1.0])
    rng = np.random.default_rng(3)
    delta = rng.uniform(-0.1, 0.1, (host.nCells, 3))
    f.dpdx[:] = gv + delta
    f.p[:] = host.a["cellCenter"] @ gv
    f.u[:], f.v[:], f.w[:] = 0.0, 0.0, 0.0
    for g in (f.dudx, f.dvdx, f.dwdx):
        g[:] = 0
    f.den[:] = 1.0
    _, out = run("mass", po, m, f, m.state)
    a = host.a
    inner = a["faceBnd"] < 0
    ip, inn = a["faceCell1"][inner], a["faceCell2"][inner]
    Cip, Cin = a["cellCenter"][ip], a["cellCenter"][inn]
    xpac, xnac = a["faceXpac"][inner], a["faceXnac"][inner]
    dp_ip = f.dpdx[ip]
    pip = f.p[ip] + (dp_ip * (xpac - Cip)).sum(1)
    pin = f.p[inn] + (dp_ip * (xpac - Cin)).sum(1)
    lam = a["faceLambda"][inner][:, None]
    dpn = (f.dpdx[inn] * lam + f.dpdx[ip] * (1 - lam)) * (xnac - xpac)
    apv = -out["rface"].reshape(-1, 2)[inner, 0]
    expect = -apv * ((pin - pip) - dpn.sum(1))
    assert np.allclose(out["massFlux"][inner], expect, rtol=1e-10, atol=1e-13)
    # the neighbour's gradient (Dolfyn's form) would give a different answer
    pin_dolfyn = f.p[inn] + (f.dpdx[inn] * (xpac - Cin)).sum(1)
    alt = -apv * ((pin_dolfyn - pip) - dpn.sum(1))
    assert not np.allclose(alt, out["massFlux"][inner], rtol=1e-6)


def test_mass_typed_boundaries(typed):
    m = po.OracleMesh(typed)
    f = W.Fields(typed, seed=3)
    f.den[:] = 1.0
    su0 = f.su.copy()
    _, out = run("mass", po, m, f, m.state)
    a = typed.a
    t = a["regType"][a["bndRegion"]]
    mfb = out["massFlux"][a["bndFace"]]
    Sb = a["faceNorm"][a["bndFace"]]
    # INLET: rho_region uvw_region . S with uvw = (1,0,0) on x = 0 (S = -A x)
    assert np.allclose(mfb[t == ch.RTYPE_INLET], -H * H, rtol=1e-14)
    assert (mfb[(t == ch.RTYPE_WALL) | (t == ch.RTYPE_SYMP)] == 0).all()
    ip = a["faceCell1"][a["bndFace"]]
    uo = np.stack([f.u, f.v, f.w], 1)[ip[t == ch.RTYPE_OUTLET]]
    mo = (uo * Sb[t == ch.RTYPE_OUTLET]).sum(1)
    assert np.allclose(mfb[t == ch.RTYPE_OUTLET], np.where(mo < 0, 1e-18, mo), rtol=1e-13)
    # su gets -mf once per INLET face of its cell
    dsu = np.zeros_like(su0)
    np.add.at(dsu, ip[t == ch.RTYPE_INLET], -mfb[t == ch.RTYPE_INLET])
    assert np.allclose(out["su"] - su0, dsu, atol=1e-14)
    assert out["counters"].tolist() == [N * N, N * N, 2 * N * N, 2 * N * N]


def test_mass_boundary_loops(typed):
    m = po.OracleMesh(typed)
    f = W.Fields(typed, seed=4)
    a = typed.a
    t = a["regType"][a["bndRegion"]]
    _, o1 = run("mass_bl1", po, m, f, m.state)
    assert np.isclose(o1["flowin"][0], f.massFlux[a["bndFace"][t == ch.RTYPE_INLET]].sum(), rtol=1e-13)
    _, o3 = run("mass_bl3", po, m, f, m.state)
    assert np.isclose(o3["ratearea"][0], -W.MASS_LOOPS["flowin"] / 1.0, rtol=1e-13)  # outlet area = 1
    f.denB[:] = 2.0
    _, o4 = run("mass_bl4", po, m, f, m.state)
    out_f = a["bndFace"][t == ch.RTYPE_OUTLET]
    assert np.allclose(o4["massFlux"][out_f], 1.0 * H * H * 1.0, rtol=1e-14)  # ratearea * A * splvl
    assert np.isclose(o4["flowout"][0], 1.0, rtol=1e-12)
    ob = np.nonzero(t == ch.RTYPE_OUTLET)[0]
    assert np.allclose(o4["uB"][ob], 1.0 / 2.0, rtol=1e-14) and (o4["vB"][ob] == 0).all()
    f2 = W.Fields(typed, seed=5)
    fr0 = f2.flowRegion.copy()
    _, o2 = run("mass_bl2", po, m, f2, m.state)
    assert np.isclose(o2["flowRegion"][1] - fr0[1], f2.massFlux[out_f].sum(), rtol=1e-12)
    assert np.array_equal(np.delete(o2["flowRegion"], 1), np.delete(fr0, 1))
    f3 = W.Fields(typed, seed=6)
    ff = f3.flowFact.copy()
    mf0 = f3.massFlux.copy()
    su0 = f3.su.copy()
    _, o5 = run("mass_bl5", po, m, f3, m.state)
    assert np.allclose(o5["massFlux"][out_f], mf0[out_f] * ff[1], rtol=1e-15)
    assert np.isclose(o5["flowout2"][0], 0.25 + (mf0[out_f] * ff[1]).sum(), rtol=1e-12)
    dsu = np.zeros_like(su0)
    np.add.at(dsu, a["faceCell1"][out_f], -mf0[out_f] * ff[1])
    assert np.allclose(o5["su"] - su0, dsu, atol=1e-14)
    f4 = W.Fields(typed, seed=7)
    fr = f4.flowRegion.copy()
    _, o6 = run("mass_region", po, m, f4, m.state)
    expect = np.where(a["regType"] == ch.RTYPE_OUTLET, -(a["regSplvl"] * 0.7) / fr, 0.0)
    assert np.allclose(o6["flowFact"], expect, rtol=1e-15)


def _visac(p, vis):
    # ivar = 1 is none of VarT/VarTE/VarED: (VisLam + (x - VisLam)/Sigma_s)/Schmidt
    return (p["vislam"] + (vis - p["vislam"]) / p["Sigma_s"]) / p["Schmidt"]


def test_scalar_diffusion_closed_form(cube):
    """Constant dPhidx = g, constant VisEff, rlencos = 0, GammaBlend = 0:
    interior faces add +-Visac g.S, which cancels over every closed interior
    cell; a cell with DEFAULT boundary faces keeps -Visac g.(sum of its
    boundary S)."""
    m = po.OracleMesh(cube)
    f = W.Fields(cube, seed=8)
    gvec = np.array([0.2, 0.5, -0.3])
    f.dphidx[:] = gvec
    p = dict(W.SCALAR, GammaBlend=0.0)
    su0 = f.su.copy()
    rc = po.FluxScalarDolfynFaceLoop(m, f.phi, f.phiB, f.visEff, f.au, f.su, f.den, f.massFlux, f.te, f.cpB,
                                     f.visEffB, f.rface, f.dphidx, p["SolveTurb"], p["SolveEnthalpy"], p["vislam"],
                                     p["ivar"], p["VarT"], p["Sigma_T"], p["Prandtl"], p["VarTE"], p["Sigma_k"],
                                     p["VarED"], p["Sigma_e"], p["Sigma_s"], p["Schmidt"], 0.0, p["Small"],
                                     p["Large"], p["TMCmu"])
    assert rc == ch.E_SUCCESS
    vis = _visac(p, 1e-3)
    a = cube.a
    bsum = np.zeros((cube.nCells, 3))
    np.add.at(bsum, a["faceCell1"][a["bndFace"]], a["faceNorm"][a["bndFace"]])
    expect = -vis * (bsum @ gvec)
    assert np.allclose(f.su - su0, expect, rtol=1e-9, atol=1e-9 * abs(vis) * H * H)
    # RFace with rlencos = 0: (-max(mf,0), min(mf,0)) on interior faces
    inner = a["faceBnd"] < 0
    rf = f.rface.reshape(-1, 2)[inner]
    mf = f.massFlux[inner]
    assert np.array_equal(rf[:, 0], -np.maximum(mf, 0)) and np.array_equal(rf[:, 1], np.minimum(mf, 0))


def test_scalar_upwind_blend_closed_form(cube):
    """GammaBlend only, constant phi = c, carried PhiFace = 0 (no assigning
    faces): each interior face moves gamma*c*mf from cell2 to cell1."""
    m = po.OracleMesh(cube)
    f = W.Fields(cube, seed=9)
    c = 1.25
    f.phi[:] = c
    f.dphidx[:] = 0
    p = W.SCALAR
    su0 = f.su.copy()
    run("scalar", po, m, f, m.state)
    a = cube.a
    inner = a["faceBnd"] < 0
    d = np.zeros(cube.nCells)
    np.add.at(d, a["faceCell1"][inner], p["GammaBlend"] * c * f.massFlux[inner])
    np.add.at(d, a["faceCell2"][inner], -p["GammaBlend"] * c * f.massFlux[inner])
    assert np.allclose(f.su - su0, d, rtol=1e-11, atol=1e-13)


def test_scalar_carried_phiface(typed):
    """PhiFace is carried from the last face that assigned it, in face order:
    OUTLET faces set it to the extrapolated cell value, INLET faces to 0
    (ivar = 1), and interior faces use it in fce = mf * PhiFace."""
    m = po.OracleMesh(typed)
    f = W.Fields(typed, seed=10)
    c = 0.8
    f.phi[:] = c
    f.dphidx[:] = 0
    f.massFlux[:] = 1.0
    su0 = f.su.copy()
    run("scalar", po, m, f, m.state)
    a = typed.a
    t = a["regType"][a["bndRegion"]]
    # walk faces in order, tracking the carried value
    carried = 0.0
    phif = np.zeros(typed.nFaces)
    for i in range(typed.nFaces):
        b = a["faceBnd"][i]
        if b < 0:
            phif[i] = carried
        elif t[b] == ch.RTYPE_OUTLET:
            carried = c
        elif t[b] == ch.RTYPE_INLET:
            carried = 0.0
    inner = a["faceBnd"] < 0
    gam = W.SCALAR["GammaBlend"]
    blend = gam * (1.0 * phif[inner] - c)  # fce - fci with mf = 1
    d = np.zeros(typed.nCells)
    np.add.at(d, a["faceCell1"][inner], -blend)
    np.add.at(d, a["faceCell2"][inner], blend)
    # boundary faces: INLET adds -f*0 with f = -VisFace + min(mf,0) = 0 (rlencos 0, mf > 0),
    # OUTLET adds fde - fdi = 0 (zero gradient): no contribution
    assert np.allclose(f.su - su0, d, rtol=1e-11, atol=1e-12)
    assert (phif[inner] == c).any() and (phif[inner] == 0).any()


def test_uvw_stress_closed_form(cube):
    """Constant velocity-gradient tensor G (rows du, dv, dw), gammaBlend = 0,
    rlencos = 0: interior faces add +-visac ((G + G^T) S), which cancels over
    closed cells; DEFAULT boundary cells keep -visac (G + G^T) sum(S_bnd)."""
    m = po.OracleMesh(cube)
    f = W.Fields(cube, seed=11)
    G = np.array([[0.3, -0.2, 0.7], [0.1, 0.4, -0.5], [0.9, 0.2, -0.6]])
    f.dudx[:], f.dvdx[:], f.dwdx[:] = G[0], G[1], G[2]
    s0 = np.stack([f.su.copy(), f.sv.copy(), f.sw.copy()], 1)
    rc = po.FluxUVWDolfynFaceLoop1(m, 0.0, 1e-18, f.u, f.v, f.w, f.uB, f.vB, f.wB, f.visEff, f.visEffB, f.massFlux,
                                   f.dudx, f.dvdx, f.dwdx, f.rface, f.su, f.sv, f.sw, f.au, f.av, f.aw)
    assert rc == ch.E_SUCCESS
    a = cube.a
    bsum = np.zeros((cube.nCells, 3))
    np.add.at(bsum, a["faceCell1"][a["bndFace"]], a["faceNorm"][a["bndFace"]])
    expect = -1e-3 * bsum @ (G + G.T).T
    got = np.stack([f.su, f.sv, f.sw], 1) - s0
    assert np.allclose(got, expect, rtol=1e-9, atol=1e-15)


def test_uvw_boundaries(typed):
    """WALL: shear force = visac * u_t / dn * A with u_t the tangential velocity
    relative to the wall; SYMP: normal stress only; INLET/OUTLET set the
    boundary velocity."""
    m = po.OracleMesh(typed)
    f = W.Fields(typed, seed=12)
    run("uvw", po, m, f, m.state)
    a = typed.a
    t = a["regType"][a["bndRegion"]]
    ip = a["faceCell1"][a["bndFace"]]
    # WALL boundaries: shear parallel to the wall (y-normal), y component 0
    w = t == ch.RTYPE_WALL
    shear = m.arr["bndShear"][w]
    assert (shear[:, 1] == 0).all()
    uw = a["regUVW"][a["bndRegion"][w]]
    ut = np.stack([f.u, f.v, f.w], 1)[ip[w]] - uw
    ut[:, 1] = 0
    assert np.allclose(shear, 1e-3 * ut / (0.5 * H) * H * H, rtol=1e-12)
    # INLET boundary velocity = region uvw; WALL = wall velocity
    assert np.array_equal(np.stack([f.uB, f.vB, f.wB], 1)[t == ch.RTYPE_INLET],
                          a["regUVW"][a["bndRegion"][t == ch.RTYPE_INLET]])
    assert np.array_equal(np.stack([f.uB, f.vB, f.wB], 1)[w], uw)
    # region force tangent: zeroed, then the WALL shears summed per region
    po.FluxUVWDolfynRegionLoop1(m)
    assert (m.arr["regForceTangent"] == 0).all()
    po.FluxUVWDolfynBndsLoop1(m)
    for r in (2, 3):
        assert np.allclose(m.arr["regForceTangent"][r], m.arr["bndShear"][a["bndRegion"] == r].sum(0), rtol=1e-12)
    assert (m.arr["regForceTangent"][[0, 1, 4, 5]] == 0).all()


def test_viscosity_loops(typed):
    m = po.OracleMesh(typed)
    f = W.Fields(typed, seed=13)
    q = W.VISC
    # cell loop 1 with ED below `small`: relaxes towards visLam
    f.ed[:5] = 0.0
    v0 = f.visEff.copy()
    run("visc_cell1", po, m, f, m.state)
    c1 = q["cell1"]
    assert np.allclose(f.visEff[:5], v0[:5] + c1["visURF"] * (c1["visLam"] - v0[:5]), rtol=1e-15)
    vn = c1["visLam"] + c1["cmu"] * f.den[5:] * f.te[5:] ** 2 / f.ed[5:]
    assert np.allclose(f.visEff[5:], v0[5:] + c1["visURF"] * (vn - v0[5:]), rtol=1e-13)
    # region loop: the standard wall's ylog is the fixed point y = ln(E y)/kappa
    po.calculateViscosityDolfynRegionLoop(q["region"]["kappa"], m)
    y = m.arr["regYLog"][2]
    assert abs(y - np.log(9.0 * y) / 1.1) < 2e-3 and y > 1
    assert m.arr["regYLog"][3] == 0.0  # non-standard wall
    # boundary loop
    m.arr["regYLog"][2] = 11.0
    b = q["boundary"]
    f2 = W.Fields(typed, seed=14)
    run("visc_boundary", po, m, f2, m.state)
    a = typed.a
    t = a["regType"][a["bndRegion"]]
    ip = a["faceCell1"][a["bndFace"]]
    assert np.array_equal(f2.visEffB[t == ch.RTYPE_OUTLET], f2.visEff[ip[t == ch.RTYPE_OUTLET]])
    inlet = b["visLam"] + 1.0 * b["tmcmu"] * 1e-3 ** 2 / (1e-4 + b["small"])
    assert np.allclose(f2.visEffB[t == ch.RTYPE_INLET], inlet, rtol=1e-15)
    wall = t == ch.RTYPE_WALL
    yp = 0.5 * H * b["tmcmu"] ** 0.25 * np.sqrt(f2.te[ip[wall]]) * f2.den[ip[wall]] / (b["visLam"] + b["small"])
    assert np.allclose(m.arr["bndYPlus"][wall], yp, rtol=1e-13)
    # cell loop 2 clamps at 1e6 * visLam
    f3 = W.Fields(typed, seed=15)
    f3.visEff[:3] = 1e9
    run("visc_cell2", po, m, f3, m.state)
    assert (f3.visEff[:3] == 1e6 * 1.1).all() and (f3.visEff[3:] == 1e-3).all()


def test_float_instantiation_tracks_double():
    host = ch.HostMesh(6, 5, 4, typed=True, face_extras=True, seed=2)
    for k in ("gradient", "mass", "scalar", "uvw", "visc_boundary"):
        outs = []
        for prec in (64, 32):
            m = po.OracleMesh(host, prec)
            f = W.Fields(host, seed=21).cast(np.float64 if prec == 64 else np.float32)
            outs.append(run(k, po, m, f, m.state)[1])
        for name in outs[0]:
            # infinities (u+ = ln(1) = 0 on the non-standard wall) must coincide
            assert compare(outs[1][name], outs[0][name])[1] < 1e-4, (k, name)

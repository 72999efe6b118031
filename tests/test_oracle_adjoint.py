"""The oracle's tangent and adjoint of gradient (nGradient = 1) + scalar face loop (CPU).

Pins, each chosen so that a plausible slip (a dropped term, a swapped cell,
a wrong sign) fails:
  * the tangent equals central finite differences of the forward oracle
    (GradientPhiGaussDolfyn then FluxScalarDolfynFaceLoop) along seeded
    directions: the chain is linear in phi and piecewise linear in MassFlux,
    so the differences are exact up to rounding once faces with |mf| < 1e-5
    are left out of the MassFlux directions;
  * the adjoint satisfies the dot-product identity <ybar, J xdot> =
    <J^T ybar, xdot> against that tangent;
  * with GammaBlend = 0 the mass flux only reaches RFace, so mfbar is the
    RFace cotangent alone: -rbar[2f] where mf > 0, rbar[2f+1] where mf < 0;
  * a typed mesh (non-DEFAULT regions) is refused.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import workload as W


def _forward(m, host, f, phi, mf, p):
    C, F = host.nCells, host.nFaces
    g = np.zeros((C, 3))
    go = np.zeros((C, 3))
    assert po.GradientPhiGaussDolfyn(m, 1, phi, f.phiB.copy(), g, go) == 95
    su = np.zeros(C)
    au = np.zeros(C)
    rf = np.zeros(2 * F)
    rc = po.FluxScalarDolfynFaceLoop(m, phi, f.phiB.copy(), f.visEff, au, su, f.den, mf, f.te, f.cpB, f.visEffB, rf,
                                     g, p["SolveTurb"], p["SolveEnthalpy"], p["vislam"], p["ivar"], p["VarT"],
                                     p["Sigma_T"], p["Prandtl"], p["VarTE"], p["Sigma_k"], p["VarED"], p["Sigma_e"],
                                     p["Sigma_s"], p["Schmidt"], p["GammaBlend"], p["Small"], p["Large"], p["TMCmu"])
    assert rc == 95
    return su, rf


@pytest.fixture(scope="module")
def case():
    host = ch.HostMesh(9, 8, 7, seed=5, permute_cells=True, permute_faces=True)
    m = po.OracleMesh(host, 64)
    f = W.Fields(host, seed=21)
    return host, m, f


def _direction(rng, host, f, with_mf=True):
    pd = rng.normal(size=host.nCells)
    md = rng.normal(size=host.nFaces) if with_mf else np.zeros(host.nFaces)
    md[np.abs(f.massFlux) < 1e-5] = 0.0
    md[host.a["faceBnd"] > -1] = 0.0
    n = np.sqrt(pd @ pd + md @ md)
    return pd / n, md / n


def test_tangent_matches_central_differences(case):
    host, m, f = case
    p = dict(W.SCALAR)
    rng = np.random.default_rng(1)
    eps = 1e-6
    for _ in range(6):
        pd, md = _direction(rng, host, f)
        rc, sd, rd = po.scalar_tangent(m, f.phi, f.visEff, f.massFlux, pd, md, p)
        assert rc == 95
        sp, rp = _forward(m, host, f, f.phi + eps * pd, f.massFlux + eps * md, p)
        sm, rm = _forward(m, host, f, f.phi - eps * pd, f.massFlux - eps * md, p)
        fd_s, fd_r = (sp - sm) / (2 * eps), (rp - rm) / (2 * eps)
        assert np.max(np.abs(fd_s - sd)) <= 1e-6 * max(1.0, np.max(np.abs(sd)))
        assert np.max(np.abs(fd_r - rd)) <= 1e-6 * max(1.0, np.max(np.abs(rd)))


def test_adjoint_dot_product(case):
    host, m, f = case
    p = dict(W.SCALAR)
    rng = np.random.default_rng(2)
    for _ in range(4):
        pd, md = _direction(rng, host, f)
        sb = rng.normal(size=host.nCells)
        rb = rng.normal(size=2 * host.nFaces)
        _, sd, rd = po.scalar_tangent(m, f.phi, f.visEff, f.massFlux, pd, md, p)
        rc, pb, mb = po.scalar_adjoint(m, f.phi, f.visEff, f.massFlux, sb, rb, p)
        assert rc == 95
        lhs = sb @ sd + rb @ rd
        rhs = pb @ pd + mb @ md
        assert abs(lhs - rhs) <= 1e-12 * (np.abs(sb) @ np.abs(sd) + np.abs(rb) @ np.abs(rd))


def test_mass_flux_cotangent_without_blend(case):
    host, m, f = case
    p = dict(W.SCALAR, GammaBlend=0.0)
    rng = np.random.default_rng(3)
    sb = rng.normal(size=host.nCells)
    rb = rng.normal(size=2 * host.nFaces)
    rc, _, mb = po.scalar_adjoint(m, f.phi, f.visEff, f.massFlux, sb, rb, p)
    assert rc == 95
    mf = f.massFlux
    interior = host.a["faceBnd"] < 0
    want = np.where(mf > 0, -rb[0::2], np.where(mf < 0, rb[1::2], 0.0))
    want[~interior] = 0.0
    assert np.array_equal(np.abs(mb), np.abs(want)) and np.allclose(mb, want, rtol=0, atol=0)


def test_typed_mesh_refused():
    host = ch.HostMesh(4, typed=True)
    m = po.OracleMesh(host, 64)
    f = W.Fields(host, seed=1)
    rc, _, _ = po.scalar_adjoint(m, f.phi, f.visEff, f.massFlux, np.zeros(host.nCells), np.zeros(2 * host.nFaces),
                                 dict(W.SCALAR))
    assert rc == 109

"""The oracle's tangent and adjoint of gradient (nGradient sweeps) + scalar face loop on any mesh (CPU).

or_scalar_tangent_full / or_scalar_adjoint_full extend the nGradient = 1,
all-DEFAULT chain to the gradient's correction sweeps and to typed
boundaries: INLET and OUTLET source terms, the PhiFace an OUTLET, INLET or
enthalpy WALL face leaves for the faces after it, and the WALL's Hcoef term.
The reference has neither; the pins are what the mathematics fixes:
  * the tangent equals central differences of the forward oracle
    (GradientPhiGaussDolfyn with nGradient sweeps, then
    FluxScalarDolfynFaceLoop). The chain is at most bilinear in
    (phi, MassFlux) (mf * PhiFace, min/max(mf, 0) * phi), so central
    differences are exact up to rounding once faces with |mf| < 1e-5 leave
    the MassFlux directions;
  * the adjoint satisfies the dot-product identity against that tangent;
  * on an all-DEFAULT mesh with nGradient = 1 it is the nGradient = 1 adjoint,
    bit for bit.
Meshes: typed and relabelled hexes with non-zero xpac/xnac/rlencos, typed
jittered tets, and a DEFAULT relabelled hex mesh; the scalar loop with INLET
assigning PhiFace, with INLET carrying it (ivar = VarT), and with the enthalpy
wall (SolveEnthalpy, ivar = VarT: flux and non-flux WALL regions).
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import workload as W

MESHES = {
    "typed_perm": dict(nx=7, ny=6, nz=5, typed=True, permute_cells=True, permute_faces=True, face_extras=True,
                       seed=11),
    "tets_typed": dict(nx=4, ny=3, nz=3, tets=True, jitter=0.2, typed=True, permute_cells=True, permute_faces=True,
                       face_extras=True, seed=5),
    "plain": dict(nx=9, ny=8, nz=7, seed=5, permute_cells=True, permute_faces=True),
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
        m = po.OracleMesh(host, 64)
        f = W.Fields(host, seed=21)
        _cases[name] = (host, m, f)
    return _cases[name]


def gradient(m, host, f, phi, nG):
    g = np.zeros((host.nCells, 3))
    go = np.zeros((host.nCells, 3))
    assert po.GradientPhiGaussDolfyn(m, nG, np.array(phi), f.phiB.copy(), g, go) == 95
    return g


def forward(m, host, f, phi, mf, p, nG):
    C, F = host.nCells, host.nFaces
    g = gradient(m, host, f, phi, nG)
    su = np.zeros(C)
    au = np.zeros(C)
    rf = np.zeros(2 * F)
    rc = po.FluxScalarDolfynFaceLoop(m, np.array(phi), f.phiB.copy(), f.visEff.copy(), au, su, f.den.copy(),
                                     np.array(mf), f.te.copy(), f.cpB.copy(), f.visEffB.copy(), rf, g,
                                     p["SolveTurb"], p["SolveEnthalpy"], p["vislam"], p["ivar"], p["VarT"],
                                     p["Sigma_T"], p["Prandtl"], p["VarTE"], p["Sigma_k"], p["VarED"], p["Sigma_e"],
                                     p["Sigma_s"], p["Schmidt"], p["GammaBlend"], p["Small"], p["Large"], p["TMCmu"])
    assert rc == 95
    return su, rf


def direction(rng, host, f):
    pd = rng.normal(size=host.nCells)
    md = rng.normal(size=host.nFaces)
    md[np.abs(f.massFlux) < 1e-5] = 0.0
    n = np.sqrt(pd @ pd + md @ md)
    return pd / n, md / n


def tangent(m, host, f, p, nG, pd, md):
    g = gradient(m, host, f, f.phi, nG)
    rc, sd, rd = po.scalar_tangent_full(m, nG, f.phi, f.visEff, f.massFlux, g, f.te, f.den, f.cpB, f.visEffB, pd,
                                        md, p)
    assert rc == 95
    return sd, rd


def adjoint(m, host, f, p, nG, sb, rb):
    g = gradient(m, host, f, f.phi, nG)
    rc, pb, mb = po.scalar_adjoint_full(m, nG, f.phi, f.visEff, f.massFlux, g, f.te, f.den, f.cpB, f.visEffB, sb,
                                        rb, p)
    assert rc == 95
    return pb, mb


@pytest.mark.parametrize("nG", [1, 2, 3])
@pytest.mark.parametrize("variant", ["scalar", "scalar_T", "scalar_wall"])
@pytest.mark.parametrize("name", list(MESHES))
def test_tangent_matches_central_differences(name, variant, nG):
    host, m, f = case(name)
    p = params(variant)
    rng = np.random.default_rng(100 * nG + len(variant))
    eps = 1e-6
    for _ in range(3):
        pd, md = direction(rng, host, f)
        sd, rd = tangent(m, host, f, p, nG, pd, md)
        sp, rp = forward(m, host, f, f.phi + eps * pd, f.massFlux + eps * md, p, nG)
        sm, rm = forward(m, host, f, f.phi - eps * pd, f.massFlux - eps * md, p, nG)
        fd_s, fd_r = (sp - sm) / (2 * eps), (rp - rm) / (2 * eps)
        assert np.max(np.abs(fd_s - sd)) <= 1e-6 * max(1.0, np.max(np.abs(sd))), (name, variant, nG)
        assert np.max(np.abs(fd_r - rd)) <= 1e-6 * max(1.0, np.max(np.abs(rd))), (name, variant, nG)


@pytest.mark.parametrize("nG", [1, 2, 3])
@pytest.mark.parametrize("variant", ["scalar", "scalar_T", "scalar_wall"])
@pytest.mark.parametrize("name", list(MESHES))
def test_adjoint_dot_product(name, variant, nG):
    host, m, f = case(name)
    p = params(variant)
    rng = np.random.default_rng(7 + nG)
    for _ in range(3):
        pd, md = direction(rng, host, f)
        sb = rng.normal(size=host.nCells)
        rb = rng.normal(size=2 * host.nFaces)
        sd, rd = tangent(m, host, f, p, nG, pd, md)
        pb, mb = adjoint(m, host, f, p, nG, sb, rb)
        lhs = sb @ sd + rb @ rd
        rhs = pb @ pd + mb @ md
        scale = np.abs(sb * sd).sum() + np.abs(rb * rd).sum()
        assert abs(lhs - rhs) <= 1e-12 * scale, (name, variant, nG, lhs, rhs)


def test_typed_terms_are_live():
    """The typed branches change the adjoint: the enthalpy wall adds Hcoef terms and an INLET that carries
    PhiFace couples faces after it, so the three variants give three different PhiBar."""
    host, m, f = case("typed_perm")
    rng = np.random.default_rng(3)
    sb = rng.normal(size=host.nCells)
    rb = rng.normal(size=2 * host.nFaces)
    pbs = [adjoint(m, host, f, params(v), 2, sb, rb)[0] for v in ("scalar", "scalar_T", "scalar_wall")]
    assert not np.array_equal(pbs[0], pbs[1]) and not np.array_equal(pbs[1], pbs[2])


def test_default_mesh_one_sweep_is_the_nGradient_1_adjoint():
    host, m, f = case("plain")
    p = params("scalar")
    rng = np.random.default_rng(4)
    sb = rng.normal(size=host.nCells)
    rb = rng.normal(size=2 * host.nFaces)
    rc, pb1, mb1 = po.scalar_adjoint(m, f.phi, f.visEff, f.massFlux, sb, rb, p)
    assert rc == 95
    pb, mb = adjoint(m, host, f, p, 1, sb, rb)
    assert np.array_equal(pb.view(np.uint64), pb1.view(np.uint64))
    assert np.array_equal(mb.view(np.uint64), mb1.view(np.uint64))

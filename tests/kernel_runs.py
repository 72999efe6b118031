"""Runs each fvm kernel through an API module with the reference's signatures.

``api`` is either cupcfd_hip.fvm (the HIP library, drop-in mode) or
pyoracle (the CPU restatement); both take the same arguments, so the same
driver feeds identical inputs to both and returns every output array,
including the mesh state the kernel mutates.
"""
from __future__ import annotations

import numpy as np

from cupcfd_hip import workload as W

KERNELS = ["gradient", "gradient3", "mass", "mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5", "mass_region",
           "scalar", "scalar_t", "scalar_wall", "uvw", "uvw_region1", "uvw_bnds1", "visc_cell1", "visc_region",
           "visc_boundary", "visc_cell2"]


def run(name, api, mesh, f, state_get):
    """Run kernel ``name``; ``f`` is a Fields (already cast, copied by the caller).

    Returns (eCode, {output name: array}). ``state_get()`` returns the mesh state
    after the call (the oracle's own copy, or the host view the drop-in wrote)."""
    C = f.phi.shape[0]
    dt = f.phi.dtype
    out = {}
    if name in ("gradient", "gradient3"):
        g = np.full((C, 3), 7.0, dt)
        go = np.full((C, 3), 7.0, dt)
        rc = api.GradientPhiGaussDolfyn(mesh, 1 if name == "gradient" else 3, f.phi, f.phiB, g, go)
        out = dict(dPhidx=g, dPhidxo=go)
    elif name == "mass":
        cnt = np.zeros(4, np.int32)
        c = [cnt[i:i + 1] for i in range(4)]
        p = W.MASS
        rc = api.FluxMassDolfynFaceLoop(mesh, f.dudx, f.dvdx, f.dwdx, f.dpdx, f.den, f.denB, f.u, f.v, f.w,
                                        f.massFlux, f.p, f.ar, f.su, f.rface, p["small"], c[0], c[1], c[2], c[3],
                                        1, 1, 1, 1, f.te, f.teB, f.ed, f.edB, f.visEff, f.visEffB, f.t, f.tB)
        out = dict(massFlux=f.massFlux, rface=f.rface, su=f.su, denB=f.denB, teB=f.teB, edB=f.edB,
                   visEffB=f.visEffB, tB=f.tB, counters=cnt)
    elif name == "mass_bl1":
        fl = np.zeros(1, dt)
        rc = api.FluxMassDolfynBoundaryLoop1(mesh, f.massFlux, fl)
        out = dict(flowin=fl)
    elif name == "mass_bl2":
        fl = np.zeros(1, dt)
        rc = api.FluxMassDolfynBoundaryLoop2(mesh, f.massFlux, f.flowRegion, fl)
        out = dict(flowout=fl, flowRegion=f.flowRegion)
    elif name == "mass_bl3":
        r = np.zeros(1, dt)
        rc = api.FluxMassDolfynBoundaryLoop3(mesh, W.MASS_LOOPS["flowin"], r)
        out = dict(ratearea=r)
    elif name == "mass_bl4":
        fl = np.zeros(1, dt)
        rc = api.FluxMassDolfynBoundaryLoop4(mesh, f.massFlux, f.uB, f.vB, f.wB, f.denB, W.MASS_LOOPS["ratearea"],
                                             fl)
        out = dict(massFlux=f.massFlux, uB=f.uB, vB=f.vB, wB=f.wB, flowout=fl)
    elif name == "mass_bl5":
        fl = np.full(1, 0.25, dt)
        q = W.MASS_LOOPS
        rc = api.FluxMassDolfynBoundaryLoop5(mesh, f.massFlux, f.su, f.uB, f.vB, f.wB, q["fact"], q["solveU"],
                                             q["solveV"], q["solveW"], f.flowFact, fl)
        out = dict(massFlux=f.massFlux, su=f.su, uB=f.uB, vB=f.vB, wB=f.wB, flowout2=fl)
    elif name == "mass_region":
        rc = api.FluxMassDolfynRegionLoop(mesh, f.flowFact, f.flowRegion, 0.7)
        out = dict(flowFact=f.flowFact)
    elif name in ("scalar", "scalar_t", "scalar_wall"):
        p = dict(W.SCALAR)
        if name == "scalar_t":      # ivar == VarT without enthalpy: INLET keeps the carried PhiFace
            p["ivar"] = p["VarT"]
        if name == "scalar_wall":   # enthalpy wall-function branch
            p["ivar"] = p["VarT"]
            p["SolveEnthalpy"] = 1
        rc = api.FluxScalarDolfynFaceLoop(mesh, f.phi, f.phiB, f.visEff, f.au, f.su, f.den, f.massFlux, f.te, f.cpB,
                                          f.visEffB, f.rface, f.dphidx, p["SolveTurb"], p["SolveEnthalpy"],
                                          p["vislam"], p["ivar"], p["VarT"], p["Sigma_T"], p["Prandtl"], p["VarTE"],
                                          p["Sigma_k"], p["VarED"], p["Sigma_e"], p["Sigma_s"], p["Schmidt"],
                                          p["GammaBlend"], p["Small"], p["Large"], p["TMCmu"])
        out = dict(Su=f.su, Au=f.au, PhiB=f.phiB, RFace=f.rface)
    elif name == "uvw":
        q = W.UVW
        rc = api.FluxUVWDolfynFaceLoop1(mesh, q["gammaBlend"], q["small"], f.u, f.v, f.w, f.uB, f.vB, f.wB,
                                        f.visEff, f.visEffB, f.massFlux, f.dudx, f.dvdx, f.dwdx, f.rface, f.su, f.sv,
                                        f.sw, f.au, f.av, f.aw)
        out = dict(su=f.su, sv=f.sv, sw=f.sw, au=f.au, av=f.av, aw=f.aw, uB=f.uB, vB=f.vB, wB=f.wB,
                   massFlux=f.massFlux, rFace=f.rface)
    elif name == "uvw_region1":
        rc = api.FluxUVWDolfynRegionLoop1(mesh)
    elif name == "uvw_bnds1":
        rc = api.FluxUVWDolfynBndsLoop1(mesh)
    elif name == "visc_cell1":
        q = W.VISC["cell1"]
        rc = api.calculateViscosityDolfynCellLoop1(mesh, q["small"], q["cmu"], q["visURF"], q["visLam"], f.te, f.ed,
                                                   f.den, f.visEff)
        out = dict(visEff=f.visEff)
    elif name == "visc_region":
        rc = api.calculateViscosityDolfynRegionLoop(W.VISC["region"]["kappa"], mesh)
    elif name == "visc_boundary":
        q = W.VISC["boundary"]
        rc = api.calculateViscosityDolfynBoundaryLoop(mesh, q["tmcmu"], q["large"], q["small"], q["kappa"],
                                                      q["visLam"], f.te, f.den, f.visEff, f.visEffB)
        out = dict(visEffB=f.visEffB)
    elif name == "visc_cell2":
        rc = api.calculateViscosityDolfynCellLoop2(mesh, W.VISC["cell2"]["visLam"], f.visEff, f.visEffB)
        out = dict(visEff=f.visEff, visEffB=f.visEffB)
    else:
        raise KeyError(name)
    for k, v in state_get().items():
        out["mesh." + k] = v
    return rc, out


def seed_state(host, rng_seed=5):
    """A non-trivial mesh state (y+, u+, shear, force tangents) so kernels that
    read it are exercised."""
    rng = np.random.default_rng(rng_seed)
    B, R = host.nBoundaries, host.nRegions
    return {"bndYPlus": rng.uniform(1, 30, B), "bndUPlus": rng.uniform(1, 20, B), "bndQ": np.zeros(B),
            "bndH": np.zeros(B), "bndT": np.zeros(B), "bndShear": rng.uniform(-1, 1, (B, 3)),
            "regYLog": np.full(R, 11.0), "regForceTangent": rng.uniform(-1, 1, (R, 3))}


def compare(a, b):
    """(bit-identical?, max-norm relative difference ||a-b||inf / ||b||inf).

    NaN and infinity positions must coincide; -0 and +0 count as equal."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.size == 0:
        return True, 0.0
    na, nb = np.isnan(a), np.isnan(b)
    assert (na == nb).all(), "NaN positions differ"
    a, b = a[~na], b[~nb]
    inf = np.isinf(a) | np.isinf(b)
    assert (a[inf] == b[inf]).all(), "infinities differ"
    a, b = a[~inf], b[~inf]
    same = bool(np.array_equal(a, b))
    if a.size == 0:
        return same, 0.0
    scale = float(np.abs(b).max())
    diff = float(np.abs(a - b).max())
    return same, (diff / scale if scale > 0 else diff)

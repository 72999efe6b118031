"""ORACLE — test infrastructure only.

ctypes wrapper over liboracle_fvm.so (oracle_fvm.c, the sequential CPU
restatement of CUP-CFD's fvm kernels). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; the product never does.

Functions take the same numpy arrays as cupcfd_hip.fvm (mesh numbering,
(n, 3) vectors) and an ``OracleMesh`` built from a HostMesh's arrays. The
oracle mesh holds its own copy of the mutable mesh state so that oracle and
GPU runs can be compared side by side.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CUPCFD_ORACLE_LIBRARY names another build of it (the sanitizer build of oracle/Makefile)
LIB = os.environ.get("CUPCFD_ORACLE_LIBRARY") or os.path.join(HERE, "liboracle_fvm.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ct.CDLL(LIB)
        _lib.or_set_carry_init.argtypes = [ct.c_double]
        _lib.or_set_carry_init.restype = None
    return _lib


def _fields():
    P = ct.c_void_p
    names = ["faceCell1", "faceCell2", "faceBnd", "faceLambda", "faceNorm", "faceCenter", "faceXpac", "faceXnac",
             "faceRLencos", "faceArea", "cellCenter", "cellVolume", "bndFace", "bndRegion", "bndDistance",
             "bndYPlus", "bndUPlus", "bndShear", "bndQ", "bndH", "bndT", "regType", "regStd", "regFlux",
             "regAdiab", "regYLog", "regELog", "regDensity", "regTurbKE", "regTurbDiss", "regSplvl", "regDen",
             "regR", "regT", "regForceTangent", "regUVW"]
    return [("nCells", ct.c_int), ("nFaces", ct.c_int), ("nBoundaries", ct.c_int), ("nRegions", ct.c_int)] + \
        [(n, P) for n in names], names


class _OrMesh(ct.Structure):
    _fields_, NAMES = _fields()


FLOAT_KEYS = {"faceLambda", "faceNorm", "faceCenter", "faceXpac", "faceXnac", "faceRLencos", "faceArea",
              "cellCenter", "cellVolume", "bndDistance", "bndYPlus", "bndUPlus", "bndShear", "bndQ", "bndH", "bndT",
              "regYLog", "regELog", "regDensity", "regTurbKE", "regTurbDiss", "regSplvl", "regDen", "regR", "regT",
              "regForceTangent", "regUVW"}
STATE_KEYS = ("bndYPlus", "bndUPlus", "bndQ", "bndH", "bndT", "bndShear", "regYLog", "regForceTangent")


class OracleMesh:
    """Arrays of a HostMesh (``host.a``) in precision 64 or 32, state copied."""

    def __init__(self, host, precision=64):
        self.precision = precision
        self.dtype = np.float64 if precision == 64 else np.float32
        self.sfx = "f64" if precision == 64 else "f32"
        self.arr = {}
        for k in _OrMesh.NAMES:
            a = host.a[k]
            if k in FLOAT_KEYS:
                a = np.ascontiguousarray(a, dtype=self.dtype).copy()
            else:
                a = np.ascontiguousarray(a).copy()
            self.arr[k] = a
        self.s = _OrMesh(host.nCells, host.nFaces, host.nBoundaries, host.nRegions,
                         *[self.arr[k].ctypes.data for k in _OrMesh.NAMES])
        self.nCells, self.nFaces, self.nBoundaries, self.nRegions = (host.nCells, host.nFaces, host.nBoundaries,
                                                                    host.nRegions)

    def state(self):
        return {k: self.arr[k].copy() for k in STATE_KEYS}

    def set_state(self, st):
        for k, v in st.items():
            self.arr[k][...] = v

    @property
    def ref(self):
        return ct.byref(self.s)


def set_carry_init(v):
    lib().or_set_carry_init(float(v))


def _P(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


def _N(a, per=1):
    return ct.c_int(0 if a is None else a.size // per)


def _A(a):
    return [_P(a), _N(a)]


def _V(a):
    return [_P(a), ct.c_int(0 if a is None else a.shape[0])]


def _fn(m, name):
    f = getattr(lib(), f"or_{name}_{m.sfx}")
    f.restype = ct.c_int
    return f


def _T(m, x):
    return ct.c_double(x) if m.precision == 64 else ct.c_float(x)


def _I(x):
    return ct.c_int(int(x))


def GradientPhiGaussDolfyn(m, nGradient, phiCell, phiBoundary, dPhidxCell, dPhidxoCell):
    return _fn(m, "gradient_phi_gauss_dolfyn")(m.ref, _I(nGradient), *_A(phiCell), *_A(phiBoundary),
                                               *_V(dPhidxCell), *_V(dPhidxoCell))


def FluxMassDolfynFaceLoop(m, dudx, dvdx, dwdx, dpdx, denCell, denBoundary, uCell, vCell, wCell, massFlux, p, ar, su,
                           rface, small, icinl, icout, icsym, icwal, solveTurbEnergy, solveTurbDiss, solveVisc,
                           solveEnthalpy, teCell, teBoundary, edCell, edBoundary, viseffCell, viseffBoundary, tCell,
                           tBoundary):
    return _fn(m, "flux_mass_face_loop")(
        m.ref, *_V(dudx), *_V(dvdx), *_V(dwdx), *_V(dpdx), *_A(denCell), *_A(denBoundary), *_A(uCell), *_A(vCell),
        *_A(wCell), *_A(massFlux), *_A(p), *_A(ar), *_A(su), *_A(rface), _T(m, small), _P(icinl), _P(icout),
        _P(icsym), _P(icwal), _I(solveTurbEnergy), _I(solveTurbDiss), _I(solveVisc), _I(solveEnthalpy),
        *_A(teCell), *_A(teBoundary), *_A(edCell), *_A(edBoundary), *_A(viseffCell), *_A(viseffBoundary),
        *_A(tCell), *_A(tBoundary))


def FluxMassDolfynBoundaryLoop1(m, massFlux, flowin):
    return _fn(m, "flux_mass_boundary_loop1")(m.ref, *_A(massFlux), _P(flowin))


def FluxMassDolfynBoundaryLoop2(m, massFlux, flowRegion, flowout):
    return _fn(m, "flux_mass_boundary_loop2")(m.ref, *_A(massFlux), *_A(flowRegion), _P(flowout))


def FluxMassDolfynBoundaryLoop3(m, flowin, ratearea):
    return _fn(m, "flux_mass_boundary_loop3")(m.ref, _T(m, flowin), _P(ratearea))


def FluxMassDolfynBoundaryLoop4(m, massFlux, uBoundary, vBoundary, wBoundary, denBoundary, ratearea, flowout):
    return _fn(m, "flux_mass_boundary_loop4")(m.ref, *_A(massFlux), *_A(uBoundary), *_A(vBoundary),
                                              *_A(wBoundary), *_A(denBoundary), _T(m, ratearea), _P(flowout))


def FluxMassDolfynBoundaryLoop5(m, massFlux, su, uBoundary, vBoundary, wBoundary, fact, solveU, solveV, solveW,
                                flowFact, flowout2):
    return _fn(m, "flux_mass_boundary_loop5")(m.ref, *_A(massFlux), *_A(su), *_A(uBoundary), *_A(vBoundary),
                                              *_A(wBoundary), _T(m, fact), _I(solveU), _I(solveV), _I(solveW),
                                              *_A(flowFact), _P(flowout2))


def FluxMassDolfynRegionLoop(m, flowFact, flowRegion, flowIn):
    return _fn(m, "flux_mass_region_loop")(m.ref, *_A(flowFact), *_A(flowRegion), _T(m, flowIn))


def FluxScalarDolfynFaceLoop(m, PhiCell, PhiBoundary, VisEff, Au, Su, Den, MassFlux, TE, CpBoundary, visEffBoundary,
                             RFace, dPhidx, SolveTurb, SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE,
                             Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, Small, Large, TMCmu):
    T = lambda x: _T(m, x)  # noqa: E731
    return _fn(m, "flux_scalar_face_loop")(
        m.ref, *_A(PhiCell), *_A(PhiBoundary), *_A(VisEff), *_A(Au), *_A(Su), *_A(Den), *_A(MassFlux), *_A(TE),
        *_A(CpBoundary), *_A(visEffBoundary), *_A(RFace), *_V(dPhidx), _I(SolveTurb), _I(SolveEnthalpy), T(vislam),
        _I(ivar), _I(VarT), T(Sigma_T), T(Prandtl), _I(VarTE), T(Sigma_k), _I(VarED), T(Sigma_e), T(Sigma_s),
        T(Schmidt), T(GammaBlend), T(Small), T(Large), T(TMCmu))


def FluxUVWDolfynFaceLoop1(m, gammaBlend, small, uCell, vCell, wCell, uBoundary, vBoundary, wBoundary, visEffCell,
                           visEffBoundary, massFlux, dudx, dvdx, dwdx, rFace, su, sv, sw, au, av, aw):
    return _fn(m, "flux_uvw_face_loop1")(
        m.ref, _T(m, gammaBlend), _T(m, small), *_A(uCell), *_A(vCell), *_A(wCell), *_A(uBoundary), *_A(vBoundary),
        *_A(wBoundary), *_A(visEffCell), *_A(visEffBoundary), *_A(massFlux), *_V(dudx), *_V(dvdx), *_V(dwdx),
        *_A(rFace), *_A(su), *_A(sv), *_A(sw), *_A(au), *_A(av), *_A(aw))


def FluxUVWDolfynRegionLoop1(m):
    f = _fn(m, "flux_uvw_region_loop1")
    f.restype = None
    f(m.ref)
    return 95


def FluxUVWDolfynBndsLoop1(m):
    f = _fn(m, "flux_uvw_bnds_loop1")
    f.restype = None
    f(m.ref)
    return 95


def calculateViscosityDolfynCellLoop1(m, small, cmu, visURF, visLam, TE, ED, den, visEff):
    T = lambda x: _T(m, x)  # noqa: E731
    return _fn(m, "viscosity_cell_loop1")(m.ref, T(small), T(cmu), T(visURF), T(visLam), *_A(TE), *_A(ED), *_A(den),
                                          *_A(visEff))


def calculateViscosityDolfynRegionLoop(kappa, m):
    f = _fn(m, "viscosity_region_loop")
    f.restype = None
    f(_T(m, kappa), m.ref)
    return 95


def calculateViscosityDolfynBoundaryLoop(m, tmcmu, large, small, kappa, visLam, TE, den, visEffCell, visEffBoundary):
    T = lambda x: _T(m, x)  # noqa: E731
    return _fn(m, "viscosity_boundary_loop")(m.ref, T(tmcmu), T(large), T(small), T(kappa), T(visLam), *_A(TE),
                                             *_A(den), *_A(visEffCell), *_A(visEffBoundary))


def calculateViscosityDolfynCellLoop2(m, visLam, visEffCell, visEffBoundary):
    return _fn(m, "viscosity_cell_loop2")(m.ref, _T(m, visLam), *_A(visEffCell), *_A(visEffBoundary))


class ScalarParams(ct.Structure):
    """or_scalar_params: the scalar face loop's parameters its tangent and adjoint read."""
    _fields_ = [("SolveTurb", ct.c_int), ("ivar", ct.c_int), ("VarT", ct.c_int), ("VarTE", ct.c_int),
                ("VarED", ct.c_int), ("vislam", ct.c_double), ("Sigma_T", ct.c_double), ("Prandtl", ct.c_double),
                ("Sigma_k", ct.c_double), ("Sigma_e", ct.c_double), ("Sigma_s", ct.c_double),
                ("Schmidt", ct.c_double), ("GammaBlend", ct.c_double), ("SolveEnthalpy", ct.c_int),
                ("Small", ct.c_double), ("TMCmu", ct.c_double)]

    @classmethod
    def of(cls, p):
        return cls(*[p[n] for n, _ in cls._fields_])


def scalar_tangent(m, phi, visEff, massFlux, phidot, mfdot, params):
    """(sudot, rfacedot): tangent of gradient (nGradient = 1) + scalar face loop w.r.t. (phi, massFlux)."""
    assert m.precision == 64
    sd = np.zeros(m.nCells)
    rd = np.zeros(2 * m.nFaces)
    P = ScalarParams.of(params)
    rc = _fn(m, "scalar_tangent")(m.ref, _P(phi), _P(visEff), _P(massFlux), _P(phidot), _P(mfdot), ct.byref(P),
                                  _P(sd), _P(rd))
    return rc, sd, rd


def scalar_adjoint(m, phi, visEff, massFlux, subar, rbar, params):
    """(phibar, mfbar): adjoint of gradient (nGradient = 1) + scalar face loop w.r.t. (phi, massFlux)."""
    assert m.precision == 64
    pb = np.zeros(m.nCells)
    mb = np.zeros(m.nFaces)
    P = ScalarParams.of(params)
    rc = _fn(m, "scalar_adjoint")(m.ref, _P(phi), _P(visEff), _P(massFlux), _P(subar), _P(rbar), ct.byref(P),
                                  _P(pb), _P(mb))
    return rc, pb, mb


def scalar_tangent_full(m, nGradient, phi, visEff, massFlux, dPhidx, TE, Den, CpB, visEffB, phidot, mfdot, params):
    """(rc, sudot, rfacedot): tangent of gradient (nGradient sweeps) + scalar face loop w.r.t. (phi, massFlux),
    on any mesh; dPhidx is the forward gradient the scalar loop read."""
    assert m.precision == 64
    sd = np.zeros(m.nCells)
    rd = np.zeros(2 * m.nFaces)
    P = ScalarParams.of(params)
    keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (phi, visEff, massFlux, dPhidx, TE, Den, CpB,
                                                                visEffB, phidot, mfdot)]
    rc = _fn(m, "scalar_tangent_full")(m.ref, ct.c_int(nGradient), *[_P(x) for x in keep], ct.byref(P), _P(sd),
                                       _P(rd))
    return rc, sd, rd


def scalar_adjoint_full(m, nGradient, phi, visEff, massFlux, dPhidx, TE, Den, CpB, visEffB, subar, rbar, params):
    """(rc, phibar, mfbar): adjoint of gradient (nGradient sweeps) + scalar face loop w.r.t. (phi, massFlux),
    on any mesh; dPhidx is the forward gradient the scalar loop read."""
    assert m.precision == 64
    pb = np.zeros(m.nCells)
    mb = np.zeros(m.nFaces)
    P = ScalarParams.of(params)
    keep = [np.ascontiguousarray(x, dtype=np.float64) for x in (phi, visEff, massFlux, dPhidx, TE, Den, CpB,
                                                                visEffB, subar, rbar)]
    rc = _fn(m, "scalar_adjoint_full")(m.ref, ct.c_int(nGradient), *[_P(x) for x in keep], ct.byref(P), _P(pb),
                                       _P(mb))
    return rc, pb, mb

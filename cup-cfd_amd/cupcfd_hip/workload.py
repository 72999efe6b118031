"""Seeded synthetic inputs for the fvm kernels, and BenchmarkKernels' parameter values.

Parameter values follow src/benchmarks/implementation/component/BenchmarkKernels.cpp
(gradient :145, mass face loop :248-256, mass boundary/region loops :423, :474,
:537-540, :599, scalar :677-695, UVW :761-762, viscosity :952-955, :1010,
:1042-1046, :1096). The reference fills its arrays from an unseeded
randomUniform(1e-6, 1e-2) and leaves several uninitialised; here every array
is drawn from one seeded numpy stream instead, with the field shapes named
for each workload:
  phi(x) = sin(2 pi x) cos(2 pi y) + 0.1 U(-1,1) at cell centres (phiB at
  boundary face centres), velocity U(-1,1)^3, density 1, pressure U(-1,1),
  ar U(0.5,1.5), effective viscosity 1e-3, gradients U(-1,1)^3,
  turbulence fields U(1e-6,1e-2).
"""
from __future__ import annotations

import numpy as np

MASS = dict(small=1e-18, solveTurbEnergy=0, solveTurbDiss=0, solveVisc=1, solveEnthalpy=0)
MASS_LOOPS = dict(flowin=1.1, ratearea=1.0, fact=1.1, solveU=1, solveV=1, solveW=1, flowIn=0.0)
SCALAR = dict(SolveTurb=1, SolveEnthalpy=0, vislam=0.001, ivar=1, VarT=2, Sigma_T=1.4, Prandtl=1.2e-3, VarTE=3,
              Sigma_k=1.5, VarED=4, Sigma_e=1.6, Sigma_s=1.7, Schmidt=1.4e-3, GammaBlend=1.1, Small=1e-18,
              Large=1e18, TMCmu=1.8)
UVW = dict(gammaBlend=0.9, small=1e-18)
VISC = dict(cell1=dict(small=1e-18, cmu=1e-4, visURF=0.232, visLam=0.00547), region=dict(kappa=1.1),
            boundary=dict(tmcmu=1.0, large=1e18, small=1e-18, kappa=1.1, visLam=1.2), cell2=dict(visLam=1.1))


class Fields:
    """All kernel inputs for one mesh, float64, in the mesh's numbering."""

    def __init__(self, host, seed=1234, viscosity=1e-3):
        rng = np.random.default_rng(seed)
        C, F, B, R = host.nCells, host.nFaces, host.nBoundaries, host.nRegions
        cc = host.a["cellCenter"]
        fc = host.a["faceCenter"][host.a["bndFace"]] if B else np.zeros((0, 3))
        U = lambda lo, hi, *shape: rng.uniform(lo, hi, size=shape)  # noqa: E731
        two_pi = 2.0 * np.pi
        self.phi = np.sin(two_pi * cc[:, 0]) * np.cos(two_pi * cc[:, 1]) + 0.1 * U(-1, 1, C)
        self.phiB = np.sin(two_pi * fc[:, 0]) * np.cos(two_pi * fc[:, 1]) + 0.1 * U(-1, 1, B)
        self.u, self.v, self.w = U(-1, 1, C), U(-1, 1, C), U(-1, 1, C)
        self.den = np.ones(C)
        self.p = U(-1, 1, C)
        self.ar = U(0.5, 1.5, C)
        self.visEff = np.full(C, viscosity)
        self.visEffB = np.full(B, viscosity)
        self.dudx, self.dvdx, self.dwdx, self.dpdx = (U(-1, 1, C, 3) for _ in range(4))
        self.dphidx = U(-1, 1, C, 3)
        self.massFlux = U(-1, 1, F)
        self.te, self.ed, self.t = U(1e-6, 1e-2, C), U(1e-6, 1e-2, C), U(1e-6, 1e-2, C)
        self.teB, self.edB, self.tB = U(1e-6, 1e-2, B), U(1e-6, 1e-2, B), U(1e-6, 1e-2, B)
        self.denB = np.ones(B)
        self.cpB = U(0.5, 1.5, B)
        self.uB, self.vB, self.wB = U(-1, 1, B), U(-1, 1, B), U(-1, 1, B)
        self.su, self.sv, self.sw = U(-1, 1, C), U(-1, 1, C), U(-1, 1, C)
        self.au, self.av, self.aw = U(0.5, 1.5, C), U(0.5, 1.5, C), U(0.5, 1.5, C)
        self.rface = np.zeros(2 * F)
        self.flowRegion = U(0.5, 1.5, R)
        self.flowFact = U(0.5, 1.5, R)

    def cast(self, dtype):
        """An independent copy in ``dtype`` (never aliases this object's arrays)."""
        out = Fields.__new__(Fields)
        for k, v in self.__dict__.items():
            out.__dict__[k] = np.array(v, dtype=dtype, order="C", copy=True)
        return out

    def copy(self):
        return self.cast(next(iter(self.__dict__.values())).dtype)

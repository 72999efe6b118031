"""Python mirror of namespace cupcfd::fvm (include/fvm/implementation/component/*Kernels.h).

Each function keeps the reference's name and argument order; the
(pointer, length) pairs collapse to one numpy array whose length is passed
through. Arrays are host numpy arrays in the mesh's numbering (drop-in mode,
synchronous); 3-vectors are (n, 3) arrays. Every function returns the
eCodes value of the C ABI (E_SUCCESS = 95).

``resident`` holds the same kernels over device tensors in the device mesh's
internal order (cupcfd_dev_*), launched on a HIP stream.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

__all__ = [
    "GradientPhiGaussDolfyn", "FluxMassDolfynFaceLoop", "FluxMassDolfynBoundaryLoop1",
    "FluxMassDolfynBoundaryLoop2", "FluxMassDolfynBoundaryLoop3", "FluxMassDolfynBoundaryLoop4",
    "FluxMassDolfynBoundaryLoop5", "FluxMassDolfynRegionLoop", "FluxScalarDolfynFaceLoop",
    "FluxUVWDolfynFaceLoop1", "FluxUVWDolfynRegionLoop1", "FluxUVWDolfynBndsLoop1",
    "calculateViscosityDolfynCellLoop1", "calculateViscosityDolfynRegionLoop",
    "calculateViscosityDolfynBoundaryLoop", "calculateViscosityDolfynCellLoop2", "resident",
]


def _P(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], "C-contiguous numpy arrays expected"
    return a.ctypes.data_as(ct.c_void_p)


def _N(a, per=1):
    return ct.c_int32(0 if a is None else a.size // per)


def _I(x):
    return ct.c_int32(int(x))


def _A(m, a):  # array + length
    return [_P(a), _N(a)]


def _V(m, a):  # 3-vector array + element count
    if a is None:
        return [m.V3(0, 0), ct.c_int32(0)]
    return [m.v3(a), ct.c_int32(a.shape[0])]


def _raw(m, name, *args):
    from . import lib
    return getattr(lib(), f"cupcfd_fvm_{name}_{m.sfx}")(*args)


def GradientPhiGaussDolfyn(mesh, nGradient, phiCell, phiBoundary, dPhidxCell, dPhidxoCell):
    """GradientKernels.h:48-54"""
    return _raw(mesh, "gradient_phi_gauss_dolfyn", mesh.handle, _I(nGradient), *_A(mesh, phiCell),
                *_A(mesh, phiBoundary), *_V(mesh, dPhidxCell), *_V(mesh, dPhidxoCell))


def FluxMassDolfynFaceLoop(mesh, dudx, dvdx, dwdx, dpdx, denCell, denBoundary, uCell, vCell, wCell, massFlux, p, ar,
                           su, rface, small, icinl, icout, icsym, icwal, solveTurbEnergy, solveTurbDiss, solveVisc,
                           solveEnthalpy, teCell, teBoundary, edCell, edBoundary, viseffCell, viseffBoundary, tCell,
                           tBoundary):
    """MassKernels.h:48-81. icinl..icwal are 1-element int32 arrays (incremented in place)."""
    T = mesh.T
    return _raw(mesh, "flux_mass_face_loop", mesh.handle, *_V(mesh, dudx), *_V(mesh, dvdx), *_V(mesh, dwdx),
                *_V(mesh, dpdx), *_A(mesh, denCell), *_A(mesh, denBoundary), *_A(mesh, uCell), *_A(mesh, vCell),
                *_A(mesh, wCell), *_A(mesh, massFlux), *_A(mesh, p), *_A(mesh, ar), *_A(mesh, su),
                *_A(mesh, rface), T(small), _P(icinl), _P(icout), _P(icsym), _P(icwal), _I(solveTurbEnergy),
                _I(solveTurbDiss), _I(solveVisc), _I(solveEnthalpy), *_A(mesh, teCell), *_A(mesh, teBoundary),
                *_A(mesh, edCell), *_A(mesh, edBoundary), *_A(mesh, viseffCell), *_A(mesh, viseffBoundary),
                *_A(mesh, tCell), *_A(mesh, tBoundary))


def FluxMassDolfynBoundaryLoop1(mesh, massFlux, flowin):
    """MassKernels.h:90-94. flowin: 1-element array (output)."""
    return _raw(mesh, "flux_mass_boundary_loop1", mesh.handle, *_A(mesh, massFlux), _P(flowin))


def FluxMassDolfynBoundaryLoop2(mesh, massFlux, flowRegion, flowout):
    """MassKernels.h:103-108"""
    return _raw(mesh, "flux_mass_boundary_loop2", mesh.handle, *_A(mesh, massFlux), *_A(mesh, flowRegion),
                _P(flowout))


def FluxMassDolfynBoundaryLoop3(mesh, flowin, ratearea):
    """MassKernels.h:117-120"""
    return _raw(mesh, "flux_mass_boundary_loop3", mesh.handle, mesh.T(flowin), _P(ratearea))


def FluxMassDolfynBoundaryLoop4(mesh, massFlux, uBoundary, vBoundary, wBoundary, denBoundary, ratearea, flowout):
    """MassKernels.h:129-138"""
    return _raw(mesh, "flux_mass_boundary_loop4", mesh.handle, *_A(mesh, massFlux), *_A(mesh, uBoundary),
                *_A(mesh, vBoundary), *_A(mesh, wBoundary), *_A(mesh, denBoundary), mesh.T(ratearea), _P(flowout))


def FluxMassDolfynBoundaryLoop5(mesh, massFlux, su, uBoundary, vBoundary, wBoundary, fact, solveU, solveV, solveW,
                                flowFact, flowout2):
    """MassKernels.h:147-157"""
    return _raw(mesh, "flux_mass_boundary_loop5", mesh.handle, *_A(mesh, massFlux), *_A(mesh, su),
                *_A(mesh, uBoundary), *_A(mesh, vBoundary), *_A(mesh, wBoundary), mesh.T(fact), _I(solveU),
                _I(solveV), _I(solveW), *_A(mesh, flowFact), _P(flowout2))


def FluxMassDolfynRegionLoop(mesh, flowFact, flowRegion, flowIn):
    """MassKernels.h:166-171"""
    return _raw(mesh, "flux_mass_region_loop", mesh.handle, *_A(mesh, flowFact), *_A(mesh, flowRegion),
                mesh.T(flowIn))


def FluxScalarDolfynFaceLoop(mesh, PhiCell, PhiBoundary, VisEff, Au, Su, Den, MassFlux, TE, CpBoundary,
                             visEffBoundary, RFace, dPhidx, SolveTurb, SolveEnthalpy, vislam, ivar, VarT, Sigma_T,
                             Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, Small, Large,
                             TMCmu):
    """ScalarKernels.h:47-67"""
    T = mesh.T
    return _raw(mesh, "flux_scalar_face_loop", mesh.handle, *_A(mesh, PhiCell), *_A(mesh, PhiBoundary),
                *_A(mesh, VisEff), *_A(mesh, Au), *_A(mesh, Su), *_A(mesh, Den), *_A(mesh, MassFlux),
                *_A(mesh, TE), *_A(mesh, CpBoundary), *_A(mesh, visEffBoundary), *_A(mesh, RFace),
                *_V(mesh, dPhidx), _I(SolveTurb), _I(SolveEnthalpy), T(vislam), _I(ivar), _I(VarT), T(Sigma_T),
                T(Prandtl), _I(VarTE), T(Sigma_k), _I(VarED), T(Sigma_e), T(Sigma_s), T(Schmidt), T(GammaBlend),
                T(Small), T(Large), T(TMCmu))


def FluxUVWDolfynFaceLoop1(mesh, gammaBlend, small, uCell, vCell, wCell, uBoundary, vBoundary, wBoundary, visEffCell,
                           visEffBoundary, massFlux, dudx, dvdx, dwdx, rFace, su, sv, sw, au, av, aw):
    """UVWKernels.h:43-67"""
    T = mesh.T
    return _raw(mesh, "flux_uvw_face_loop1", mesh.handle, T(gammaBlend), T(small), *_A(mesh, uCell),
                *_A(mesh, vCell), *_A(mesh, wCell), *_A(mesh, uBoundary), *_A(mesh, vBoundary), *_A(mesh, wBoundary),
                *_A(mesh, visEffCell), *_A(mesh, visEffBoundary), *_A(mesh, massFlux), *_V(mesh, dudx),
                *_V(mesh, dvdx), *_V(mesh, dwdx), *_A(mesh, rFace), *_A(mesh, su), *_A(mesh, sv), *_A(mesh, sw),
                *_A(mesh, au), *_A(mesh, av), *_A(mesh, aw))


def FluxUVWDolfynRegionLoop1(mesh):
    """UVWKernels.h:72-73"""
    return _raw(mesh, "flux_uvw_region_loop1", mesh.handle)


def FluxUVWDolfynBndsLoop1(mesh):
    """UVWKernels.h:78-79"""
    return _raw(mesh, "flux_uvw_bnds_loop1", mesh.handle)


def calculateViscosityDolfynCellLoop1(mesh, small, cmu, visURF, visLam, TE, ED, den, visEff):
    """ViscosityKernels.h:38-48"""
    T = mesh.T
    return _raw(mesh, "viscosity_cell_loop1", mesh.handle, T(small), T(cmu), T(visURF), T(visLam), *_A(mesh, TE),
                *_A(mesh, ED), *_A(mesh, den), *_A(mesh, visEff))


def calculateViscosityDolfynRegionLoop(kappa, mesh):
    """ViscosityKernels.h:50-51"""
    return _raw(mesh, "viscosity_region_loop", mesh.T(kappa), mesh.handle)


def calculateViscosityDolfynBoundaryLoop(mesh, tmcmu, large, small, kappa, visLam, TE, den, visEffCell,
                                         visEffBoundary):
    """ViscosityKernels.h:53-60"""
    T = mesh.T
    return _raw(mesh, "viscosity_boundary_loop", mesh.handle, T(tmcmu), T(large), T(small), T(kappa), T(visLam),
                *_A(mesh, TE), *_A(mesh, den), *_A(mesh, visEffCell), *_A(mesh, visEffBoundary))


def calculateViscosityDolfynCellLoop2(mesh, visLam, visEffCell, visEffBoundary):
    """ViscosityKernels.h:62-67"""
    return _raw(mesh, "viscosity_cell_loop2", mesh.handle, mesh.T(visLam), *_A(mesh, visEffCell),
                *_A(mesh, visEffBoundary))


class resident:
    """cupcfd_dev_* over torch device tensors (internal order, packed xyz vectors).

    ``stream`` is a torch.cuda.Stream or None (the current stream)."""

    @staticmethod
    def _s(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return ct.c_void_p(s.cuda_stream)

    @staticmethod
    def _t(t):
        return None if t is None else ct.c_void_p(t.data_ptr())

    @staticmethod
    def call(mesh, name, *args):
        from . import lib, check
        fn = f"cupcfd_dev_{name}_{mesh.sfx}"
        check(fn, getattr(lib(), fn)(*args))

    @classmethod
    def gradient(cls, mesh, nGradient, phi, phiB, g, go, stream=None):
        cls.call(mesh, "gradient_phi_gauss_dolfyn", mesh.handle, _I(nGradient), cls._t(phi), cls._t(phiB),
                 cls._t(g), cls._t(go), cls._s(stream))

    @classmethod
    def mass(cls, mesh, dudx, dvdx, dwdx, dpdx, den, denB, u, v, w, mf, p, ar, su, rface, small=1e-18,
             counters=None, flags=(0, 0, 1, 0), teC=None, teB=None, edC=None, edB=None, visC=None, visB=None,
             tC=None, tB=None, stream=None):
        cnt = counters if counters is not None else np.zeros(4, np.int32)
        t = cls._t
        base = cnt.ctypes.data
        cls.call(mesh, "flux_mass_face_loop", mesh.handle, t(dudx), t(dvdx), t(dwdx), t(dpdx), t(den), t(denB),
                 t(u), t(v), t(w), t(mf), t(p), t(ar), t(su), t(rface), mesh.T(small), ct.c_void_p(base),
                 ct.c_void_p(base + 4), ct.c_void_p(base + 8), ct.c_void_p(base + 12), *[_I(x) for x in flags],
                 t(teC), t(teB), t(edC), t(edB), t(visC), t(visB), t(tC), t(tB), cls._s(stream))
        return cnt

    @classmethod
    def scalar(cls, mesh, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx, params,
               stream=None):
        t = cls._t
        T = mesh.T
        p = params
        cls.call(mesh, "flux_scalar_face_loop", mesh.handle, t(phi), t(phiB), t(visEff), t(au), t(su), t(den),
                 t(mf), t(te), t(cpB), t(visEffB), t(rface), t(dphidx), _I(p["SolveTurb"]), _I(p["SolveEnthalpy"]),
                 T(p["vislam"]), _I(p["ivar"]), _I(p["VarT"]), T(p["Sigma_T"]), T(p["Prandtl"]), _I(p["VarTE"]),
                 T(p["Sigma_k"]), _I(p["VarED"]), T(p["Sigma_e"]), T(p["Sigma_s"]), T(p["Schmidt"]),
                 T(p["GammaBlend"]), T(p["Small"]), T(p["Large"]), T(p["TMCmu"]), cls._s(stream))

    @classmethod
    def uvw(cls, mesh, gammaBlend, small, u, v, w, uB, vB, wB, visEff, visEffB, mf, dudx, dvdx, dwdx, rface, su, sv,
            sw, au, av, aw, stream=None):
        t = cls._t
        cls.call(mesh, "flux_uvw_face_loop1", mesh.handle, mesh.T(gammaBlend), mesh.T(small), t(u), t(v), t(w),
                 t(uB), t(vB), t(wB), t(visEff), t(visEffB), t(mf), t(dudx), t(dvdx), t(dwdx), t(rface), t(su),
                 t(sv), t(sw), t(au), t(av), t(aw), cls._s(stream))

    @classmethod
    def gradient_batched(cls, mesh, nGradient, K, phi, phiB, g, go, stream=None):
        """K scalars: phi (C, K), phiB (B, K), g and go (C, 3, K), scalar-innermost."""
        cls.call(mesh, "gradient_batched", mesh.handle, _I(nGradient), _I(K), cls._t(phi), cls._t(phiB),
                 cls._t(g), cls._t(go), cls._s(stream))

    @classmethod
    def scalar_batched(cls, mesh, K, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx, params,
                       stream=None):
        """K scalars: phi, au, su (C, K), phiB (B, K), dphidx (C, 3, K); the rest shared."""
        t = cls._t
        T = mesh.T
        p = params
        cls.call(mesh, "flux_scalar_batched", mesh.handle, _I(K), t(phi), t(phiB), t(visEff), t(au), t(su), t(den),
                 t(mf), t(te), t(cpB), t(visEffB), t(rface), t(dphidx), _I(p["SolveTurb"]), _I(p["SolveEnthalpy"]),
                 T(p["vislam"]), _I(p["ivar"]), _I(p["VarT"]), T(p["Sigma_T"]), T(p["Prandtl"]), _I(p["VarTE"]),
                 T(p["Sigma_k"]), _I(p["VarED"]), T(p["Sigma_e"]), T(p["Sigma_s"]), T(p["Schmidt"]),
                 T(p["GammaBlend"]), T(p["Small"]), T(p["Large"]), T(p["TMCmu"]), cls._s(stream))

    @classmethod
    def scalar_adjoint(cls, mesh, phi, visEff, mf, subar, rbar, params, phibar, mfbar, stream=None, noalloc=False):
        """Adjoint of gradient (nGradient = 1) + scalar face loop w.r.t. (phi, mf): writes phibar (C) and
        mfbar (F) from the cotangents subar (C) and rbar (2F). FP64 meshes whose regions are all DEFAULT.
        ``noalloc`` calls cupcfd_dev_scalar_adjoint_noalloc_f64 (fixed library scratch, no device
        allocation, meshes of up to 2,097,152 cells; same bits)."""
        from . import lib, check
        t = cls._t
        p = params
        D = ct.c_double
        fn = "cupcfd_dev_scalar_adjoint_noalloc_f64" if noalloc else "cupcfd_dev_scalar_adjoint_f64"
        check(fn, getattr(lib(), fn)(
            mesh.handle, t(phi), t(visEff), t(mf), t(subar), t(rbar), _I(p["SolveTurb"]), D(p["vislam"]),
            _I(p["ivar"]), _I(p["VarT"]), D(p["Sigma_T"]), D(p["Prandtl"]), _I(p["VarTE"]), D(p["Sigma_k"]),
            _I(p["VarED"]), D(p["Sigma_e"]), D(p["Sigma_s"]), D(p["Schmidt"]), D(p["GammaBlend"]), t(phibar),
            t(mfbar), cls._s(stream)))

    @classmethod
    def gradient_batched_f32s(cls, mesh, nGradient, K, phi, phiB, g, go, stream=None):
        """Batched gradient with float32 storage (phi, phiB, g, go float32 tensors) on a 64-bit mesh: double
        geometry and arithmetic, the final gradient rounded to float."""
        from . import lib, check
        assert mesh.precision == 64
        check("cupcfd_dev_gradient_batched_f32s", lib().cupcfd_dev_gradient_batched_f32s(
            mesh.handle, _I(nGradient), _I(K), cls._t(phi), cls._t(phiB), cls._t(g), cls._t(go), cls._s(stream)))

    @classmethod
    def scalar_batched_f32s(cls, mesh, K, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx,
                            params, stream=None):
        """Batched scalar flux with float32 storage on a 64-bit mesh: phi, phiB, au, su, dphidx float32 (laid out as
        for scalar_batched); the shared arrays and all arithmetic float64, Au/Su/PhiB rounded to float once."""
        from . import lib, check
        assert mesh.precision == 64
        t = cls._t
        p = params
        D = ct.c_double
        check("cupcfd_dev_flux_scalar_batched_f32s", lib().cupcfd_dev_flux_scalar_batched_f32s(
            mesh.handle, _I(K), t(phi), t(phiB), t(visEff), t(au), t(su), t(den), t(mf), t(te), t(cpB), t(visEffB),
            t(rface), t(dphidx), _I(p["SolveTurb"]), _I(p["SolveEnthalpy"]), D(p["vislam"]), _I(p["ivar"]),
            _I(p["VarT"]), D(p["Sigma_T"]), D(p["Prandtl"]), _I(p["VarTE"]), D(p["Sigma_k"]), _I(p["VarED"]),
            D(p["Sigma_e"]), D(p["Sigma_s"]), D(p["Schmidt"]), D(p["GammaBlend"]), D(p["Small"]), D(p["Large"]),
            D(p["TMCmu"]), cls._s(stream)))

    @classmethod
    def gradient_batched_set(cls, meshset, K, phi_ptrs, phiB_ptrs, g_ptrs, go_ptrs=None, stream=None):
        """Batched gradient (nGradient = 1) on every mesh of a DeviceMeshSet in one persistent launch.
        *_ptrs are int64 device tensors holding each member's device array address (data_ptr())."""
        from . import lib, check
        fn = f"cupcfd_dev_gradient_batched_set_{meshset.sfx}"
        check(fn, getattr(lib(), fn)(meshset.handle, _I(K), cls._t(phi_ptrs), cls._t(phiB_ptrs), cls._t(g_ptrs),
                                     cls._t(go_ptrs), cls._s(stream)))

    @classmethod
    def scalar_adjoint_batched(cls, mesh, K, phi, visEff, mf, subar, rbar, params, phibar, mfbar, stream=None):
        """Batched adjoint: phi, subar, phibar (C, K) scalar-innermost; visEff, mf, rbar shared; mfbar (F) sums
        the K scalars' terms after the RFace term."""
        from . import lib, check
        t = cls._t
        p = params
        D = ct.c_double
        check("cupcfd_dev_scalar_adjoint_batched_f64", lib().cupcfd_dev_scalar_adjoint_batched_f64(
            mesh.handle, _I(K), t(phi), t(visEff), t(mf), t(subar), t(rbar), _I(p["SolveTurb"]), D(p["vislam"]),
            _I(p["ivar"]), _I(p["VarT"]), D(p["Sigma_T"]), D(p["Prandtl"]), _I(p["VarTE"]), D(p["Sigma_k"]),
            _I(p["VarED"]), D(p["Sigma_e"]), D(p["Sigma_s"]), D(p["Schmidt"]), D(p["GammaBlend"]), t(phibar),
            t(mfbar), cls._s(stream)))

    @classmethod
    def scalar_adjoint_batched_f32s(cls, mesh, K, phi, visEff, mf, subar, rbar, params, phibar, mfbar, stream=None):
        """Batched adjoint with float32 storage of phi, subar and phibar (the rest float64), on a 64-bit mesh."""
        from . import lib, check
        assert mesh.precision == 64
        t = cls._t
        p = params
        D = ct.c_double
        check("cupcfd_dev_scalar_adjoint_batched_f32s", lib().cupcfd_dev_scalar_adjoint_batched_f32s(
            mesh.handle, _I(K), t(phi), t(visEff), t(mf), t(subar), t(rbar), _I(p["SolveTurb"]), D(p["vislam"]),
            _I(p["ivar"]), _I(p["VarT"]), D(p["Sigma_T"]), D(p["Prandtl"]), _I(p["VarTE"]), D(p["Sigma_k"]),
            _I(p["VarED"]), D(p["Sigma_e"]), D(p["Sigma_s"]), D(p["Schmidt"]), D(p["GammaBlend"]), t(phibar),
            t(mfbar), cls._s(stream)))

    # the boundary, region and viscosity loops; scalar outputs (flowin, flowout, ratearea, ...) are one-value
    # device tensors
    @classmethod
    def mass_bl1(cls, mesh, mf, flowin, stream=None):
        """MassKernels.h:90-94"""
        cls.call(mesh, "flux_mass_boundary_loop1", mesh.handle, cls._t(mf), cls._t(flowin), cls._s(stream))

    @classmethod
    def mass_bl2(cls, mesh, mf, flowRegion, flowout, stream=None):
        """MassKernels.h:103-108"""
        t = cls._t
        cls.call(mesh, "flux_mass_boundary_loop2", mesh.handle, t(mf), t(flowRegion), t(flowout), cls._s(stream))

    @classmethod
    def mass_bl3(cls, mesh, flowin, ratearea, stream=None):
        """MassKernels.h:117-120 (flowin: a host value)"""
        cls.call(mesh, "flux_mass_boundary_loop3", mesh.handle, mesh.T(flowin), cls._t(ratearea), cls._s(stream))

    @classmethod
    def mass_bl4(cls, mesh, mf, uB, vB, wB, denB, ratearea, flowout, stream=None):
        """MassKernels.h:129-138 (ratearea: a host value)"""
        t = cls._t
        cls.call(mesh, "flux_mass_boundary_loop4", mesh.handle, t(mf), t(uB), t(vB), t(wB), t(denB),
                 mesh.T(ratearea), t(flowout), cls._s(stream))

    @classmethod
    def mass_bl5(cls, mesh, mf, su, uB, vB, wB, fact, solveU, solveV, solveW, flowFact, flowout2, stream=None):
        """MassKernels.h:147-157"""
        t = cls._t
        cls.call(mesh, "flux_mass_boundary_loop5", mesh.handle, t(mf), t(su), t(uB), t(vB), t(wB), mesh.T(fact),
                 _I(solveU), _I(solveV), _I(solveW), t(flowFact), t(flowout2), cls._s(stream))

    @classmethod
    def mass_region(cls, mesh, flowFact, flowRegion, flowIn, stream=None):
        """MassKernels.h:166-171 (flowIn: a host value)"""
        cls.call(mesh, "flux_mass_region_loop", mesh.handle, cls._t(flowFact), cls._t(flowRegion), mesh.T(flowIn),
                 cls._s(stream))

    @classmethod
    def uvw_region1(cls, mesh, stream=None):
        """UVWKernels.h:72-73"""
        cls.call(mesh, "flux_uvw_region_loop1", mesh.handle, cls._s(stream))

    @classmethod
    def uvw_bnds1(cls, mesh, stream=None):
        """UVWKernels.h:78-79"""
        cls.call(mesh, "flux_uvw_bnds_loop1", mesh.handle, cls._s(stream))

    @classmethod
    def visc_cell1(cls, mesh, small, cmu, visURF, visLam, te, ed, den, visEff, stream=None):
        """ViscosityKernels.h:38-48"""
        t, T = cls._t, mesh.T
        cls.call(mesh, "viscosity_cell_loop1", mesh.handle, T(small), T(cmu), T(visURF), T(visLam), t(te), t(ed),
                 t(den), t(visEff), cls._s(stream))

    @classmethod
    def visc_region(cls, mesh, kappa, stream=None):
        """ViscosityKernels.h:50-51 (the mesh is the second argument there)"""
        from . import lib, check
        fn = f"cupcfd_dev_viscosity_region_loop_{mesh.sfx}"
        check(fn, getattr(lib(), fn)(mesh.T(kappa), mesh.handle, cls._s(stream)))

    @classmethod
    def visc_boundary(cls, mesh, tmcmu, large, small, kappa, visLam, te, den, visEff, visEffB, stream=None):
        """ViscosityKernels.h:53-60"""
        t, T = cls._t, mesh.T
        cls.call(mesh, "viscosity_boundary_loop", mesh.handle, T(tmcmu), T(large), T(small), T(kappa), T(visLam),
                 t(te), t(den), t(visEff), t(visEffB), cls._s(stream))

    @classmethod
    def visc_cell2(cls, mesh, visLam, visEff, visEffB, stream=None):
        """ViscosityKernels.h:62-67"""
        cls.call(mesh, "viscosity_cell_loop2", mesh.handle, mesh.T(visLam), cls._t(visEff), cls._t(visEffB),
                 cls._s(stream))

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

    ``stream`` is a torch.cuda.Stream or None (the current stream). Every tensor is checked before the call:
    on the GPU, contiguous, of the mesh's precision (float32 for the per-scalar arrays of the _f32s calls)
    and at least as long as the mesh needs (cells, 3 x cells, faces, 2 x faces, boundaries or regions, times
    K for the batched per-scalar arrays); a violation raises CupCfdError (E_ARRAY_SIZE_MISMATCH for a short
    tensor, E_CONFIG_INVALID_VALUE otherwise) and launches nothing."""

    @staticmethod
    def _s(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return ct.c_void_p(s.cuda_stream)

    @staticmethod
    def _t(t):
        return None if t is None else ct.c_void_p(t.data_ptr())

    @staticmethod
    def _chk(t, n, dtype, name, gpu=True):
        """ctypes pointer of tensor t after checking that it is a contiguous CUDA tensor of ``dtype`` with at
        least n elements (None passes through: the C ABI decides whether the argument may be absent)."""
        if t is None:
            return None
        import torch
        from . import CupCfdError, E_ARRAY_SIZE_MISMATCH
        bad = 90  # E_CONFIG_INVALID_VALUE
        if not isinstance(t, torch.Tensor):
            raise CupCfdError(f"{name}: expected a torch tensor, got {type(t).__name__}", bad)
        if t.dtype != dtype:
            raise CupCfdError(f"{name}: dtype {t.dtype}, expected {dtype}", bad)
        if not t.is_contiguous():
            raise CupCfdError(f"{name}: tensor is not contiguous", bad)
        if t.numel() < n:
            raise CupCfdError(f"{name}: {t.numel()} elements, the mesh needs {n}", E_ARRAY_SIZE_MISMATCH)
        if gpu and not t.is_cuda:
            raise CupCfdError(f"{name}: tensor is not on the GPU", bad)
        return ct.c_void_p(t.data_ptr())

    @classmethod
    def _args(cls, mesh, spec, K=1, f32=()):
        """Checked pointers for [(tensor, kind, name), ...]; names in f32 are float32 per-scalar arrays times K,
        kinds ending in 'K' are per-scalar (times K)."""
        import torch
        size = {"c": mesh.nCells, "c3": 3 * mesh.nCells, "f": mesh.nFaces, "f2": 2 * mesh.nFaces,
                "b": mesh.nBoundaries, "r": mesh.nRegions, "1": 1}
        native = torch.float64 if mesh.precision == 64 else torch.float32
        out = []
        for t, kind, name in spec:  # layout and sizes of every argument first, then the device
            n = size[kind.rstrip("K")] * (K if kind.endswith("K") else 1)
            out.append(cls._chk(t, n, torch.float32 if name in f32 else native, name, gpu=False))
        for t, _, name in spec:
            if t is not None and not t.is_cuda:
                from . import CupCfdError
                raise CupCfdError(f"{name}: tensor is not on the GPU", 90)
        return out

    @staticmethod
    def call(mesh, name, *args):
        from . import lib, check
        fn = f"cupcfd_dev_{name}_{mesh.sfx}"
        check(fn, getattr(lib(), fn)(*args))

    @classmethod
    def gradient(cls, mesh, nGradient, phi, phiB, g, go, stream=None):
        a = cls._args(mesh, [(phi, "c", "phi"), (phiB, "b", "phiB"), (g, "c3", "dPhidx"), (go, "c3", "dPhidxo")])
        cls.call(mesh, "gradient_phi_gauss_dolfyn", mesh.handle, _I(nGradient), *a, cls._s(stream))

    @classmethod
    def mass(cls, mesh, dudx, dvdx, dwdx, dpdx, den, denB, u, v, w, mf, p, ar, su, rface, small=1e-18,
             counters=None, flags=(0, 0, 1, 0), teC=None, teB=None, edC=None, edB=None, visC=None, visB=None,
             tC=None, tB=None, stream=None):
        cnt = counters if counters is not None else np.zeros(4, np.int32)
        assert isinstance(cnt, np.ndarray) and cnt.dtype == np.int32 and cnt.size >= 4 and cnt.flags["C_CONTIGUOUS"]
        a = cls._args(mesh, [(dudx, "c3", "dudx"), (dvdx, "c3", "dvdx"), (dwdx, "c3", "dwdx"), (dpdx, "c3", "dpdx"),
                             (den, "c", "den"), (denB, "b", "denB"), (u, "c", "u"), (v, "c", "v"), (w, "c", "w"),
                             (mf, "f", "massFlux"), (p, "c", "p"), (ar, "c", "ar"), (su, "c", "su"),
                             (rface, "f2", "rface")])
        e = cls._args(mesh, [(teC, "c", "teCell"), (teB, "b", "teBoundary"), (edC, "c", "edCell"),
                             (edB, "b", "edBoundary"), (visC, "c", "viseffCell"), (visB, "b", "viseffBoundary"),
                             (tC, "c", "tCell"), (tB, "b", "tBoundary")])
        base = cnt.ctypes.data
        cls.call(mesh, "flux_mass_face_loop", mesh.handle, *a, mesh.T(small), ct.c_void_p(base),
                 ct.c_void_p(base + 4), ct.c_void_p(base + 8), ct.c_void_p(base + 12), *[_I(x) for x in flags],
                 *e, cls._s(stream))
        return cnt

    @classmethod
    def _scalar_spec(cls, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx, K=None):
        k = "K" if K is not None else ""
        return [(phi, "c" + k, "phi"), (phiB, "b" + k, "phiB"), (visEff, "c", "visEff"), (au, "c" + k, "au"),
                (su, "c" + k, "su"), (den, "c", "den"), (mf, "f", "massFlux"), (te, "c", "te"), (cpB, "b", "cpB"),
                (visEffB, "b", "visEffB"), (rface, "f2", "rface"), (dphidx, "c3" + k, "dphidx")]

    @staticmethod
    def _scalar_params(T, p):
        return [_I(p["SolveTurb"]), _I(p["SolveEnthalpy"]), T(p["vislam"]), _I(p["ivar"]), _I(p["VarT"]),
                T(p["Sigma_T"]), T(p["Prandtl"]), _I(p["VarTE"]), T(p["Sigma_k"]), _I(p["VarED"]), T(p["Sigma_e"]),
                T(p["Sigma_s"]), T(p["Schmidt"]), T(p["GammaBlend"]), T(p["Small"]), T(p["Large"]), T(p["TMCmu"])]

    @classmethod
    def scalar(cls, mesh, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx, params,
               stream=None):
        a = cls._args(mesh, cls._scalar_spec(phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx))
        cls.call(mesh, "flux_scalar_face_loop", mesh.handle, *a, *cls._scalar_params(mesh.T, params), cls._s(stream))

    @classmethod
    def uvw(cls, mesh, gammaBlend, small, u, v, w, uB, vB, wB, visEff, visEffB, mf, dudx, dvdx, dwdx, rface, su, sv,
            sw, au, av, aw, stream=None):
        a = cls._args(mesh, [(u, "c", "u"), (v, "c", "v"), (w, "c", "w"), (uB, "b", "uB"), (vB, "b", "vB"),
                             (wB, "b", "wB"), (visEff, "c", "visEff"), (visEffB, "b", "visEffB"),
                             (mf, "f", "massFlux"), (dudx, "c3", "dudx"), (dvdx, "c3", "dvdx"), (dwdx, "c3", "dwdx"),
                             (rface, "f2", "rface"), (su, "c", "su"), (sv, "c", "sv"), (sw, "c", "sw"),
                             (au, "c", "au"), (av, "c", "av"), (aw, "c", "aw")])
        cls.call(mesh, "flux_uvw_face_loop1", mesh.handle, mesh.T(gammaBlend), mesh.T(small), *a, cls._s(stream))

    @classmethod
    def gradient_batched(cls, mesh, nGradient, K, phi, phiB, g, go, stream=None):
        """K scalars: phi (C, K), phiB (B, K), g and go (C, 3, K), scalar-innermost."""
        a = cls._args(mesh, [(phi, "cK", "phi"), (phiB, "bK", "phiB"), (g, "c3K", "dPhidx"), (go, "c3K", "dPhidxo")],
                      K=K)
        cls.call(mesh, "gradient_batched", mesh.handle, _I(nGradient), _I(K), *a, cls._s(stream))

    @classmethod
    def scalar_batched(cls, mesh, K, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx, params,
                       stream=None):
        """K scalars: phi, au, su (C, K), phiB (B, K), dphidx (C, 3, K); the rest shared."""
        a = cls._args(mesh, cls._scalar_spec(phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx,
                                             K=K), K=K)
        cls.call(mesh, "flux_scalar_batched", mesh.handle, _I(K), *a, *cls._scalar_params(mesh.T, params),
                 cls._s(stream))

    @staticmethod
    def _adjoint_params(p):
        D = ct.c_double
        return [_I(p["SolveTurb"]), D(p["vislam"]), _I(p["ivar"]), _I(p["VarT"]), D(p["Sigma_T"]), D(p["Prandtl"]),
                _I(p["VarTE"]), D(p["Sigma_k"]), _I(p["VarED"]), D(p["Sigma_e"]), D(p["Sigma_s"]), D(p["Schmidt"]),
                D(p["GammaBlend"])]

    @classmethod
    def _adjoint_args(cls, mesh, phi, visEff, mf, subar, rbar, phibar, mfbar, K=None, f32=()):
        k = "K" if K is not None else ""
        return cls._args(mesh, [(phi, "c" + k, "phi"), (visEff, "c", "visEff"), (mf, "f", "massFlux"),
                                (subar, "c" + k, "subar"), (rbar, "f2", "rbar"), (phibar, "c" + k, "phibar"),
                                (mfbar, "f", "mfbar")], K=K or 1, f32=f32)

    @classmethod
    def scalar_adjoint(cls, mesh, phi, visEff, mf, subar, rbar, params, phibar, mfbar, stream=None, noalloc=False):
        """Adjoint of gradient (nGradient = 1) + scalar face loop w.r.t. (phi, mf): writes phibar (C) and
        mfbar (F) from the cotangents subar (C) and rbar (2F). FP64 meshes whose regions are all DEFAULT.
        ``noalloc`` calls cupcfd_dev_scalar_adjoint_noalloc_f64 (fixed library scratch, no device
        allocation, meshes of up to 2,097,152 cells; same bits)."""
        from . import lib, check
        a = cls._adjoint_args(mesh, phi, visEff, mf, subar, rbar, phibar, mfbar)
        fn = "cupcfd_dev_scalar_adjoint_noalloc_f64" if noalloc else "cupcfd_dev_scalar_adjoint_f64"
        check(fn, getattr(lib(), fn)(mesh.handle, *a[:5], *cls._adjoint_params(params), *a[5:], cls._s(stream)))

    @classmethod
    def scalar_adjoint_full(cls, mesh, nGradient, phi, visEff, mf, dphidx, te, den, cpB, visEffB, subar, rbar,
                            params, phibar, mfbar, stream=None):
        """Adjoint of gradient (nGradient sweeps) + scalar face loop w.r.t. (phi, mf) on any FP64 mesh
        (cupcfd_dev_scalar_adjoint_full_f64): dphidx is the forward gradient the scalar loop read, te / den /
        cpB / visEffB its other inputs (None allowed on all-DEFAULT meshes)."""
        from . import lib, check
        a = cls._args(mesh, [(phi, "c", "phi"), (visEff, "c", "visEff"), (mf, "f", "massFlux"),
                             (dphidx, "c3", "dPhidx"), (te, "c", "TE"), (den, "c", "Den"), (cpB, "b", "CpBoundary"),
                             (visEffB, "b", "visEffBoundary"), (subar, "c", "subar"), (rbar, "f2", "rbar"),
                             (phibar, "c", "phibar"), (mfbar, "f", "mfbar")])
        D = ct.c_double
        p = params
        check("cupcfd_dev_scalar_adjoint_full_f64", lib().cupcfd_dev_scalar_adjoint_full_f64(
            mesh.handle, _I(nGradient), *a[:10], _I(p["SolveTurb"]), _I(p["SolveEnthalpy"]), D(p["vislam"]),
            _I(p["ivar"]), _I(p["VarT"]), D(p["Sigma_T"]), D(p["Prandtl"]), _I(p["VarTE"]), D(p["Sigma_k"]),
            _I(p["VarED"]), D(p["Sigma_e"]), D(p["Sigma_s"]), D(p["Schmidt"]), D(p["GammaBlend"]), D(p["Small"]),
            D(p["TMCmu"]), *a[10:], cls._s(stream)))

    @classmethod
    def gradient_batched_f32s(cls, mesh, nGradient, K, phi, phiB, g, go, stream=None):
        """Batched gradient with float32 storage (phi, phiB, g, go float32 tensors) on a 64-bit mesh: double
        geometry and arithmetic, the final gradient rounded to float."""
        from . import lib, check
        assert mesh.precision == 64
        a = cls._args(mesh, [(phi, "cK", "phi"), (phiB, "bK", "phiB"), (g, "c3K", "dPhidx"), (go, "c3K", "dPhidxo")],
                      K=K, f32=("phi", "phiB", "dPhidx", "dPhidxo"))
        check("cupcfd_dev_gradient_batched_f32s", lib().cupcfd_dev_gradient_batched_f32s(
            mesh.handle, _I(nGradient), _I(K), *a, cls._s(stream)))

    @classmethod
    def scalar_batched_f32s(cls, mesh, K, phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx,
                            params, stream=None):
        """Batched scalar flux with float32 storage on a 64-bit mesh: phi, phiB, au, su, dphidx float32 (laid out as
        for scalar_batched); the shared arrays and all arithmetic float64, Au/Su/PhiB rounded to float once."""
        from . import lib, check
        assert mesh.precision == 64
        a = cls._args(mesh, cls._scalar_spec(phi, phiB, visEff, au, su, den, mf, te, cpB, visEffB, rface, dphidx,
                                             K=K), K=K, f32=("phi", "phiB", "au", "su", "dphidx"))
        check("cupcfd_dev_flux_scalar_batched_f32s", lib().cupcfd_dev_flux_scalar_batched_f32s(
            mesh.handle, _I(K), *a, *cls._scalar_params(ct.c_double, params), cls._s(stream)))

    @classmethod
    def gradient_batched_set(cls, meshset, K, phi_ptrs, phiB_ptrs, g_ptrs, go_ptrs=None, stream=None):
        """Batched gradient (nGradient = 1) on every mesh of a DeviceMeshSet in one persistent launch.
        *_ptrs are int64 device tensors holding each member's device array address (data_ptr())."""
        import torch
        from . import lib, check
        n = len(meshset.meshes)  # the pointer tables hold one address per member mesh
        ptrs = [cls._chk(t, n, torch.int64, name)
                for t, name in ((phi_ptrs, "phi_ptrs"), (phiB_ptrs, "phiB_ptrs"), (g_ptrs, "g_ptrs"),
                                (go_ptrs, "go_ptrs"))]
        fn = f"cupcfd_dev_gradient_batched_set_{meshset.sfx}"
        check(fn, getattr(lib(), fn)(meshset.handle, _I(K), *ptrs, cls._s(stream)))

    @classmethod
    def scalar_adjoint_batched(cls, mesh, K, phi, visEff, mf, subar, rbar, params, phibar, mfbar, stream=None):
        """Batched adjoint: phi, subar, phibar (C, K) scalar-innermost; visEff, mf, rbar shared; mfbar (F) sums
        the K scalars' terms after the RFace term."""
        from . import lib, check
        a = cls._adjoint_args(mesh, phi, visEff, mf, subar, rbar, phibar, mfbar, K=K)
        check("cupcfd_dev_scalar_adjoint_batched_f64", lib().cupcfd_dev_scalar_adjoint_batched_f64(
            mesh.handle, _I(K), *a[:5], *cls._adjoint_params(params), *a[5:], cls._s(stream)))

    @classmethod
    def scalar_adjoint_batched_f32s(cls, mesh, K, phi, visEff, mf, subar, rbar, params, phibar, mfbar, stream=None):
        """Batched adjoint with float32 storage of phi, subar and phibar (the rest float64), on a 64-bit mesh."""
        from . import lib, check
        assert mesh.precision == 64
        a = cls._adjoint_args(mesh, phi, visEff, mf, subar, rbar, phibar, mfbar, K=K, f32=("phi", "subar", "phibar"))
        check("cupcfd_dev_scalar_adjoint_batched_f32s", lib().cupcfd_dev_scalar_adjoint_batched_f32s(
            mesh.handle, _I(K), *a[:5], *cls._adjoint_params(params), *a[5:], cls._s(stream)))

    @classmethod
    def residual_norms(cls, mesh, fields, norms=None, stream=None):
        """max |x| and sqrt(sum x^2) of each cell field (1..8 tensors of nCells values, internal order; float64,
        or float32 storage), accumulated in double. Returns the (nfields, 2) float64 device tensor it wrote."""
        import torch
        from . import lib, check
        fields = list(fields)
        dt = fields[0].dtype if fields else torch.float64
        if dt not in (torch.float64, torch.float32):
            from . import CupCfdError
            raise CupCfdError("residual_norms: fields must be float64 or float32", 90)
        ptrs = [cls._chk(t, mesh.nCells, dt, f"fields[{i}]") for i, t in enumerate(fields)]
        if norms is None:
            norms = torch.empty((len(fields), 2), dtype=torch.float64, device="cuda")
        out = cls._chk(norms, 2 * len(fields), torch.float64, "norms")
        fn = "cupcfd_dev_residual_norms_f64" if dt == torch.float64 else "cupcfd_dev_residual_norms_f32"
        arr = (ct.c_void_p * max(1, len(ptrs)))(*[p.value for p in ptrs])
        check(fn, getattr(lib(), fn)(mesh.handle, _I(len(ptrs)), arr, out, cls._s(stream)))
        return norms

    # the boundary, region and viscosity loops; scalar outputs (flowin, flowout, ratearea, ...) are one-value
    # device tensors
    @classmethod
    def mass_bl1(cls, mesh, mf, flowin, stream=None):
        """MassKernels.h:90-94"""
        a = cls._args(mesh, [(mf, "f", "massFlux"), (flowin, "1", "flowin")])
        cls.call(mesh, "flux_mass_boundary_loop1", mesh.handle, *a, cls._s(stream))

    @classmethod
    def mass_bl2(cls, mesh, mf, flowRegion, flowout, stream=None):
        """MassKernels.h:103-108"""
        a = cls._args(mesh, [(mf, "f", "massFlux"), (flowRegion, "r", "flowRegion"), (flowout, "1", "flowout")])
        cls.call(mesh, "flux_mass_boundary_loop2", mesh.handle, *a, cls._s(stream))

    @classmethod
    def mass_bl3(cls, mesh, flowin, ratearea, stream=None):
        """MassKernels.h:117-120 (flowin: a host value)"""
        a = cls._args(mesh, [(ratearea, "1", "ratearea")])
        cls.call(mesh, "flux_mass_boundary_loop3", mesh.handle, mesh.T(flowin), *a, cls._s(stream))

    @classmethod
    def mass_bl4(cls, mesh, mf, uB, vB, wB, denB, ratearea, flowout, stream=None):
        """MassKernels.h:129-138 (ratearea: a host value)"""
        a = cls._args(mesh, [(mf, "f", "massFlux"), (uB, "b", "uB"), (vB, "b", "vB"), (wB, "b", "wB"),
                             (denB, "b", "denB"), (flowout, "1", "flowout")])
        cls.call(mesh, "flux_mass_boundary_loop4", mesh.handle, *a[:5], mesh.T(ratearea), a[5], cls._s(stream))

    @classmethod
    def mass_bl5(cls, mesh, mf, su, uB, vB, wB, fact, solveU, solveV, solveW, flowFact, flowout2, stream=None):
        """MassKernels.h:147-157"""
        a = cls._args(mesh, [(mf, "f", "massFlux"), (su, "c", "su"), (uB, "b", "uB"), (vB, "b", "vB"),
                             (wB, "b", "wB"), (flowFact, "r", "flowFact"), (flowout2, "1", "flowout2")])
        cls.call(mesh, "flux_mass_boundary_loop5", mesh.handle, *a[:5], mesh.T(fact), _I(solveU), _I(solveV),
                 _I(solveW), *a[5:], cls._s(stream))

    @classmethod
    def mass_region(cls, mesh, flowFact, flowRegion, flowIn, stream=None):
        """MassKernels.h:166-171 (flowIn: a host value)"""
        a = cls._args(mesh, [(flowFact, "r", "flowFact"), (flowRegion, "r", "flowRegion")])
        cls.call(mesh, "flux_mass_region_loop", mesh.handle, *a, mesh.T(flowIn), cls._s(stream))

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
        T = mesh.T
        a = cls._args(mesh, [(te, "c", "TE"), (ed, "c", "ED"), (den, "c", "den"), (visEff, "c", "visEff")])
        cls.call(mesh, "viscosity_cell_loop1", mesh.handle, T(small), T(cmu), T(visURF), T(visLam), *a,
                 cls._s(stream))

    @classmethod
    def visc_region(cls, mesh, kappa, stream=None):
        """ViscosityKernels.h:50-51 (the mesh is the second argument there)"""
        from . import lib, check
        fn = f"cupcfd_dev_viscosity_region_loop_{mesh.sfx}"
        check(fn, getattr(lib(), fn)(mesh.T(kappa), mesh.handle, cls._s(stream)))

    @classmethod
    def visc_boundary(cls, mesh, tmcmu, large, small, kappa, visLam, te, den, visEff, visEffB, stream=None):
        """ViscosityKernels.h:53-60"""
        T = mesh.T
        a = cls._args(mesh, [(te, "c", "TE"), (den, "c", "den"), (visEff, "c", "visEff"), (visEffB, "b", "visEffB")])
        cls.call(mesh, "viscosity_boundary_loop", mesh.handle, T(tmcmu), T(large), T(small), T(kappa), T(visLam),
                 *a, cls._s(stream))

    @classmethod
    def visc_cell2(cls, mesh, visLam, visEff, visEffB, stream=None):
        """ViscosityKernels.h:62-67"""
        a = cls._args(mesh, [(visEff, "c", "visEff"), (visEffB, "b", "visEffB")])
        cls.call(mesh, "viscosity_cell_loop2", mesh.handle, mesh.T(visLam), *a, cls._s(stream))

// cupcfd_hip_fvm.hpp — header-only C++ shim with the call shape of namespace
// cupcfd::fvm (include/fvm/implementation/component/*Kernels.h in CUP-CFD),
// forwarding to the drop-in C ABI of cupcfd_hip.h.
//
// Each function template takes the mesh object, the same (pointer, length)
// pairs, scalar parameters and flags, in the reference's order, and returns
// the int value of cupcfd::error::eCodes. The mesh may be any type exposing
// CupCfdSoAMesh's public members (CupCfdSoAMesh.h:71-154: faceCell1ID,
// faceNorm, boundaryRegionID, regionType, ..., and properties.lTCells /
// lFaces / lBoundaries / lRegions); 3-vector arrays may be any type with a
// `cmp[3]` member (EuclideanVector<T,3> / EuclideanPoint<T,3>, whose leading
// vptr is skipped by passing the element stride).
//
// The device mesh for a host mesh is built on first use and cached by the
// mesh's address; cupcfd_hip::shim::forget(mesh) drops it (call it after
// CupCfdSoAMesh::finalize changes the mesh). The build runs on the GPU
// (cupcfd_hip_mesh_create_device_*) unless cupcfd_hip::shim::device_repack()
// is set to false, which selects the host repack; both give identical meshes. Kernels that set mesh state
// (boundary shear, Q/H/T, y+/u+, region yLog and force tangent) write it
// back into the mesh's vectors before returning, as the reference does.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "cupcfd_hip.h"

namespace cupcfd_hip {
namespace shim {

template <class T> struct Abi;
template <> struct Abi<double> {
    using V3 = cupcfd_v3_f64;
    using View = cupcfd_mesh_view_f64;
    static int create(const View *v, int32_t flags, bool on_device, cupcfd_hip_mesh **out) {
        return on_device ? cupcfd_hip_mesh_create_device_f64(v, flags, out)
                         : cupcfd_hip_mesh_create_f64(v, flags, out);
    }
};
template <> struct Abi<float> {
    using V3 = cupcfd_v3_f32;
    using View = cupcfd_mesh_view_f32;
    static int create(const View *v, int32_t flags, bool on_device, cupcfd_hip_mesh **out) {
        return on_device ? cupcfd_hip_mesh_create_device_f32(v, flags, out)
                         : cupcfd_hip_mesh_create_f32(v, flags, out);
    }
};

// 3-vector array descriptor: element i component k at &a[i].cmp[k]
template <class T, class Vec> typename Abi<T>::V3 v3(Vec *a) {
    if (!a) return typename Abi<T>::V3{nullptr, 0};
    return typename Abi<T>::V3{const_cast<T *>(&a[0].cmp[0]), (int64_t)sizeof(Vec)};
}
template <class T, class VecVector> typename Abi<T>::V3 v3(VecVector &v) {
    return v3<T>(v.empty() ? (typename VecVector::value_type *)nullptr : v.data());
}

template <class B> std::vector<uint8_t> bytes(const B &bits) {  // std::vector<bool> -> uint8_t
    std::vector<uint8_t> out(bits.size());
    for (size_t i = 0; i < bits.size(); i++) out[i] = bits[i] ? 1 : 0;
    return out;
}

struct Entry {
    cupcfd_hip_mesh *dev = nullptr;
    std::vector<uint8_t> std_, flux, adiab;
    int64_t nCells = -1, nFaces = -1, nBoundaries = -1, nRegions = -1;
    ~Entry() {
        if (dev) cupcfd_hip_mesh_destroy(dev);
    }
};

inline std::map<const void *, std::unique_ptr<Entry>> &cache() {
    static std::map<const void *, std::unique_ptr<Entry>> c;
    return c;
}

inline void forget(const void *mesh) { cache().erase(mesh); }

// true (the default): device meshes are repacked on the GPU; false: on the host
inline bool &device_repack() {
    static bool on = true;
    return on;
}

// Device mesh of `mesh` (built and cached on first use); nullptr on failure.
template <class T, class Mesh> cupcfd_hip_mesh *device(Mesh &mesh, int *rc_out = nullptr,
                                                        int32_t flags = CUPCFD_MESH_REORDER) {
    auto &e = cache()[(const void *)&mesh];
    const auto &p = mesh.properties;
    if (e && e->dev && e->nCells == p.lTCells && e->nFaces == p.lFaces && e->nBoundaries == p.lBoundaries &&
        e->nRegions == p.lRegions) {
        if (rc_out) *rc_out = CUPCFD_E_SUCCESS;
        return e->dev;
    }
    e.reset(new Entry());
    e->std_ = bytes(mesh.regionStd);
    e->flux = bytes(mesh.regionFlux);
    e->adiab = bytes(mesh.regionAdiab);
    typename Abi<T>::View v{};
    v.nCells = (int32_t)p.lTCells;
    v.nFaces = (int32_t)p.lFaces;
    v.nBoundaries = (int32_t)p.lBoundaries;
    v.nRegions = (int32_t)p.lRegions;
    v.faceCell1ID = (const int32_t *)mesh.faceCell1ID.data();
    v.faceCell2ID = (const int32_t *)mesh.faceCell2ID.data();
    v.faceBoundaryID = (const int32_t *)mesh.faceBoundaryID.data();
    v.faceLambda = mesh.faceLambda.data();
    v.faceRLencos = mesh.faceRLencos.data();
    v.faceArea = mesh.faceArea.data();
    v.faceNorm = v3<T>(mesh.faceNorm);
    v.faceCenter = v3<T>(mesh.faceCenter);
    v.faceXpac = v3<T>(mesh.faceXpac);
    v.faceXnac = v3<T>(mesh.faceXnac);
    v.cellCenter = v3<T>(mesh.cellCenter);
    v.cellVolume = mesh.cellVolume.data();
    v.boundaryFaceID = (const int32_t *)mesh.boundaryFaceID.data();
    v.boundaryRegionID = (const int32_t *)mesh.boundaryRegionID.data();
    v.boundaryDistance = mesh.boundaryDistance.data();
    v.boundaryYPlus = mesh.boundaryYPlus.data();
    v.boundaryUPlus = mesh.boundaryUPlus.data();
    v.boundaryQ = mesh.boundaryQ.data();
    v.boundaryH = mesh.boundaryH.data();
    v.boundaryT = mesh.boundaryT.data();
    v.boundaryShear = v3<T>(mesh.boundaryShear);
    v.regionType = (const int32_t *)mesh.regionType.data();
    v.regionStd = e->std_.data();
    v.regionFlux = e->flux.data();
    v.regionAdiab = e->adiab.data();
    v.regionYLog = mesh.regionYLog.data();
    v.regionELog = mesh.regionELog.data();
    v.regionDensity = mesh.regionDensity.data();
    v.regionTurbKE = mesh.regionTurbKE.data();
    v.regionTurbDiss = mesh.regionTurbDiss.data();
    v.regionSplvl = mesh.regionSplvl.data();
    v.regionDen = mesh.regionDen.data();
    v.regionR = mesh.regionR.data();
    v.regionT = mesh.regionT.data();
    v.regionForceTangent = v3<T>(mesh.regionForceTangent);
    v.regionUVW = v3<T>(mesh.regionUVW);
    static_assert(sizeof(mesh.faceCell1ID[0]) == 4, "int32 ids expected (I = L = int)");
    static_assert(sizeof(mesh.regionType[0]) == 4, "RType stored as a 4-byte enum");
    int rc = Abi<T>::create(&v, flags, device_repack(), &e->dev);
    if (rc_out) *rc_out = rc;
    if (rc != CUPCFD_E_SUCCESS) {
        e->dev = nullptr;
        return nullptr;
    }
    e->nCells = p.lTCells;
    e->nFaces = p.lFaces;
    e->nBoundaries = p.lBoundaries;
    e->nRegions = p.lRegions;
    return e->dev;
}

}  // namespace shim

namespace fvm {

#define CUPCFD_HIP_DEV(T)                               \
    int rc_ = CUPCFD_E_SUCCESS;                         \
    cupcfd_hip_mesh *dev = shim::device<T>(mesh, &rc_); \
    if (!dev) return rc_;

template <class T> struct Sfx;

// GradientKernels.h:48-54
template <class M, class I, class T, class Vec>
int GradientPhiGaussDolfyn(M &mesh, I nGradient, T *phiCell, I nPhiCell, T *phiBoundary, I nPhiBoundary,
                           Vec *dPhidxCell, I nDPhidxCell, Vec *dPhidxoCell, I nDPhidxoCell) {
    CUPCFD_HIP_DEV(T)
    if constexpr (std::is_same<T, double>::value)
        return cupcfd_fvm_gradient_phi_gauss_dolfyn_f64(dev, nGradient, phiCell, nPhiCell, phiBoundary, nPhiBoundary,
                                                        shim::v3<T>(dPhidxCell), nDPhidxCell,
                                                        shim::v3<T>(dPhidxoCell), nDPhidxoCell);
    else
        return cupcfd_fvm_gradient_phi_gauss_dolfyn_f32(dev, nGradient, phiCell, nPhiCell, phiBoundary, nPhiBoundary,
                                                        shim::v3<T>(dPhidxCell), nDPhidxCell,
                                                        shim::v3<T>(dPhidxoCell), nDPhidxoCell);
}

#define CUPCFD_HIP_PICK(name, ...)                                                  \
    if constexpr (std::is_same<T, double>::value) return name##_f64(__VA_ARGS__);  \
    else return name##_f32(__VA_ARGS__);

// MassKernels.h:48-81
template <class M, class I, class T, class Vec>
int FluxMassDolfynFaceLoop(M &mesh, Vec *dudx, I nDudx, Vec *dvdx, I nDvdx, Vec *dwdx, I nDwdx, Vec *dpdx, I nDpdx,
                           T *denCell, I nDenCell, T *denBoundary, I nDenBoundary, T *uCell, I nUCell, T *vCell,
                           I nVCell, T *wCell, I nWCell, T *massFlux, I nMassFlux, T *p, I nP, T *ar, I nAr, T *su,
                           I nSu, T *rface, I nRFace, T small, I *icinl, I *icout, I *icsym, I *icwal,
                           bool solveTurbEnergy, bool solveTurbDiss, bool solveVisc, bool solveEnthalpy, T *teCell,
                           I nTeCell, T *teBoundary, I nTeBoundary, T *edCell, I nEdCell, T *edBoundary,
                           I nEdBoundary, T *viseffCell, I nViseffCell, T *viseffBoundary, I nViseffBoundary,
                           T *tCell, I nTCell, T *tBoundary, I nTBoundary) {
    CUPCFD_HIP_DEV(T)
    static_assert(sizeof(I) == 4, "int32 counters expected");
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_face_loop, dev, shim::v3<T>(dudx), nDudx, shim::v3<T>(dvdx), nDvdx,
                    shim::v3<T>(dwdx), nDwdx, shim::v3<T>(dpdx), nDpdx, denCell, nDenCell, denBoundary,
                    nDenBoundary, uCell, nUCell, vCell, nVCell, wCell, nWCell, massFlux, nMassFlux, p, nP, ar, nAr,
                    su, nSu, rface, nRFace, small, (int32_t *)icinl, (int32_t *)icout, (int32_t *)icsym,
                    (int32_t *)icwal, solveTurbEnergy, solveTurbDiss, solveVisc, solveEnthalpy, teCell, nTeCell,
                    teBoundary, nTeBoundary, edCell, nEdCell, edBoundary, nEdBoundary, viseffCell, nViseffCell,
                    viseffBoundary, nViseffBoundary, tCell, nTCell, tBoundary, nTBoundary)
}

// MassKernels.h:90-94
template <class M, class I, class T> int FluxMassDolfynBoundaryLoop1(M &mesh, T *massFlux, I nMassFlux, T *flowin) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_boundary_loop1, dev, massFlux, nMassFlux, flowin)
}

// MassKernels.h:103-108
template <class M, class I, class T>
int FluxMassDolfynBoundaryLoop2(M &mesh, T *massFlux, I nMassFlux, T *flowRegion, I nFlowRegion, T *flowout) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_boundary_loop2, dev, massFlux, nMassFlux, flowRegion, nFlowRegion, flowout)
}

// MassKernels.h:117-120
template <class M, class T> int FluxMassDolfynBoundaryLoop3(M &mesh, T flowin, T *ratearea) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_boundary_loop3, dev, flowin, ratearea)
}

// MassKernels.h:129-138
template <class M, class I, class T>
int FluxMassDolfynBoundaryLoop4(M &mesh, T *massFlux, I nMassFlux, T *uBoundary, I nUBoundary, T *vBoundary,
                                I nVboundary, T *wBoundary, I nWBoundary, T *denBoundary, I nDenBoundary, T ratearea,
                                T *flowout) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_boundary_loop4, dev, massFlux, nMassFlux, uBoundary, nUBoundary, vBoundary,
                    nVboundary, wBoundary, nWBoundary, denBoundary, nDenBoundary, ratearea, flowout)
}

// MassKernels.h:147-157
template <class M, class I, class T>
int FluxMassDolfynBoundaryLoop5(M &mesh, T *massFlux, I nMassFlux, T *su, I nSu, T *uBoundary, I nUBoundary,
                                T *vBoundary, I nVBoundary, T *wBoundary, I nWBoundary, T fact, bool solveU,
                                bool solveV, bool solveW, T *flowFact, I nFlowFact, T *flowout2) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_boundary_loop5, dev, massFlux, nMassFlux, su, nSu, uBoundary, nUBoundary,
                    vBoundary, nVBoundary, wBoundary, nWBoundary, fact, solveU, solveV, solveW, flowFact, nFlowFact,
                    flowout2)
}

// MassKernels.h:166-171
template <class M, class I, class T>
int FluxMassDolfynRegionLoop(M &mesh, T *flowFact, I nFlowFact, T *flowRegion, I nFlowRegion, T flowIn) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_mass_region_loop, dev, flowFact, nFlowFact, flowRegion, nFlowRegion, flowIn)
}

// ScalarKernels.h:47-67
template <class M, class I, class T, class Vec>
int FluxScalarDolfynFaceLoop(M &mesh, T *PhiCell, I nPhiCell, T *PhiBoundary, I nPhiBoundary, T *VisEff, I nVisEff,
                             T *Au, I nAu, T *Su, I nSu, T *Den, I nDen, T *MassFlux, I nMassFlux, T *TE, I nTE,
                             T *CpBoundary, I nCpBoundary, T *visEffBoundary, I nVisEffBoundary, T *RFace, I nRFace,
                             Vec *dPhidx, int ndPhidx, bool SolveTurb, bool SolveEnthalpy, T vislam, int ivar,
                             int VarT, T Sigma_T, T Prandtl, int VarTE, T Sigma_k, int VarED, T Sigma_e, T Sigma_s,
                             T Schmidt, T GammaBlend, T Small, T Large, T TMCmu) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_scalar_face_loop, dev, PhiCell, nPhiCell, PhiBoundary, nPhiBoundary, VisEff,
                    nVisEff, Au, nAu, Su, nSu, Den, nDen, MassFlux, nMassFlux, TE, nTE, CpBoundary, nCpBoundary,
                    visEffBoundary, nVisEffBoundary, RFace, nRFace, shim::v3<T>(dPhidx), ndPhidx, SolveTurb,
                    SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s,
                    Schmidt, GammaBlend, Small, Large, TMCmu)
}

// UVWKernels.h:43-67
template <class M, class I, class T, class Vec>
int FluxUVWDolfynFaceLoop1(M &mesh, T gammaBlend, T small, T *uCell, I nUCell, T *vCell, I nVCell, T *wCell,
                           I nWCell, T *uBoundary, I nUBoundary, T *vBoundary, I nVBoundary, T *wBoundary,
                           I nWBoundary, T *visEffCell, I nVisEffCell, T *visEffBoundary, I nVisEffBoundary,
                           T *massFlux, I nMassFlux, Vec *dudx, I nDudx, Vec *dvdx, I nDvdx, Vec *dwdx, I nDwdx,
                           T *rFace, I nRFace, T *su, I nSu, T *sv, I nSv, T *sw, I nSw, T *au, I nAu, T *av, I nAv,
                           T *aw, I nAw) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_uvw_face_loop1, dev, gammaBlend, small, uCell, nUCell, vCell, nVCell, wCell,
                    nWCell, uBoundary, nUBoundary, vBoundary, nVBoundary, wBoundary, nWBoundary, visEffCell,
                    nVisEffCell, visEffBoundary, nVisEffBoundary, massFlux, nMassFlux, shim::v3<T>(dudx), nDudx,
                    shim::v3<T>(dvdx), nDvdx, shim::v3<T>(dwdx), nDwdx, rFace, nRFace, su, nSu, sv, nSv, sw, nSw, au,
                    nAu, av, nAv, aw, nAw)
}

// UVWKernels.h:72-73 (void in the reference; the shim returns the eCodes value)
template <class T, class M> int FluxUVWDolfynRegionLoop1(M &mesh) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_uvw_region_loop1, dev)
}

// UVWKernels.h:78-79
template <class T, class M> int FluxUVWDolfynBndsLoop1(M &mesh) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_flux_uvw_bnds_loop1, dev)
}

// ViscosityKernels.h:38-48
template <class M, class I, class T>
int calculateViscosityDolfynCellLoop1(M &mesh, T small, T cmu, T visURF, T visLam, T *TE, I nTE, T *ED, I nED,
                                      T *den, I nDen, T *visEff, I nVisEff) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_viscosity_cell_loop1, dev, small, cmu, visURF, visLam, TE, nTE, ED, nED, den, nDen,
                    visEff, nVisEff)
}

// ViscosityKernels.h:50-51
template <class M, class T> int calculateViscosityDolfynRegionLoop(T kappa, M &mesh) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_viscosity_region_loop, kappa, dev)
}

// ViscosityKernels.h:53-60
template <class M, class I, class T>
int calculateViscosityDolfynBoundaryLoop(M &mesh, T tmcmu, T large, T small, T kappa, T visLam, T *TE, I nTE, T *den,
                                         I nDen, T *visEffCell, I nVisEffCell, T *visEffBoundary,
                                         I nVisEffBoundary) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_viscosity_boundary_loop, dev, tmcmu, large, small, kappa, visLam, TE, nTE, den, nDen,
                    visEffCell, nVisEffCell, visEffBoundary, nVisEffBoundary)
}

// ViscosityKernels.h:62-67
template <class M, class I, class T>
int calculateViscosityDolfynCellLoop2(M &mesh, T visLam, T *visEffCell, I nVisEffCell, T *visEffBoundary,
                                      I nVisEffBoundary) {
    CUPCFD_HIP_DEV(T)
    CUPCFD_HIP_PICK(cupcfd_fvm_viscosity_cell_loop2, dev, visLam, visEffCell, nVisEffCell, visEffBoundary,
                    nVisEffBoundary)
}

#undef CUPCFD_HIP_PICK
#undef CUPCFD_HIP_DEV

}  // namespace fvm
}  // namespace cupcfd_hip

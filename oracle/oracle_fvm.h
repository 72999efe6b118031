/*
 * ORACLE — test infrastructure only.
 *
 * A sequential CPU restatement of CUP-CFD's Dolfyn-derived finite-volume
 * kernels (the include/fvm/implementation/component .ipp files). It exists to check
 * the HIP path; nothing in the product links or calls it. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Mesh layout: plain structure-of-arrays mirroring the std::vector members of
 * CupCfdSoAMesh (include/geometry/mesh/implementation/component/CupCfdSoAMesh.h:71-154),
 * with every 3-vector packed as xyz triples (the reference's vptr is dropped).
 *
 * Every kernel returns a cupcfd::error::eCodes value (include/error/Error.h).
 */
#ifndef CUPCFD_ORACLE_FVM_H
#define CUPCFD_ORACLE_FVM_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_E_SUCCESS 95
#define OR_E_INVALID_INDEX 68
#define OR_E_NOT_IMPLEMENTED 109

#define OR_RTYPE_DEFAULT 0
#define OR_RTYPE_INLET 1
#define OR_RTYPE_OUTLET 2
#define OR_RTYPE_SYMP 3
#define OR_RTYPE_WALL 4

#define OR_MESH_FIELDS(T)                                                        \
    int nCells, nFaces, nBoundaries, nRegions;                                   \
    const int *faceCell1, *faceCell2, *faceBnd;                                  \
    const T *faceLambda, *faceNorm, *faceCenter, *faceXpac, *faceXnac;           \
    const T *faceRLencos, *faceArea;                                             \
    const T *cellCenter, *cellVolume;                                            \
    const int *bndFace, *bndRegion;                                              \
    const T *bndDistance;                                                        \
    T *bndYPlus, *bndUPlus, *bndShear, *bndQ, *bndH, *bndT;                      \
    const int *regType;                                                          \
    const unsigned char *regStd, *regFlux, *regAdiab;                            \
    T *regYLog;                                                                  \
    const T *regELog, *regDensity, *regTurbKE, *regTurbDiss, *regSplvl, *regDen; \
    const T *regR, *regT;                                                        \
    T *regForceTangent;                                                          \
    const T *regUVW;

typedef struct { OR_MESH_FIELDS(double) } or_mesh_f64;
typedef struct { OR_MESH_FIELDS(float) } or_mesh_f32;

/* Value taken by the face-to-face carried locals (PhiFace, uFace/vFace/wFace,
 * visFace) and by yplustmp before their first assignment. The reference
 * leaves them uninitialised (undefined behaviour); this restatement and the
 * HIP path both start them at this value. Default 0. */
void or_set_carry_init(double v);

#define OR_DECL(T, S)                                                                              \
int or_gradient_phi_gauss_dolfyn_##S(or_mesh_##S *m, int nGradient,                                \
    T *phiCell, int nPhiCell, T *phiBoundary, int nPhiBoundary,                                    \
    T *dPhidxCell, int nDPhidxCell, T *dPhidxoCell, int nDPhidxoCell);                             \
int or_flux_mass_face_loop_##S(or_mesh_##S *m,                                                     \
    T *dudx, int nDudx, T *dvdx, int nDvdx, T *dwdx, int nDwdx, T *dpdx, int nDpdx,                \
    T *denCell, int nDenCell, T *denBoundary, int nDenBoundary,                                    \
    T *uCell, int nUCell, T *vCell, int nVCell, T *wCell, int nWCell,                              \
    T *massFlux, int nMassFlux, T *p, int nP, T *ar, int nAr, T *su, int nSu,                      \
    T *rface, int nRFace, T small, int *icinl, int *icout, int *icsym, int *icwal,                 \
    int solveTurbEnergy, int solveTurbDiss, int solveVisc, int solveEnthalpy,                      \
    T *teCell, int nTeCell, T *teBoundary, int nTeBoundary,                                        \
    T *edCell, int nEdCell, T *edBoundary, int nEdBoundary,                                        \
    T *viseffCell, int nViseffCell, T *viseffBoundary, int nViseffBoundary,                        \
    T *tCell, int nTCell, T *tBoundary, int nTBoundary);                                           \
int or_flux_mass_boundary_loop1_##S(or_mesh_##S *m, T *massFlux, int nMassFlux, T *flowin);        \
int or_flux_mass_boundary_loop2_##S(or_mesh_##S *m, T *massFlux, int nMassFlux,                    \
    T *flowRegion, int nFlowRegion, T *flowout);                                                   \
int or_flux_mass_boundary_loop3_##S(or_mesh_##S *m, T flowin, T *ratearea);                        \
int or_flux_mass_boundary_loop4_##S(or_mesh_##S *m, T *massFlux, int nMassFlux,                    \
    T *uBoundary, int nUBoundary, T *vBoundary, int nVBoundary, T *wBoundary, int nWBoundary,      \
    T *denBoundary, int nDenBoundary, T ratearea, T *flowout);                                     \
int or_flux_mass_boundary_loop5_##S(or_mesh_##S *m, T *massFlux, int nMassFlux, T *su, int nSu,    \
    T *uBoundary, int nUBoundary, T *vBoundary, int nVBoundary, T *wBoundary, int nWBoundary,      \
    T fact, int solveU, int solveV, int solveW, T *flowFact, int nFlowFact, T *flowout2);          \
int or_flux_mass_region_loop_##S(or_mesh_##S *m, T *flowFact, int nFlowFact,                       \
    T *flowRegion, int nFlowRegion, T flowIn);                                                     \
int or_flux_scalar_face_loop_##S(or_mesh_##S *m,                                                   \
    T *PhiCell, int nPhiCell, T *PhiBoundary, int nPhiBoundary, T *VisEff, int nVisEff,            \
    T *Au, int nAu, T *Su, int nSu, T *Den, int nDen, T *MassFlux, int nMassFlux,                  \
    T *TE, int nTE, T *CpBoundary, int nCpBoundary, T *visEffBoundary, int nVisEffBoundary,        \
    T *RFace, int nRFace, T *dPhidx, int ndPhidx, int SolveTurb, int SolveEnthalpy,                \
    T vislam, int ivar, int VarT, T Sigma_T, T Prandtl, int VarTE, T Sigma_k, int VarED,           \
    T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large, T TMCmu);                     \
int or_flux_uvw_face_loop1_##S(or_mesh_##S *m, T gammaBlend, T small,                              \
    T *uCell, int nUCell, T *vCell, int nVCell, T *wCell, int nWCell,                              \
    T *uBoundary, int nUBoundary, T *vBoundary, int nVBoundary, T *wBoundary, int nWBoundary,      \
    T *visEffCell, int nVisEffCell, T *visEffBoundary, int nVisEffBoundary,                        \
    T *massFlux, int nMassFlux, T *dudx, int nDudx, T *dvdx, int nDvdx, T *dwdx, int nDwdx,        \
    T *rFace, int nRFace, T *su, int nSu, T *sv, int nSv, T *sw, int nSw,                          \
    T *au, int nAu, T *av, int nAv, T *aw, int nAw);                                               \
void or_flux_uvw_region_loop1_##S(or_mesh_##S *m);                                                 \
void or_flux_uvw_bnds_loop1_##S(or_mesh_##S *m);                                                   \
int or_viscosity_cell_loop1_##S(or_mesh_##S *m, T small, T cmu, T visURF, T visLam,                \
    T *TE, int nTE, T *ED, int nED, T *den, int nDen, T *visEff, int nVisEff);                     \
void or_viscosity_region_loop_##S(T kappa, or_mesh_##S *m);                                        \
int or_viscosity_boundary_loop_##S(or_mesh_##S *m, T tmcmu, T large, T small, T kappa, T visLam,   \
    T *TE, int nTE, T *den, int nDen, T *visEffCell, int nVisEffCell,                              \
    T *visEffBoundary, int nVisEffBoundary);                                                       \
int or_viscosity_cell_loop2_##S(or_mesh_##S *m, T visLam, T *visEffCell, int nVisEffCell,          \
    T *visEffBoundary, int nVisEffBoundary);

OR_DECL(double, f64)
OR_DECL(float, f32)

#undef OR_DECL

/* Parameters of the scalar face loop that its tangent and adjoint read. */
typedef struct {
    int SolveTurb, ivar, VarT, VarTE, VarED;
    double vislam, Sigma_T, Prandtl, Sigma_k, Sigma_e, Sigma_s, Schmidt, GammaBlend;
    int SolveEnthalpy;
    double Small, TMCmu;
} or_scalar_params;

/* Tangent (sudot, rfacedot) and adjoint (phibar, mfbar) of GradientPhiGaussDolfyn
 * (nGradient = 1) followed by FluxScalarDolfynFaceLoop, w.r.t. (phi, MassFlux),
 * on all-DEFAULT meshes (oracle_adjoint.inc). Outputs are overwritten.
 * Returns OR_E_NOT_IMPLEMENTED on a mesh with a non-DEFAULT region. */
int or_scalar_tangent_f64(or_mesh_f64 *m, const double *phi, const double *VisEff, const double *MassFlux,
                          const double *phidot, const double *mfdot, const or_scalar_params *p, double *sudot,
                          double *rfacedot);
int or_scalar_adjoint_f64(or_mesh_f64 *m, const double *phi, const double *VisEff, const double *MassFlux,
                          const double *subar, const double *rbar, const or_scalar_params *p, double *phibar,
                          double *mfbar);

/* The same chain with nGradient >= 1 gradient sweeps, on any mesh: boundary
 * faces, the PhiFace carried from INLET / OUTLET / WALL faces and the
 * enthalpy wall included (oracle_adjoint.inc). dPhidx is the forward
 * gradient the scalar loop read; TE, Den, CpB and visEffB its other inputs.
 * Cotangents are on Su and RFace (PhiBoundary, Au and the mesh state are
 * not differentiated). */
int or_scalar_tangent_full_f64(or_mesh_f64 *m, int nGradient, const double *phi, const double *VisEff,
                               const double *MassFlux, const double *dPhidx, const double *TE, const double *Den,
                               const double *CpB, const double *visEffB, const double *phidot,
                               const double *mfdot, const or_scalar_params *p, double *sudot, double *rfacedot);
int or_scalar_adjoint_full_f64(or_mesh_f64 *m, int nGradient, const double *phi, const double *VisEff,
                               const double *MassFlux, const double *dPhidx, const double *TE, const double *Den,
                               const double *CpB, const double *visEffB, const double *subar,
                               const double *rbar, const or_scalar_params *p, double *phibar, double *mfbar);

#ifdef __cplusplus
}
#endif

#endif

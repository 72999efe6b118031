/*
 * cupcfd_hip.h — C ABI of the MI355X (gfx950) finite-volume kernel library.
 *
 * It replaces the 16 function templates of namespace cupcfd::fvm
 * (include/fvm/implementation/component/{Gradient,Mass,Scalar,UVW,Viscosity}Kernels.h
 * in CUP-CFD) for a MiniSoA mesh (CupCfdSoAMesh<int,T,int>, T = double or float).
 *
 * Conventions
 *   - Every entry point returns an int equal to a cupcfd::error::eCodes value
 *     (include/error/Error.h): CUPCFD_E_SUCCESS (95) on success. HIP runtime
 *     failures map to CUPCFD_E_ERROR. Nothing here throws.
 *   - 3-vectors are described by cupcfd_v3_*: component k of element i lives
 *     at ((char*)base + i*stride) + k*sizeof(T). stride == 0 means packed xyz.
 *     The reference's EuclideanVector<double,3> (vptr + cmp[3]) is
 *     {&v[0].cmp[0], 32}; EuclideanVector<float,3> is {&v[0].cmp[0], 24}.
 *   - Two calling modes per kernel:
 *       cupcfd_fvm_*  "drop-in": host arrays in the mesh's own numbering, with
 *                     the reference's (pointer, length) pairs; synchronous:
 *                     results (and mesh state the kernel mutates) are visible
 *                     on return, as in the reference.
 *       cupcfd_dev_*  "resident": device arrays in the device mesh's internal
 *                     numbering (cell and face orders from
 *                     cupcfd_hip_mesh_cell_order / _face_order; boundary and
 *                     region arrays keep the mesh's numbering), 3-vectors
 *                     packed xyz, launched asynchronously on `stream`
 *                     (a hipStream_t, NULL = default stream). Mesh state stays
 *                     on the device until cupcfd_hip_mesh_pull_state.
 *   - Carried locals. The reference reads PhiFace (scalar loop),
 *     uFace/vFace/wFace/visFace (UVW loop) and yplustmp (viscosity region
 *     loop) before any assignment on meshes whose boundary regions do not
 *     assign them. This library starts them at the mesh's carry-init value
 *     (default 0, cupcfd_hip_mesh_set_carry_init) and otherwise carries them
 *     from face to face in ascending face order exactly as the reference does.
 */
#ifndef CUPCFD_HIP_H
#define CUPCFD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cupcfd::error::eCodes values used at this boundary (include/error/Error.h) */
#define CUPCFD_E_FILE_MISSING 10
#define CUPCFD_E_ARRAY_SIZE_MISMATCH 25
#define CUPCFD_E_INVALID_INDEX 68
#define CUPCFD_E_CONFIG_INVALID_VALUE 90
#define CUPCFD_E_SUCCESS 95
#define CUPCFD_E_ERROR 98
#define CUPCFD_E_NULL_PTR 99
#define CUPCFD_E_INVALID_FILE_FORMAT 105
#define CUPCFD_E_NOT_IMPLEMENTED 109

/* cupcfd::geometry::mesh::RType (UnstructuredMeshInterface.h:58-65) */
#define CUPCFD_RTYPE_DEFAULT 0
#define CUPCFD_RTYPE_INLET 1
#define CUPCFD_RTYPE_OUTLET 2
#define CUPCFD_RTYPE_SYMP 3
#define CUPCFD_RTYPE_WALL 4

typedef struct cupcfd_v3_f64 { double *base; int64_t stride; } cupcfd_v3_f64;
typedef struct cupcfd_v3_f32 { float *base; int64_t stride; } cupcfd_v3_f32;

/*
 * A borrowed view of a MiniSoA mesh: one pointer per std::vector member of
 * CupCfdSoAMesh (CupCfdSoAMesh.h:71-154) that the fvm kernels read or write,
 * plus UnstructuredMeshProperties' local counts (lTCells, lFaces,
 * lBoundaries, lRegions). std::vector<bool> members (regionStd/Flux/Adiab)
 * are passed as uint8_t arrays. Pointers marked "mutable" are mesh state
 * that kernels set (setBoundaryShear/Q/H/T/YPlus/UPlus,
 * setRegionYLog/ForceTangent); drop-in calls write them back here.
 */
#define CUPCFD_MESH_VIEW_FIELDS(T, V3)                                                    \
    int32_t nCells, nFaces, nBoundaries, nRegions;                                        \
    const int32_t *faceCell1ID, *faceCell2ID, *faceBoundaryID;                            \
    const T *faceLambda, *faceRLencos, *faceArea;                                         \
    V3 faceNorm, faceCenter, faceXpac, faceXnac;                                          \
    V3 cellCenter;                                                                        \
    const T *cellVolume;                                                                  \
    const int32_t *boundaryFaceID, *boundaryRegionID;                                     \
    const T *boundaryDistance;                                                            \
    T *boundaryYPlus, *boundaryUPlus, *boundaryQ, *boundaryH, *boundaryT; /* mutable */   \
    V3 boundaryShear;                                                     /* mutable */   \
    const int32_t *regionType;                                                            \
    const uint8_t *regionStd, *regionFlux, *regionAdiab;                                  \
    T *regionYLog;                                                        /* mutable */   \
    const T *regionELog, *regionDensity, *regionTurbKE, *regionTurbDiss;                  \
    const T *regionSplvl, *regionDen, *regionR, *regionT;                                 \
    V3 regionForceTangent;                                                /* mutable */   \
    V3 regionUVW;

typedef struct cupcfd_mesh_view_f64 { CUPCFD_MESH_VIEW_FIELDS(double, cupcfd_v3_f64) } cupcfd_mesh_view_f64;
typedef struct cupcfd_mesh_view_f32 { CUPCFD_MESH_VIEW_FIELDS(float, cupcfd_v3_f32) } cupcfd_mesh_view_f32;

/* ---------------------------------------------------------------------------
 * Host mesh sources (run before the kernels; host C++).
 * ------------------------------------------------------------------------- */
typedef struct cupcfd_host_mesh cupcfd_host_mesh;

/*
 * Structured hex grid with MeshStructGenSource's numbering
 * (src/geometry/mesh/implementation/source/MeshStructGenSource.cpp:33-1259:
 * cell label x + y*nx + z*nx*ny, faces YZ then XZ then XY, boundary labels,
 * cell1 = low-side cell, cell2 = high-side cell or -1, lambda 0.5 / -1) as
 * assembled by UnstructuredMeshInterface::addData on one rank. Geometry
 * follows the reference's own computation bit for bit: face areas and
 * centres from Heron triangles of the face's four vertices
 * (Polygon3D.cpp:31-55), normals by the clockwise test from cell1's centre
 * (MeshStructGenSource.cpp:887-944), cell volumes as five tetrahedra of the
 * hexahedron (Hexahedron.ipp:30-68).
 *   regions = 0: one RTYPE_DEFAULT region, boundary distance 0 (what the
 *                reference builds); regions = 1: six typed regions, x-low
 *                INLET, x-high OUTLET, y-low/y-high WALL, z-low/z-high SYMP,
 *                boundary distance = half the cell width normal to the face.
 *   face_extras = 0: xpac = xnac = rlencos = 0 (what the reference builds);
 *                1: seeded non-zero xpac/xnac/rlencos so those terms are live.
 *   permute_cells / permute_faces: seeded Fisher-Yates relabelling of cell /
 *                face ids (boundary ids keep StructGen's labels).
 *   tets = 1:    split every hex into 6 tetrahedra around its main diagonal
 *                (12 faces per hex interior) with vertices jittered by
 *                jitter * spacing (boundary vertices stay on their planes).
 */
typedef struct cupcfd_structgen_params {
    int32_t nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    uint64_t seed;
    int32_t permute_cells, permute_faces;
    int32_t regions;
    int32_t face_extras;
    int32_t tets;
    double jitter;
} cupcfd_structgen_params;

int cupcfd_structgen_create(const cupcfd_structgen_params *params, cupcfd_host_mesh **out);
/*
 * A mesh read from one of the reference's HDF5 mesh files (MeshHDF5Source,
 * src/geometry/mesh/implementation/source/MeshHDF5Source.cpp: 1-based labels,
 * /cell, /face, /bnd groups), assembled as UnstructuredMeshInterface::addData
 * assembles it on one rank: cells in label order, the faces the cells list in
 * label order, their boundaries in label order, one RTYPE_DEFAULT region per
 * region label, xpac = xnac = rlencos = 0, real data converted to double.
 * libhdf5 is loaded at run time (CUPCFD_HDF5_LIBRARY, else libhdf5.so[.103]
 * on the loader path, else /opt/conda/lib/libhdf5.so.103); without it this
 * returns CUPCFD_E_NOT_IMPLEMENTED. A path that does not exist gives
 * CUPCFD_E_FILE_MISSING; a file libhdf5 cannot open, or missing or
 * mis-shaped datasets, give CUPCFD_E_INVALID_FILE_FORMAT; labels out of
 * range give CUPCFD_E_INVALID_INDEX.
 */
int cupcfd_hdf5_mesh_create(const char *path, cupcfd_host_mesh **out);
int cupcfd_host_mesh_view_f64(cupcfd_host_mesh *mesh, cupcfd_mesh_view_f64 *out);
/* float copies of every array, owned by the host mesh */
int cupcfd_host_mesh_view_f32(cupcfd_host_mesh *mesh, cupcfd_mesh_view_f32 *out);
void cupcfd_host_mesh_destroy(cupcfd_host_mesh *mesh);

/* ---------------------------------------------------------------------------
 * Device mesh: the view repacked once into HBM (cell→face CSR whose rows are
 * sorted by the mesh's face id, so per-cell accumulation follows the
 * reference's sequential face order bit for bit, with no atomics).
 * ------------------------------------------------------------------------- */
typedef struct cupcfd_hip_mesh cupcfd_hip_mesh;

/* Renumber cells along a Morton curve of their centres (and faces by owner)
 * for locality. Results are returned in the mesh's numbering either way. */
#define CUPCFD_MESH_REORDER 1

/* Repacks `view` (CupCfdSoAMesh::finalize must have run: CupCfdSoAMesh.cpp:842-888).
 * The view's pointers are kept: drop-in calls read and write the mutable
 * mesh state through them, so they must outlive the device mesh. */
int cupcfd_hip_mesh_create_f64(const cupcfd_mesh_view_f64 *view, int32_t flags, cupcfd_hip_mesh **out);
int cupcfd_hip_mesh_create_f32(const cupcfd_mesh_view_f32 *view, int32_t flags, cupcfd_hip_mesh **out);
/* The same repack run on the GPU (stable radix sorts and scans over the
 * uploaded MiniSoA arrays) instead of the host: every device array and every
 * order it reports are bit-identical to cupcfd_hip_mesh_create_*'s, and the
 * errors are the same. It replaces the host-side map building of
 * CupCfdSoAMesh::updateCellFaceMap (CupCfdSoAMesh.cpp:549-668), which the
 * reference runs in finalize. */
int cupcfd_hip_mesh_create_device_f64(const cupcfd_mesh_view_f64 *view, int32_t flags, cupcfd_hip_mesh **out);
int cupcfd_hip_mesh_create_device_f32(const cupcfd_mesh_view_f32 *view, int32_t flags, cupcfd_hip_mesh **out);
/* One 64-bit hash per array of the device mesh and per host-side order it
 * keeps, in a fixed order (cupcfd_hip_mesh_digest_field names entry i), so two
 * builds can be compared field by field. count receives the number of
 * entries; capacity smaller than that returns CUPCFD_E_ARRAY_SIZE_MISMATCH.
 * Synchronous; no counterpart in the reference. */
int cupcfd_hip_mesh_digest(const cupcfd_hip_mesh *mesh, uint64_t *digest, int32_t capacity, int32_t *count);
const char *cupcfd_hip_mesh_digest_field(int32_t i);
/* Generate the mesh cupcfd_structgen_create(params) describes on the GPU and
 * repack it there: no per-cell or per-face array crosses PCIe (the
 * relabelling's two permutations are drawn on the host). The device mesh is
 * bit-identical to cupcfd_hip_mesh_create_*(view of cupcfd_structgen_create's
 * mesh, flags), and the errors are the generator's. host_out (may be NULL)
 * receives a host copy equal to cupcfd_structgen_create's mesh; the device
 * mesh then keeps that copy's view, as cupcfd_hip_mesh_create_* keeps the
 * caller's. With host_out NULL the mesh keeps no view: its state starts as
 * the generator leaves it and stays on the device (push/pull state do
 * nothing). The caller destroys both objects, the device mesh first. It
 * replaces MeshStructGenSource plus the finalize-time map building of
 * CupCfdSoAMesh (MeshStructGenSource.cpp:33-1259, CupCfdSoAMesh.cpp:549-668). */
int cupcfd_hip_mesh_generate_f64(const cupcfd_structgen_params *params, int32_t flags, cupcfd_host_mesh **host_out,
                                 cupcfd_hip_mesh **out);
int cupcfd_hip_mesh_generate_f32(const cupcfd_structgen_params *params, int32_t flags, cupcfd_host_mesh **host_out,
                                 cupcfd_hip_mesh **out);
int cupcfd_hip_mesh_destroy(cupcfd_hip_mesh *mesh);
int cupcfd_hip_mesh_counts(const cupcfd_hip_mesh *mesh, int32_t *nCells, int32_t *nFaces,
                           int32_t *nBoundaries, int32_t *nRegions);
/* internal → mesh numbering: order[k] is the mesh id at internal position k */
int cupcfd_hip_mesh_cell_order(const cupcfd_hip_mesh *mesh, int32_t *order);
int cupcfd_hip_mesh_face_order(const cupcfd_hip_mesh *mesh, int32_t *order);
int cupcfd_hip_mesh_set_carry_init(cupcfd_hip_mesh *mesh, double value);
/* How the boundary loops add their per-boundary values (FluxMassDolfyn
 * BoundaryLoop1-5's flow and area sums, FluxUVWDolfynBndsLoop1's force
 * tangents; no counterpart in the reference, whose loops are sequential):
 * sequential != 0 (the default) adds them one after another in boundary
 * order, as the reference does, so the sums are its bits; sequential == 0
 * adds them in a fixed-shape parallel tree (one workgroup: a strided partial
 * sum per thread, then a fixed reduction tree), which is much faster for long
 * boundary lists and gives the same bits on every call, within a few ulps of
 * the sum's magnitude of the sequential result rather than equal to it. */
int cupcfd_hip_mesh_set_ordered_sums(cupcfd_hip_mesh *mesh, int32_t sequential);
/* The mesh's cell-tile layout (diagnostic; no counterpart in the reference):
 * info[0..5] = number of 64-cell tiles, most face slots of any tile, most CSR
 * entries of any tile, most slots of any tile whose face has a cell outside
 * the tile or is a boundary face, most slots between two cells of one tile,
 * and 1 when the UVW tile runs its split LDS layout on this mesh (FP64
 * meshes whose tiles have at most 96 and 144 slots of the two kinds: every
 * 4x4x4-cell hex tile), else 0. */
int cupcfd_hip_mesh_tile_layout(const cupcfd_hip_mesh *mesh, int32_t info[6]);
/* Gather schedule of the per-cell face loops (gradient, scalar, UVW): 64 for
 * tiles of 64 cells whose values are staged in LDS and whose faces are each
 * evaluated once; lanes evaluating one cell's faces in parallel (4, 6 or 8; a
 * row longer than that falls back to one thread per cell); 1 for one
 * thread per cell; or 2 for colour groups (gradient only; the scalar and UVW
 * loops then take the automatic schedule): one launch per colour of the
 * mesh's order-preserving face colouring (cupcfd_hip_mesh_face_colours), in
 * which each face adds its terms straight to its cells. The colouring is
 * built on the schedule's first call, which allocates and synchronises. 0 (default) takes tiles when the mesh's tiles fit (at most
 * 32767 face slots per tile and 64 KiB of LDS), else the row length of
 * all-tet / all-hex meshes, else 8. The mass-flux face loop takes the same
 * tiles (each tile evaluating the interior faces whose cell1 it holds) for 0
 * and 64, and one thread per face for the other schedules. Every schedule
 * gives bit-identical results. */
int cupcfd_hip_mesh_set_schedule(cupcfd_hip_mesh *mesh, int32_t lanes_per_cell);
/* The mesh's order-preserving face colouring, built on the device on first
 * use: colour(f) = 1 + the largest colour among the faces before f (in the
 * mesh's face order) that share a cell with it, 0 if none. So no two faces of
 * one colour share a cell, and each cell's faces have rising colours in face
 * order: colour-by-colour accumulation adds a cell's terms in the
 * reference's order. colour (nFaces entries, mesh face numbering; may be
 * NULL) and ncolours (may be NULL) receive it. Synchronous; no counterpart in
 * the reference (its loops are sequential). */
int cupcfd_hip_mesh_face_colours(cupcfd_hip_mesh *mesh, int32_t *colour, int32_t *ncolours);
/* Choose the gather schedule of the gradient, scalar and UVW loops separately
 * by timing each candidate (cell tiles, the lane-per-face width the mesh
 * allows, one thread per cell, and colour groups for the gradient) on scratch
 * inputs on a private stream, and that of the mass-flux face loop (cell tiles
 * or one thread per face) the same way; the
 * mesh's boundary and region state is left as it was. schedules (may be NULL)
 * receives the three gather choices (the mass-flux choice is kept on the mesh
 * and in the cache file). A later cupcfd_hip_mesh_set_schedule replaces
 * them. The choice is cached per (GPU architecture, precision, mesh class:
 * row lengths, tile applicability and LDS width, typed regions, cell count
 * to the nearest power of two), so a later call on a mesh of the same class
 * takes it without timing; with the environment variable
 * CUPCFD_HIP_TUNE_CACHE naming a file, the cache is read from and written to
 * that file ("key gradient scalar uvw mass" per line; a line without the
 * mass entry is read too), so it also persists across
 * processes. Synchronous; no counterpart in the reference. */
int cupcfd_hip_mesh_autotune(cupcfd_hip_mesh *mesh, int32_t *schedules);
/* Forget the process's autotune cache (the file, if any, is left alone). */
int cupcfd_hip_autotune_forget(void);
/* copy mutable mesh state view → device (push) or device → view (pull) */
int cupcfd_hip_mesh_push_state(cupcfd_hip_mesh *mesh);
int cupcfd_hip_mesh_pull_state(cupcfd_hip_mesh *mesh);
/* resident-mode layout helpers: mesh-numbered device arrays ⇄ internal order,
 * ncomp interleaved components per element (1 for scalars, 3 for vectors,
 * 2 for rface) */
int cupcfd_dev_cells_to_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                 int32_t elem_bytes, void *stream);
int cupcfd_dev_cells_from_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                   int32_t elem_bytes, void *stream);
int cupcfd_dev_faces_to_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                 int32_t elem_bytes, void *stream);
int cupcfd_dev_faces_from_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                   int32_t elem_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Kernels. Declared once per precision by CUPCFD_FVM_API(T, V3, S).
 * Reference interfaces replaced (include/fvm/implementation/component/):
 *   gradient_phi_gauss_dolfyn   GradientPhiGaussDolfyn           GradientKernels.h:48-54
 *   flux_mass_face_loop         FluxMassDolfynFaceLoop           MassKernels.h:48-81
 *   flux_mass_boundary_loop1    FluxMassDolfynBoundaryLoop1      MassKernels.h:90-94
 *   flux_mass_boundary_loop2    FluxMassDolfynBoundaryLoop2      MassKernels.h:103-108
 *   flux_mass_boundary_loop3    FluxMassDolfynBoundaryLoop3      MassKernels.h:117-120
 *   flux_mass_boundary_loop4    FluxMassDolfynBoundaryLoop4      MassKernels.h:129-138
 *   flux_mass_boundary_loop5    FluxMassDolfynBoundaryLoop5      MassKernels.h:147-157
 *   flux_mass_region_loop       FluxMassDolfynRegionLoop         MassKernels.h:166-171
 *   flux_scalar_face_loop       FluxScalarDolfynFaceLoop         ScalarKernels.h:47-67
 *   flux_uvw_face_loop1         FluxUVWDolfynFaceLoop1           UVWKernels.h:43-67
 *   flux_uvw_region_loop1       FluxUVWDolfynRegionLoop1         UVWKernels.h:72-73
 *   flux_uvw_bnds_loop1         FluxUVWDolfynBndsLoop1           UVWKernels.h:78-79
 *   viscosity_cell_loop1        calculateViscosityDolfynCellLoop1    ViscosityKernels.h:38-48
 *   viscosity_region_loop       calculateViscosityDolfynRegionLoop   ViscosityKernels.h:50-51
 *   viscosity_boundary_loop     calculateViscosityDolfynBoundaryLoop ViscosityKernels.h:53-60
 *   viscosity_cell_loop2        calculateViscosityDolfynCellLoop2    ViscosityKernels.h:62-67
 * Argument order, meaning and lengths follow the reference declarations;
 * bool parameters are int32_t (0 / non-zero). The mass face loop's counters
 * icinl..icwal are host pointers in both modes.
 * ------------------------------------------------------------------------- */
#define CUPCFD_FVM_API(T, V3, S)                                                                   \
/* ---- drop-in: host arrays ---- */                                                              \
int cupcfd_fvm_gradient_phi_gauss_dolfyn_##S(cupcfd_hip_mesh *mesh, int32_t nGradient,             \
    const T *phiCell, int32_t nPhiCell, const T *phiBoundary, int32_t nPhiBoundary,                \
    V3 dPhidxCell, int32_t nDPhidxCell, V3 dPhidxoCell, int32_t nDPhidxoCell);                     \
int cupcfd_fvm_flux_mass_face_loop_##S(cupcfd_hip_mesh *mesh,                                      \
    V3 dudx, int32_t nDudx, V3 dvdx, int32_t nDvdx, V3 dwdx, int32_t nDwdx, V3 dpdx, int32_t nDpdx,\
    const T *denCell, int32_t nDenCell, T *denBoundary, int32_t nDenBoundary,                      \
    const T *uCell, int32_t nUCell, const T *vCell, int32_t nVCell, const T *wCell, int32_t nWCell,\
    T *massFlux, int32_t nMassFlux, const T *p, int32_t nP, const T *ar, int32_t nAr,              \
    T *su, int32_t nSu, T *rface, int32_t nRFace, T small,                                         \
    int32_t *icinl, int32_t *icout, int32_t *icsym, int32_t *icwal,                                \
    int32_t solveTurbEnergy, int32_t solveTurbDiss, int32_t solveVisc, int32_t solveEnthalpy,      \
    const T *teCell, int32_t nTeCell, T *teBoundary, int32_t nTeBoundary,                          \
    const T *edCell, int32_t nEdCell, T *edBoundary, int32_t nEdBoundary,                          \
    const T *viseffCell, int32_t nViseffCell, T *viseffBoundary, int32_t nViseffBoundary,          \
    const T *tCell, int32_t nTCell, T *tBoundary, int32_t nTBoundary);                             \
int cupcfd_fvm_flux_mass_boundary_loop1_##S(cupcfd_hip_mesh *mesh, const T *massFlux,              \
    int32_t nMassFlux, T *flowin);                                                                 \
int cupcfd_fvm_flux_mass_boundary_loop2_##S(cupcfd_hip_mesh *mesh, const T *massFlux,              \
    int32_t nMassFlux, T *flowRegion, int32_t nFlowRegion, T *flowout);                            \
int cupcfd_fvm_flux_mass_boundary_loop3_##S(cupcfd_hip_mesh *mesh, T flowin, T *ratearea);         \
int cupcfd_fvm_flux_mass_boundary_loop4_##S(cupcfd_hip_mesh *mesh, T *massFlux, int32_t nMassFlux, \
    T *uBoundary, int32_t nUBoundary, T *vBoundary, int32_t nVBoundary, T *wBoundary,              \
    int32_t nWBoundary, const T *denBoundary, int32_t nDenBoundary, T ratearea, T *flowout);       \
int cupcfd_fvm_flux_mass_boundary_loop5_##S(cupcfd_hip_mesh *mesh, T *massFlux, int32_t nMassFlux, \
    T *su, int32_t nSu, T *uBoundary, int32_t nUBoundary, T *vBoundary, int32_t nVBoundary,        \
    T *wBoundary, int32_t nWBoundary, T fact, int32_t solveU, int32_t solveV, int32_t solveW,      \
    const T *flowFact, int32_t nFlowFact, T *flowout2);                                            \
int cupcfd_fvm_flux_mass_region_loop_##S(cupcfd_hip_mesh *mesh, T *flowFact, int32_t nFlowFact,    \
    const T *flowRegion, int32_t nFlowRegion, T flowIn);                                           \
int cupcfd_fvm_flux_scalar_face_loop_##S(cupcfd_hip_mesh *mesh,                                    \
    const T *PhiCell, int32_t nPhiCell, T *PhiBoundary, int32_t nPhiBoundary,                      \
    const T *VisEff, int32_t nVisEff, T *Au, int32_t nAu, T *Su, int32_t nSu,                      \
    const T *Den, int32_t nDen, const T *MassFlux, int32_t nMassFlux, const T *TE, int32_t nTE,    \
    const T *CpBoundary, int32_t nCpBoundary, const T *visEffBoundary, int32_t nVisEffBoundary,    \
    T *RFace, int32_t nRFace, V3 dPhidx, int32_t ndPhidx, int32_t SolveTurb,                       \
    int32_t SolveEnthalpy, T vislam, int32_t ivar, int32_t VarT, T Sigma_T, T Prandtl,             \
    int32_t VarTE, T Sigma_k, int32_t VarED, T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend,        \
    T Small, T Large, T TMCmu);                                                                    \
int cupcfd_fvm_flux_uvw_face_loop1_##S(cupcfd_hip_mesh *mesh, T gammaBlend, T small,               \
    const T *uCell, int32_t nUCell, const T *vCell, int32_t nVCell, const T *wCell, int32_t nWCell,\
    T *uBoundary, int32_t nUBoundary, T *vBoundary, int32_t nVBoundary, T *wBoundary,              \
    int32_t nWBoundary, const T *visEffCell, int32_t nVisEffCell, const T *visEffBoundary,         \
    int32_t nVisEffBoundary, T *massFlux, int32_t nMassFlux, V3 dudx, int32_t nDudx,               \
    V3 dvdx, int32_t nDvdx, V3 dwdx, int32_t nDwdx, T *rFace, int32_t nRFace,                      \
    T *su, int32_t nSu, T *sv, int32_t nSv, T *sw, int32_t nSw,                                    \
    T *au, int32_t nAu, T *av, int32_t nAv, T *aw, int32_t nAw);                                   \
int cupcfd_fvm_flux_uvw_region_loop1_##S(cupcfd_hip_mesh *mesh);                                   \
int cupcfd_fvm_flux_uvw_bnds_loop1_##S(cupcfd_hip_mesh *mesh);                                     \
int cupcfd_fvm_viscosity_cell_loop1_##S(cupcfd_hip_mesh *mesh, T small, T cmu, T visURF,           \
    T visLam, const T *TE, int32_t nTE, const T *ED, int32_t nED, const T *den, int32_t nDen,      \
    T *visEff, int32_t nVisEff);                                                                   \
int cupcfd_fvm_viscosity_region_loop_##S(T kappa, cupcfd_hip_mesh *mesh);                          \
int cupcfd_fvm_viscosity_boundary_loop_##S(cupcfd_hip_mesh *mesh, T tmcmu, T large, T small,       \
    T kappa, T visLam, const T *TE, int32_t nTE, const T *den, int32_t nDen,                       \
    const T *visEffCell, int32_t nVisEffCell, T *visEffBoundary, int32_t nVisEffBoundary);         \
int cupcfd_fvm_viscosity_cell_loop2_##S(cupcfd_hip_mesh *mesh, T visLam, T *visEffCell,            \
    int32_t nVisEffCell, T *visEffBoundary, int32_t nVisEffBoundary);                              \
/* ---- resident: device arrays, internal numbering, async on stream ---- */                      \
int cupcfd_dev_gradient_phi_gauss_dolfyn_##S(cupcfd_hip_mesh *mesh, int32_t nGradient,             \
    const T *phiCell, const T *phiBoundary, T *dPhidxCell, T *dPhidxoCell, void *stream);          \
int cupcfd_dev_flux_mass_face_loop_##S(cupcfd_hip_mesh *mesh,                                      \
    const T *dudx, const T *dvdx, const T *dwdx, const T *dpdx, const T *denCell, T *denBoundary,  \
    const T *uCell, const T *vCell, const T *wCell, T *massFlux, const T *p, const T *ar,          \
    T *su, T *rface, T small, int32_t *icinl, int32_t *icout, int32_t *icsym, int32_t *icwal,      \
    int32_t solveTurbEnergy, int32_t solveTurbDiss, int32_t solveVisc, int32_t solveEnthalpy,      \
    const T *teCell, T *teBoundary, const T *edCell, T *edBoundary, const T *viseffCell,           \
    T *viseffBoundary, const T *tCell, T *tBoundary, void *stream);                                \
int cupcfd_dev_flux_mass_boundary_loop1_##S(cupcfd_hip_mesh *mesh, const T *massFlux, T *flowin,   \
    void *stream);                                                                                 \
int cupcfd_dev_flux_mass_boundary_loop2_##S(cupcfd_hip_mesh *mesh, const T *massFlux,              \
    T *flowRegion, T *flowout, void *stream);                                                      \
int cupcfd_dev_flux_mass_boundary_loop3_##S(cupcfd_hip_mesh *mesh, T flowin, T *ratearea,          \
    void *stream);                                                                                 \
int cupcfd_dev_flux_mass_boundary_loop4_##S(cupcfd_hip_mesh *mesh, T *massFlux, T *uBoundary,      \
    T *vBoundary, T *wBoundary, const T *denBoundary, T ratearea, T *flowout, void *stream);       \
int cupcfd_dev_flux_mass_boundary_loop5_##S(cupcfd_hip_mesh *mesh, T *massFlux, T *su,             \
    T *uBoundary, T *vBoundary, T *wBoundary, T fact, int32_t solveU, int32_t solveV,              \
    int32_t solveW, const T *flowFact, T *flowout2, void *stream);                                 \
int cupcfd_dev_flux_mass_region_loop_##S(cupcfd_hip_mesh *mesh, T *flowFact, const T *flowRegion,  \
    T flowIn, void *stream);                                                                       \
int cupcfd_dev_flux_scalar_face_loop_##S(cupcfd_hip_mesh *mesh,                                    \
    const T *PhiCell, T *PhiBoundary, const T *VisEff, T *Au, T *Su, const T *Den,                 \
    const T *MassFlux, const T *TE, const T *CpBoundary, const T *visEffBoundary, T *RFace,        \
    const T *dPhidx, int32_t SolveTurb, int32_t SolveEnthalpy, T vislam, int32_t ivar,             \
    int32_t VarT, T Sigma_T, T Prandtl, int32_t VarTE, T Sigma_k, int32_t VarED, T Sigma_e,        \
    T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large, T TMCmu, void *stream);                  \
int cupcfd_dev_flux_uvw_face_loop1_##S(cupcfd_hip_mesh *mesh, T gammaBlend, T small,               \
    const T *uCell, const T *vCell, const T *wCell, T *uBoundary, T *vBoundary, T *wBoundary,      \
    const T *visEffCell, const T *visEffBoundary, T *massFlux, const T *dudx, const T *dvdx,       \
    const T *dwdx, T *rFace, T *su, T *sv, T *sw, T *au, T *av, T *aw, void *stream);              \
int cupcfd_dev_flux_uvw_region_loop1_##S(cupcfd_hip_mesh *mesh, void *stream);                     \
int cupcfd_dev_flux_uvw_bnds_loop1_##S(cupcfd_hip_mesh *mesh, void *stream);                       \
int cupcfd_dev_viscosity_cell_loop1_##S(cupcfd_hip_mesh *mesh, T small, T cmu, T visURF,           \
    T visLam, const T *TE, const T *ED, const T *den, T *visEff, void *stream);                    \
int cupcfd_dev_viscosity_region_loop_##S(T kappa, cupcfd_hip_mesh *mesh, void *stream);            \
int cupcfd_dev_viscosity_boundary_loop_##S(cupcfd_hip_mesh *mesh, T tmcmu, T large, T small,       \
    T kappa, T visLam, const T *TE, const T *den, const T *visEffCell, T *visEffBoundary,          \
    void *stream);                                                                                 \
int cupcfd_dev_viscosity_cell_loop2_##S(cupcfd_hip_mesh *mesh, T visLam, T *visEffCell,            \
    T *visEffBoundary, void *stream);

CUPCFD_FVM_API(double, cupcfd_v3_f64, f64)
CUPCFD_FVM_API(float, cupcfd_v3_f32, f32)

/* ---------------------------------------------------------------------------
 * Batched kernels: K scalars on one mesh, resident mode (device arrays,
 * internal numbering, async on stream). No counterpart in the reference: the
 * result for scalar k is, bit for bit, that of K separate calls of
 *   gradient_batched      GradientPhiGaussDolfyn     GradientKernels.h:48-54
 *   flux_scalar_batched   FluxScalarDolfynFaceLoop   ScalarKernels.h:47-67
 * on scalar k's arrays. Per-scalar arrays are scalar-innermost:
 *   cell scalars     phiCell, Au, Su            [c*K + k]
 *   boundary values  phiBoundary, PhiBoundary   [b*K + k]
 *   cell vectors     dPhidxCell, dPhidxoCell, dPhidx   [(c*3 + j)*K + k]
 * Shared by the K scalars (one array, as in the single-scalar entry points):
 * VisEff, Den, MassFlux, TE, CpBoundary, visEffBoundary, and RFace (written
 * once: it does not depend on phi). Mesh state written by the scalar loop
 * (boundary Q/H/T) is left as scalar K-1's call leaves it. K >= 1; returns
 * E_CONFIG_INVALID_VALUE for K < 1.
 * ------------------------------------------------------------------------- */
#define CUPCFD_BATCH_API(T, S)                                                                     \
int cupcfd_dev_gradient_batched_##S(cupcfd_hip_mesh *mesh, int32_t nGradient, int32_t K,           \
    const T *phiCell, const T *phiBoundary, T *dPhidxCell, T *dPhidxoCell, void *stream);          \
int cupcfd_dev_flux_scalar_batched_##S(cupcfd_hip_mesh *mesh, int32_t K,                           \
    const T *PhiCell, T *PhiBoundary, const T *VisEff, T *Au, T *Su, const T *Den,                 \
    const T *MassFlux, const T *TE, const T *CpBoundary, const T *visEffBoundary, T *RFace,        \
    const T *dPhidx, int32_t SolveTurb, int32_t SolveEnthalpy, T vislam, int32_t ivar,             \
    int32_t VarT, T Sigma_T, T Prandtl, int32_t VarTE, T Sigma_k, int32_t VarED, T Sigma_e,        \
    T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large, T TMCmu, void *stream);

CUPCFD_BATCH_API(double, f64)
CUPCFD_BATCH_API(float, f32)

/* Batched gradient with FP32 storage and FP64 arithmetic, on a 64-bit mesh
 * (no counterpart in the reference). phiCell, phiBoundary, dPhidxCell and
 * dPhidxoCell are float, laid out as cupcfd_dev_gradient_batched_f64's; the
 * geometry, every product and sum, and the intermediate gradients of
 * nGradient > 1 are double, and only the final gradient is rounded to float.
 * The result is therefore cupcfd_dev_gradient_batched_f64's on the inputs
 * widened to double, rounded to float (round to nearest even). */
int cupcfd_dev_gradient_batched_f32s(cupcfd_hip_mesh *mesh, int32_t nGradient, int32_t K,
    const float *phiCell, const float *phiBoundary, float *dPhidxCell, float *dPhidxoCell, void *stream);

/* Batched scalar flux with FP32 storage and FP64 arithmetic, on a 64-bit mesh
 * (no counterpart in the reference). The per-scalar arrays PhiCell,
 * PhiBoundary, Au, Su and dPhidx are float, laid out as
 * cupcfd_dev_flux_scalar_batched_f64's; the shared arrays (VisEff, Den,
 * MassFlux, TE, CpBoundary, visEffBoundary, RFace), the geometry and every
 * product and sum are double. Each Au, Su and PhiBoundary entry is widened,
 * updated in double and rounded to float once, so the result is
 * cupcfd_dev_flux_scalar_batched_f64's on the inputs widened to double,
 * rounded to float (round to nearest even). */
int cupcfd_dev_flux_scalar_batched_f32s(cupcfd_hip_mesh *mesh, int32_t K, const float *PhiCell,
    float *PhiBoundary, const double *VisEff, float *Au, float *Su, const double *Den,
    const double *MassFlux, const double *TE, const double *CpBoundary,
    const double *visEffBoundary, double *RFace, const float *dPhidx, int32_t SolveTurb,
    int32_t SolveEnthalpy, double vislam, int32_t ivar, int32_t VarT, double Sigma_T,
    double Prandtl, int32_t VarTE, double Sigma_k, int32_t VarED, double Sigma_e,
    double Sigma_s, double Schmidt, double GammaBlend, double Small, double Large, double TMCmu,
    void *stream);

/* ---------------------------------------------------------------------------
 * Mesh sets: one persistent launch over many meshes of one precision (no
 * counterpart in the reference; a rank owning several partitions would make
 * one call per partition). The set holds the members' device descriptors; it
 * must be destroyed before any member, and a member's schedule settings do
 * not apply to set launches.
 * cupcfd_dev_gradient_batched_set_*: GradientPhiGaussDolfyn with nGradient = 1
 * for K scalars on every member, bit for bit what
 * cupcfd_dev_gradient_batched_* gives per member. phiCell .. dPhidxoCell are
 * DEVICE arrays of n device pointers (member i's arrays, laid out as for the
 * batched entry point); dPhidxoCell may be NULL. The work (blocks of cells of
 * every member) is split into contiguous ranges over a fixed grid of
 * workgroups that each leave after their last range.
 * ------------------------------------------------------------------------- */
typedef struct cupcfd_hip_meshset cupcfd_hip_meshset;
int cupcfd_hip_meshset_create(cupcfd_hip_mesh *const *meshes, int32_t n, cupcfd_hip_meshset **out);
int cupcfd_hip_meshset_destroy(cupcfd_hip_meshset *set);
int cupcfd_dev_gradient_batched_set_f64(cupcfd_hip_meshset *set, int32_t K, const double *const *phiCell,
    const double *const *phiBoundary, double *const *dPhidxCell, double *const *dPhidxoCell, void *stream);
int cupcfd_dev_gradient_batched_set_f32(cupcfd_hip_meshset *set, int32_t K, const float *const *phiCell,
    const float *const *phiBoundary, float *const *dPhidxCell, float *const *dPhidxoCell, void *stream);

/* ---------------------------------------------------------------------------
 * Adjoint (resident mode). No counterpart in the reference: the reverse mode
 * of GradientPhiGaussDolfyn (GradientKernels.h:48-54, nGradient = 1) followed
 * by FluxScalarDolfynFaceLoop (ScalarKernels.h:47-67), with respect to
 * (PhiCell, MassFlux). Given the cotangents SuBar[C] on Su and RFaceBar[2F]
 * on RFace, writes PhiBar[C] and MassFluxBar[F] (overwritten, internal
 * numbering). The scalar loop's parameters have the same meaning as there;
 * the derivative of min/max(mf, 0) at mf = 0 is taken as 0. Meshes whose
 * regions are all DEFAULT (the reference StructGen's) only: other meshes
 * return E_NOT_IMPLEMENTED. Sums are taken per cell over its faces in
 * ascending face id, so results are deterministic.
 * ------------------------------------------------------------------------- */
/* Batched adjoint: K scalars with the batched layout (PhiCell, SuBar, PhiBar
 * [c*K + k]) sharing VisEff, MassFlux and RFace. PhiBar for scalar k is, bit
 * for bit, cupcfd_dev_scalar_adjoint_f64's for that scalar; MassFluxBar is the
 * RFace cotangent term, then the K scalars' terms added in order k = 0..K-1
 * (K = 1 gives cupcfd_dev_scalar_adjoint_f64's result). */
int cupcfd_dev_scalar_adjoint_batched_f64(cupcfd_hip_mesh *mesh, int32_t K, const double *PhiCell,
    const double *VisEff, const double *MassFlux, const double *SuBar, const double *RFaceBar,
    int32_t SolveTurb, double vislam, int32_t ivar, int32_t VarT, double Sigma_T, double Prandtl,
    int32_t VarTE, double Sigma_k, int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt,
    double GammaBlend, double *PhiBar, double *MassFluxBar, void *stream);
/* The batched adjoint with FP32 storage of the per-scalar arrays (PhiCell,
 * SuBar, PhiBar float, laid out as in cupcfd_dev_scalar_adjoint_batched_f64);
 * VisEff, MassFlux, RFaceBar, MassFluxBar and all arithmetic are double. The
 * result is cupcfd_dev_scalar_adjoint_batched_f64's on the inputs widened to
 * double, with PhiBar rounded to float once (MassFluxBar exactly). No
 * counterpart in the reference. */
int cupcfd_dev_scalar_adjoint_batched_f32s(cupcfd_hip_mesh *mesh, int32_t K, const float *PhiCell,
    const double *VisEff, const double *MassFlux, const float *SuBar, const double *RFaceBar,
    int32_t SolveTurb, double vislam, int32_t ivar, int32_t VarT, double Sigma_T, double Prandtl,
    int32_t VarTE, double Sigma_k, int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt,
    double GammaBlend, float *PhiBar, double *MassFluxBar, void *stream);
int cupcfd_dev_scalar_adjoint_f64(cupcfd_hip_mesh *mesh, const double *PhiCell, const double *VisEff,
    const double *MassFlux, const double *SuBar, const double *RFaceBar, int32_t SolveTurb,
    double vislam, int32_t ivar, int32_t VarT, double Sigma_T, double Prandtl, int32_t VarTE,
    double Sigma_k, int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt, double GammaBlend,
    double *PhiBar, double *MassFluxBar, void *stream);
/* The adjoint of GradientPhiGaussDolfyn with nGradient >= 1 sweeps
 * (GradientKernels.ipp:38-136, its correction sweeps included) followed by
 * FluxScalarDolfynFaceLoop (ScalarKernels.ipp:36-463) w.r.t. (PhiCell,
 * MassFlux), on any mesh: the INLET and OUTLET source terms, the PhiFace that
 * OUTLET, INLET and enthalpy WALL faces carry to the faces after them, and the
 * WALL's Hcoef term are transposed. No counterpart in the reference. dPhidx
 * is the forward gradient the scalar loop read; TE, Den, CpBoundary and
 * visEffBoundary are its other inputs (all four, and dPhidx, may be NULL on
 * an all-DEFAULT mesh, where nothing reads them). Cotangents are on Su (SuBar,
 * nCells) and RFace (RFaceBar, 2 nFaces); PhiBoundary, Au and the mesh state
 * are not differentiated. Writes PhiBar (nCells) and MassFluxBar (nFaces).
 * Device arrays in internal numbering (boundary arrays in mesh numbering), as
 * for the other resident calls. With nGradient = 1 on an all-DEFAULT mesh the
 * result is cupcfd_dev_scalar_adjoint_f64's, bit for bit. */
int cupcfd_dev_scalar_adjoint_full_f64(cupcfd_hip_mesh *mesh, int32_t nGradient, const double *PhiCell,
    const double *VisEff, const double *MassFlux, const double *dPhidx, const double *TE, const double *Den,
    const double *CpBoundary, const double *visEffBoundary, const double *SuBar, const double *RFaceBar,
    int32_t SolveTurb, int32_t SolveEnthalpy, double vislam, int32_t ivar, int32_t VarT, double Sigma_T,
    double Prandtl, int32_t VarTE, double Sigma_k, int32_t VarED, double Sigma_e, double Sigma_s,
    double Schmidt, double GammaBlend, double Small, double TMCmu, double *PhiBar, double *MassFluxBar,
    void *stream);
/* cupcfd_dev_scalar_adjoint_f64 with no device allocation: no counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its pass-A
 * scratch (32 bytes per cell) lives in a fixed 64 MiB device array of the
 * library, one per GPU, held until the process ends; meshes of more than
 * 2,097,152 cells return CUPCFD_E_CONFIG_INVALID_VALUE and launch nothing.
 * Results are bit for bit those of cupcfd_dev_scalar_adjoint_f64. Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
int cupcfd_dev_scalar_adjoint_noalloc_f64(cupcfd_hip_mesh *mesh, const double *PhiCell, const double *VisEff,
    const double *MassFlux, const double *SuBar, const double *RFaceBar, int32_t SolveTurb,
    double vislam, int32_t ivar, int32_t VarT, double Sigma_T, double Prandtl, int32_t VarTE,
    double Sigma_k, int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt, double GammaBlend,
    double *PhiBar, double *MassFluxBar, void *stream);

/* Residual norms of cell fields, for a solver's convergence test after a
 * sweep (no counterpart in the reference, whose kernels leave their residuals
 * in Su / su / sv / sw for the caller). fields: a host array of nfields (1..8)
 * device pointers, each to the mesh's nCells values in internal order
 * (double for _f64; float for _f32, on a 32- or 64-bit mesh). norms: a device
 * array of 2 * nfields doubles receiving, per field, max |x| and
 * sqrt(sum x^2), accumulated in double in an order fixed by the launch shape
 * (the same bits on every call). NaN propagates. Asynchronous on stream; the
 * first call on a mesh sizes a workspace, later calls allocate nothing. */
int cupcfd_dev_residual_norms_f64(cupcfd_hip_mesh *mesh, int32_t nfields, const double *const *fields,
    double *norms, void *stream);
int cupcfd_dev_residual_norms_f32(cupcfd_hip_mesh *mesh, int32_t nfields, const float *const *fields,
    double *norms, void *stream);

/* Library version string and the gfx target it was built for. */
const char *cupcfd_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CUPCFD_HIP_H */

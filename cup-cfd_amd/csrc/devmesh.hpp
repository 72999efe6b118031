// Device mesh: the MiniSoA view repacked into HBM for the gfx950 kernels.
//
// Layout (internal numbering):
//   cells  : Morton order of centres when CUPCFD_MESH_REORDER is set, else the
//            mesh's own order.
//   faces  : interior faces first (positions [0, Fi)), then boundary faces;
//            each group sorted by (internal owner cell, mesh face id).
//   fconn  : int2 (c1, nb) per face; nb >= 0 is cell2, nb < 0 is ~boundary id.
//   fgeoS  : (Sx, Sy, Sz, lambda)     fgeoX : (Xf, area)
//   fgeoP  : (xpac, xnac.x)           fgeoN : (xnac.y, xnac.z)
//   frl    : rlencos
//            so each kernel streams exactly the face geometry it uses (the
//            gradient and interior scalar/UVW faces read only fgeoS and frl).
//   cgeo   : (Cc, volume) per cell, plus a separate volume stream.
//   row/ent: cell→face CSR; ent = ((internal face << 1) | role, other) with
//            role 1 when the cell is the face's cell2 ("in") and `other` the
//            face's other cell, or ~boundary id. Each row is sorted by the
//            mesh's face id, which is the order the reference's sequential
//            face loop visits them, so a per-cell gather reproduces the
//            reference's accumulation order exactly without atomics.
// Boundary and region arrays keep the mesh's numbering.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "cupcfd_hip.h"

struct cupcfd_hip_mesh;

namespace cupcfd_hip {
// cell tiles of the LDS-staged gathers: kTileCells consecutive cells per tile,
// one workgroup of kTileBlock = 4 * kTileCells threads per tile (phase A
// stages one element of a cell field per thread)
constexpr int kTileCells = 64;
constexpr int kTileBlock = 4 * kTileCells;
}  // namespace cupcfd_hip

namespace cupcfd_hip {

template <class T> struct alignas(4 * sizeof(T)) V4 { T x, y, z, w; };
template <class T> struct alignas(2 * sizeof(T)) V2 { T x, y; };

template <class T> struct DevMesh {
    int C, F, Fi, B, R, nnz;
    int maxrow;  // longest cell→face row (faces per cell)
    const int2 *fconn;
    const V4<T> *fgeoS, *fgeoX, *fgeoP;
    const V2<T> *fgeoN;
    const T *frl;
    const V4<T> *cgeo;
    const T *cvol;
    const int *row;
    const int2 *ent;
    int rowlen;  // faces per cell when every row has the same length, else 0
    // boundaries (mesh numbering)
    const int *bface, *bregion, *bcell;
    const T *bdist;
    T *byplus, *buplus, *bq, *bh, *bt, *bshear;
    // regions
    const int *rtype;
    const uint8_t *rstd, *rflux, *radiab;
    T *rylog, *rforcet;
    const T *relog, *rdensity, *rturbke, *rturbdiss, *rsplvl, *rden, *rr, *rt, *ruvw;
    // carried-value sources (typed meshes only; null otherwise):
    // interior face -> last INLET/OUTLET face before it (UVW uFace..wFace),
    // SYMP face     -> last interior/INLET/OUTLET face before it (UVW visFace)
    const int *carry_src;
    int typed;
    T carry_init;
    // boundary sums in the reference's sequential order (0, the default) or as
    // a fixed-shape parallel tree (1; cupcfd_hip_mesh_set_ordered_sums)
    int tree_sums;
    // cells that own at least one boundary face
    const int *bcells;
    int nbc;
    // boundaries of each region type, ascending id
    const int *blist[5];
    int nblist[5];
    // cells owning OUTLET boundaries, with those boundaries in ascending id
    const int *ocells, *orow, *oent;
    int noc;
    // cell tiles for the LDS-staged gathers: tile t holds cells
    // [t*tcells, (t+1)*tcells); its face slots are its owned interior faces,
    // its owned boundary faces, then its halo faces (interior faces whose
    // cell2 is in the tile and cell1 is not). tinfo[t] = (first owned interior
    // face, count, first owned boundary face, count); halo faces of tile t are
    // thf[thp[t] .. thp[t+1]) with their (cell1, cell2) in thc; eslot[e] is
    // the slot of CSR entry e, with bit 15 set when the entry's cell is the
    // face's cell2.
    int tcells, ntiles, tslots;  // tslots: most slots of any tile
    int tnnz;                    // most CSR entries of any tile
    const int4 *tinfo;
    const int *thp, *thf;
    const int2 *thc;
    const uint16_t *eslot;
};

// Host-side owner of a device mesh.
template <class T> struct Mesh {
    DevMesh<T> d{};
    std::vector<void *> allocs;
    // view kept for drop-in state mirroring
    cupcfd_mesh_view_f64 view64{};
    cupcfd_mesh_view_f32 view32{};
    std::vector<int32_t> cell_order, face_order;      // internal -> mesh
    std::vector<int32_t> cell_rank, face_rank;        // mesh -> internal
    std::vector<int32_t> bregion_h;                   // region per boundary
    std::vector<uint8_t> radiab_h;                    // regionAdiab per region
    std::vector<int32_t> fconn_h;                     // 2F, host copy for carry precompute
    std::vector<int32_t> btype_h;                     // type per boundary
    int32_t ntype[5] = {0, 0, 0, 0, 0};               // boundary faces per region type
    // gather schedule: lanes per cell (0 = automatic, 1 = one thread per cell, 4, 6, 8; 64 = cell tiles)
    int schedule_lpc = 0;
    // per-kernel schedules chosen by cupcfd_hip_mesh_autotune (gradient, scalar, UVW, mass flux; -1 =
    // schedule_lpc)
    int schedule_k[4] = {-1, -1, -1, -1};
    // scalar carry sources per (inlet assigns, wall assigns) combination
    int *sc_src[4] = {nullptr, nullptr, nullptr, nullptr};
    // per carry source (boundary face k = g - Fi): the faces that read the
    // PhiFace it leaves, ascending mesh face id (sc_rd[key][sc_rd_off[key][k]
    // .. sc_rd_off[key][k + 1])), built on first use by the adjoint
    int *sc_rd_off[4] = {nullptr, nullptr, nullptr, nullptr};
    int *sc_rd[4] = {nullptr, nullptr, nullptr, nullptr};
    int sc_nrd[4] = {0, 0, 0, 0};
    // boundaries grouped by region, ascending id within each region (region r:
    // rbnd[rbnd_off[r] .. rbnd_off[r + 1])), built on first use by the
    // per-region ordered sums
    int *rbnd = nullptr, *rbnd_off = nullptr;
    // device copies of cell_order / face_order for the drop-in staging,
    // uploaded on first use
    int *cell_order_d = nullptr, *face_order_d = nullptr;
    // order-preserving face colouring, built on first use of the colour-group
    // schedule (-1: not built): colour(f) = 1 + the largest colour of the face
    // before f in the rows of f's cells, so colours rise along every row and no
    // two faces of one colour share a cell. Faces of colour q (internal ids,
    // ascending) are colour_faces[colour_off[q] .. colour_off[q + 1]).
    int ncolours = -1;
    const int *colour_of = nullptr, *colour_faces = nullptr;
    std::vector<int> colour_off;
    // The UVW tile's split LDS layout (tile_split_build, at creation): slots
    // whose face has a cell outside the tile, or is a boundary face, take LDS
    // columns 0 .. split_na - 1 of a 16-row table (the outside cell, then the
    // slot's ten terms); faces between two cells of the tile take columns
    // kSplitA + rank of a 9-row table. uvw_col[t * tslots + sl] is slot sl's
    // column, uvw_eslot[e] is eslot[e] with the slot replaced by its column.
    // split_na / split_nb: the most slots of each kind in any tile; uvw_split
    // is 1 when every tile fits the layout (FP64 meshes only).
    uint8_t *uvw_col = nullptr;
    uint16_t *uvw_eslot = nullptr;
    int split_na = 0, split_nb = 0, uvw_split = 0;
    // workspace pool for drop-in staging and multi-sweep scratch. A slot that
    // must grow gets a new buffer; the old one is retired, not freed, until
    // the mesh is destroyed, so a graph captured with the old buffer never
    // replays into freed memory.
    std::vector<void *> ws;
    std::vector<size_t> ws_size;
    std::vector<void *> ws_retired;
    void *ws_get(int slot, size_t bytes);
    ~Mesh();
};

template <class T> int *scalar_carry_src(Mesh<T> &M, bool inlet_assigns, bool wall_assigns);
// M.uvw_col / M.uvw_eslot and M.uvw_split (on the null stream, synchronous);
// a mesh whose tiles do not fit the split layout keeps uvw_split = 0
template <class T> void tile_split_build(Mesh<T> &M);
// M.rbnd / M.rbnd_off, built on first call; false if they cannot be allocated
template <class T> bool region_boundaries(Mesh<T> &M);
// Device copy of M.cell_order (faces = false) or M.face_order, uploaded on first call; null on failure
template <class T> const int *device_order(Mesh<T> &M, bool faces);
// Drop-in staging on the device (null stream, asynchronous): dst[k*ncomp + j]
// = raw[order[k]*stride + j] (mesh order -> internal), and raw[order[k]*ncomp
// + j] = src[k*ncomp + j] (internal -> mesh order), for k < n, j < ncomp
template <class T> int stage_gather(T *dst, const T *raw, const int *order, int n, int ncomp, int64_t stride);
template <class T> int stage_scatter(T *raw, const T *src, const int *order, int n, int ncomp);
// Synchronous copies between pageable host memory and the device (the
// drop-in entry points' arrays): large ones are pipelined through pinned
// buffers by a few host threads, small ones take hipMemcpy. The device to host
// copy first waits for the null stream. CUPCFD_E_SUCCESS or CUPCFD_E_ERROR.
int staged_h2d(void *dev, const void *host, size_t bytes);
int staged_d2h(void *host, const void *dev, size_t bytes);
int sync_state_any(struct ::cupcfd_hip_mesh *h, bool push);

// Counts and pointers a MiniSoA view must carry before either repack reads it.
template <class View> int view_check(const View &v) {
    const int C = v.nCells, F = v.nFaces, B = v.nBoundaries, R = v.nRegions;
    if (C < 0 || F < 0 || B < 0 || R < 0) return CUPCFD_E_CONFIG_INVALID_VALUE;
    if ((int64_t)F >= ((int64_t)1 << 30)) return CUPCFD_E_CONFIG_INVALID_VALUE;
    if ((C && !v.cellVolume) || (F && (!v.faceCell1ID || !v.faceCell2ID || !v.faceBoundaryID || !v.faceLambda ||
                                      !v.faceArea || !v.faceRLencos || !v.faceNorm.base || !v.faceCenter.base ||
                                      !v.faceXpac.base || !v.faceXnac.base)) ||
        (C && !v.cellCenter.base) || (B && (!v.boundaryFaceID || !v.boundaryRegionID || !v.boundaryDistance)) ||
        (R && !v.regionType))
        return CUPCFD_E_NULL_PTR;
    return CUPCFD_E_SUCCESS;
}

// Device scratch carved from a few large allocations (256-byte aligned), all
// freed together: a build's many temporaries cost a handful of hipMalloc /
// hipFree calls instead of one pair each.
struct DeviceChunks {
    std::vector<void *> bufs;
    char *cur = nullptr;
    size_t left = 0, min_chunk = size_t(1) << 20;
    DeviceChunks() = default;
    DeviceChunks(const DeviceChunks &) = delete;
    DeviceChunks &operator=(const DeviceChunks &) = delete;
    ~DeviceChunks() {
        for (void *p : bufs) (void)hipFree(p);
    }
    hipError_t get(size_t bytes, void **out) {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        if (bytes > left) {
            const size_t sz = std::max(bytes, min_chunk);
            void *p = nullptr;
            if (hipError_t e = hipMalloc(&p, sz); e != hipSuccess) return e;
            bufs.push_back(p);
            cur = (char *)p;
            left = sz;
        }
        *out = cur;
        cur += bytes;
        left -= bytes;
        return hipSuccess;
    }
};

// Build M's face colouring on the device from its CSR (no-op once built).
template <class T> int build_face_colours(Mesh<T> &M);

// The readers of each scalar carry source (Mesh::sc_rd_off / sc_rd) for one
// (INLET assigns, WALL assigns) combination: interior faces, and INLET faces
// when INLET does not assign; built once per mesh and combination.
template <class T>
int scalar_carry_readers(Mesh<T> &M, bool inlet_assigns, bool wall_assigns, const int **off, const int **rd,
                         int *nrd);

// The same repack as the host build, run on the device: every array of the
// DevMesh and every host-side mirror of Mesh comes out bit-identical. With
// `resident` the view's per-cell, per-face and per-boundary arrays are device
// pointers (3-vectors packed) and its region arrays host pointers.
template <class T, class View> int build_device(Mesh<T> &M, const View &v, int flags, bool resident = false);

}  // namespace cupcfd_hip

struct cupcfd_hip_mesh {
    int precision;  // 64 or 32
    cupcfd_hip::Mesh<double> *m64;
    cupcfd_hip::Mesh<float> *m32;
};

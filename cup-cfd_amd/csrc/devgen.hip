// Mesh generation on the GPU: StructGen hexes and the jittered Kuhn-tet mesh
// generated in HBM, then repacked there (build_device on device-resident
// arrays), so no per-cell or per-face array crosses PCIe. Every array equals
// what the host generators (mesh_gen.cpp, tetgen.cpp) produce, bit for bit:
//   - the same double expressions, evaluated without contraction;
//   - SplitMix64 draws taken by index: the stream is counter-based (draw n
//     mixes state seed + (n+1)*0x9E3779B97F4A7C15), so draw n of the host's
//     sequential stream can be computed by any thread;
//   - the tet face matching is two stable radix sorts of the host's key
//     (lowest two vertex ids, highest vertex id), with ties in slot order,
//     and the face and boundary ids are prefix counts in slot order.
// The Fisher-Yates relabelling is sequential by construction: its two
// permutations are drawn on the host and applied on the device.
// StructGen numbering follows src/geometry/mesh/implementation/source/MeshStructGenSource.cpp
// as assembled on one rank by UnstructuredMeshInterface::addData
// (include/geometry/mesh/interface/component/UnstructuredMeshInterface.ipp:1464-1862).
#include <algorithm>
#include <cstring>
#include <new>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "devmesh.hpp"
#include "host_mesh.hpp"
#include "sg_geom.hpp"

namespace cupcfd_hip {

namespace {

constexpr int kB = 256;

inline int grid_for(int64_t n) {
    const int64_t g = (n + kB - 1) / kB;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

inline unsigned bits_for(int64_t x) {  // bits that hold every value in [0, x]
    unsigned b = 1;
    while (b < 64 && (((int64_t)1 << b) <= x)) b++;
    return b;
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)(n); i += (int64_t)gridDim.x * blockDim.x)

// draw n of SplitMix64(seed), as a uniform double in [0, 1) (top 53 bits)
__device__ inline double sm64_uniform(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// ---- StructGen ------------------------------------------------------------

struct Grid {
    int nx, ny, nz, regions;
    double mn[3], dd[3];
    int64_t nYZ, nXZ, nBYZ, nBXZ;
};

// cell centre as getCellCenter computes it (MeshStructGenSource.cpp:330-340),
// volume from the reference's five-tetrahedron split around that centre
__global__ void k_sg_cells(Grid g, sg::HexPlan plan, int C, double *cc, double *vol) {
    const double half[3] = {g.dd[0] / 2, g.dd[1] / 2, g.dd[2] / 2};
    GRID_STRIDE(c, C) {
        const int ijk[3] = {(int)(c % g.nx), (int)((c / g.nx) % g.ny), (int)(c / ((int64_t)g.nx * g.ny))};
        sg::P3 ctr;
        for (int k = 0; k < 3; k++) {
            const double l = g.mn[k] + (ijk[k] * g.dd[k]), h = g.mn[k] + (ijk[k] + 1) * g.dd[k];
            ctr.c[k] = l + ((h - l) / 2.0);
            cc[3 * c + k] = ctr.c[k];
        }
        vol[c] = sg::hex_volume(plan, ctr, half);
    }
}

__device__ inline int32_t sg_label(int x, int y, int z, int xs, int ys) { return x + y * xs + z * xs * ys; }

// faces in label order: YZ (x-normal), XZ (y-normal), XY (z-normal)
// (getFaceCell1Labels / getFaceCell2Labels / getFaceBoundaryLabels, MeshStructGenSource.cpp:518-821)
__global__ void k_sg_faces(Grid g, int F, int *c1, int *c2, int *fb, double *lam, double *area, double *nrm,
                           double *ctr, int *bface, int *bregion, double *bdist) {
    GRID_STRIDE(L, F) {
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        int axis, n, a, b, c;
        int64_t r;
        if (L < g.nYZ) {
            axis = 0;
            r = L;
            n = (int)(r % (nx + 1));
            b = (int)((r / (nx + 1)) % ny);
            c = (int)(r / ((int64_t)(nx + 1) * ny));
            a = n;
        } else if (L < g.nYZ + g.nXZ) {
            axis = 1;
            r = L - g.nYZ;
            a = (int)(r % nx);
            n = (int)((r / nx) % (ny + 1));
            c = (int)(r / ((int64_t)nx * (ny + 1)));
            b = n;
        } else {
            axis = 2;
            r = L - g.nYZ - g.nXZ;
            a = (int)(r % nx);
            b = (int)((r / nx) % ny);
            n = (int)(r / ((int64_t)nx * ny));
            c = n;
        }
        const int nmax = axis == 0 ? nx : (axis == 1 ? ny : nz);
        const bool bnd = (n == 0 || n == nmax);
        int lo[3] = {a, b, c};
        if (n > 0) lo[axis] = n - 1;
        const int f = (int)L;
        c1[f] = sg_label(lo[0], lo[1], lo[2], nx, ny);
        if (!bnd) {
            int hi[3] = {lo[0], lo[1], lo[2]};
            hi[axis] = n;
            c2[f] = sg_label(hi[0], hi[1], hi[2], nx, ny);
            fb[f] = -1;
            lam[f] = 0.5;
        } else {
            c2[f] = -1;
            lam[f] = -1.0;
            const int side = (n == 0) ? 0 : 1;
            int32_t bl;
            int reg;
            if (axis == 0) {
                bl = sg_label(side, b, c, 2, ny);
                reg = side ? REG_OUTLET : REG_INLET;
            } else if (axis == 1) {
                bl = (int32_t)g.nBYZ + sg_label(a, side, c, nx, 2);
                reg = side ? REG_WALL_HI : REG_WALL_LO;
            } else {
                bl = (int32_t)(g.nBYZ + g.nBXZ) + sg_label(a, b, side, nx, ny);
                reg = side ? REG_SYMP_HI : REG_SYMP_LO;
            }
            fb[f] = bl;
            bface[bl] = f;
            bregion[bl] = g.regions ? reg : 0;
            bdist[bl] = g.regions ? g.dd[axis] / 2.0 : 0.0;
        }
        // geometry from the face's four vertices in getFaceVerticesLabelsCSR's
        // order (MeshStructGenSource.cpp:1010-1085)
        int node[3] = {a, b, c};
        node[axis] = n;
        const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;
        sg::P3 v[4];
        for (int i = 0; i < 4; i++) {
            int q[3] = {node[0], node[1], node[2]};
            q[u] += (i == 1 || i == 2) ? 1 : 0;
            q[w] += (i >= 2) ? 1 : 0;
            for (int k = 0; k < 3; k++) v[i].c[k] = g.mn[k] + (g.dd[k] * q[k]);
        }
        double ar;
        sg::P3 mid;
        sg::quad_area_centroid(v, &ar, &mid);
        // cell1's centre, as k_sg_cells computes it
        sg::P3 c1c;
        for (int k = 0; k < 3; k++) {
            const double l = g.mn[k] + (lo[k] * g.dd[k]), h = g.mn[k] + (lo[k] + 1) * g.dd[k];
            c1c.c[k] = l + ((h - l) / 2.0);
        }
        const sg::P3 S = sg::quad_normal(v, c1c);
        area[f] = ar;
        for (int k = 0; k < 3; k++) {
            ctr[3 * (int64_t)f + k] = mid.c[k];
            nrm[3 * (int64_t)f + k] = S.c[k];
        }
    }
}

// seeded xpac / xnac around the face centre and rlencos = area / |Cc2 - Cc1|
// (|Xf - Cc1| on boundaries): draws 6f + 2k (xpac) and 6f + 2k + 1 (xnac)
__global__ void k_face_extras(int F, uint64_t seed, double h, const int *c1, const int *c2, const double *cc,
                              const double *ctr, const double *area, double *xp, double *xn, double *rl) {
    GRID_STRIDE(f, F) {
        for (int k = 0; k < 3; k++) {
            xp[3 * f + k] = ctr[3 * f + k] + 0.05 * h * (2.0 * sm64_uniform(seed, 6 * f + 2 * k) - 1.0);
            xn[3 * f + k] = ctr[3 * f + k] + 0.05 * h * (2.0 * sm64_uniform(seed, 6 * f + 2 * k + 1) - 1.0);
        }
        const int ip = c1[f], in = c2[f];
        const double *a = cc + 3 * (int64_t)ip;
        const double *b = in >= 0 ? cc + 3 * (int64_t)in : ctr + 3 * f;
        double d2 = 0;
        for (int k = 0; k < 3; k++) d2 += (b[k] - a[k]) * (b[k] - a[k]);
        rl[f] = area[f] / sqrt(d2);
    }
}

// ---- tetrahedra -------------------------------------------------------------

struct TetGrid {
    int nx, ny, nz, regions;
    double mn[3], dd[3];
    double jitter;
    uint64_t seed;
};

__device__ inline void tsub(double *o, const double *a, const double *b) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
__device__ inline double tdot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void tcross(double *o, const double *a, const double *b) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double tdist(const double *a, const double *b) {
    double d[3];
    tsub(d, a, b);
    return sqrt(tdot(d, d));
}

// vertices in vertex-id order; draw 3v + k jitters component k of interior vertex v
__global__ void k_tet_verts(TetGrid g, int64_t Vn, double *X) {
    GRID_STRIDE(v, Vn) {
        const int n3[3] = {g.nx, g.ny, g.nz};
        const int ijk[3] = {(int)(v % (g.nx + 1)), (int)((v / (g.nx + 1)) % (g.ny + 1)),
                            (int)(v / ((int64_t)(g.nx + 1) * (g.ny + 1)))};
        for (int k = 0; k < 3; k++) {
            double x = g.mn[k] + ijk[k] * g.dd[k];
            const double u = 2.0 * sm64_uniform(g.seed, 3 * v + k) - 1.0;
            if (ijk[k] > 0 && ijk[k] < n3[k]) x += g.jitter * g.dd[k] * u;
            X[3 * v + k] = x;
        }
    }
}

__constant__ int kKuhn[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// tet 6h + t walks from hex corner (0,0,0) to (1,1,1) along the axes in permutation t
__global__ void k_tet_vertices_of(TetGrid g, int C, uint32_t *tv) {
    GRID_STRIDE(c, C) {
        const int64_t h = c / 6;
        const int t = (int)(c % 6);
        const int x = (int)(h % g.nx), y = (int)((h / g.nx) % g.ny), z = (int)(h / ((int64_t)g.nx * g.ny));
        auto vid = [&](int i, int j, int k) -> uint32_t {
            return (uint32_t)(i + (int64_t)(g.nx + 1) * (j + (int64_t)(g.ny + 1) * k));
        };
        int q[3] = {x, y, z};
        uint32_t *w = tv + 4 * c;
        w[0] = vid(q[0], q[1], q[2]);
        q[kKuhn[t][0]]++;
        w[1] = vid(q[0], q[1], q[2]);
        q[kKuhn[t][1]]++;
        w[2] = vid(q[0], q[1], q[2]);
        w[3] = vid(x + 1, y + 1, z + 1);
    }
}

// the triangle of slot s = 4 tet + l (the face opposite local vertex l), vertex ids ascending
__device__ inline void tet_tri(const uint32_t *tv, int64_t s, uint32_t *f3) {
    const int64_t c = s / 4;
    const int l = (int)(s % 4);
    int n = 0;
    for (int q = 0; q < 4; q++)
        if (q != l) f3[n++] = tv[4 * c + q];
    uint32_t t;
    if (f3[1] < f3[0]) { t = f3[0]; f3[0] = f3[1]; f3[1] = t; }
    if (f3[2] < f3[1]) { t = f3[1]; f3[1] = f3[2]; f3[2] = t; }
    if (f3[1] < f3[0]) { t = f3[0]; f3[0] = f3[1]; f3[1] = t; }
}

__global__ void k_tet_key_hi(int64_t S, const uint32_t *tv, uint32_t *key, int *val) {
    GRID_STRIDE(s, S) {
        uint32_t f3[3];
        tet_tri(tv, s, f3);
        key[s] = f3[2];
        val[s] = (int)s;
    }
}

__global__ void k_tet_key_lo(int64_t S, const uint32_t *tv, const int *slot, uint64_t *key) {
    GRID_STRIDE(i, S) {
        uint32_t f3[3];
        tet_tri(tv, slot[i], f3);
        key[i] = ((uint64_t)f3[0] << 32) | f3[1];
    }
}

// twin[s]: the other slot holding the same triangle, or -1; a triangle in
// three or more tets flags an error
__global__ void k_tet_twin(int64_t S, const uint32_t *tv, const uint64_t *klo, const int *slot, int *twin, int *err) {
    GRID_STRIDE(i, S) {
        uint32_t a[3];
        tet_tri(tv, slot[i], a);
        bool prev = false, next = false;
        if (i > 0 && klo[i - 1] == klo[i]) {
            uint32_t b[3];
            tet_tri(tv, slot[i - 1], b);
            prev = b[2] == a[2];
        }
        if (i + 1 < S && klo[i + 1] == klo[i]) {
            uint32_t b[3];
            tet_tri(tv, slot[i + 1], b);
            next = b[2] == a[2];
        }
        if (prev && next) atomicOr(err, 1);
        if (next && !prev) {
            twin[slot[i]] = slot[i + 1];
            twin[slot[i + 1]] = slot[i];
        } else if (!prev && !next) {
            twin[slot[i]] = -1;
        }
    }
}

// a slot numbers a face when it is the triangle's lower slot; boundary slots number boundaries
__global__ void k_tet_marks(int64_t S, const int *twin, int *isface, int *isbnd) {
    GRID_STRIDE(s, S) {
        const int t2 = twin[s];
        isface[s] = (t2 < 0 || t2 > s) ? 1 : 0;
        isbnd[s] = t2 < 0 ? 1 : 0;
    }
}

__global__ void k_tet_cells(int C, const uint32_t *tv, const double *X, double *cc, double *vol) {
    GRID_STRIDE(c, C) {
        const double *v[4];
        for (int q = 0; q < 4; q++) v[q] = X + 3 * (int64_t)tv[4 * c + q];
        for (int k = 0; k < 3; k++) cc[3 * c + k] = (v[0][k] + v[1][k] + v[2][k] + v[3][k]) / 4.0;
        double e1[3], e2[3], e3[3], cr[3];
        tsub(e1, v[1], v[0]);
        tsub(e2, v[2], v[0]);
        tsub(e3, v[3], v[0]);
        tcross(cr, e2, e3);
        vol[c] = fabs(tdot(e1, cr)) / 6.0;
    }
}

__global__ void k_tet_faces(TetGrid g, int64_t S, const uint32_t *tv, const double *X, const double *cc,
                            const int *twin, const int *fid, const int *bid, int *c1o, int *c2o, int *fbo,
                            double *lam, double *area, double *nrm, double *ctr, int *bface, int *bregion,
                            double *bdist) {
    GRID_STRIDE(s, S) {
        const int t2 = twin[s];
        if (t2 >= 0 && t2 < s) continue;  // numbered at its lower slot
        const int f = fid[s];
        const int c1 = (int)(s / 4), l = (int)(s % 4);
        const double *v3[3];
        uint32_t vv[3];
        int n = 0;
        for (int q = 0; q < 4; q++)
            if (q != l) {
                vv[n] = tv[4 * (int64_t)c1 + q];
                v3[n] = X + 3 * (int64_t)vv[n];
                n++;
            }
        double xf[3], e1[3], e2[3], Sv[3], d1[3];
        for (int k = 0; k < 3; k++) xf[k] = (v3[0][k] + v3[1][k] + v3[2][k]) / 3.0;
        tsub(e1, v3[1], v3[0]);
        tsub(e2, v3[2], v3[0]);
        tcross(Sv, e1, e2);
        for (int k = 0; k < 3; k++) Sv[k] *= 0.5;
        const double *cc1 = cc + 3 * (int64_t)c1;
        tsub(d1, xf, cc1);
        if (tdot(Sv, d1) < 0)
            for (int k = 0; k < 3; k++) Sv[k] = -Sv[k];
        const double la = tdist(v3[0], v3[1]), lb = tdist(v3[1], v3[2]), lc = tdist(v3[2], v3[0]);
        const double hs = (la + lb + lc) / 2.0;
        const double hp = hs * (hs - la) * (hs - lb) * (hs - lc);
        c1o[f] = c1;
        area[f] = sqrt(0.0 < hp ? hp : 0.0);
        for (int k = 0; k < 3; k++) {
            nrm[3 * (int64_t)f + k] = Sv[k];
            ctr[3 * (int64_t)f + k] = xf[k];
        }
        if (t2 >= 0) {
            const int c2 = t2 / 4;
            c2o[f] = c2;
            fbo[f] = -1;
            const double dP = tdist(xf, cc1), dN = tdist(xf, cc + 3 * (int64_t)c2);
            lam[f] = dP / (dP + dN);
        } else {
            const int b = bid[s];
            c2o[f] = -1;
            lam[f] = -1.0;
            fbo[f] = b;
            bface[b] = f;
            // the box side the triangle lies on: the axis whose index is shared by its 3 vertices
            const int n3[3] = {g.nx, g.ny, g.nz};
            int reg = 0;
            for (int k = 0; k < 3; k++) {
                int idx[3];
                for (int q = 0; q < 3; q++) {
                    const uint32_t w = vv[q];
                    idx[q] = k == 0 ? (int)(w % (g.nx + 1))
                                    : k == 1 ? (int)((w / (g.nx + 1)) % (g.ny + 1))
                                             : (int)(w / ((uint32_t)(g.nx + 1) * (g.ny + 1)));
                }
                if (idx[0] == idx[1] && idx[1] == idx[2] && (idx[0] == 0 || idx[0] == n3[k])) {
                    const bool hi = idx[0] == n3[k];
                    reg = k == 0 ? (hi ? REG_OUTLET : REG_INLET)
                                 : k == 1 ? (hi ? REG_WALL_HI : REG_WALL_LO) : (hi ? REG_SYMP_HI : REG_SYMP_LO);
                }
            }
            bregion[b] = g.regions ? reg : 0;
            const double an = sqrt(tdot(Sv, Sv));
            bdist[b] = g.regions ? fabs(tdot(d1, Sv)) / an : 0.0;
        }
    }
}

// ---- relabelling --------------------------------------------------------------

__global__ void k_perm_cells(int C, const int *pc, const double *cc, const double *vol, double *cco, double *volo) {
    GRID_STRIDE(c, C) {
        const int64_t p = pc[c];
        for (int k = 0; k < 3; k++) cco[3 * p + k] = cc[3 * c + k];
        volo[p] = vol[c];
    }
}

__global__ void k_remap_cells(int F, const int *pc, int *c1, int *c2) {
    GRID_STRIDE(f, F) {
        c1[f] = pc[c1[f]];
        if (c2[f] >= 0) c2[f] = pc[c2[f]];
    }
}

template <class X> __global__ void k_perm(int n, int w, const int *p, const X *src, X *dst) {
    GRID_STRIDE(i, (int64_t)n * w) {
        const int64_t f = i / w;
        dst[(int64_t)p[f] * w + (i - f * w)] = src[i];
    }
}

__global__ void k_remap(int n, const int *p, int *v) {
    GRID_STRIDE(i, n) v[i] = p[v[i]];
}

__global__ void k_to_f32(int64_t n, const double *s, float *d) {
    GRID_STRIDE(i, n) d[i] = (float)s[i];
}

// Fisher-Yates over SplitMix64, as the host relabelling draws it: new id of old id i
std::vector<int32_t> fisher_yates(SplitMix64 &rng, int n) {
    std::vector<int32_t> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int)rng.below((uint64_t)i + 1)]);
    return perm;
}

// device allocations freed when generation ends; first HIP error sticks
struct Pool {
    DeviceChunks bufs;
    void *prim = nullptr;
    size_t prim_bytes = 0;
    hipError_t err = hipSuccess;
    ~Pool() {
        if (prim) (void)hipFree(prim);
    }
    void note(hipError_t e) {
        if (err == hipSuccess && e != hipSuccess) err = e;
    }
    template <class X> X *alloc(size_t n) {
        if (err != hipSuccess) return nullptr;
        void *p = nullptr;
        note(bufs.get(std::max<size_t>(n, 1) * sizeof(X), &p));
        return err == hipSuccess ? (X *)p : nullptr;
    }
    template <class Fn> void prim_call(Fn &&fn) {
        if (err != hipSuccess) return;
        size_t bytes = 0;
        note(fn(nullptr, bytes));
        if (err != hipSuccess) return;
        if (bytes > prim_bytes) {
            if (prim) (void)hipFree(prim);
            prim = nullptr;
            prim_bytes = 0;
            note(hipMalloc(&prim, bytes));
            if (err != hipSuccess) return;
            prim_bytes = bytes;
        }
        note(fn(prim, bytes));
    }
    template <class K, class... A> void launch(int64_t n, K kernel, A... args) {
        if (err != hipSuccess || n <= 0) return;
        hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kB), 0, 0, args...);
        note(hipGetLastError());
    }
};

// the generated MiniSoA arrays, double geometry, mesh numbering
struct GenMesh {
    int C = 0, F = 0, B = 0;
    int *c1 = nullptr, *c2 = nullptr, *fb = nullptr, *bface = nullptr, *bregion = nullptr;
    double *lam = nullptr, *rl = nullptr, *area = nullptr, *nrm = nullptr, *ctr = nullptr, *xp = nullptr,
           *xn = nullptr, *cc = nullptr, *vol = nullptr, *bdist = nullptr;
    void alloc(Pool &P, int C_, int F_, int B_) {
        C = C_, F = F_, B = B_;
        c1 = P.alloc<int>(F), c2 = P.alloc<int>(F), fb = P.alloc<int>(F);
        lam = P.alloc<double>(F), rl = P.alloc<double>(F), area = P.alloc<double>(F);
        nrm = P.alloc<double>(3 * (size_t)F), ctr = P.alloc<double>(3 * (size_t)F);
        xp = P.alloc<double>(3 * (size_t)F), xn = P.alloc<double>(3 * (size_t)F);
        cc = P.alloc<double>(3 * (size_t)C), vol = P.alloc<double>(C);
        bface = P.alloc<int>(B), bregion = P.alloc<int>(B), bdist = P.alloc<double>(B);
        // arrays the generators leave at their defaults (resize_mesh): xpac, xnac, rlencos 0
        for (void *p : {(void *)xp, (void *)xn})
            if (p && F) P.note(hipMemsetAsync(p, 0, 3 * (size_t)F * sizeof(double), 0));
        if (rl && F) P.note(hipMemsetAsync(rl, 0, (size_t)F * sizeof(double), 0));
    }
};

int gen_structgen(const cupcfd_structgen_params &p, Pool &P, GenMesh &G) {
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    Grid g;
    g.nx = p.nx, g.ny = p.ny, g.nz = p.nz, g.regions = p.regions != 0;
    g.mn[0] = p.xmin, g.mn[1] = p.ymin, g.mn[2] = p.zmin;
    // spatial division (MeshStructGenSource.cpp:44-47)
    g.dd[0] = (p.xmax - p.xmin) / double(p.nx);
    g.dd[1] = (p.ymax - p.ymin) / double(p.ny);
    g.dd[2] = (p.zmax - p.zmin) / double(p.nz);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int64_t C64 = (int64_t)nx * ny * nz;
    g.nYZ = (int64_t)ny * nz * (nx + 1), g.nXZ = (int64_t)nx * nz * (ny + 1);
    const int64_t nXY = (int64_t)nx * ny * (nz + 1);
    const int64_t F64 = g.nYZ + g.nXZ + nXY;
    g.nBYZ = (int64_t)ny * nz * 2, g.nBXZ = (int64_t)nx * nz * 2;
    const int64_t B64 = g.nBYZ + g.nBXZ + (int64_t)nx * ny * 2;
    if (F64 >= (int64_t)1 << 30) return CUPCFD_E_CONFIG_INVALID_VALUE;
    P.bufs.min_chunk = std::max<size_t>(size_t(1) << 20, 256 * (size_t)F64);
    G.alloc(P, (int)C64, (int)F64, (int)B64);
    const double half[3] = {g.dd[0] / 2, g.dd[1] / 2, g.dd[2] / 2};
    P.launch(G.C, k_sg_cells, g, sg::hex_plan(half), G.C, G.cc, G.vol);
    P.launch(G.F, k_sg_faces, g, G.F, G.c1, G.c2, G.fb, G.lam, G.area, G.nrm, G.ctr, G.bface, G.bregion, G.bdist);
    if (p.face_extras)
        P.launch(G.F, k_face_extras, G.F, (uint64_t)(p.seed ^ 0xFACEE47A5ull),
                 std::min(g.dd[0], std::min(g.dd[1], g.dd[2])), (const int *)G.c1, (const int *)G.c2,
                 (const double *)G.cc, (const double *)G.ctr, (const double *)G.area, G.xp, G.xn, G.rl);
    return P.err == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

int gen_tets(const cupcfd_structgen_params &p, Pool &P, GenMesh &G) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    if (nx < 1 || ny < 1 || nz < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    if (p.jitter < 0.0 || p.jitter >= 0.5) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const int64_t H = (int64_t)nx * ny * nz;
    const int64_t C64 = 6 * H;
    const int64_t Vn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    if (4 * C64 >= ((int64_t)1 << 31) || Vn >= ((int64_t)1 << 31)) return CUPCFD_E_CONFIG_INVALID_VALUE;
    TetGrid g;
    g.nx = nx, g.ny = ny, g.nz = nz, g.regions = p.regions != 0;
    g.mn[0] = p.xmin, g.mn[1] = p.ymin, g.mn[2] = p.zmin;
    g.dd[0] = (p.xmax - p.xmin) / nx, g.dd[1] = (p.ymax - p.ymin) / ny, g.dd[2] = (p.zmax - p.zmin) / nz;
    g.jitter = p.jitter;
    g.seed = p.seed ^ 0x7E7A7E7A5EEDull;
    const int C = (int)C64;
    const int64_t S = 4 * C64;
    P.bufs.min_chunk = std::max<size_t>(size_t(1) << 20, 160 * (size_t)S);

    double *X = P.alloc<double>(3 * (size_t)Vn);
    uint32_t *tv = P.alloc<uint32_t>((size_t)S);
    P.launch(Vn, k_tet_verts, g, Vn, X);
    P.launch(C, k_tet_vertices_of, g, C, tv);
    // match triangles: sort slots by (lowest two vertices, highest vertex), ties in slot order
    uint32_t *khi = P.alloc<uint32_t>(S), *khi_s = P.alloc<uint32_t>(S);
    int *slot0 = P.alloc<int>(S), *slot1 = P.alloc<int>(S), *slot2 = P.alloc<int>(S);
    uint64_t *klo = P.alloc<uint64_t>(S), *klo_s = P.alloc<uint64_t>(S);
    P.launch(S, k_tet_key_hi, S, (const uint32_t *)tv, khi, slot0);
    P.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::radix_sort_pairs(tmp, bytes, khi, khi_s, slot0, slot1, (size_t)S, 0, bits_for(Vn),
                                         (hipStream_t)0);
    });
    P.launch(S, k_tet_key_lo, S, (const uint32_t *)tv, (const int *)slot1, klo);
    P.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::radix_sort_pairs(tmp, bytes, klo, klo_s, slot1, slot2, (size_t)S, 0, 32 + bits_for(Vn),
                                         (hipStream_t)0);
    });
    int *twin = P.alloc<int>(S), *err = P.alloc<int>(2);
    if (P.err == hipSuccess) P.note(hipMemsetAsync(err, 0, 2 * sizeof(int), 0));
    P.launch(S, k_tet_twin, S, (const uint32_t *)tv, (const uint64_t *)klo_s, (const int *)slot2, twin, err);
    int *isface = P.alloc<int>(S + 1), *isbnd = P.alloc<int>(S + 1), *fid = P.alloc<int>(S + 1),
        *bid = P.alloc<int>(S + 1);
    if (P.err == hipSuccess) P.note(hipMemsetAsync(isface + S, 0, sizeof(int), 0));
    if (P.err == hipSuccess) P.note(hipMemsetAsync(isbnd + S, 0, sizeof(int), 0));
    P.launch(S, k_tet_marks, S, (const int *)twin, isface, isbnd);
    P.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::exclusive_scan(tmp, bytes, isface, fid, 0, (size_t)S + 1, rocprim::plus<int>(), (hipStream_t)0);
    });
    P.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::exclusive_scan(tmp, bytes, isbnd, bid, 0, (size_t)S + 1, rocprim::plus<int>(), (hipStream_t)0);
    });
    int cnt[3] = {0, 0, 0};
    if (P.err == hipSuccess) P.note(hipMemcpy(&cnt[0], err, sizeof(int), hipMemcpyDeviceToHost));
    if (P.err == hipSuccess) P.note(hipMemcpy(&cnt[1], fid + S, sizeof(int), hipMemcpyDeviceToHost));
    if (P.err == hipSuccess) P.note(hipMemcpy(&cnt[2], bid + S, sizeof(int), hipMemcpyDeviceToHost));
    if (P.err != hipSuccess) return CUPCFD_E_ERROR;
    if (cnt[0]) return CUPCFD_E_ERROR;  // a triangle in three tets: not a valid mesh
    if (cnt[1] >= (1 << 30)) return CUPCFD_E_CONFIG_INVALID_VALUE;
    G.alloc(P, C, cnt[1], cnt[2]);
    P.launch(C, k_tet_cells, C, (const uint32_t *)tv, (const double *)X, G.cc, G.vol);
    P.launch(S, k_tet_faces, g, S, (const uint32_t *)tv, (const double *)X, (const double *)G.cc, (const int *)twin,
             (const int *)fid, (const int *)bid, G.c1, G.c2, G.fb, G.lam, G.area, G.nrm, G.ctr, G.bface, G.bregion,
             G.bdist);
    if (p.face_extras)
        P.launch(G.F, k_face_extras, G.F, (uint64_t)(p.seed ^ 0xFACEE47A5ull),
                 std::min(g.dd[0], std::min(g.dd[1], g.dd[2])), (const int *)G.c1, (const int *)G.c2,
                 (const double *)G.cc, (const double *)G.ctr, (const double *)G.area, G.xp, G.xn, G.rl);
    return P.err == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

// apply the host relabelling's permutations (cells first, then faces, one stream)
int relabel_device(const cupcfd_structgen_params &p, Pool &P, GenMesh &G) {
    SplitMix64 rng(p.seed);
    const int C = G.C, F = G.F, B = G.B;
    if (p.permute_cells) {
        const std::vector<int32_t> pc = fisher_yates(rng, C);
        int *d_pc = P.alloc<int>(C);
        double *cc = P.alloc<double>(3 * (size_t)C), *vol = P.alloc<double>(C);
        if (P.err == hipSuccess && C) P.note(hipMemcpy(d_pc, pc.data(), C * sizeof(int), hipMemcpyHostToDevice));
        P.launch(C, k_perm_cells, C, (const int *)d_pc, (const double *)G.cc, (const double *)G.vol, cc, vol);
        P.launch(F, k_remap_cells, F, (const int *)d_pc, G.c1, G.c2);
        G.cc = cc, G.vol = vol;
    }
    if (p.permute_faces) {
        const std::vector<int32_t> pf = fisher_yates(rng, F);
        int *d_pf = P.alloc<int>(F);
        if (P.err == hipSuccess && F) P.note(hipMemcpy(d_pf, pf.data(), F * sizeof(int), hipMemcpyHostToDevice));
        auto perm_i = [&](int *&v) {
            int *o = P.alloc<int>(F);
            P.launch(F, k_perm<int>, F, 1, (const int *)d_pf, (const int *)v, o);
            v = o;
        };
        auto perm_d = [&](double *&v, int w) {
            double *o = P.alloc<double>((size_t)w * F);
            P.launch((int64_t)w * F, k_perm<double>, F, w, (const int *)d_pf, (const double *)v, o);
            v = o;
        };
        perm_i(G.c1);
        perm_i(G.c2);
        perm_i(G.fb);
        perm_d(G.lam, 1);
        perm_d(G.rl, 1);
        perm_d(G.area, 1);
        perm_d(G.nrm, 3);
        perm_d(G.ctr, 3);
        perm_d(G.xp, 3);
        perm_d(G.xn, 3);
        P.launch(B, k_remap, B, (const int *)d_pf, G.bface);
    }
    return P.err == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

template <class X> bool fetch(std::vector<X> &h, const X *d, size_t n) {
    h.resize(n);
    return n == 0 || hipMemcpy(h.data(), d, n * sizeof(X), hipMemcpyDeviceToHost) == hipSuccess;
}

// copy the generated arrays into a host mesh (regions already added)
bool download(const GenMesh &G, cupcfd_host_mesh &m) {
    resize_mesh(m, G.C, G.F, G.B);
    const size_t F = G.F, C = G.C, B = G.B;
    return fetch(m.faceCell1, G.c1, F) && fetch(m.faceCell2, G.c2, F) && fetch(m.faceBnd, G.fb, F) &&
           fetch(m.faceLambda, G.lam, F) && fetch(m.faceRLencos, G.rl, F) && fetch(m.faceArea, G.area, F) &&
           fetch(m.faceNorm, G.nrm, 3 * F) && fetch(m.faceCenter, G.ctr, 3 * F) && fetch(m.faceXpac, G.xp, 3 * F) &&
           fetch(m.faceXnac, G.xn, 3 * F) && fetch(m.cellCenter, G.cc, 3 * C) && fetch(m.cellVolume, G.vol, C) &&
           fetch(m.bndFace, G.bface, B) && fetch(m.bndRegion, G.bregion, B) && fetch(m.bndDistance, G.bdist, B);
}

template <class T> struct ViewOf;
template <> struct ViewOf<double> {
    using type = cupcfd_mesh_view_f64;
    using v3 = cupcfd_v3_f64;
    static int host_view(cupcfd_host_mesh *m, type *v) { return cupcfd_host_mesh_view_f64(m, v); }
    static type &kept(Mesh<double> &M) { return M.view64; }
};
template <> struct ViewOf<float> {
    using type = cupcfd_mesh_view_f32;
    using v3 = cupcfd_v3_f32;
    static int host_view(cupcfd_host_mesh *m, type *v) { return cupcfd_host_mesh_view_f32(m, v); }
    static type &kept(Mesh<float> &M) { return M.view32; }
};

template <class T>
int generate(const cupcfd_structgen_params &p, int32_t flags, cupcfd_host_mesh **host_out, Mesh<T> &M) {
    using VT = ViewOf<T>;
    using View = typename VT::type;
    using V3 = typename VT::v3;
    Pool P;
    GenMesh G;
    int rc = p.tets ? gen_tets(p, P, G) : gen_structgen(p, P, G);
    if (rc == CUPCFD_E_SUCCESS && (p.permute_cells || p.permute_faces)) rc = relabel_device(p, P, G);
    if (rc != CUPCFD_E_SUCCESS) return rc;

    // regions are a handful of host entries, as the host generators add them
    cupcfd_host_mesh *hm = new (std::nothrow) cupcfd_host_mesh();
    if (!hm) return CUPCFD_E_ERROR;
    add_regions(*hm, p.regions != 0);
    View rv{};
    VT::host_view(hm, &rv);  // region arrays of the host precision

    // the view the device repack reads: device-resident arrays of precision T
    View v{};
    v.nCells = G.C, v.nFaces = G.F, v.nBoundaries = G.B, v.nRegions = hm->nRegions;
    T *lam, *rl, *area, *nrm, *ctr, *xp, *xn, *cc, *vol, *bdist;
    if constexpr (sizeof(T) == 8) {
        lam = G.lam, rl = G.rl, area = G.area, nrm = G.nrm, ctr = G.ctr, xp = G.xp, xn = G.xn, cc = G.cc,
        vol = G.vol, bdist = G.bdist;
    } else {
        auto cvt = [&](const double *s, size_t n) {
            float *d = P.alloc<float>(n);
            P.launch((int64_t)n, k_to_f32, (int64_t)n, s, d);
            return d;
        };
        const size_t F = G.F, C = G.C, B = G.B;
        lam = cvt(G.lam, F), rl = cvt(G.rl, F), area = cvt(G.area, F), nrm = cvt(G.nrm, 3 * F);
        ctr = cvt(G.ctr, 3 * F), xp = cvt(G.xp, 3 * F), xn = cvt(G.xn, 3 * F), cc = cvt(G.cc, 3 * C);
        vol = cvt(G.vol, C), bdist = cvt(G.bdist, B);
    }
    if (P.err != hipSuccess) {
        delete hm;
        return CUPCFD_E_ERROR;
    }
    v.faceCell1ID = G.c1, v.faceCell2ID = G.c2, v.faceBoundaryID = G.fb;
    v.faceLambda = lam, v.faceRLencos = rl, v.faceArea = area;
    v.faceNorm = V3{nrm, 0}, v.faceCenter = V3{ctr, 0}, v.faceXpac = V3{xp, 0}, v.faceXnac = V3{xn, 0};
    v.cellCenter = V3{cc, 0}, v.cellVolume = vol;
    v.boundaryFaceID = G.bface, v.boundaryRegionID = G.bregion, v.boundaryDistance = bdist;
    v.regionType = rv.regionType, v.regionStd = rv.regionStd, v.regionFlux = rv.regionFlux;
    v.regionAdiab = rv.regionAdiab, v.regionELog = rv.regionELog, v.regionDensity = rv.regionDensity;
    v.regionTurbKE = rv.regionTurbKE, v.regionTurbDiss = rv.regionTurbDiss, v.regionSplvl = rv.regionSplvl;
    v.regionDen = rv.regionDen, v.regionR = rv.regionR, v.regionT = rv.regionT, v.regionUVW = rv.regionUVW;
    try {
        rc = build_device(M, v, flags, true);
    } catch (const std::bad_alloc &) {
        rc = CUPCFD_E_ERROR;
    }
    if (rc == CUPCFD_E_SUCCESS && host_out) {
        // the host copy becomes the mesh's view, as for a mesh built from one
        rc = download(G, *hm) ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
        if (rc == CUPCFD_E_SUCCESS) {
            View hv{};
            VT::host_view(hm, &hv);
            VT::kept(M) = hv;
        }
    } else if (rc == CUPCFD_E_SUCCESS) {
        // no host copy: the state starts as the host generators leave it
        // (boundary state 0, region state from the region table), and the
        // mesh keeps no view to mirror it through
        const DevMesh<T> &d = M.d;
        const size_t B = G.B, R = hm->nRegions;
        for (T *s : {d.byplus, d.buplus, d.bq, d.bh, d.bt})
            if (B && hipMemset(s, 0, B * sizeof(T)) != hipSuccess) rc = CUPCFD_E_ERROR;
        if (B && hipMemset(d.bshear, 0, 3 * B * sizeof(T)) != hipSuccess) rc = CUPCFD_E_ERROR;
        if (R && hipMemcpy(d.rylog, rv.regionYLog, R * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            rc = CUPCFD_E_ERROR;
        if (R && hipMemcpy(d.rforcet, rv.regionForceTangent.base, 3 * R * sizeof(T), hipMemcpyHostToDevice) !=
                     hipSuccess)
            rc = CUPCFD_E_ERROR;
        VT::kept(M) = View{};
    }
    if (rc == CUPCFD_E_SUCCESS && host_out) *host_out = hm;
    else delete hm;
    return rc;
}

template <class T>
int generate_entry(const cupcfd_structgen_params *params, int32_t flags, cupcfd_host_mesh **host_out,
                   cupcfd_hip_mesh **out) {
    if (!params || !out) return CUPCFD_E_NULL_PTR;
    *out = nullptr;
    if (host_out) *host_out = nullptr;
    auto *M = new (std::nothrow) Mesh<T>();
    if (!M) return CUPCFD_E_ERROR;
    cupcfd_host_mesh *hm = nullptr;
    int rc;
    try {
        rc = generate<T>(*params, flags, host_out ? &hm : nullptr, *M);
    } catch (const std::bad_alloc &) {
        rc = CUPCFD_E_ERROR;
    }
    auto *h = rc == CUPCFD_E_SUCCESS ? new (std::nothrow) cupcfd_hip_mesh{} : nullptr;
    if (h) {
        if constexpr (sizeof(T) == 8) *h = cupcfd_hip_mesh{64, M, nullptr};
        else *h = cupcfd_hip_mesh{32, nullptr, M};
        // a host copy's state (zero boundary state, the region table) goes to the device through the view
        if (hm) rc = sync_state_any(h, true);
    }
    if (!h || rc != CUPCFD_E_SUCCESS) {
        if (h) delete h;
        delete M;
        delete hm;
        return rc == CUPCFD_E_SUCCESS ? CUPCFD_E_ERROR : rc;
    }
    tile_split_build(*M);
    if (host_out) *host_out = hm;
    *out = h;
    return CUPCFD_E_SUCCESS;
}

}  // namespace

}  // namespace cupcfd_hip

using namespace cupcfd_hip;

extern "C" int cupcfd_hip_mesh_generate_f64(const cupcfd_structgen_params *params, int32_t flags,
                                            cupcfd_host_mesh **host_out, cupcfd_hip_mesh **out) {
    return generate_entry<double>(params, flags, host_out, out);
}

extern "C" int cupcfd_hip_mesh_generate_f32(const cupcfd_structgen_params *params, int32_t flags,
                                            cupcfd_host_mesh **host_out, cupcfd_hip_mesh **out) {
    return generate_entry<float>(params, flags, host_out, out);
}

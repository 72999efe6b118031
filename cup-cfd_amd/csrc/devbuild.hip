// Device-side mesh repack: the host repack of devmesh.cpp as HIP kernels plus
// rocPRIM stable radix sorts and scans. It produces the same DevMesh arrays
// bit for bit, and the same host-side mirrors (orders, ranks, boundary types).
// The reference builds its cell→face map on the host at the same point
// (CupCfdSoAMesh::updateCellFaceMap,
// src/geometry/mesh/implementation/component/CupCfdSoAMesh.cpp:549-668).
//
// Why each step matches the host build exactly:
//   cell order  std::stable_sort by Morton key == stable LSD radix sort of
//               (key, cell id) pairs; the key arithmetic is the same double
//               expression (built without contraction), and min/max of the
//               bounding box are order-independent.
//   face order  the host's counting sort by (boundary?, owner) bucket is
//               stable == radix sort of (bucket, face id); its stable sort
//               of each owner tile's interior run by neighbour group ==
//               a stable radix sort of the interior run by (tile, group).
//   CSR rows    slot 2f is face f's cell1 entry, 2f+1 its cell2 entry; a
//               stable sort of the 2F slots by internal cell keeps ascending
//               slot order within each row, which is the host's fill order.
//   lists       boundaries by region type, OUTLET boundaries by owning cell
//               and halo faces by tile are stable sorts of ascending ids.
//   carries     "the last face of a kind before f, in mesh face order" is an
//               exclusive max-scan of (kind ? f : -1).
//   tiles       the owned faces of cells [c0, c1) are one run of the sorted
//               bucket keys, found by binary search.
#include <climits>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <vector>

#include "devmesh.hpp"

namespace cupcfd_hip {

namespace {

constexpr int kB = 256;

// small counters gathered by the build kernels
enum {
    S_ERR = 0,     // an index outside its range
    S_FI,          // interior faces
    S_MAXROW,      // longest CSR row
    S_MINROW,      // shortest CSR row
    S_TSLOTS,      // most slots of any tile
    S_TNNZ,        // most CSR entries of any tile
    S_TFAIL,       // some tile cannot be addressed by 15-bit slots
    S_BTYPE = 7,   // boundaries per region type (5)
    S_NTYPE = 12,  // boundary faces per region type (5)
    S_COUNT = 17
};

inline int grid_for(int64_t n) {
    const int64_t g = (n + kB - 1) / kB;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

inline unsigned bits_for(int64_t x) {  // bits that hold every value in [0, x]
    unsigned b = 1;
    while (b < 32 && (((int64_t)1 << b) <= x)) b++;
    return b;
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)(n); i += (int64_t)gridDim.x * blockDim.x)

// Block-level counters: threads combine in LDS, then one global atomic per
// block and counter (a global atomic per thread serialises millions of
// updates on one address).
struct BlockCounters {
    int *s;
    __device__ explicit BlockCounters(int *lds, int n) : s(lds) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = 0;
        __syncthreads();
    }
};

__global__ void __launch_bounds__(kB) k_validate_faces(int F, int C, int B, const int *c1, const int *c2,
                                                       const int *fb, int *stat) {
    __shared__ int sc[2];
    BlockCounters bc(sc, 2);
    int fi = 0, bad = 0;
    GRID_STRIDE(f, F) {
        const int a = c1[f], b = fb[f];
        if (a < 0 || a >= C) bad = 1;
        if (b > -1) {
            if (b >= B) bad = 1;
        } else {
            const int n = c2[f];
            if (n < 0 || n >= C) bad = 1;
            else fi++;
        }
    }
    if (bad) atomicOr(&sc[0], 1);
    if (fi) atomicAdd(&sc[1], fi);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (sc[0]) atomicOr(&stat[S_ERR], 1);
        if (sc[1]) atomicAdd(&stat[S_FI], sc[1]);
    }
}

__global__ void k_validate_bnds(int B, int F, int R, const int *bf, const int *br, int *stat) {
    int bad = 0;
    GRID_STRIDE(b, B) {
        const int f = bf[b], r = br[b];
        if (f < 0 || f >= F || r < 0 || r >= R) bad = 1;
    }
    if (bad) atomicOr(&stat[S_ERR], 1);
}

// bounding box of the cell centres with the host's comparisons
// (lo = x < lo ? x : lo, hi = hi < x ? x : hi), which skip NaN coordinates
__device__ inline void box_fold(double *lo, double *hi, const double *plo, const double *phi) {
    for (int k = 0; k < 3; k++) {
        lo[k] = plo[k] < lo[k] ? plo[k] : lo[k];
        hi[k] = hi[k] < phi[k] ? phi[k] : hi[k];
    }
}

__device__ inline void box_block(double *lo, double *hi, double *out) {
    __shared__ double s[6][kB];
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double plo[3], phi[3];
            for (int k = 0; k < 3; k++) {
                plo[k] = s[k][threadIdx.x + w];
                phi[k] = s[3 + k][threadIdx.x + w];
                lo[k] = s[k][threadIdx.x];
                hi[k] = s[3 + k][threadIdx.x];
            }
            box_fold(lo, hi, plo, phi);
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = lo[k];
                s[3 + k][threadIdx.x] = hi[k];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 6; j++) out[j] = s[j][0];
}

template <class T> __global__ void __launch_bounds__(kB) k_box(int C, const T *cc, double *part) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    GRID_STRIDE(c, C) {
        double x[3];
        for (int k = 0; k < 3; k++) x[k] = (double)cc[3 * c + k];
        box_fold(lo, hi, x, x);
    }
    box_block(lo, hi, part + 6 * (int64_t)blockIdx.x);
}

__global__ void __launch_bounds__(kB) k_box_final(int nb, const double *part, double *box) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = threadIdx.x; i < nb; i += kB) box_fold(lo, hi, part + 6 * i, part + 6 * i + 3);
    box_block(lo, hi, box);
}

__device__ inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Morton key of each cell centre quantised to 21 bits per axis over the box
template <class T> __global__ void k_morton(int C, const T *cc, const double *box, uint64_t *key, int *val) {
    GRID_STRIDE(c, C) {
        uint64_t q[3];
        for (int k = 0; k < 3; k++) {
            const double lo = box[k], span = box[3 + k] - box[k];
            const double t = span > 0 ? ((double)cc[3 * c + k] - lo) / span : 0.0;
            const double r = floor(t * 2097151.0 + 0.5);
            const double m = 0.0 < r ? r : 0.0;
            q[k] = (uint64_t)(m < 2097151.0 ? m : 2097151.0);
        }
        key[c] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
        val[c] = (int)c;
    }
}

__global__ void k_iota(int n, int *v) {
    GRID_STRIDE(i, n) v[i] = (int)i;
}

__global__ void k_rank(int n, const int *order, int *rank) {
    GRID_STRIDE(k, n) rank[order[k]] = (int)k;
}

// face bucket: interior faces by internal owner, then boundary faces by owner
__global__ void k_face_bucket(int F, int C, const int *c1, const int *fb, const int *crank, unsigned *key,
                              int *val) {
    GRID_STRIDE(f, F) {
        key[f] = (unsigned)((fb[f] > -1 ? C : 0) + crank[c1[f]]);
        val[f] = (int)f;
    }
}

// interior faces regrouped within each owner tile: same-tile faces first,
// then the faces shared with each other tile, by that tile (key (t1, group))
__global__ void k_face_group(int Fi, int TC, unsigned gb, const int *forder, const int *c1, const int *c2,
                             const int *crank, uint64_t *key, int *val) {
    GRID_STRIDE(k, Fi) {
        const int f = forder[k];
        const int t1 = crank[c1[f]] / TC, t2 = crank[c2[f]] / TC;
        key[k] = ((uint64_t)t1 << gb) | (uint64_t)(t2 == t1 ? 0 : t2 + 1);
        val[k] = f;
    }
}

template <class T>
__global__ void k_face_streams(int F, const int *order, const int *c1, const int *c2, const int *fb, const int *crank,
                               const T *lam, const T *rl, const T *area, const T *nrm, const T *ctr, const T *xp,
                               const T *xn, int2 *fconn, V4<T> *fS, V4<T> *fX, V4<T> *fP, V2<T> *fN, T *frl) {
    GRID_STRIDE(k, F) {
        const int f = order[k], b = fb[f];
        fconn[k] = make_int2(crank[c1[f]], b > -1 ? ~b : crank[c2[f]]);
        const int64_t f3 = 3 * (int64_t)f;
        fS[k] = V4<T>{nrm[f3], nrm[f3 + 1], nrm[f3 + 2], lam[f]};
        fX[k] = V4<T>{ctr[f3], ctr[f3 + 1], ctr[f3 + 2], area[f]};
        fP[k] = V4<T>{xp[f3], xp[f3 + 1], xp[f3 + 2], xn[f3]};
        fN[k] = V2<T>{xn[f3 + 1], xn[f3 + 2]};
        frl[k] = rl[f];
    }
}

template <class T>
__global__ void k_cell_streams(int C, const int *order, const T *cc, const T *vol, V4<T> *cg, T *cv) {
    GRID_STRIDE(k, C) {
        const int c = order[k];
        cg[k] = V4<T>{cc[3 * (int64_t)c], cc[3 * (int64_t)c + 1], cc[3 * (int64_t)c + 2], vol[c]};
        cv[k] = vol[c];
    }
}

// CSR slots: 2f = face f in its cell1's row, 2f+1 = in its cell2's row
// (key C = no entry: boundary faces have no cell2); cnt[c] counts row c
__global__ void k_csr_slots(int F, int C, const int *c1, const int *c2, const int *fb, const int *crank,
                            unsigned *key, int *val, int *cnt) {
    GRID_STRIDE(f, F) {
        const int a = crank[c1[f]];
        key[2 * f] = (unsigned)a;
        val[2 * f] = (int)(2 * f);
        atomicAdd(&cnt[a], 1);
        if (fb[f] > -1) {
            key[2 * f + 1] = (unsigned)C;
        } else {
            const int n = crank[c2[f]];
            key[2 * f + 1] = (unsigned)n;
            atomicAdd(&cnt[n], 1);
        }
        val[2 * f + 1] = (int)(2 * f + 1);
    }
}

__global__ void k_csr_ent(int nnz, const int *slot, const int *frank, const int *c1, const int *c2, const int *fb,
                          const int *crank, int2 *ent) {
    GRID_STRIDE(e, nnz) {
        const int s = slot[e], f = s >> 1, role = s & 1;
        int other;
        if (role) {
            other = crank[c1[f]];
        } else {
            const int b = fb[f];
            other = b > -1 ? ~b : crank[c2[f]];
        }
        ent[e] = make_int2((frank[f] << 1) | role, other);
    }
}

__global__ void __launch_bounds__(kB) k_row_stats(int C, const int *row, int *stat) {
    __shared__ int sc[2];
    if (threadIdx.x == 0) sc[0] = 0, sc[1] = INT_MAX;
    __syncthreads();
    int mx = 0, mn = INT_MAX;
    GRID_STRIDE(c, C) {
        const int l = row[c + 1] - row[c];
        mx = l > mx ? l : mx;
        mn = l < mn ? l : mn;
    }
    atomicMax(&sc[0], mx);
    atomicMin(&sc[1], mn);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(&stat[S_MAXROW], sc[0]);
        atomicMin(&stat[S_MINROW], sc[1]);
    }
}

__device__ inline void flush_types(const int *sc, int *dst) {
    __syncthreads();
    if (threadIdx.x < 5 && sc[threadIdx.x]) atomicAdd(&dst[threadIdx.x], sc[threadIdx.x]);
}

__global__ void __launch_bounds__(kB) k_bnd(int B, const int *bf, const int *br, const int *rtype, const int *frank,
                                            const int *c1, const int *crank, int *bface, int *bcell, unsigned *tkey,
                                            int *val, int *stat) {
    __shared__ int sc[5];
    BlockCounters bc(sc, 5);
    GRID_STRIDE(b, B) {
        const int f = bf[b], t = rtype[br[b]];
        bface[b] = frank[f];
        bcell[b] = crank[c1[f]];
        tkey[b] = (unsigned)t;
        val[b] = (int)b;
        atomicAdd(&sc[t], 1);
    }
    flush_types(sc, stat + S_BTYPE);
}

__global__ void __launch_bounds__(kB) k_ntype(int F, const int *fb, const int *br, const int *rtype, int *stat) {
    __shared__ int sc[5];
    BlockCounters bc(sc, 5);
    GRID_STRIDE(f, F) {
        const int b = fb[f];
        if (b > -1) atomicAdd(&sc[rtype[br[b]]], 1);
    }
    flush_types(sc, stat + S_NTYPE);
}

__global__ void k_owner_flags(int n, const int2 *fconn, unsigned char *flag) {
    GRID_STRIDE(i, n) flag[fconn[i].x] = 1;
}

__global__ void k_gather_key(int n, const int *idx, const int *src, unsigned *key) {
    GRID_STRIDE(i, n) key[i] = (unsigned)src[idx[i]];
}

__global__ void k_carry_marks(int F, const int *fb, const int *br, const int *rtype, int *io, int *vis) {
    GRID_STRIDE(f, F) {
        const int b = fb[f];
        const int t = b > -1 ? rtype[br[b]] : -1;
        const bool inout = t == CUPCFD_RTYPE_INLET || t == CUPCFD_RTYPE_OUTLET;
        io[f] = inout ? (int)f : -1;
        vis[f] = (b <= -1 || inout) ? (int)f : -1;
    }
}

__global__ void k_carry(int F, const int *fb, const int *br, const int *rtype, const int *frank, const int *last_io,
                        const int *last_vis, int *carry) {
    GRID_STRIDE(f, F) {
        const int b = fb[f];
        int src = -1;
        if (b <= -1) src = last_io[f];
        else if (rtype[br[b]] == CUPCFD_RTYPE_SYMP) src = last_vis[f];
        carry[frank[f]] = src >= 0 ? frank[src] : -1;
    }
}

__device__ inline int lower_bound_u(const unsigned *a, int n, unsigned x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__global__ void k_tinfo(int NT, int C, int F, int Fi, int TC, const unsigned *fkey, int4 *tinfo) {
    GRID_STRIDE(t, NT) {
        const int c0 = (int)t * TC, c1 = min(C, c0 + TC);
        const int i0 = lower_bound_u(fkey, F, (unsigned)c0), i1 = lower_bound_u(fkey, F, (unsigned)c1);
        const int b0 = lower_bound_u(fkey, F, (unsigned)(C + c0)), b1 = lower_bound_u(fkey, F, (unsigned)(C + c1));
        tinfo[t] = make_int4(i0, i1 - i0, b0, b1 - b0);
    }
}

// halo faces: interior faces whose cell2 lies in another tile than cell1 (key NT = none)
__global__ void k_halo_keys(int Fi, int TC, int NT, const int2 *fconn, unsigned *key, int *val, int *cnt) {
    GRID_STRIDE(k, Fi) {
        const int2 c = fconn[k];
        const int t2 = c.y / TC;
        if (t2 != c.x / TC) {
            key[k] = (unsigned)t2;
            atomicAdd(&cnt[t2], 1);
        } else {
            key[k] = (unsigned)NT;
        }
        val[k] = (int)k;
    }
}

__global__ void k_thc(int H, const int *thf, const int2 *fconn, int2 *thc) {
    GRID_STRIDE(h, H) thc[h] = fconn[thf[h]];
}

__global__ void k_tile_stats(int NT, int C, int TC, const int4 *tinfo, const int *thp, const int *row,
                             const int2 *fconn, int *stat) {
    GRID_STRIDE(t, NT) {
        const int4 ti = tinfo[t];
        const int nslot = ti.y + ti.w + (thp[t + 1] - thp[t]);
        const int c0 = (int)t * TC, c1 = min(C, c0 + TC);
        atomicMax(&stat[S_TSLOTS], nslot);
        atomicMax(&stat[S_TNNZ], row[c1] - row[c0]);
        bool ok = nslot >= 1 && nslot <= 32767;  // as the host repack: 15-bit slots, none empty
        for (int k = ti.x; k < ti.x + ti.y; k++)
            if (fconn[k].x < c0 || fconn[k].x >= c1) ok = false;
        for (int k = ti.z; k < ti.z + ti.w; k++)
            if (fconn[k].x < c0 || fconn[k].x >= c1) ok = false;
        if (!ok) atomicOr(&stat[S_TFAIL], 1);
    }
}

// slot of each CSR entry within its row's tile; every slot must decode back
// to the entry's face the way the tile kernels decode it
__global__ void k_eslot(int nnz, int TC, int Fi, const unsigned *ecell, const int2 *ent, const int4 *tinfo,
                        const int *thp, const int *thf, const int2 *fconn, uint16_t *eslot, int *stat) {
    GRID_STRIDE(e, nnz) {
        const int t = (int)ecell[e] / TC;
        const int4 ti = tinfo[t];
        const int h0 = thp[t], h1 = thp[t + 1];
        const int nslot = ti.y + ti.w + (h1 - h0);
        const int k = ent[e].x >> 1, role = ent[e].x & 1;
        int slot;
        if (k >= Fi) {
            slot = ti.y + (k - ti.z);
        } else if (fconn[k].x / TC == t) {
            slot = k - ti.x;
        } else {
            int lo = h0, hi = h1;
            while (lo < hi) {
                const int m = (lo + hi) >> 1;
                if (thf[m] < k) lo = m + 1;
                else hi = m;
            }
            slot = ti.y + ti.w + (lo - h0);
        }
        int back;
        if (slot < ti.y) back = ti.x + slot;
        else if (slot < ti.y + ti.w) back = ti.z + (slot - ti.y);
        else if (slot < nslot) back = thf[h0 + (slot - ti.y - ti.w)];
        else back = -1;
        if (back != k) atomicOr(&stat[S_TFAIL], 1);
        eslot[e] = (uint16_t)(slot | (role << 15));
    }
}

// Device allocations of one build: `keep` ones belong to the mesh, the rest
// are freed when the build ends. The first HIP error sticks.
struct Work {
    std::vector<void *> &keep;
    DeviceChunks tmp;
    void *prim = nullptr;
    size_t prim_bytes = 0;
    hipError_t err = hipSuccess;
    explicit Work(std::vector<void *> &k) : keep(k) {}
    ~Work() {
        if (prim) (void)hipFree(prim);
    }
    void note(hipError_t e) {
        if (err == hipSuccess && e != hipSuccess) err = e;
    }
    template <class X> X *alloc(size_t n, bool persistent) {
        if (err != hipSuccess) return nullptr;
        if (n == 0) n = 1;  // keep every pointer valid
        void *p = nullptr;
        if (persistent) {
            note(hipMalloc(&p, n * sizeof(X)));
            if (err == hipSuccess) keep.push_back(p);
        } else {
            note(tmp.get(n * sizeof(X), &p));
        }
        return err == hipSuccess ? (X *)p : nullptr;
    }
    template <class X> X *upload(const X *src, size_t n, bool persistent) {
        X *d = alloc<X>(n, persistent);
        if (d && src && n) note(hipMemcpy(d, src, n * sizeof(X), hipMemcpyHostToDevice));
        return d;
    }
    // an input array of the view: read in place when the view is device-resident
    // (copied only when the mesh keeps it), else uploaded
    bool resident = false;
    template <class X> X *input(const X *src, size_t n, bool persistent) {
        if (!resident) return upload(src, n, persistent);
        if (!persistent) return const_cast<X *>(src);
        X *d = alloc<X>(n, true);
        if (d && n) note(hipMemcpyAsync(d, src, n * sizeof(X), hipMemcpyDeviceToDevice, 0));
        return d;
    }
    template <class X> void download(std::vector<X> &h, const X *d, size_t n) {
        h.resize(n);
        if (err == hipSuccess && n) note(hipMemcpy(h.data(), d, n * sizeof(X), hipMemcpyDeviceToHost));
    }
    template <class X> void zero(X *d, size_t n) {
        if (err == hipSuccess && n) note(hipMemsetAsync(d, 0, n * sizeof(X), 0));
    }
    // a rocPRIM call: size query, shared temporary storage, then the call
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

template <class V3> const auto *v3at(const V3 &v, int64_t i, int k) {
    using S = std::remove_pointer_t<decltype(v.base)>;
    const int64_t stride = v.stride ? v.stride : (int64_t)(3 * sizeof(S));
    return (const S *)((const char *)v.base + i * stride) + k;
}

// a 3-vector array, packed xyz on the device (a device-resident view is packed)
template <class T, class V3> T *upload3(Work &W, const V3 &v, int64_t n, bool persistent) {
    if (v.stride == 0 || v.stride == (int64_t)(3 * sizeof(T)) || W.resident)
        return W.input((const T *)v.base, 3 * n, persistent);
    std::vector<T> h(3 * (size_t)n);
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) h[3 * i + k] = *v3at(v, i, k);
    return W.upload(h.data(), 3 * n, persistent);
}

}  // namespace

template <class T, class View> int build_device(Mesh<T> &M, const View &v, int flags, bool resident) {
    if (int rc = view_check(v); rc != CUPCFD_E_SUCCESS) return rc;
    const int C = v.nCells, F = v.nFaces, B = v.nBoundaries, R = v.nRegions;
    const int TC = kTileCells;
    Work W(M.allocs);
    W.resident = resident;
    // the temporaries take about this much: uploads (host views), sort buffers, marks
    W.tmp.min_chunk = std::max<size_t>(size_t(1) << 20, 96 * ((size_t)F + C + B));
    int st[S_COUNT];

    // ---- validation: every index in range, then region types --------------
    int *c1 = W.input(v.faceCell1ID, F, false);
    int *c2 = W.input(v.faceCell2ID, F, false);
    int *fb = W.input(v.faceBoundaryID, F, false);
    int *bf = W.input(v.boundaryFaceID, B, false);
    int *br = W.input(v.boundaryRegionID, B, true);
    int *stat;
    {
        int init[S_COUNT] = {};
        init[S_MINROW] = INT_MAX;
        stat = W.upload(init, S_COUNT, false);
    }
    W.launch(F, k_validate_faces, F, C, B, (const int *)c1, (const int *)c2, (const int *)fb, stat);
    W.launch(B, k_validate_bnds, B, F, R, (const int *)bf, (const int *)br, stat);
    if (W.err == hipSuccess) W.note(hipMemcpy(st, stat, sizeof(st), hipMemcpyDeviceToHost));
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;
    if (st[S_ERR]) return CUPCFD_E_INVALID_INDEX;
    for (int r = 0; r < R; r++)
        if (v.regionType[r] < 0 || v.regionType[r] > 4) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const int Fi = st[S_FI];
    const int nnz = F + Fi;
    int *rtype = W.upload(v.regionType, R, true);

    // ---- cell order -------------------------------------------------------
    T *cc = upload3<T>(W, v.cellCenter, C, false);
    T *vol = W.input(v.cellVolume, C, false);
    int *corder = W.alloc<int>(C, false), *crank = W.alloc<int>(C, false);
    if ((flags & CUPCFD_MESH_REORDER) && C > 1) {
        const int nb = std::min(grid_for(C), 1024);
        double *part = W.alloc<double>(6 * (size_t)nb, false), *box = W.alloc<double>(6, false);
        if (W.err == hipSuccess) {
            hipLaunchKernelGGL(k_box<T>, dim3(nb), dim3(kB), 0, 0, C, (const T *)cc, part);
            hipLaunchKernelGGL(k_box_final, dim3(1), dim3(kB), 0, 0, nb, (const double *)part, box);
            W.note(hipGetLastError());
        }
        uint64_t *key = W.alloc<uint64_t>(C, false), *skey = W.alloc<uint64_t>(C, false);
        int *val = W.alloc<int>(C, false);
        W.launch(C, k_morton<T>, C, (const T *)cc, (const double *)box, key, val);
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, key, skey, val, corder, (size_t)C, 0, 63, (hipStream_t)0);
        });
    } else {
        W.launch(C, k_iota, C, corder);
    }
    W.launch(C, k_rank, C, (const int *)corder, crank);

    // ---- face order: (boundary?, owner) buckets, mesh order within, then ----
    // ---- interior faces grouped by neighbour tile within each owner tile ---
    unsigned *fkey = W.alloc<unsigned>(F, false), *fskey = W.alloc<unsigned>(F, false);
    int *fval = W.alloc<int>(F, false), *forder = W.alloc<int>(F, false), *frank = W.alloc<int>(F, false);
    W.launch(F, k_face_bucket, F, C, (const int *)c1, (const int *)fb, (const int *)crank, fkey, fval);
    if (F)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, fkey, fskey, fval, forder, (size_t)F, 0,
                                             bits_for(2 * (int64_t)C), (hipStream_t)0);
        });
    if (Fi) {
        const int NTg = (C + kTileCells - 1) / kTileCells;
        const unsigned gb = bits_for(NTg);
        uint64_t *gkey = W.alloc<uint64_t>(Fi, false), *gskey = W.alloc<uint64_t>(Fi, false);
        int *gval = W.alloc<int>(Fi, false);
        W.launch(Fi, k_face_group, Fi, kTileCells, gb, (const int *)forder, (const int *)c1, (const int *)c2,
                 (const int *)crank, gkey, gval);
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, gkey, gskey, gval, forder, (size_t)Fi, 0, 2 * gb,
                                             (hipStream_t)0);
        });
    }
    W.launch(F, k_rank, F, (const int *)forder, frank);

    // ---- face and cell streams ---------------------------------------------
    DevMesh<T> &d = M.d;
    int2 *fconn = W.alloc<int2>(F, true);
    V4<T> *fS = W.alloc<V4<T>>(F, true), *fX = W.alloc<V4<T>>(F, true), *fP = W.alloc<V4<T>>(F, true);
    V2<T> *fN = W.alloc<V2<T>>(F, true);
    T *frl = W.alloc<T>(F, true);
    {
        T *lam = W.input(v.faceLambda, F, false), *rl = W.input(v.faceRLencos, F, false);
        T *area = W.input(v.faceArea, F, false);
        T *nrm = upload3<T>(W, v.faceNorm, F, false), *ctr = upload3<T>(W, v.faceCenter, F, false);
        T *xp = upload3<T>(W, v.faceXpac, F, false), *xn = upload3<T>(W, v.faceXnac, F, false);
        W.launch(F, k_face_streams<T>, F, (const int *)forder, (const int *)c1, (const int *)c2, (const int *)fb,
                 (const int *)crank, (const T *)lam, (const T *)rl, (const T *)area, (const T *)nrm, (const T *)ctr,
                 (const T *)xp, (const T *)xn, fconn, fS, fX, fP, fN, frl);
    }
    V4<T> *cg = W.alloc<V4<T>>(C, true);
    T *cv = W.alloc<T>(C, true);
    W.launch(C, k_cell_streams<T>, C, (const int *)corder, (const T *)cc, (const T *)vol, cg, cv);

    // ---- cell→face CSR, rows in mesh face order -----------------------------
    int *rcnt = W.alloc<int>((size_t)C + 1, false);
    int *row = W.alloc<int>((size_t)C + 1, true);
    unsigned *skey = W.alloc<unsigned>(2 * (size_t)F, false), *sskey = W.alloc<unsigned>(2 * (size_t)F, false);
    int *sval = W.alloc<int>(2 * (size_t)F, false), *ssval = W.alloc<int>(2 * (size_t)F, false);
    int2 *ent = W.alloc<int2>(nnz, true);
    W.zero(rcnt, (size_t)C + 1);
    W.launch(F, k_csr_slots, F, C, (const int *)c1, (const int *)c2, (const int *)fb, (const int *)crank, skey, sval,
             rcnt);
    if (F)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, skey, sskey, sval, ssval, 2 * (size_t)F, 0, bits_for(C),
                                             (hipStream_t)0);
        });
    W.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::exclusive_scan(tmp, bytes, rcnt, row, 0, (size_t)C + 1, rocprim::plus<int>(), (hipStream_t)0);
    });
    W.launch(nnz, k_csr_ent, nnz, (const int *)ssval, (const int *)frank, (const int *)c1, (const int *)c2,
             (const int *)fb, (const int *)crank, ent);
    W.launch(C, k_row_stats, C, (const int *)row, stat);

    // ---- boundaries / regions ----------------------------------------------
    int *bface = W.alloc<int>(B, true), *bcell = W.alloc<int>(B, true), *blist = W.alloc<int>(B, true);
    unsigned *tkey = W.alloc<unsigned>(B, false), *tskey = W.alloc<unsigned>(B, false);
    int *tval = W.alloc<int>(B, false);
    W.launch(B, k_bnd, B, (const int *)bf, (const int *)br, (const int *)rtype, (const int *)frank, (const int *)c1,
             (const int *)crank, bface, bcell, tkey, tval, stat);
    if (B)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, tkey, tskey, tval, blist, (size_t)B, 0, 3, (hipStream_t)0);
        });
    W.launch(F, k_ntype, F, (const int *)fb, (const int *)br, (const int *)rtype, stat);
    T *bdist = W.input(v.boundaryDistance, B, true);

    // cells owning a boundary face, ascending
    unsigned char *has = W.alloc<unsigned char>(C, false);
    int *bcells_t = W.alloc<int>(C, false);
    unsigned *nsel = W.alloc<unsigned>(2, false);
    W.zero(has, C);
    W.zero(nsel, 2);
    W.launch(F - Fi, k_owner_flags, F - Fi, (const int2 *)fconn + Fi, has);
    if (C)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::select(tmp, bytes, rocprim::counting_iterator<int>(0), has, bcells_t, nsel, (size_t)C,
                                   (hipStream_t)0);
        });

    // boundary counts per type give the OUTLET run of the by-type list
    if (W.err == hipSuccess) W.note(hipMemcpy(st, stat, sizeof(st), hipMemcpyDeviceToHost));
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;
    int boff[5], off = 0;
    for (int t = 0; t < 5; t++) {
        boff[t] = off;
        off += st[S_BTYPE + t];
    }
    const int nout = st[S_BTYPE + CUPCFD_RTYPE_OUTLET];
    int *oent = W.alloc<int>(nout, true);
    unsigned *okey = W.alloc<unsigned>(nout, false), *oskey = W.alloc<unsigned>(nout, false);
    int *ocell_t = W.alloc<int>(nout, false), *ocnt_t = W.alloc<int>(nout, false);
    if (nout) {
        const int *olist = blist + boff[CUPCFD_RTYPE_OUTLET];
        W.launch(nout, k_gather_key, nout, olist, (const int *)bcell, okey);
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, okey, oskey, olist, oent, (size_t)nout, 0, bits_for(C),
                                             (hipStream_t)0);
        });
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::run_length_encode(tmp, bytes, oskey, (size_t)nout, (unsigned *)ocell_t, ocnt_t, nsel + 1,
                                              (hipStream_t)0);
        });
    }

    // ---- carried-value sources (typed meshes) -------------------------------
    bool typed = false;
    for (int r = 0; r < R; r++)
        if (v.regionType[r] != CUPCFD_RTYPE_DEFAULT) typed = true;
    int *carry = nullptr;
    if (typed) {
        carry = W.alloc<int>(F, true);
        int *io = W.alloc<int>(F, false), *vis = W.alloc<int>(F, false);
        int *lio = W.alloc<int>(F, false), *lvis = W.alloc<int>(F, false);
        W.launch(F, k_carry_marks, F, (const int *)fb, (const int *)br, (const int *)rtype, io, vis);
        if (F) {
            W.prim_call([&](void *tmp, size_t &bytes) {
                return rocprim::exclusive_scan(tmp, bytes, io, lio, -1, (size_t)F, rocprim::maximum<int>(),
                                               (hipStream_t)0);
            });
            W.prim_call([&](void *tmp, size_t &bytes) {
                return rocprim::exclusive_scan(tmp, bytes, vis, lvis, -1, (size_t)F, rocprim::maximum<int>(),
                                               (hipStream_t)0);
            });
        }
        W.launch(F, k_carry, F, (const int *)fb, (const int *)br, (const int *)rtype, (const int *)frank,
                 (const int *)lio, (const int *)lvis, carry);
    }

    // ---- cell tiles ---------------------------------------------------------
    const int NT = (C + TC - 1) / TC;
    int4 *tinfo = W.alloc<int4>(NT, true);
    int *thp = W.alloc<int>((size_t)NT + 1, true);
    int *hcnt = W.alloc<int>((size_t)NT + 1, false);
    unsigned *hkey = W.alloc<unsigned>(Fi, false), *hskey = W.alloc<unsigned>(Fi, false);
    int *hval = W.alloc<int>(Fi, false), *hsval = W.alloc<int>(Fi, false);
    W.launch(NT, k_tinfo, NT, C, F, Fi, TC, (const unsigned *)fskey, tinfo);
    W.zero(hcnt, (size_t)NT + 1);
    W.launch(Fi, k_halo_keys, Fi, TC, NT, (const int2 *)fconn, hkey, hval, hcnt);
    if (Fi)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, hkey, hskey, hval, hsval, (size_t)Fi, 0, bits_for(NT),
                                             (hipStream_t)0);
        });
    W.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::exclusive_scan(tmp, bytes, hcnt, thp, 0, (size_t)NT + 1, rocprim::plus<int>(), (hipStream_t)0);
    });

    // list lengths fix the exact sizes of the remaining arrays
    unsigned sel[2] = {0, 0};
    int H = 0;
    if (W.err == hipSuccess) W.note(hipMemcpy(sel, nsel, sizeof(sel), hipMemcpyDeviceToHost));
    if (W.err == hipSuccess) W.note(hipMemcpy(&H, thp + NT, sizeof(int), hipMemcpyDeviceToHost));
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;
    const int nbc = (int)sel[0], noc = nout ? (int)sel[1] : 0;
    int *bcells = W.alloc<int>(nbc, true);
    int *ocells = W.alloc<int>(noc, true), *orow = W.alloc<int>((size_t)noc + 1, true);
    int *thf = W.alloc<int>(H, true);
    int2 *thc = W.alloc<int2>(H, true);
    uint16_t *eslot = W.alloc<uint16_t>(nnz, true);
    if (W.err == hipSuccess && nbc) W.note(hipMemcpyAsync(bcells, bcells_t, nbc * sizeof(int), hipMemcpyDeviceToDevice, 0));
    if (W.err == hipSuccess && noc) W.note(hipMemcpyAsync(ocells, ocell_t, noc * sizeof(int), hipMemcpyDeviceToDevice, 0));
    W.zero(orow, 1);
    if (noc)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::inclusive_scan(tmp, bytes, ocnt_t, orow + 1, (size_t)noc, rocprim::plus<int>(),
                                           (hipStream_t)0);
        });
    if (W.err == hipSuccess && H) W.note(hipMemcpyAsync(thf, hsval, H * sizeof(int), hipMemcpyDeviceToDevice, 0));
    W.launch(H, k_thc, H, (const int *)thf, (const int2 *)fconn, thc);
    W.launch(NT, k_tile_stats, NT, C, TC, (const int4 *)tinfo, (const int *)thp, (const int *)row,
             (const int2 *)fconn, stat);
    W.launch(nnz, k_eslot, nnz, TC, Fi, (const unsigned *)sskey, (const int2 *)ent, (const int4 *)tinfo,
             (const int *)thp, (const int *)thf, (const int2 *)fconn, eslot, stat);

    // ---- host-side mirrors --------------------------------------------------
    W.download(M.cell_order, corder, C);
    W.download(M.cell_rank, crank, C);
    W.download(M.face_order, forder, F);
    W.download(M.face_rank, frank, F);
    {
        std::vector<int2> fc;
        W.download(fc, fconn, F);
        M.fconn_h.resize(2 * (size_t)F);
        if (F) std::memcpy(M.fconn_h.data(), fc.data(), 2 * (size_t)F * sizeof(int32_t));
    }
    {
        std::vector<unsigned> bt;
        W.download(bt, tkey, B);
        M.btype_h.assign(bt.begin(), bt.end());
    }
    if (W.err == hipSuccess) W.note(hipMemcpy(st, stat, sizeof(st), hipMemcpyDeviceToHost));
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;
    for (int t = 0; t < 5; t++) M.ntype[t] = st[S_NTYPE + t];
    if (resident) W.download(M.bregion_h, (const int32_t *)br, B);
    else M.bregion_h.assign(v.boundaryRegionID, v.boundaryRegionID + B);
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;

    // ---- region arrays (a handful of entries: built on the host) -------------
    std::vector<uint8_t> rstd(R), rflux(R), radiab(R);
    for (int r = 0; r < R; r++) {
        rstd[r] = v.regionStd ? v.regionStd[r] : 0;
        rflux[r] = v.regionFlux ? v.regionFlux[r] : 0;
        radiab[r] = v.regionAdiab ? v.regionAdiab[r] : 0;
    }
    std::vector<T> rbuf;
    auto put_r = [&](const T *src) {
        rbuf.assign(R, (T)0);
        if (src)
            for (int r = 0; r < R; r++) rbuf[r] = src[r];
        return (const T *)W.upload(rbuf.data(), R, true);
    };

    d.C = C;
    d.F = F;
    d.Fi = Fi;
    d.B = B;
    d.R = R;
    d.nnz = nnz;
    d.maxrow = st[S_MAXROW];
    d.rowlen = C && st[S_MINROW] == st[S_MAXROW] ? st[S_MAXROW] : 0;
    d.fconn = fconn;
    d.fgeoS = fS;
    d.fgeoX = fX;
    d.fgeoP = fP;
    d.fgeoN = fN;
    d.frl = frl;
    d.cgeo = cg;
    d.cvol = cv;
    d.row = row;
    d.ent = ent;
    d.bface = bface;
    d.bregion = br;
    d.bcell = bcell;
    d.bdist = bdist;
    d.byplus = W.alloc<T>(B, true);
    d.buplus = W.alloc<T>(B, true);
    d.bq = W.alloc<T>(B, true);
    d.bh = W.alloc<T>(B, true);
    d.bt = W.alloc<T>(B, true);
    d.bshear = W.alloc<T>(3 * (size_t)B, true);
    d.rtype = rtype;
    d.rstd = W.upload(rstd.data(), R, true);
    d.rflux = W.upload(rflux.data(), R, true);
    d.radiab = W.upload(radiab.data(), R, true);
    d.rylog = W.alloc<T>(R, true);
    d.rforcet = W.alloc<T>(3 * (size_t)R, true);
    d.relog = put_r(v.regionELog);
    d.rdensity = put_r(v.regionDensity);
    d.rturbke = put_r(v.regionTurbKE);
    d.rturbdiss = put_r(v.regionTurbDiss);
    d.rsplvl = put_r(v.regionSplvl);
    d.rden = put_r(v.regionDen);
    d.rr = put_r(v.regionR);
    d.rt = put_r(v.regionT);
    {
        std::vector<T> u(3 * (size_t)R, (T)0);
        if (v.regionUVW.base)
            for (int r = 0; r < R; r++)
                for (int k = 0; k < 3; k++) u[3 * r + k] = *v3at(v.regionUVW, r, k);
        d.ruvw = W.upload(u.data(), 3 * (size_t)R, true);
    }
    d.carry_src = carry;
    d.typed = typed ? 1 : 0;
    d.carry_init = (T)0;
    d.bcells = bcells;
    d.nbc = nbc;
    for (int t = 0; t < 5; t++) {
        d.blist[t] = blist + boff[t];
        d.nblist[t] = st[S_BTYPE + t];
    }
    d.ocells = ocells;
    d.orow = orow;
    d.oent = oent;
    d.noc = noc;
    d.tcells = TC;
    d.ntiles = st[S_TFAIL] ? 0 : NT;
    d.tslots = st[S_TSLOTS];
    d.tnnz = st[S_TNNZ];
    d.tinfo = tinfo;
    d.thp = thp;
    d.thf = thf;
    d.thc = thc;
    d.eslot = eslot;
    M.radiab_h = radiab;
    if (W.err == hipSuccess) W.note(hipDeviceSynchronize());
    return W.err == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

// ---- order-preserving face colouring -----------------------------------------

namespace {

__global__ void k_entry_cell(int C, const int *row, int *ecell) {
    GRID_STRIDE(c, C) {
        for (int e = row[c]; e < row[c + 1]; e++) ecell[e] = (int)c;
    }
}

// colour(face of entry e) >= colour(face of entry e-1 in the same row) + 1.
// Colours only rise and never pass their fixed point (the longest chain of
// faces ending at the face along shared rows), so repeating this until
// nothing changes reaches exactly that fixed point in any update order.
__global__ void __launch_bounds__(kB) k_colour_relax(int nnz, const int *row, const int *ecell, const int2 *ent,
                                                     int *colour, int *changed) {
    __shared__ int sc[1];
    BlockCounters bc(sc, 1);
    int any = 0;
    GRID_STRIDE(e, nnz) {
        if (e == row[ecell[e]]) continue;
        const int p = ent[e - 1].x >> 1, k = ent[e].x >> 1;
        const int cand = __hip_atomic_load(&colour[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
        if (cand > __hip_atomic_load(&colour[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomicMax(&colour[k], cand);
            any = 1;
        }
    }
    if (any) atomicOr(&sc[0], 1);
    __syncthreads();
    if (threadIdx.x == 0 && sc[0]) atomicOr(changed, 1);
}

__global__ void __launch_bounds__(kB) k_colour_stats(int F, const int *colour, unsigned *key, int *val, int *cmax) {
    __shared__ int sc[1];
    BlockCounters bc(sc, 1);
    int mx = 0;
    GRID_STRIDE(k, F) {
        key[k] = (unsigned)colour[k];
        val[k] = (int)k;
        mx = colour[k] > mx ? colour[k] : mx;
    }
    atomicMax(&sc[0], mx);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(cmax, sc[0]);
}

__global__ void k_colour_count(int F, const unsigned *key, int *cnt) {
    GRID_STRIDE(k, F) atomicAdd(&cnt[key[k]], 1);
}

}  // namespace

template <class T> int build_face_colours(Mesh<T> &M) {
    if (M.ncolours >= 0) return CUPCFD_E_SUCCESS;
    const DevMesh<T> &d = M.d;
    const int F = d.F, C = d.C, nnz = d.nnz;
    Work W(M.allocs);
    W.tmp.min_chunk = std::max<size_t>(size_t(1) << 20, 24 * ((size_t)F + nnz));
    int *colour = W.alloc<int>(F, true);
    int *ecell = W.alloc<int>(nnz, false);
    int *flag = W.alloc<int>(2, false);
    W.zero(colour, F);
    W.launch(C, k_entry_cell, C, d.row, ecell);
    // each pass lengthens every chain it has not finished by at least one face
    for (int it = 0; it <= F && W.err == hipSuccess; it++) {
        W.zero(flag, 1);
        W.launch(nnz, k_colour_relax, nnz, d.row, (const int *)ecell, d.ent, colour, flag);
        int changed = 0;
        if (W.err == hipSuccess) W.note(hipMemcpy(&changed, flag, sizeof(int), hipMemcpyDeviceToHost));
        if (!changed) break;
    }
    unsigned *key = W.alloc<unsigned>(F, false), *skey = W.alloc<unsigned>(F, false);
    int *val = W.alloc<int>(F, false);
    int *faces = W.alloc<int>(F, true);
    W.zero(flag, 2);
    W.launch(F, k_colour_stats, F, (const int *)colour, key, val, flag + 1);
    int cmax = 0;
    if (W.err == hipSuccess) W.note(hipMemcpy(&cmax, flag + 1, sizeof(int), hipMemcpyDeviceToHost));
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;
    const int ncol = F ? cmax + 1 : 0;
    if (F)
        W.prim_call([&](void *tmp, size_t &bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, key, skey, val, faces, (size_t)F, 0, bits_for(ncol),
                                             (hipStream_t)0);
        });
    int *cnt = W.alloc<int>((size_t)ncol + 1, false), *off = W.alloc<int>((size_t)ncol + 1, false);
    W.zero(cnt, (size_t)ncol + 1);
    W.launch(F, k_colour_count, F, (const unsigned *)key, cnt);
    W.prim_call([&](void *tmp, size_t &bytes) {
        return rocprim::exclusive_scan(tmp, bytes, cnt, off, 0, (size_t)ncol + 1, rocprim::plus<int>(),
                                       (hipStream_t)0);
    });
    W.download(M.colour_off, (const int *)off, (size_t)ncol + 1);
    if (W.err == hipSuccess) W.note(hipDeviceSynchronize());
    if (W.err != hipSuccess) return CUPCFD_E_ERROR;
    M.colour_of = colour;
    M.colour_faces = faces;
    M.ncolours = ncol;
    return CUPCFD_E_SUCCESS;
}

template int build_face_colours<double>(Mesh<double> &);
template int build_face_colours<float>(Mesh<float> &);

template int build_device<double, cupcfd_mesh_view_f64>(Mesh<double> &, const cupcfd_mesh_view_f64 &, int, bool);
template int build_device<float, cupcfd_mesh_view_f32>(Mesh<float> &, const cupcfd_mesh_view_f32 &, int, bool);

}  // namespace cupcfd_hip

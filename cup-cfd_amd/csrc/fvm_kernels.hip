// gfx950 kernels for CUP-CFD's Dolfyn-derived finite-volume loops
// (include/fvm/implementation/component/*Kernels.ipp) and the resident-mode
// C ABI (cupcfd_dev_*).
//
// Design
//   * Face loops that scatter into both cells of a face (gradient, scalar,
//     UVW) run as per-cell gathers over the cell→face CSR, whose rows are in
//     the reference's face order: each cell applies its faces' updates in
//     exactly the sequence the sequential loop does, so results are
//     bit-identical with no atomics and no colouring. Each face is evaluated
//     by both of its cells (same inputs, same operations, same bits).
//   * The mass-flux face loop writes only per-face outputs for interior faces
//     and runs face-parallel; its boundary faces run per owning cell.
//   * Loop-carried locals of the reference (PhiFace, uFace/vFace/wFace,
//     visFace) are recomputed from the face that last assigned them, found by
//     a per-face index built with the mesh.
//   * Sums over boundaries (mass boundary loops, force tangents) keep the
//     reference's sequential order with one wavefront: 64 lanes load, every
//     lane accumulates the 64 values in order through cross-lane reads.
//   * Arithmetic mirrors the reference expression by expression; the file is
//     compiled with -ffp-contract=off so no multiply-add is fused.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "devmesh.hpp"

#pragma clang fp contract(off)

namespace cupcfd_hip {

constexpr int kBlock = 256;

template <class T> struct v3 { T x, y, z; };

template <class T> __device__ __forceinline__ v3<T> ld3(const T *__restrict__ p, int i) {
    return v3<T>{p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
}
// 3-vector i of an array of vectors whose components are `ks` apart
// (components of one scalar in a scalar-innermost batch of ks scalars)
template <class T> __device__ __forceinline__ v3<T> ld3s(const T *p, int i, int ks);
// a strided 3-vector of storage type S, widened to T
template <class T, class S> __device__ __forceinline__ v3<T> ld3s_as(const S *p, int i, int ks) {
    const v3<S> v = ld3s(p, i, ks);
    return v3<T>{(T)v.x, (T)v.y, (T)v.z};
}
template <class T> __device__ __forceinline__ v3<T> ld3s(const T *p, int i, int ks) {
    const size_t b = (size_t)3 * i * ks;
    return v3<T>{p[b], p[b + ks], p[b + 2 * (size_t)ks]};
}
// a 3-vector of storage type S widened to T. Batched kernels keep gathered
// values in their storage type until first use: a conversion right after a
// load inside a branch makes the wave wait for that load before it issues the
// next, so a row's gathers would run one memory round trip at a time.
template <class T, class S> __device__ __forceinline__ v3<T> widen3(v3<S> v) {
    return v3<T>{(T)v.x, (T)v.y, (T)v.z};
}
template <class T> __device__ __forceinline__ void st3(T *__restrict__ p, int i, v3<T> v) {
    p[3 * (size_t)i] = v.x;
    p[3 * (size_t)i + 1] = v.y;
    p[3 * (size_t)i + 2] = v.z;
}
template <class T> __device__ __forceinline__ v3<T> xyz(const V4<T> &a) { return v3<T>{a.x, a.y, a.z}; }

// Per-face outputs written once per sweep (RFace, the mass flux) and a tile's
// final cell sums: non-temporal stores, which keep them from evicting the cell
// values that neighbouring faces and tiles read again from L2. (Non-temporal
// loads of the face records were measured slower, and so were non-temporal
// stores of the gradient, whose outputs the next kernels read back.)
template <class T> __device__ __forceinline__ void st_stream(T *p, T v) { __builtin_nontemporal_store(v, p); }

// EuclideanVector::dotProduct (EuclideanVector.cpp:83-90): from 0.0, in order
template <class T> __device__ __forceinline__ T dot3(v3<T> a, v3<T> b) {
    T d = 0.0;
    d += (a.x * b.x);
    d += (a.y * b.y);
    d += (a.z * b.z);
    return d;
}
__device__ __forceinline__ double sqrtw(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrtw(float x) { return sqrtf(x); }
// EuclideanVector::length (EuclideanVector.cpp:93-99): a double
template <class T> __device__ __forceinline__ double length3(v3<T> a) {
    T d = dot3(a, a);
    double l = sqrtw(d);
    return l;
}
// EuclideanVector::normalise (EuclideanVector.cpp:102-110)
template <class T> __device__ __forceinline__ v3<T> normalise3(v3<T> v) {
    T s = (T)length3(v);
    if (s != (T)0) {
        s = (T)(1.0) / s;
        v.x *= s;
        v.y *= s;
        v.z *= s;
    }
    return v;
}
// (a * fa) + (b * fb) componentwise
template <class T> __device__ __forceinline__ v3<T> lerp3(v3<T> a, T fa, v3<T> b, T fb) {
    return v3<T>{a.x * fa + b.x * fb, a.y * fa + b.y * fb, a.z * fa + b.z * fb};
}
template <class T> __device__ __forceinline__ v3<T> sub3(v3<T> a, v3<T> b) {
    return v3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
// std::min / std::max
template <class T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// Ordered sums (the boundary loops' flow sums, region flows and force
// tangents): for each component c < NV, init_c + v_c(0) + v_c(1) + ... +
// v_c(n-1), added left to right as the reference's sequential loops do, so
// the result is the reference's, bit for bit. One workgroup of kSeqBlock
// threads: wave c adds component c, reading each value from LDS at a
// wave-uniform address (a broadcast, so the chain of dependent adds is the
// only serial work), while the other waves gather the next chunk of values
// into the other half of a double buffer. Returns component `wave` of the
// sum to the waves c < NV.
// ---------------------------------------------------------------------------
constexpr int kSeqBlock = 1024;
constexpr int kSeqLds = 3072;  // values per half of the double buffer (all components)

template <class T, int NV, class Get> __device__ T block_ordered_sum(T init, int n, Get get, T *buf) {
    constexpr int CH = kSeqLds / NV;  // items per chunk
    static_assert(CH % 64 == 0 && NV * 64 < kSeqBlock, "chunk of whole waves; staging waves left over");
    const int tid = threadIdx.x, wave = tid >> 6;
    auto stage = [&](int base, T *dst, int t0, int nt) {
        const int cnt = min(CH, n - base);
        for (int i = t0; i < cnt; i += nt) {
            T v[NV];
            get(base + i, v);
            for (int c = 0; c < NV; c++) dst[c * CH + i] = v[c];
        }
    };
    T acc = init;
    if (n > 0) stage(0, buf, tid, kSeqBlock);
    __syncthreads();
    int k = 0;
    for (int base = 0; base < n; base += CH, k ^= 1) {
        if (wave >= NV) {
            if (base + CH < n) stage(base + CH, buf + (k ^ 1) * NV * CH, tid - NV * 64, kSeqBlock - NV * 64);
        } else {
            const T *col = buf + k * NV * CH + wave * CH;
            const int cnt = min(CH, n - base);
#pragma unroll 32
            for (int i = 0; i < cnt; i++) acc = acc + col[i];
        }
        __syncthreads();
    }
    return acc;
}

// The same sums as a fixed-shape parallel tree (the mesh's tree_sums mode):
// thread i adds values i, i + kSeqBlock, ... in order, each wave combines its
// threads with a fixed xor tree, and one thread per component adds the waves
// in order; init is added last. The same bits on every call, not the
// sequential sum's.
template <class T, int NV, class Get> __device__ T block_tree_sum(T init, int n, Get get, T *buf) {
    constexpr int NW = kSeqBlock / 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    T p[NV];
    for (int c = 0; c < NV; c++) p[c] = 0;
    for (int i = tid; i < n; i += kSeqBlock) {
        T v[NV];
        get(i, v);
        for (int c = 0; c < NV; c++) p[c] = p[c] + v[c];
    }
    for (int c = 0; c < NV; c++) {
        for (int o = 32; o >= 1; o >>= 1) p[c] = p[c] + __shfl_xor(p[c], o, 64);
        if (lane == 0) buf[c * NW + wave] = p[c];
    }
    __syncthreads();
    if (tid < NV) {
        T s = buf[tid * NW];
        for (int w = 1; w < NW; w++) s = s + buf[tid * NW + w];
        buf[NV * NW + tid] = s;
    }
    __syncthreads();
    return wave < NV ? init + buf[NV * NW + wave] : init;
}

template <class T, int NV, class Get> __device__ T block_sum(int tree, T init, int n, Get get, T *buf) {
    return tree ? block_tree_sum<T, NV>(init, n, get, buf) : block_ordered_sum<T, NV>(init, n, get, buf);
}

// ===========================================================================
// Per-cell gathers: the face loops that scatter into both cells of a face
// (gradient, scalar, UVW) are evaluated per cell over its CSR row.
//
// Each face contributes a fixed sequence of additions to each accumulated
// cell output: the reference's `x = ((x - a) + b) - c` is `((x + (-a)) + b)
// + (-c)`, bit for bit, since IEEE negation is exact. A face therefore yields
// a few "terms"; a step with no third operation uses -0.0, for which x + -0.0
// == x holds exactly for every x. Folding the terms of a cell's faces in row
// order (the reference's face order) reproduces the sequential loop.
//
// Two launch shapes share the per-face term functions:
//   serial: one thread per cell, folding as it walks the row (any row length);
//   vector: LPC lanes per cell, lane j evaluating the j-th face of the row so
//           all of a cell's face loads are in flight at once; the terms go
//           through LDS and one lane per output component folds them in order.
// Both use XCD-aware block numbering: the GPU hands workgroups to its 8 XCDs
// round-robin, so logical block (b % 8) * per + b / 8 gives each XCD (and its
// private L2) one contiguous range of the locality-ordered cells.
// ===========================================================================
constexpr int kXcd = 8;
// cells per block of the lane-per-face gathers (block = lanes per cell x 32)
constexpr int kCellsPerBlock = 32;
__device__ __forceinline__ int xcd_block(int b, int per) { return (b % kXcd) * per + b / kXcd; }

template <class T> __device__ __forceinline__ T nzero() { return (T)(-0.0); }

// CSR entries are (face << 1 | role, other): `other` is the face's other cell,
// or ~boundary id. entry_cells returns the face's (cell1, cell2 or ~boundary).
__device__ __forceinline__ int2 entry_cells(int c, int2 en) {
    if (en.y < 0) return make_int2(c, en.y);
    return (en.x & 1) ? make_int2(en.y, c) : make_int2(c, en.y);
}

// Row of cell c: computed for meshes whose cells all have the same number of
// faces (no row-pointer load), read from the CSR otherwise
template <class T> __device__ __forceinline__ void row_range(const DevMesh<T> &m, int c, int &e0, int &n) {
    if (m.rowlen) {
        e0 = c * m.rowlen;
        n = m.rowlen;
    } else {
        e0 = m.row[c];
        n = m.row[c + 1] - e0;
    }
}

// ===========================================================================
// a1  GradientPhiGaussDolfyn — GradientKernels.ipp:38-136
// SWEEP0: dPhidxo is still the zero field of the first sweep, so the
// face-centre correction dot(0, Xf - xac) adds +0 and is skipped together
// with its centre and gradient loads.
// ===========================================================================
template <class T> struct GradArgs {
    const T *phi, *phiB, *grado;
    T *out1, *out2;
};

// the cell values an interior face of the gradient reads
template <class T> struct GradCell {
    T phi;
    v3<T> go, cc;
};
template <class T, bool SWEEP0>
__device__ __forceinline__ GradCell<T> grad_cell(const DevMesh<T> &m, const GradArgs<T> &a, int c) {
    GradCell<T> r;
    r.phi = a.phi[c];
    if (!SWEEP0) {
        r.go = ld3(a.grado, c);
        r.cc = xyz(m.cgeo[c]);
    }
    return r;
}

// phiF * S of interior face f, the same for both of its cells
template <class T, bool SWEEP0>
__device__ __forceinline__ void gradient_interior(const V4<T> &xf, const V4<T> &s, const GradCell<T> &P,
                                                  const GradCell<T> &N, T t[3]) {
    const T facn = s.w;
    const T facp = 1.0 - facn;
    T phiF = (N.phi * facn) + (P.phi * facp);
    if (!SWEEP0) {
        const v3<T> xac = lerp3(N.cc, facn, P.cc, facp);
        const v3<T> g = lerp3(N.go, facn, P.go, facp);
        phiF += dot3(g, sub3(xyz(xf), xac));
    }
    t[0] = phiF * s.x;
    t[1] = phiF * s.y;
    t[2] = phiF * s.z;
}

template <class T>
__device__ __forceinline__ void gradient_boundary(const GradArgs<T> &a, int ib, const V4<T> &s, T t[3]) {
    const T phiF = a.phiB[ib];
    t[0] = phiF * s.x;
    t[1] = phiF * s.y;
    t[2] = phiF * s.z;
}

template <class T, bool SWEEP0>
__device__ __forceinline__ void gradient_terms(const DevMesh<T> &m, const GradArgs<T> &a, int c, int2 en, T t[3]) {
    const int f = en.x >> 1;
    const int2 cn = entry_cells(c, en);
    const V4<T> s = m.fgeoS[f];
    if (cn.y >= 0)
        gradient_interior<T, SWEEP0>(SWEEP0 ? s : m.fgeoX[f], s, grad_cell<T, SWEEP0>(m, a, cn.x),
                                     grad_cell<T, SWEEP0>(m, a, cn.y), t);
    else
        gradient_boundary(a, ~cn.y, s, t);
}

template <class T>
__device__ __forceinline__ void gradient_store(const GradArgs<T> &a, int c, int k, T acc, T vol) {
    const T fact = 1.0 / vol;
    const T r = acc * fact;
    a.out1[3 * (size_t)c + k] = r;
    if (a.out2) a.out2[3 * (size_t)c + k] = r;
}

template <class T, bool SWEEP0>
__global__ __launch_bounds__(kBlock) void k_gradient(DevMesh<T> m, GradArgs<T> a, int per, int nblk) {
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int c = lb * kBlock + threadIdx.x;
    if (c >= m.C) return;
    T acc[3] = {0, 0, 0};
    int e0, n;
    row_range(m, c, e0, n);
    for (int e = e0; e < e0 + n; e++) {
        T t[3];
        gradient_terms<T, SWEEP0>(m, a, c, m.ent[e], t);
        acc[0] = acc[0] + t[0];
        acc[1] = acc[1] + t[1];
        acc[2] = acc[2] + t[2];
    }
    const T vol = m.cvol[c];
    for (int k = 0; k < 3; k++) gradient_store(a, c, k, acc[k], vol);
}

template <class T, bool SWEEP0, int LPC>
__global__ __launch_bounds__(LPC * 32) void k_gradient_vec(DevMesh<T> m, GradArgs<T> a, int per, int nblk) {
    constexpr int CPB = kCellsPerBlock, NT = 3, P = LPC + 1;
    __shared__ T lds[NT * CPB * P];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cl = threadIdx.x / LPC, j = threadIdx.x % LPC;
    const int c = lb * CPB + cl;
    int e0 = 0, n = 0;
    if (c < m.C) row_range(m, c, e0, n);
    if (j < n) {
        T t[NT];
        gradient_terms<T, SWEEP0>(m, a, c, m.ent[e0 + j], t);
        for (int k = 0; k < NT; k++) lds[(k * CPB + cl) * P + j] = t[k];
    }
    __syncthreads();
    if (c < m.C && j < 3) {
        T acc = 0;
        for (int i = 0; i < n; i++) acc = acc + lds[(j * CPB + cl) * P + i];
        gradient_store(a, c, j, acc, m.cvol[c]);
    }
}

// Colour-group schedule: one launch per colour of the mesh's order-preserving
// face colouring. Within a colour no two faces share a cell, so each face adds
// its terms to its cells' accumulators directly; colours rise along every
// cell's row, so a cell receives its terms in row order, from the zero the
// reference starts from: the gather's sums, bit for bit.
template <class T> __global__ void k_fill_zero(size_t n, T *p) {
    const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
    if (i < n) p[i] = (T)0;
}

template <class T, bool SWEEP0>
__global__ __launch_bounds__(kBlock) void k_gradient_colour(DevMesh<T> m, GradArgs<T> a, const int *faces, int n,
                                                            T *acc) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int f = faces[i];
    const int2 cn = m.fconn[f];
    const V4<T> s = m.fgeoS[f];
    T t[3];
    if (cn.y >= 0) {
        gradient_interior<T, SWEEP0>(SWEEP0 ? s : m.fgeoX[f], s, grad_cell<T, SWEEP0>(m, a, cn.x),
                                     grad_cell<T, SWEEP0>(m, a, cn.y), t);
        for (int k = 0; k < 3; k++) acc[3 * (size_t)cn.y + k] = acc[3 * (size_t)cn.y + k] + t[k];
    } else {
        gradient_boundary(a, ~cn.y, s, t);
    }
    for (int k = 0; k < 3; k++) acc[3 * (size_t)cn.x + k] = acc[3 * (size_t)cn.x + k] + t[k];
}

template <class T> __global__ __launch_bounds__(kBlock) void k_gradient_colour_store(DevMesh<T> m, GradArgs<T> a,
                                                                                      const T *acc) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.C) return;
    const T vol = m.cvol[c];
    for (int k = 0; k < 3; k++) gradient_store(a, c, k, acc[3 * (size_t)c + k], vol);
}

// ---------------------------------------------------------------------------
// Tile schedule (all three gathers): block b works on tile t = xcd_block(b),
// the cells [64t, 64t + 64), in three phases separated by barriers:
//  A. every global load of the block that does not depend on another block:
//     the tile's cell values (coalesced, into LDS), the CSR slots of its rows
//     (into LDS), each thread's first face slot with that face's values and
//     the values of the face's cell outside the tile (held in registers), and
//     each fold item's starting value;
//  B. each face slot of the tile (owned interior faces, owned boundary faces,
//     halo faces) is evaluated once, its terms in the cell1 view kept in LDS;
//     per-face outputs are written by the owning tile only;
//  C. each (cell, output component) folds its row in face order, reading the
//     terms of slot eslot[e] (negated for the cell2 side where the reference's
//     cell2 update is the cell1 update with every term's sign flipped).
// ---------------------------------------------------------------------------
static_assert(kTileBlock >= 4 * kTileCells, "phase A gives each thread at most one element of a cell field");

struct TileSlots {
    int4 ti;  // first owned interior face, count, first owned boundary face, count
    int h0, nh, nslot;
};
template <class T> __device__ __forceinline__ TileSlots tile_slots(const DevMesh<T> &m, int t) {
    TileSlots r;
    r.ti = m.tinfo[t];
    r.h0 = m.thp[t];
    r.nh = m.thp[t + 1] - r.h0;
    r.nslot = r.ti.y + r.ti.w + r.nh;
    return r;
}
// face of slot sl with its (cell1, cell2 or ~boundary); `owned` is false for halo slots
template <class T>
__device__ __forceinline__ int slot_conn(const DevMesh<T> &m, const TileSlots &ts, int sl, int2 &cn, bool &owned) {
    owned = true;
    int f;
    if (sl < ts.ti.y) {
        f = ts.ti.x + sl;
    } else if (sl < ts.ti.y + ts.ti.w) {
        f = ts.ti.z + (sl - ts.ti.y);
    } else {
        const int h = ts.h0 + (sl - ts.ti.y - ts.ti.w);
        owned = false;
        cn = m.thc[h];
        return m.thf[h];
    }
    cn = m.fconn[f];
    return f;
}
// slot_conn without branches: both loads go through selected pointers, so
// they issue with the other loads of a phase's first round (sl < ts.nslot)
template <class T>
__device__ __forceinline__ int slot_conn_flat(const DevMesh<T> &m, const TileSlots &ts, int sl, int2 &cn,
                                              bool &owned) {
    const int nown = ts.ti.y + ts.ti.w;
    owned = sl < nown;
    const int fo = sl < ts.ti.y ? ts.ti.x + sl : ts.ti.z + (sl - ts.ti.y);
    const int h = ts.h0 + (sl - nown);
    cn = *(owned ? m.fconn + fo : m.thc + h);
    const int fh = *(owned ? m.thp : m.thf + h);  // thp: any valid address when the slot is owned
    return owned ? fo : fh;
}
__device__ __forceinline__ bool in_tile(int c, int c0, int nc) { return (unsigned)(c - c0) < (unsigned)nc; }
// the cell of a slot's face outside the tile, or -1 (a slot has at most one)
__device__ __forceinline__ int outside_cell(int2 cn, int c0, int nc) {
    if (!in_tile(cn.x, c0, nc)) return cn.x;
    if (cn.y >= 0 && !in_tile(cn.y, c0, nc)) return cn.y;
    return -1;
}
// phase-A staging of one cell field, of a packed 3-vector field (3 LDS
// fields) and of the cell centres (3 LDS fields): one element per thread
template <class T> __device__ __forceinline__ void stage1(T *l, const T *p, int c0, int nc) {
    const int i = threadIdx.x;
    if (i < nc) l[i] = p[c0 + i];
}
template <class T> __device__ __forceinline__ void stage3(T *l, const T *p, int c0, int nc, int ld = kTileCells) {
    const int i = threadIdx.x;
    if (i < 3 * nc) l[(i % 3) * ld + i / 3] = p[3 * (size_t)c0 + i];
}
template <class T>
__device__ __forceinline__ void stage_centre(T *l, const V4<T> *g, int c0, int nc, int ld = kTileCells) {
    const int i = threadIdx.x;
    if (i < 4 * nc && (i & 3) != 3) l[(i & 3) * ld + (i >> 2)] = ((const T *)(g + c0))[i];
}
// the CSR slots of the tile's rows (src: m.eslot, or a column map of it) into
// LDS; returns the tile's first entry
template <class T>
__device__ __forceinline__ int stage_eslots(const DevMesh<T> &m, const uint16_t *src, uint16_t *le, int c0, int nc) {
    int e0, e1;
    if (m.rowlen) {
        e0 = c0 * m.rowlen;
        e1 = (c0 + nc) * m.rowlen;
    } else {
        e0 = m.row[c0];
        e1 = m.row[c0 + nc];
    }
    for (int i = threadIdx.x; i < e1 - e0; i += kTileBlock) le[i] = src[e0 + i];
    return e0;
}
template <class T> __device__ __forceinline__ int stage_eslots(const DevMesh<T> &m, uint16_t *le, int c0, int nc) {
    return stage_eslots(m, m.eslot, le, c0, nc);
}

template <class T, bool SWEEP0>
__global__ __launch_bounds__(kTileBlock) void k_gradient_tile(DevMesh<T> m, GradArgs<T> a, int per, int nblk) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TC = kTileCells, NF = SWEEP0 ? 1 : 7;  // phi (+ dPhidxo xyz, centre xyz)
    const int S = m.tslots;
    T *lc = (T *)smem;                         // [NF][TC]
    T *lt = lc + NF * TC;                      // [3][S]
    T *lo = lt + 3 * S;                        // [TC][3]: the tile's gradients, packed for the stores
    uint16_t *le = (uint16_t *)(lo + 3 * TC);  // [tnnz]
    const int t = xcd_block(blockIdx.x, per);
    if (t >= nblk) return;
    const int tid = threadIdx.x;
    const int c0 = t * TC, nc = min(TC, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    // A. First round: every load that depends on no other load, at clamped
    // indices so that all of them are unconditional and issue before the
    // first wait; their LDS writes follow. Thread tid stages phi of cell
    // tid % TC (tid < TC) and, after the first sweep, element tid of the
    // packed previous gradient and of the centres; CSR slot codes tid and
    // tid + kTileBlock; its slot's connectivity; and its fold item's row and
    // volume.
    const int ci = min(tid & (TC - 1), nc - 1);
    const T x1 = a.phi[c0 + ci];
    T gd = 0, xc = 0;
    if (!SWEEP0) {
        gd = a.grado[3 * (size_t)c0 + min(tid, 3 * nc - 1)];
        xc = ((const T *)(m.cgeo + c0))[min(tid, 4 * nc - 1)];
    }
    int eb, ee;
    if (m.rowlen) {
        eb = c0 * m.rowlen;
        ee = (c0 + nc) * m.rowlen;
    } else {
        eb = m.row[c0];
        ee = m.row[c0 + nc];
    }
    const int ne = ee - eb;  // >= 1: a tile whose cells have no faces takes no tile schedule
    const uint16_t q0 = m.eslot[eb + min(tid, ne - 1)], q1 = m.eslot[eb + min(tid + kTileBlock, ne - 1)];
    const int k = tid / TC, cl = tid % TC;
    const bool fold = tid < 3 * TC && cl < nc;
    int e0, n;
    row_range(m, c0 + ci, e0, n);
    const T vol = m.cvol[c0 + ci];
    const bool live = tid < ts.nslot;
    const int sl0 = min(tid, ts.nslot - 1);
    int2 cn;
    bool owned;
    int f = slot_conn_flat(m, ts, sl0, cn, owned);
    if (tid < nc) lc[tid] = x1;
    if (!SWEEP0) {
        if (tid < 3 * nc) lc[(1 + tid % 3) * TC + tid / 3] = gd;
        if (tid < 4 * nc && (tid & 3) != 3) lc[(4 + (tid & 3)) * TC + (tid >> 2)] = xc;
    }
    if (tid < ne) le[tid] = q0;
    if (tid + kTileBlock < ne) le[tid + kTileBlock] = q1;
    for (int i = tid + 2 * kTileBlock; i < ne; i += kTileBlock) le[i] = m.eslot[eb + i];
    int out = -1;
    V4<T> s{}, xf{};
    GradCell<T> o{};
    // second round: slot sl's face values and its cell outside the tile (cn
    // and f already set)
    auto rest = [&]() {
        s = m.fgeoS[f];
        out = -1;
        if (cn.y >= 0) {
            if (!SWEEP0) xf = m.fgeoX[f];
            out = outside_cell(cn, c0, nc);
            if (out >= 0) o = grad_cell<T, SWEEP0>(m, a, out);
        }
    };
    // a later slot sl of this thread (tiles of more than kTileBlock slots)
    auto load = [&](int sl) {
        f = slot_conn(m, ts, sl, cn, owned);
        rest();
    };
    if (live) rest();
    __syncthreads();
    // B
    auto cell = [&](int c) -> GradCell<T> {
        if (c == out) return o;
        const int l = c - c0;
        GradCell<T> r;
        r.phi = lc[l];
        if (!SWEEP0) {
            r.go = v3<T>{lc[1 * TC + l], lc[2 * TC + l], lc[3 * TC + l]};
            r.cc = v3<T>{lc[4 * TC + l], lc[5 * TC + l], lc[6 * TC + l]};
        }
        return r;
    };
    for (int sl = tid; sl < ts.nslot; sl += kTileBlock) {
        if (sl != tid) load(sl);
        T tt[3];
        if (cn.y >= 0) gradient_interior<T, SWEEP0>(xf, s, cell(cn.x), cell(cn.y), tt);
        else gradient_boundary(a, ~cn.y, s, tt);
        for (int q = 0; q < 3; q++) lt[q * S + sl] = tt[q];
    }
    __syncthreads();
    // C: fold, then (FP64) store the tile's 3 * nc consecutive values with
    // consecutive threads (the folds are component-major)
    if constexpr (sizeof(T) == 8) {
        if (fold) {
            T acc = 0;
            for (int e = e0; e < e0 + n; e++) acc = acc + lt[k * S + (le[e - eb] & 0x7fff)];
            const T fact = 1.0 / vol;
            lo[cl * 3 + k] = acc * fact;
        }
        __syncthreads();
        for (int i = tid; i < 3 * nc; i += kTileBlock) {
            const T r = lo[i];
            a.out1[3 * (size_t)c0 + i] = r;
            if (a.out2) a.out2[3 * (size_t)c0 + i] = r;
        }
    } else if (fold) {
        T acc = 0;
        for (int e = e0; e < e0 + n; e++) acc = acc + lt[k * S + (le[e - eb] & 0x7fff)];
        gradient_store(a, c0 + cl, k, acc, vol);
    }
}

// ===========================================================================
// a2  FluxMassDolfynFaceLoop — MassKernels.ipp:36-309
// ===========================================================================
template <class T> struct MassArgs {
    const T *dudx, *dvdx, *dwdx, *dpdx, *den, *u, *v, *w, *p, *ar;
    T *denB, *massFlux, *su, *rface;
    T small;
    int solveTurbEnergy, solveTurbDiss, solveVisc, solveEnthalpy;
    const T *teCell, *edCell, *viseffCell, *tCell;
    T *teB, *edB, *viseffB, *tB;
};

// the cell values an interior face of the mass loop reads
template <class T> struct MassCell {
    v3<T> du, dv, dw, dp;
    T den, u, v, w, p, ar;
    V4<T> g;  // centre, volume
};
template <class T> __device__ __forceinline__ MassCell<T> mass_cell(const DevMesh<T> &m, const MassArgs<T> &a, int c) {
    return MassCell<T>{ld3(a.dudx, c), ld3(a.dvdx, c), ld3(a.dwdx, c), ld3(a.dpdx, c), a.den[c], a.u[c],
                       a.v[c],         a.w[c],         a.p[c],         a.ar[c],         m.cgeo[c]};
}

// Interior face f (MassKernels.ipp:137-195) between ip = cell1 (P) and
// in = cell2 (N): mass flux and the two RFace entries.
template <class T>
__device__ __forceinline__ void mass_face(const MassArgs<T> &a, int f, const V4<T> &s, const V4<T> &xf,
                                          const V4<T> &pa, const V2<T> &na, const MassCell<T> &N,
                                          const MassCell<T> &P) {
    const T facn = s.w;
    const T facp = 1.0 - facn;
    const v3<T> dudxac = lerp3(N.du, facn, P.du, facp);
    const v3<T> dvdxac = lerp3(N.dv, facn, P.dv, facp);
    const v3<T> dwdxac = lerp3(N.dw, facn, P.dw, facp);
    const T denf = N.den * facn + P.den * facp;
    const v3<T> cin = xyz(N.g), cip = xyz(P.g);
    const v3<T> xac = lerp3(cin, facn, cip, facp);
    const v3<T> delta = sub3(xyz(xf), xac);
    const T uFace = N.u * facn + P.u * facp + dot3(dudxac, delta);
    const T vFace = N.v * facn + P.v * facp + dot3(dvdxac, delta);
    const T wFace = N.w * facn + P.w * facp + dot3(dwdxac, delta);
    const T mf = denf * (uFace * s.x + vFace * s.y + wFace * s.z);
    const v3<T> xnorm = normalise3(v3<T>{s.x, s.y, s.z});
    const v3<T> xpac = xyz(pa);
    const v3<T> xnac{pa.w, na.x, na.y};
    const v3<T> gp_ip = P.dp;
    const T pip = P.p + dot3(gp_ip, sub3(xpac, cip));
    const T pin = N.p + dot3(gp_ip, sub3(xpac, cin));  // dpdx[ip] (MassKernels.ipp:169)
    const v3<T> xpn = sub3(xnac, xpac);
    const v3<T> xpn2 = sub3(cin, cip);
    const T apv1 = P.den * P.ar;
    const T apv2 = N.den * N.ar;
    T apv = apv2 * facn + apv1 * facp;
    const T factv = N.g.w * facn + P.g.w * facp;
    apv *= xf.w * factv / dot3(xpn2, xnorm);
    const v3<T> gp_in = N.dp;
    const T dpx = (gp_in.x * facn + gp_ip.x * facp) * xpn.x;
    const T dpy = (gp_in.y * facn + gp_ip.y * facp) * xpn.y;
    const T dpz = (gp_in.z * facn + gp_ip.z * facp) * xpn.z;
    const T fact = apv;
    st_stream(a.rface + 2 * (size_t)f, -fact);
    st_stream(a.rface + 2 * (size_t)f + 1, -fact);
    st_stream(a.massFlux + f, mf - fact * ((pin - pip) - dpx - dpy - dpz));
}

template <class T> __device__ void mass_interior(const DevMesh<T> &m, const MassArgs<T> &a, int f) {
    const int2 cn = m.fconn[f];
    const V4<T> s = m.fgeoS[f];
    const MassCell<T> N = mass_cell(m, a, cn.y), P = mass_cell(m, a, cn.x);
    mass_face(a, f, s, m.fgeoX[f], m.fgeoP[f], m.fgeoN[f], N, P);
}

// boundary faces of one owning cell, in face order (MassKernels.ipp:197-305)
template <class T> __device__ void mass_boundary_cell(const DevMesh<T> &m, const MassArgs<T> &a, int c) {
    int e0, n;
    row_range(m, c, e0, n);
    bool touched = false;
    T su = 0;
    for (int e = e0; e < e0 + n; e++) {
        const int2 en = m.ent[e];
        const int f = en.x >> 1;
        const int2 cn = entry_cells(c, en);
        if (cn.y >= 0) continue;
        const int ib = ~cn.y;
        const int ir = m.bregion[ib];
        const int it = m.rtype[ir];
        const int ip = cn.x;
        const V4<T> s = m.fgeoS[f];
        const v3<T> S{s.x, s.y, s.z};
        if (it == CUPCFD_RTYPE_INLET) {
            const v3<T> uIn = ld3(m.ruvw, ir);
            const T dens = m.rden[ir];
            const T mf = dens * dot3(uIn, S);
            a.massFlux[f] = mf;
            if (!touched) {
                su = a.su[ip];
                touched = true;
            }
            su = su - mf;
        } else if (it == CUPCFD_RTYPE_OUTLET) {
            const T denf = a.den[ip];
            a.denB[ib] = denf;
            T mf = denf * (a.u[ip] * s.x + a.v[ip] * s.y + a.w[ip] * s.z);
            if (mf < 0.0) {
                mf = a.small;
                if (a.solveTurbEnergy) a.teB[ib] = a.teCell[ip];
                if (a.solveTurbDiss) a.edB[ib] = a.edCell[ip];
                if (a.solveVisc) a.viseffB[ib] = a.viseffCell[ip];
                if (a.solveEnthalpy) a.tB[ib] = a.tCell[ip];
            }
            a.massFlux[f] = mf;
        } else if (it == CUPCFD_RTYPE_SYMP || it == CUPCFD_RTYPE_WALL) {
            a.massFlux[f] = 0.0;
        }
    }
    if (touched) a.su[c] = su;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_mass(DevMesh<T> m, MassArgs<T> a, int per, int nblk_interior) {
    // blocks [0, 8 per): interior faces, XCD-aware; the rest: boundary cells
    const int b = blockIdx.x;
    if (b < kXcd * per) {
        const int lb = xcd_block(b, per);
        if (lb >= nblk_interior) return;
        const int f = lb * kBlock + threadIdx.x;
        if (f < m.Fi) mass_interior(m, a, f);
    } else {
        const int k = (b - kXcd * per) * kBlock + threadIdx.x;
        if (k < m.nbc) mass_boundary_cell(m, a, m.bcells[k]);
    }
}

// a2 on cell tiles: a workgroup stages the values of its tile's cells in LDS
// with coalesced loads and evaluates the interior faces the tile owns (one
// contiguous face range, every owner in the tile); a face's cell2 outside
// the tile is gathered into registers. mass_face on the same values as
// k_mass, so the same bits, with a fraction of k_mass's scattered gathers.
// Blocks from kXcd * per on: the boundary cells, as in k_mass.
template <class T>
__global__ __launch_bounds__(kTileBlock) void k_mass_tile(DevMesh<T> m, MassArgs<T> a, int per, int nblk) {
    constexpr int TC = kTileCells;
    // rows: dudx, dvdx, dwdx, dpdx (x, y, z each), den, u, v, w, p, ar, centre x, y, z, volume
    __shared__ T lc[22 * TC];
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    if (b >= kXcd * per) {
        const int k = (b - kXcd * per) * kTileBlock + tid;
        if (k < m.nbc) mass_boundary_cell(m, a, m.bcells[k]);
        return;
    }
    const int t = xcd_block(b, per);
    if (t >= nblk) return;
    const int c0 = t * TC, nc = min(TC, m.C - c0);
    const int4 ti = m.tinfo[t];
    // First round: every load that depends on no other load, at clamped
    // indices so that all of them issue before the first wait. Thread tid
    // stages element tid of the four packed gradients, den/u/v/w (tid / TC)
    // and p/ar of cell tid % TC, element tid of the centres, and loads its
    // first face's connectivity and geometry.
    const int ci = min(tid & (TC - 1), nc - 1);
    const size_t g0 = 3 * (size_t)c0 + min(tid, 3 * nc - 1);
    const T gu = a.dudx[g0], gv = a.dvdx[g0], gw = a.dwdx[g0], gp = a.dpdx[g0];
    const T x1 = (tid < TC ? a.den : tid < 2 * TC ? a.u : tid < 3 * TC ? a.v : a.w)[c0 + ci];
    const T x2 = (tid < TC ? a.p : a.ar)[c0 + ci];
    const T xc = ((const T *)(m.cgeo + c0))[min(tid, 4 * nc - 1)];
    const bool live = tid < ti.y;
    const int f0 = ti.y > 0 ? ti.x + min(tid, ti.y - 1) : 0;
    const int2 cn = m.fconn[f0];
    const V4<T> s = m.fgeoS[f0], xf = m.fgeoX[f0], pa = m.fgeoP[f0];
    const V2<T> na = m.fgeoN[f0];
    if (tid < 3 * nc) {
        const int r = tid % 3, c = tid / 3;
        lc[r * TC + c] = gu;
        lc[(3 + r) * TC + c] = gv;
        lc[(6 + r) * TC + c] = gw;
        lc[(9 + r) * TC + c] = gp;
    }
    if ((tid & (TC - 1)) < nc) {
        lc[(12 + tid / TC) * TC + (tid & (TC - 1))] = x1;
        if (tid < 2 * TC) lc[(16 + tid / TC) * TC + (tid & (TC - 1))] = x2;
    }
    if (tid < 4 * nc) lc[(18 + (tid & 3)) * TC + (tid >> 2)] = xc;
    // second round: the face's cell2 when it lies outside the tile
    const bool outside = live && !in_tile(cn.y, c0, nc);
    MassCell<T> o{};
    if (outside) o = mass_cell(m, a, cn.y);
    __syncthreads();
    auto cell = [&](int c) -> MassCell<T> {
        const T *q = lc + (c - c0);
        return MassCell<T>{v3<T>{q[0], q[TC], q[2 * TC]},
                           v3<T>{q[3 * TC], q[4 * TC], q[5 * TC]},
                           v3<T>{q[6 * TC], q[7 * TC], q[8 * TC]},
                           v3<T>{q[9 * TC], q[10 * TC], q[11 * TC]},
                           q[12 * TC],
                           q[13 * TC],
                           q[14 * TC],
                           q[15 * TC],
                           q[16 * TC],
                           q[17 * TC],
                           V4<T>{q[18 * TC], q[19 * TC], q[20 * TC], q[21 * TC]}};
    };
    if (live) mass_face(a, f0, s, xf, pa, na, outside ? o : cell(cn.y), cell(cn.x));
    // tiles that own more faces than the workgroup has threads
    for (int j = tid + kTileBlock; j < ti.y; j += kTileBlock) {
        const int f = ti.x + j;
        const int2 cj = m.fconn[f];
        const MassCell<T> N = in_tile(cj.y, c0, nc) ? cell(cj.y) : mass_cell(m, a, cj.y);
        mass_face(a, f, m.fgeoS[f], m.fgeoX[f], m.fgeoP[f], m.fgeoN[f], N, cell(cj.x));
    }
}

// a3 FluxMassDolfynBoundaryLoop1 (MassKernels.ipp:311-336)
template <class T>
__global__ __launch_bounds__(kSeqBlock) void k_mass_bl1(DevMesh<T> m, const T *__restrict__ mf, T *flowin) {
    __shared__ T buf[2 * kSeqLds];
    const int *lst = m.blist[CUPCFD_RTYPE_INLET];
    const T s = block_sum<T, 1>(m.tree_sums, T(0.0), m.nblist[CUPCFD_RTYPE_INLET],
                                [&](int i, T *v) { v[0] = mf[m.bface[lst[i]]]; }, buf);
    if (threadIdx.x == 0) *flowin = s;
}

// a4 FluxMassDolfynBoundaryLoop2 (MassKernels.ipp:338-368): block 0 sums
// flowout over the OUTLET boundaries; block 1 + r adds region r's OUTLET
// boundaries to flowRegion[r], in ascending boundary id
template <class T>
__global__ __launch_bounds__(kSeqBlock) void k_mass_bl2(DevMesh<T> m, const T *__restrict__ mf, T *flowRegion,
                                                        T *flowout, const int *__restrict__ rbnd,
                                                        const int *__restrict__ roff) {
    __shared__ T buf[2 * kSeqLds];
    if (blockIdx.x == 0) {
        const int *lst = m.blist[CUPCFD_RTYPE_OUTLET];
        const T s = block_sum<T, 1>(m.tree_sums, T(0.0), m.nblist[CUPCFD_RTYPE_OUTLET],
                                    [&](int i, T *v) { v[0] = mf[m.bface[lst[i]]]; }, buf);
        if (threadIdx.x == 0) *flowout = s;
        return;
    }
    const int r = blockIdx.x - 1;
    if (m.rtype[r] != CUPCFD_RTYPE_OUTLET) return;
    const int b0 = roff[r];
    const T s = block_sum<T, 1>(m.tree_sums, flowRegion[r], roff[r + 1] - b0,
                                [&](int i, T *v) { v[0] = mf[m.bface[rbnd[b0 + i]]]; }, buf);
    if (threadIdx.x == 0) flowRegion[r] = s;
}

// a5 FluxMassDolfynBoundaryLoop3 (MassKernels.ipp:371-393)
template <class T> __global__ __launch_bounds__(kSeqBlock) void k_mass_bl3(DevMesh<T> m, T flowin, T *ratearea) {
    __shared__ T buf[2 * kSeqLds];
    const int *lst = m.blist[CUPCFD_RTYPE_OUTLET];
    const T areaout = block_sum<T, 1>(m.tree_sums, T(0.0), m.nblist[CUPCFD_RTYPE_OUTLET],
                                      [&](int i, T *v) { v[0] = m.fgeoX[m.bface[lst[i]]].w; }, buf);
    if (threadIdx.x == 0) *ratearea = -flowin / areaout;
}

// a6 FluxMassDolfynBoundaryLoop4 (MassKernels.ipp:395-454): per-boundary part
template <class T>
__global__ __launch_bounds__(kBlock) void k_mass_bl4_faces(DevMesh<T> m, T *massFlux, T *uB, T *vB, T *wB,
                                                           const T *denB, T ratearea) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m.nblist[CUPCFD_RTYPE_OUTLET]) return;
    const int ib = m.blist[CUPCFD_RTYPE_OUTLET][k];
    const int ir = m.bregion[ib];
    const int i = m.bface[ib];
    const T split = m.rsplvl[ir];
    const T area = m.fgeoX[i].w;
    const T mf = ratearea * area * split;
    massFlux[i] = mf;
    const T faceFlux = mf / denB[ib] / area;
    const V4<T> s = m.fgeoS[i];
    const v3<T> xnorm = normalise3(v3<T>{s.x, s.y, s.z});
    uB[ib] = faceFlux * xnorm.x;
    vB[ib] = faceFlux * xnorm.y;
    wB[ib] = faceFlux * xnorm.z;
}

// sum of massFlux over OUTLET boundaries in boundary order, starting from
// *acc_in (0 for loop 4, the caller's flowout2 for loop 5)
template <class T>
__global__ __launch_bounds__(kSeqBlock) void k_outlet_sum(DevMesh<T> m, const T *__restrict__ mf, T *acc,
                                                          int from_zero) {
    __shared__ T buf[2 * kSeqLds];
    const int *lst = m.blist[CUPCFD_RTYPE_OUTLET];
    const T a0 = from_zero ? T(0.0) : *acc;
    const T s = block_sum<T, 1>(m.tree_sums, a0, m.nblist[CUPCFD_RTYPE_OUTLET],
                                [&](int i, T *v) { v[0] = mf[m.bface[lst[i]]]; }, buf);
    if (threadIdx.x == 0) *acc = s;
}

// a7 FluxMassDolfynBoundaryLoop5 (MassKernels.ipp:457-521): per OUTLET-owning
// cell, its OUTLET boundaries in ascending id
template <class T>
__global__ __launch_bounds__(kBlock) void k_mass_bl5_cells(DevMesh<T> m, T *massFlux, T *su, T *uB, T *vB, T *wB,
                                                           T fact, int solveU, int solveV, int solveW,
                                                           const T *flowFact) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m.noc) return;
    const int c = m.ocells[k];
    T s = su[c];
    for (int e = m.orow[k]; e < m.orow[k + 1]; e++) {
        const int ib = m.oent[e];
        const int ir = m.bregion[ib];
        const int i = m.bface[ib];
        const T mf = massFlux[i] * flowFact[ir];
        massFlux[i] = mf;
        if (solveU) uB[ib] = uB[ib] * fact;
        if (solveV) vB[ib] = vB[ib] * fact;
        if (solveW) wB[ib] = wB[ib] * fact;
        s = s - mf;
    }
    su[c] = s;
}

// a8 FluxMassDolfynRegionLoop (MassKernels.ipp:523-554)
template <class T> __global__ void k_mass_region(DevMesh<T> m, T *flowFact, const T *flowRegion, T flowIn) {
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= m.R) return;
    if (m.rtype[ir] == CUPCFD_RTYPE_OUTLET) {
        const T split = m.rsplvl[ir];
        flowFact[ir] = -(split * flowIn) / flowRegion[ir];
    } else {
        flowFact[ir] = 0.0;
    }
}

// ===========================================================================
// a9  FluxScalarDolfynFaceLoop — ScalarKernels.ipp:36-463
// ===========================================================================
// TS: storage type of the per-scalar arrays (phi, dPhidx, PhiB, Au, Su);
// everything else, and all arithmetic, is T
template <class T, class TS = T> struct ScalarArgs {
    const TS *phi;
    const T *visEff, *den, *massFlux, *te, *cpB, *visEffB;
    const TS *dPhidx;
    TS *phiB, *au, *su;
    T *rface;
    const int *src;  // PhiFace source per face (typed meshes), else null
    const T *cval;   // PhiFace each source face leaves (k_scalar_carry_vals), for the tile schedule; else null
    // boundary face g's terms (Su x3, Au) and whether it updates Au, at
    // bterm[4 (g - Fi)] and bflag[g - Fi] (k_scalar_bnd_terms), for the tile schedule
    const T *bterm;
    const uint8_t *bflag;
    // stride of the per-scalar arrays (phi, PhiB, Au, Su, dPhidx): 1 for one
    // scalar, K for scalar-innermost batches; write_state: this lane writes
    // the mesh's boundary Q/H/T
    int ks, write_state;
    int solveTurb, solveEnthalpy, ivar, varT, varTE, varED, inletAssigns, wallActive;
    T visLam, sigmaT, prandtl, sigmaK, sigmaE, sigmaS, schmidt, gammaBlend, small;
    T slres, cmu25;  // wall-function constants, evaluated on the host
    // the face diffusivity's divisors for this call's SolveTurb and ivar
    // (ScalarKernels.ipp:143-167): Visac = visLam + (Visac - visLam) / visD1
    // when SolveTurb, then Visac / visD2 when visHasD2
    T visD1, visD2;
    int visHasD2;
};

template <class T, class TS> __device__ __forceinline__ T scalar_visac(const ScalarArgs<T, TS> &a, T Visac) {
    if (a.solveTurb) {
        Visac = Visac - a.visLam;
        Visac = a.visLam + Visac / a.visD1;
    }
    if (a.visHasD2) Visac = Visac / a.visD2;
    return Visac;
}

// Enthalpy wall branch on a non-adiabatic region (ScalarKernels.ipp:370-445):
// returns PhiFace and the heat-transfer coefficient Hcoef and VisFace.
template <class T, class TS>
__device__ T scalar_wall_phiface(const DevMesh<T> &m, const ScalarArgs<T, TS> &a, int f, int ib, int ir, int ip,
                                 T &Hcoef, T &VisFace) {
    T PhiFlux = 0, PhiFace = 0;
    if (m.rflux[ir]) PhiFlux = m.rt[ir];
    else PhiFace = m.rt[ir];
    const T dn = m.bdist[ib];
    const T Resist = m.rr[ir];
    const T area = m.fgeoX[f].w;
    const T cp = a.cpB[ib];
    if (!a.solveTurb) {
        VisFace = a.visLam / a.prandtl / dn;
        Hcoef = 1.0 / (1.0 / VisFace + Resist * cp) * area;
    } else {
        const T Tplus = a.sigmaT * (m.buplus[ib] + a.slres);
        const T utau = a.cmu25 * sqrt((double)a.te[ip]);
        if (m.byplus[ib] < m.rylog[ir]) {
            VisFace = a.visLam / a.prandtl / dn;
            Hcoef = 1.0 / (1.0 / VisFace + Resist * cp) * area;
        } else {
            VisFace = a.den[ip] * utau / (Tplus + a.small);
            Hcoef = 1.0 / (1.0 / VisFace + Resist * cp) * area;
        }
    }
    if (m.rflux[ir]) PhiFace = (T)a.phi[(size_t)ip * a.ks] + PhiFlux / (Hcoef * cp / area);
    return PhiFace;
}

// PhiFace as boundary face g leaves it for the faces after it
template <class T, class TS>
__device__ T scalar_carry_from(const DevMesh<T> &m, const ScalarArgs<T, TS> &a, int g) {
    const int ib = ~m.fconn[g].y;
    const int ip = m.fconn[g].x;
    const int ir = m.bregion[ib];
    const int it = m.rtype[ir];
    if (it == CUPCFD_RTYPE_INLET) return (T)(0);
    if (it == CUPCFD_RTYPE_OUTLET) {
        const v3<T> Xpn = sub3(xyz(m.fgeoX[g]), xyz(m.cgeo[ip]));
        return (T)a.phi[(size_t)ip * a.ks] + dot3(ld3s_as<T>(a.dPhidx, ip, a.ks), Xpn);
    }
    T H, V;
    return scalar_wall_phiface(m, a, g, ib, ir, ip, H, V);
}

// PhiFace as carried into face f: the value the last assigning face left
template <class T, class TS> __device__ T scalar_carry(const DevMesh<T> &m, const ScalarArgs<T, TS> &a, int f) {
    if (!a.src) return m.carry_init;
    const int g = a.src[f];
    return g < 0 ? m.carry_init : scalar_carry_from(m, a, g);
}

// The value each boundary face that can carry PhiFace leaves (INLET when it
// assigns, OUTLET, WALL when it assigns and its region is not adiabatic; the
// faces a.src names), for the tile schedule to read with one load per face:
// cval[g - Fi] for boundary face g.
template <class T> __global__ __launch_bounds__(kBlock) void k_scalar_carry_vals(DevMesh<T> m, ScalarArgs<T> a,
                                                                               T *cval) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m.F - m.Fi) return;
    const int g = m.Fi + k;
    const int ib = ~m.fconn[g].y;
    const int ir = m.bregion[ib];
    const int it = m.rtype[ir];
    const bool source = it == CUPCFD_RTYPE_OUTLET || (it == CUPCFD_RTYPE_INLET && a.inletAssigns) ||
                        (it == CUPCFD_RTYPE_WALL && a.wallActive && !m.radiab[ir]);
    cval[k] = source ? scalar_carry_from(m, a, g) : m.carry_init;
}

// the cell values an interior face of the scalar loop reads
template <class T> struct ScalarCell {
    T phi, vis;
    v3<T> g, cc;
};
template <class T>
__device__ __forceinline__ ScalarCell<T> scalar_cell(const DevMesh<T> &m, const ScalarArgs<T> &a, int c) {
    return ScalarCell<T>{a.phi[c], a.visEff[c], ld3(a.dPhidx, c), xyz(m.cgeo[c])};
}

// Interior face f of the scalar loop (ScalarKernels.ipp:135-217): terms of
// the cell1 update Su[ip] = ((Su[ip] - blend) + fde1) - fdi; the cell2 update
// is the same three terms negated. Writes RFace when `owner`.
template <class T, class TS>
__device__ __forceinline__ void scalar_interior(const DevMesh<T> &m, const ScalarArgs<T, TS> &a, int f, const V4<T> &s,
                                                T mf, T rl,
                                                const ScalarCell<T> &P, const ScalarCell<T> &N, bool owner,
                                                T PhiFace, T t[3]) {
    const v3<T> S{s.x, s.y, s.z};
    const T facn = s.w;
    const T facp = 1.0 - facn;
    T Visac = N.vis * facn + P.vis * facp;
    Visac = scalar_visac(a, Visac);
    const v3<T> g = lerp3(N.g, facn, P.g, facp);
    const v3<T> Xpn = sub3(N.cc, P.cc);
    const T VisFace = Visac * rl;
    const T fce = mf * PhiFace;
    const T fde1 = Visac * dot3(g, S);
    const T fci = fmin((double)mf, 0.0) * (double)N.phi + fmax((double)mf, 0.0) * (double)P.phi;
    const T fdi = VisFace * dot3(g, Xpn);
    const T blend = a.gammaBlend * (fce - fci);
    t[0] = -blend;
    t[1] = fde1;
    t[2] = -fdi;
    if (owner) {
        st_stream(a.rface + 2 * (size_t)f, (T)(-VisFace - fmax((double)mf, 0.0)));
        st_stream(a.rface + 2 * (size_t)f + 1, (T)(-VisFace + fmin((double)mf, 0.0)));
    }
}

// Boundary face f of cell P = cell1 (ScalarKernels.ipp:219-447): t[0..2] the
// additions to Su, t[3] to Au; returns whether Au is updated (INLET, enthalpy
// WALL). Writes PhiB and the mesh's Q/H/T.
template <class T, class TS>
__device__ __forceinline__ bool scalar_boundary(const DevMesh<T> &m, const ScalarArgs<T, TS> &a, int f, const V4<T> &s,
                                                int ip, const ScalarCell<T> &P, int ib, T t[4]) {
    const v3<T> S{s.x, s.y, s.z};
    t[0] = t[1] = t[2] = t[3] = nzero<T>();
    const int ir = m.bregion[ib];
    const int it = m.rtype[ir];
    if (it == CUPCFD_RTYPE_INLET) {
        const T mf = a.massFlux[f];
        const T PhiFace = a.inletAssigns ? (T)(0) : scalar_carry(m, a, f);
        const T Visac = scalar_visac(a, a.visEffB[ib]);
        const v3<T> Xpn = sub3(xyz(m.fgeoX[f]), P.cc);
        const T VisFace = Visac * m.frl[f];
        const T fde = Visac * dot3(P.g, S);
        const T fdi = VisFace * dot3(P.g, Xpn);
        const T fF = -VisFace + fmin((double)mf, 0.0);
        // Au = Au - f;  Su = ((Su + fde) - fdi) - f * PhiFace
        t[0] = fde;
        t[1] = -fdi;
        t[2] = -(fF * PhiFace);
        t[3] = -fF;
        a.phiB[(size_t)ib * a.ks] = (TS)(PhiFace);
        return true;
    }
    if (it == CUPCFD_RTYPE_OUTLET) {
        const T Visac = scalar_visac(a, P.vis);
        const v3<T> Xpn = sub3(xyz(m.fgeoX[f]), P.cc);
        const T PhiFace = P.phi + dot3(P.g, Xpn);
        const T VisFace = Visac * m.frl[f];
        // Su = (Su + fde) - fdi
        t[0] = Visac * dot3(P.g, S);
        t[1] = -(VisFace * dot3(P.g, Xpn));
        a.phiB[(size_t)ib * a.ks] = (TS)(PhiFace);
        return false;
    }
    if (it == CUPCFD_RTYPE_SYMP) {
        const v3<T> ds = sub3(xyz(m.fgeoX[f]), P.cc);
        a.phiB[(size_t)ib * a.ks] = (TS)(P.phi + dot3(P.g, ds));
        return false;
    }
    if (it == CUPCFD_RTYPE_WALL && a.wallActive) {
        if (m.radiab[ir]) {
            a.phiB[(size_t)ib * a.ks] = (TS)(P.phi);
            if (a.write_state) m.bq[ib] = 0.0;
            return false;
        }
        T Hcoef, VisFace;
        const T PhiFace = scalar_wall_phiface(m, a, f, ib, ir, ip, Hcoef, VisFace);
        const T area = m.fgeoX[f].w;
        const T cp = a.cpB[ib];
        a.phiB[(size_t)ib * a.ks] = (TS)(PhiFace);
        // Au = Au + Hcoef;  Su = Su + Hcoef * PhiFace
        t[0] = Hcoef * PhiFace;
        t[3] = Hcoef;
        const T Tdif = (PhiFace - P.phi);
        const T h1 = Hcoef * cp / area;
        if (a.write_state) {
            m.bq[ib] = h1 * Tdif;
            m.bh[ib] = VisFace * cp;
            m.bt[ib] = PhiFace;
        }
        return true;
    }
    return false;
}

// Terms of CSR entry en of cell c (serial and lane-per-face schedules)
template <class T>
__device__ __forceinline__ bool scalar_terms(const DevMesh<T> &m, const ScalarArgs<T> &a, int c, int2 en, T t[4]) {
    const int f = en.x >> 1;
    const int2 cn = entry_cells(c, en);
    const V4<T> s = m.fgeoS[f];
    if (cn.y >= 0) {
        const bool in = (en.x & 1) != 0;
        scalar_interior(m, a, f, s, a.massFlux[f], m.frl[f], scalar_cell(m, a, cn.x), scalar_cell(m, a, cn.y), !in,
                        m.typed ? scalar_carry(m, a, f) : m.carry_init, t);
        if (in) {
            t[0] = -t[0];
            t[1] = -t[1];
            t[2] = -t[2];
        }
        t[3] = nzero<T>();
        return false;
    }
    return scalar_boundary(m, a, f, s, cn.x, scalar_cell(m, a, cn.x), ~cn.y, t);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scalar(DevMesh<T> m, ScalarArgs<T> a, int per, int nblk) {
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int c = lb * kBlock + threadIdx.x;
    if (c >= m.C) return;
    T su = a.su[c];
    T au = 0;
    bool au_loaded = false;
    int e0, n;
    row_range(m, c, e0, n);
    for (int e = e0; e < e0 + n; e++) {
        T t[4];
        const bool ta = scalar_terms(m, a, c, m.ent[e], t);
        su = ((su + t[0]) + t[1]) + t[2];
        if (ta) {
            if (!au_loaded) {
                au = a.au[c];
                au_loaded = true;
            }
            au = au + t[3];
        }
    }
    a.su[c] = su;
    if (au_loaded) a.au[c] = au;
}

template <class T, int LPC>
__global__ __launch_bounds__(LPC * 32) void k_scalar_vec(DevMesh<T> m, ScalarArgs<T> a, int per, int nblk) {
    constexpr int CPB = kCellsPerBlock, NT = 4, P = LPC + 1;
    __shared__ T lds[NT * CPB * P];
    __shared__ uint8_t flag[CPB * P];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cl = threadIdx.x / LPC, j = threadIdx.x % LPC;
    const int c = lb * CPB + cl;
    int e0 = 0, n = 0;
    if (c < m.C) row_range(m, c, e0, n);
    if (j < n) {
        T t[NT];
        flag[cl * P + j] = scalar_terms(m, a, c, m.ent[e0 + j], t) ? 1 : 0;
        for (int k = 0; k < NT; k++) lds[(k * CPB + cl) * P + j] = t[k];
    }
    __syncthreads();
    if (c >= m.C) return;
    if (j == 0) {
        T su = a.su[c];
        for (int i = 0; i < n; i++)
            su = ((su + lds[(0 * CPB + cl) * P + i]) + lds[(1 * CPB + cl) * P + i]) + lds[(2 * CPB + cl) * P + i];
        a.su[c] = su;
    } else if (j == 1) {
        bool any = false;
        for (int i = 0; i < n; i++) any |= flag[cl * P + i] != 0;
        if (any) {
            T au = a.au[c];
            for (int i = 0; i < n; i++)
                if (flag[cl * P + i]) au = au + lds[(3 * CPB + cl) * P + i];
            a.au[c] = au;
        }
    }
}

// The boundary faces of the scalar loop evaluated ahead of the tiles (they
// depend on no other face's result): boundary face g = Fi + k leaves its terms
// in bterm[4k .. 4k+3] and its Au flag in bflag[k], and writes PhiB and the
// mesh's Q/H/T as the face loop does. This keeps the wall-function code out of
// the tile kernel, whose registers then allow 7 waves per SIMD instead of 6.
// Meshes with no typed region need no such pass: there every boundary face
// adds -0.0 and writes nothing.
template <class T>
__global__ __launch_bounds__(kBlock) void k_scalar_bnd_terms(DevMesh<T> m, ScalarArgs<T> a, T *bterm,
                                                              uint8_t *bflag) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m.F - m.Fi) return;
    const int g = m.Fi + k;
    const int2 cn = m.fconn[g];
    T t[4] = {nzero<T>(), nzero<T>(), nzero<T>(), nzero<T>()};
    bool fl = false;
    if (m.rtype[m.bregion[~cn.y]] != CUPCFD_RTYPE_DEFAULT)  // a DEFAULT face adds -0.0 and writes nothing
        fl = scalar_boundary(m, a, g, m.fgeoS[g], cn.x, scalar_cell(m, a, cn.x), ~cn.y, t);
    *(V4<T> *)(bterm + 4 * (size_t)k) = V4<T>{t[0], t[1], t[2], t[3]};
    bflag[k] = fl ? 1 : 0;
}

template <class T, int WC>
__global__ __launch_bounds__(kTileBlock) void k_scalar_tile(DevMesh<T> m, ScalarArgs<T> a, int per, int nblk) {
    extern __shared__ __align__(16) unsigned char smem[];
    // phi, VisEff, dPhidx xyz, centre xyz: NF rows of W columns, columns
    // [0, TC) the tile's cells, column TC + sl the cell of slot sl's face
    // outside the tile. Phase B overwrites rows 0..3 of column TC + sl with the
    // slot's terms once the thread of slot sl has read that cell; a boundary
    // slot's column gets its terms, evaluated ahead by k_scalar_bnd_terms, in
    // phase A. W is WC when the mesh's tiles fit it (a compile-time row
    // stride), else TC + tslots.
    constexpr int TC = kTileCells, NF = 8;
    const int S = m.tslots, W = WC ? WC : TC + S;
    T *lx = (T *)smem;                         // [NF][W]
    uint16_t *le = (uint16_t *)(lx + NF * W);  // [tnnz]
    uint8_t *lf = (uint8_t *)(le + m.tnnz);    // [S]
    const int t = xcd_block(blockIdx.x, per);
    if (t >= nblk) return;
    const int tid = threadIdx.x;
    const int c0 = t * TC, nc = min(TC, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    // A. First round: every load that depends on no other load, at clamped
    // indices so that all of them are unconditional and issue before the
    // first wait; their LDS writes follow. Thread tid stages phi (tid < TC)
    // or VisEff (TC <= tid < 2 TC) of cell tid % TC, element tid of the packed
    // gradient and of the centres, CSR slot codes tid and tid + kTileBlock,
    // and its slot's connectivity; threads tid < nc also phase C's starting
    // value.
    const int ci = min(tid & (TC - 1), nc - 1);
    const T x1 = (tid < TC ? a.phi : a.visEff)[c0 + ci];
    const T gd = a.dPhidx[3 * (size_t)c0 + min(tid, 3 * nc - 1)];
    const T xc = ((const T *)(m.cgeo + c0))[min(tid, 4 * nc - 1)];
    int eb, ee;
    if (m.rowlen) {
        eb = c0 * m.rowlen;
        ee = (c0 + nc) * m.rowlen;
    } else {
        eb = m.row[c0];
        ee = m.row[c0 + nc];
    }
    const int ne = ee - eb;  // >= 1: a tile whose cells have no faces takes no tile schedule
    const uint16_t q0 = m.eslot[eb + min(tid, ne - 1)], q1 = m.eslot[eb + min(tid + kTileBlock, ne - 1)];
    const T su0 = a.su[c0 + ci];
    const bool live = tid < ts.nslot;
    const int sl0 = min(tid, ts.nslot - 1);
    int2 cn;
    bool owned;
    int f = slot_conn_flat(m, ts, sl0, cn, owned);
    if (tid < 2 * TC && (tid & (TC - 1)) < nc) lx[(tid / TC) * W + (tid & (TC - 1))] = x1;
    if (tid < 3 * nc) lx[(2 + tid % 3) * W + tid / 3] = gd;
    if (tid < 4 * nc && (tid & 3) != 3) lx[(5 + (tid & 3)) * W + (tid >> 2)] = xc;
    if (tid < ne) le[tid] = q0;
    if (tid + kTileBlock < ne) le[tid + kTileBlock] = q1;
    for (int i = tid + 2 * kTileBlock; i < ne; i += kTileBlock) le[i] = m.eslot[eb + i];
    int out = -1;
    V4<T> s{};
    T mf = 0, rl = 0, pf = m.carry_init;
    // second round: slot sl's face values and, for an interior face, its
    // carried PhiFace and its cell outside the tile into column TC + sl; for a
    // boundary face its terms, evaluated ahead, into column TC + sl (cn and f
    // already set)
    auto rest = [&](int sl) {
        out = -1;
        if (cn.y < 0) {
            T *q = lx + TC + sl;
            if (a.bterm) {
                const int k = f - m.Fi;
                const V4<T> bt = *(const V4<T> *)(a.bterm + 4 * (size_t)k);
                q[0] = bt.x, q[W] = bt.y, q[2 * W] = bt.z, q[3 * W] = bt.w;
                lf[sl] = a.bflag[k];
            } else {  // no typed region: every boundary face adds -0.0 and leaves Au alone
                q[0] = q[W] = q[2 * W] = q[3 * W] = nzero<T>();
                lf[sl] = 0;
            }
            return;
        }
        s = m.fgeoS[f];
        mf = a.massFlux[f];
        rl = m.frl[f];
        if (a.cval) {
            const int g = a.src[f];
            pf = g < 0 ? m.carry_init : a.cval[g - m.Fi];
        }
        out = outside_cell(cn, c0, nc);
        if (out >= 0) {
            const ScalarCell<T> o = scalar_cell(m, a, out);
            T *q = lx + TC + sl;
            q[0] = o.phi, q[W] = o.vis;
            q[2 * W] = o.g.x, q[3 * W] = o.g.y, q[4 * W] = o.g.z;
            q[5 * W] = o.cc.x, q[6 * W] = o.cc.y, q[7 * W] = o.cc.z;
        }
    };
    // a later slot sl of this thread (tiles of more than kTileBlock slots)
    auto load = [&](int sl) {
        f = slot_conn(m, ts, sl, cn, owned);
        rest(sl);
    };
    if (live) rest(tid);
    __syncthreads();
    // B
    for (int sl = tid; sl < ts.nslot; sl += kTileBlock) {
        if (sl != tid) load(sl);  // column TC + sl belongs to this thread alone
        auto cell = [&](int c) -> ScalarCell<T> {
            const T *q = lx + (c == out ? TC + sl : c - c0);
            return ScalarCell<T>{q[0], q[W], v3<T>{q[2 * W], q[3 * W], q[4 * W]},
                                 v3<T>{q[5 * W], q[6 * W], q[7 * W]}};
        };
        if (cn.y < 0) continue;  // boundary slots were filled in by load()
        T tt[3];
        scalar_interior(m, a, f, s, mf, rl, cell(cn.x), cell(cn.y), owned, pf, tt);
        T *q = lx + TC + sl;
        for (int j = 0; j < 3; j++) q[j * W] = tt[j];
        q[3 * W] = nzero<T>();
        lf[sl] = 0;
    }
    __syncthreads();
    // C: Su, then Au for cells with an Au-updating face
    if (tid < nc) {
        const int c = c0 + tid;
        T su = su0;
        int e0, n;
        row_range(m, c, e0, n);
        const T *l0 = lx + TC, *l1 = l0 + W, *l2 = l1 + W, *l3 = l2 + W;
        bool any = false;
        for (int e = e0; e < e0 + n; e++) {
            const int es = le[e - eb];
            const int sl = es & 0x7fff;
            T t0 = l0[sl], t1 = l1[sl], t2 = l2[sl];
            if (es & 0x8000) {
                t0 = -t0;
                t1 = -t1;
                t2 = -t2;
            }
            su = ((su + t0) + t1) + t2;
            any |= lf[sl] != 0;
        }
        st_stream(a.su + c, su);
        if (any) {
            T au = a.au[c];
            for (int e = e0; e < e0 + n; e++) {
                const int sl = le[e - eb] & 0x7fff;
                if (lf[sl]) au = au + l3[sl];
            }
            a.au[c] = au;
        }
    }
}

// ===========================================================================
// a10 FluxUVWDolfynFaceLoop1 — UVWKernels.ipp:38-515
// ===========================================================================
template <class T> struct UvwArgs {
    const T *u, *v, *w, *visEff, *visEffB, *dudx, *dvdx, *dwdx;
    T *uB, *vB, *wB, *massFlux, *rface, *su, *sv, *sw, *au, *av, *aw;
    T gammaBlend, small;
    const T *cuvw;  // uFace/vFace/wFace each source face leaves (k_uvw_carry_vals), for the tile schedule; else null
};

// uFace/vFace/wFace as INLET or OUTLET face g leaves them
template <class T> __device__ v3<T> uvw_carry_from(const DevMesh<T> &m, const UvwArgs<T> &a, int g) {
    const int2 cg = m.fconn[g];
    const int ir = m.bregion[~cg.y];
    if (m.rtype[ir] == CUPCFD_RTYPE_INLET) return ld3(m.ruvw, ir);
    return v3<T>{a.u[cg.x], a.v[cg.x], a.w[cg.x]};
}

// uFace/vFace/wFace carried into interior face f (set by INLET and OUTLET faces)
template <class T> __device__ v3<T> uvw_carry(const DevMesh<T> &m, const UvwArgs<T> &a, int f) {
    const T ci = m.carry_init;
    if (!m.typed) return v3<T>{ci, ci, ci};
    const int g = m.carry_src[f];
    if (g < 0) return v3<T>{ci, ci, ci};
    return uvw_carry_from(m, a, g);
}

// uFace/vFace/wFace each INLET and OUTLET face leaves (the faces carry_src
// names), for the tile schedule to read with one load per face: cuvw[3 (g -
// Fi) + k] for boundary face g.
template <class T> __global__ __launch_bounds__(kBlock) void k_uvw_carry_vals(DevMesh<T> m, UvwArgs<T> a, T *cuvw) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m.F - m.Fi) return;
    const int g = m.Fi + k;
    const int it = m.rtype[m.bregion[~m.fconn[g].y]];
    const T ci = m.carry_init;
    const v3<T> uf = (it == CUPCFD_RTYPE_INLET || it == CUPCFD_RTYPE_OUTLET) ? uvw_carry_from(m, a, g)
                                                                               : v3<T>{ci, ci, ci};
    st3(cuvw, k, uf);
}

// visFace carried into SYMP face f (set by interior, INLET and OUTLET faces)
template <class T> __device__ T uvw_visface_carry(const DevMesh<T> &m, const UvwArgs<T> &a, int f) {
    const int g = m.carry_src[f];
    if (g < 0) return m.carry_init;
    const int2 cg = m.fconn[g];
    const T rl = m.frl[g];
    if (cg.y >= 0) {
        const T facn = m.fgeoS[g].w;
        const T facp = 1.0 - facn;
        const T visac = a.visEff[cg.y] * facn + a.visEff[cg.x] * facp;
        return visac * rl;
    }
    const int ib = ~cg.y;
    const int ir = m.bregion[ib];
    const T visac = m.rtype[ir] == CUPCFD_RTYPE_INLET ? a.visEffB[ib] : a.visEff[cg.x];
    return visac * rl;
}

// the cell values an interior face of the UVW loop reads
template <class T> struct UvwCell {
    T u, v, w, vis;
    v3<T> du, dv, dw, cc;
};
template <class T>
__device__ __forceinline__ UvwCell<T> uvw_cell(const DevMesh<T> &m, const UvwArgs<T> &a, int c) {
    return UvwCell<T>{a.u[c], a.v[c], a.w[c], a.visEff[c],
                      ld3(a.dudx, c), ld3(a.dvdx, c), ld3(a.dwdx, c), xyz(m.cgeo[c])};
}

// Interior face f of the UVW loop (UVWKernels.ipp:128-230): t[3k..3k+2] the
// terms of the cell1 update su[ip] = ((su[ip] - blendU) + fude) - fudi (sv,
// sw alike); the cell2 update is the same terms negated. Writes rFace when
// `owner`.
template <class T>
__device__ __forceinline__ void uvw_interior(const DevMesh<T> &m, const UvwArgs<T> &a, int f, const V4<T> &s, T mf,
                                             T rl,
                                             const UvwCell<T> &P, const UvwCell<T> &N, bool owner, const v3<T> &uf,
                                             T t[9]) {
    const T sx = s.x, sy = s.y, sz = s.z;
    const T facn = s.w;
    const T facp = 1.0 - facn;
    const v3<T> du = lerp3(N.du, facn, P.du, facp);
    const v3<T> dv = lerp3(N.dv, facn, P.dv, facp);
    const v3<T> dw = lerp3(N.dw, facn, P.dw, facp);
    const T visac = N.vis * facn + P.vis * facp;
    const v3<T> xpn = sub3(N.cc, P.cc);
    const T visFace = visac * rl;
    const T fuce = mf * uf.x, fvce = mf * uf.y, fwce = mf * uf.z;
    const T fude1 = (du.x + du.x) * sx + (du.y + dv.x) * sy + (du.z + dw.x) * sz;
    const T fvde1 = (du.y + dv.x) * sx + (dv.y + dv.y) * sy + (dv.z + dw.y) * sz;
    const T fwde1 = (du.z + dw.x) * sx + (dw.y + dv.z) * sy + (dw.z + dw.z) * sz;
    const T fude = visac * fude1, fvde = visac * fvde1, fwde = visac * fwde1;
    const T fmn = smin(mf, (T)(0.0)), fmx = smax(mf, (T)(0.0));
    const T fuci = fmn * N.u + fmx * P.u;
    const T fvci = fmn * N.v + fmx * P.v;
    const T fwci = fmn * N.w + fmx * P.w;
    const T fudi = visFace * dot3(du, xpn);
    const T fvdi = visFace * dot3(dv, xpn);
    const T fwdi = visFace * dot3(dw, xpn);
    const T blendU = a.gammaBlend * (fuce - fuci);
    const T blendV = a.gammaBlend * (fvce - fvci);
    const T blendW = a.gammaBlend * (fwce - fwci);
    t[0] = -blendU, t[1] = fude, t[2] = -fudi;
    t[3] = -blendV, t[4] = fvde, t[5] = -fvdi;
    t[6] = -blendW, t[7] = fwde, t[8] = -fwdi;
    if (owner) {
        st_stream(a.rface + 2 * (size_t)f, -visFace - smax(mf, (T)(0.0)));
        st_stream(a.rface + 2 * (size_t)f + 1, -visFace + smin(mf, (T)(0.0)));
    }
}

// Boundary face f of cell P = cell1 (UVWKernels.ipp:232-515): t[0..8] the
// additions to su/sv/sw, t[9] to au, av and aw alike; returns whether the face
// updates au/av/aw. Writes uB/vB/wB, the OUTLET massFlux clamp and the wall
// shear.
template <class T>
__device__ __forceinline__ bool uvw_boundary(const DevMesh<T> &m, const UvwArgs<T> &a, int f, const V4<T> &s,
                                             const UvwCell<T> &P, int ib, T t[10]) {
    const T sx = s.x, sy = s.y, sz = s.z;
    for (int k = 0; k < 10; k++) t[k] = nzero<T>();
    const int ir = m.bregion[ib];
    const int it = m.rtype[ir];
    if (it == CUPCFD_RTYPE_DEFAULT) return false;
    const V4<T> xf = m.fgeoX[f];
    const v3<T> cp = P.cc;
    const T up = P.u, vp = P.v, wp = P.w;
    if (it == CUPCFD_RTYPE_INLET || it == CUPCFD_RTYPE_OUTLET) {
        const v3<T> du = P.du, dv = P.dv, dw = P.dw;
        v3<T> uf;
        T visac;
        if (it == CUPCFD_RTYPE_INLET) {
            uf = ld3(m.ruvw, ir);
            visac = a.visEffB[ib];
        } else {
            uf = v3<T>{up, vp, wp};
            visac = P.vis;
        }
        const v3<T> xpn = sub3(xyz(xf), cp);
        const T visFace = visac * m.frl[f];
        T fude = (du.x + du.x) * sx + (du.y + dv.x) * sy + (du.z + dw.x) * sz;
        T fvde = (du.y + dv.x) * sx + (dv.y + dv.y) * sy + (dv.z + dw.y) * sz;
        T fwde = (du.z + dw.x) * sx + (dw.y + dv.z) * sy + (dw.z + dw.z) * sz;
        fude = visac * fude;
        fvde = visac * fvde;
        fwde = visac * fwde;
        const T fudi = visFace * dot3(du, xpn);
        const T fvdi = visFace * dot3(dv, xpn);
        const T fwdi = visFace * dot3(dw, xpn);
        T mf = a.massFlux[f];
        if (it == CUPCFD_RTYPE_OUTLET && mf < 0.0) {
            mf = a.small;
            a.massFlux[f] = mf;
        }
        const T fF = -visFace + smin(mf, (T)(0.0));
        // au = au - f;  su = ((su - f * uFace) + fude) - fudi
        t[0] = -(fF * uf.x), t[1] = fude, t[2] = -fudi;
        t[3] = -(fF * uf.y), t[4] = fvde, t[5] = -fvdi;
        t[6] = -(fF * uf.z), t[7] = fwde, t[8] = -fwdi;
        t[9] = -fF;
        a.uB[ib] = uf.x;
        a.vB[ib] = uf.y;
        a.wB[ib] = uf.z;
        return true;
    }
    if (it == CUPCFD_RTYPE_SYMP) {
        const v3<T> xpn = sub3(xyz(xf), cp);
        const T visac = P.vis;
        const T du = dot3(P.du, xpn);
        const T dv = dot3(P.dv, xpn);
        const T dw = dot3(P.dw, xpn);
        const v3<T> us{up + du, vp + dv, wp + dw};
        const v3<T> xn = normalise3(v3<T>{(T)(0.0 - sx), (T)(0.0 - sy), (T)(0.0 - sz)});
        const T dp = dot3(us, xn);
        const v3<T> un{dp * xn.x, dp * xn.y, dp * xn.z};
        const T dn = m.bdist[ib];
        const T twoVis = (T)(2.0 * visac);
        const T area = xf.w;
        const v3<T> force{((twoVis * un.x) / dn) * area, ((twoVis * un.y) / dn) * area,
                          ((twoVis * un.z) / dn) * area};
        const T visFace = uvw_visface_carry(m, a, f);
        // au = au + visFace;  su = (su + visFace * u) - force.x
        t[0] = visFace * up, t[1] = -force.x;
        t[3] = visFace * vp, t[4] = -force.y;
        t[6] = visFace * wp, t[7] = -force.z;
        t[9] = visFace;
        a.uB[ib] = us.x;
        a.vB[ib] = us.y;
        a.wB[ib] = us.z;
        return true;
    }
    // WALL
    const v3<T> uw = ld3(m.ruvw, ir);
    const T visac = a.visEffB[ib];
    const v3<T> xpn = sub3(xyz(xf), cp);
    const T xpn_length = (T)length3(xpn);
    const T area = xf.w;
    const T coef = visac * area / xpn_length;
    const v3<T> xn = normalise3(v3<T>{sx, sy, sz});
    const v3<T> upr = sub3(v3<T>{up, vp, wp}, uw);
    const T dp = dot3(upr, xn);
    const v3<T> un{dp * xn.x, dp * xn.y, dp * xn.z};
    const v3<T> ut = sub3(upr, un);
    const T uvel = fabs((double)ut.x) + fabs((double)ut.y) + fabs((double)ut.z);
    v3<T> force{0.0, 0.0, 0.0};
    if (uvel > a.small) {
        const T dn = m.bdist[ib];
        force = v3<T>{((visac * ut.x) / dn) * area, ((visac * ut.y) / dn) * area, ((visac * ut.z) / dn) * area};
    }
    st3(m.bshear, ib, force);
    // au = au + coef;  su = (su + coef * u) - force.x
    t[0] = coef * up, t[1] = -force.x;
    t[3] = coef * vp, t[4] = -force.y;
    t[6] = coef * wp, t[7] = -force.z;
    t[9] = coef;
    a.uB[ib] = uw.x;
    a.vB[ib] = uw.y;
    a.wB[ib] = uw.z;
    return true;
}

// Terms of CSR entry en of cell c (serial and lane-per-face schedules)
template <class T>
__device__ __forceinline__ bool uvw_terms(const DevMesh<T> &m, const UvwArgs<T> &a, int c, int2 en, T t[10]) {
    const int f = en.x >> 1;
    const int2 cn = entry_cells(c, en);
    const V4<T> s = m.fgeoS[f];
    if (cn.y >= 0) {
        const bool in = (en.x & 1) != 0;
        uvw_interior(m, a, f, s, a.massFlux[f], m.frl[f], uvw_cell(m, a, cn.x), uvw_cell(m, a, cn.y), !in,
                     uvw_carry(m, a, f), t);
        if (in)
            for (int k = 0; k < 9; k++) t[k] = -t[k];
        t[9] = nzero<T>();
        return false;
    }
    return uvw_boundary(m, a, f, s, uvw_cell(m, a, cn.x), ~cn.y, t);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_uvw(DevMesh<T> m, UvwArgs<T> a, int per, int nblk) {
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int c = lb * kBlock + threadIdx.x;
    if (c >= m.C) return;
    T s[3] = {a.su[c], a.sv[c], a.sw[c]};
    T aa[3] = {0, 0, 0};
    bool a_loaded = false;
    int e0, n;
    row_range(m, c, e0, n);
    for (int e = e0; e < e0 + n; e++) {
        T t[10];
        const bool ta = uvw_terms(m, a, c, m.ent[e], t);
        for (int k = 0; k < 3; k++) s[k] = ((s[k] + t[3 * k]) + t[3 * k + 1]) + t[3 * k + 2];
        if (ta) {
            if (!a_loaded) {
                aa[0] = a.au[c];
                aa[1] = a.av[c];
                aa[2] = a.aw[c];
                a_loaded = true;
            }
            for (int k = 0; k < 3; k++) aa[k] = aa[k] + t[9];
        }
    }
    a.su[c] = s[0];
    a.sv[c] = s[1];
    a.sw[c] = s[2];
    if (a_loaded) {
        a.au[c] = aa[0];
        a.av[c] = aa[1];
        a.aw[c] = aa[2];
    }
}

template <class T, int LPC>
__global__ __launch_bounds__(LPC * 32) void k_uvw_vec(DevMesh<T> m, UvwArgs<T> a, int per, int nblk) {
    constexpr int CPB = kCellsPerBlock, NT = 10, P = LPC + 1;
    __shared__ T lds[NT * CPB * P];
    __shared__ uint8_t flag[CPB * P];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cl = threadIdx.x / LPC, j = threadIdx.x % LPC;
    const int c = lb * CPB + cl;
    int e0 = 0, n = 0;
    if (c < m.C) row_range(m, c, e0, n);
    if (j < n) {
        T t[NT];
        flag[cl * P + j] = uvw_terms(m, a, c, m.ent[e0 + j], t) ? 1 : 0;
        for (int k = 0; k < NT; k++) lds[(k * CPB + cl) * P + j] = t[k];
    }
    __syncthreads();
    if (c >= m.C) return;
    // lanes 0..2 fold su/sv/sw, lanes 3..5 fold au/av/aw (LPC >= 6), or lanes
    // 0..3 take two components each (LPC == 4)
    for (int k = j; k < 6; k += LPC) {
        if (k < 3) {
            T *dst = k == 0 ? a.su : (k == 1 ? a.sv : a.sw);
            T acc = dst[c];
            for (int i = 0; i < n; i++)
                acc = ((acc + lds[((3 * k) * CPB + cl) * P + i]) + lds[((3 * k + 1) * CPB + cl) * P + i]) +
                      lds[((3 * k + 2) * CPB + cl) * P + i];
            dst[c] = acc;
        } else {
            bool any = false;
            for (int i = 0; i < n; i++) any |= flag[cl * P + i] != 0;
            if (any) {
                T *dst = k == 3 ? a.au : (k == 4 ? a.av : a.aw);
                T acc = dst[c];
                for (int i = 0; i < n; i++)
                    if (flag[cl * P + i]) acc = acc + lds[(9 * CPB + cl) * P + i];
                dst[c] = acc;
            }
        }
    }
}

// The UVW tile. Layout (NA = 0): u v w VisEff, dudx dvdx dwdx centre (xyz
// each) in NF rows of W columns, columns [0, TC) the tile's cells, column
// TC + sl the cell of slot sl's face outside the tile. Phase B overwrites rows
// 0..9 of column TC + sl with the slot's terms once the thread of slot sl has
// read that cell. W is WC when the mesh's tiles fit it (a compile-time row
// stride, so LDS addresses become immediate offsets), else TC + tslots.
// Split layout (NA > 0, the mesh's uvw_col / uvw_eslot): a slot with a cell
// outside the tile or a boundary face takes column TC + col (col < NA) of the
// NF-row table A, which holds that cell and then the slot's ten terms; a face
// between two cells of the tile needs no cell column, and its nine terms go
// to column col - NA of the 9-row table B. On 4x4x4 hex tiles (96 + 144
// slots) that is 31.7 KB of LDS instead of 39.9 KB: five workgroups per CU
// instead of four, with 96 VGPRs (waves_per_eu(5)) for five waves per SIMD.
// The wider variants keep waves_per_eu(4): at 96 VGPRs they would spill to
// scratch.
template <class T, int WC, int NA>
__global__ __launch_bounds__(kTileBlock) __attribute__((amdgpu_waves_per_eu(WC == kTileCells + 240 ? 5 : 4))) void k_uvw_tile(
    DevMesh<T> m, UvwArgs<T> a, int per, int nblk, const uint8_t *__restrict__ ucol,
    const uint16_t *__restrict__ ueslot) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TC = kTileCells, NF = 16;
    static_assert(NA == 0 || (WC != 0 && NA < WC - TC), "split layout on a compile-time width");
    constexpr int WB = NA ? WC - TC - NA : 1;    // table B's row stride
    const int S = m.tslots;
    const int W = NA ? TC + NA : (WC ? WC : TC + S);  // table A's row stride
    T *lx = (T *)smem;                              // A: [NF][W]
    T *lb = lx + NF * W;                            // B: [9][WB]
    uint16_t *le = (uint16_t *)(lb + (NA ? 9 * WB : 0));  // [tnnz]
    uint8_t *lf = (uint8_t *)(le + m.tnnz);         // [NA or S]
    const int t = xcd_block(blockIdx.x, per);
    if (t >= nblk) return;
    const int tid = threadIdx.x;
    const int c0 = t * TC, nc = min(TC, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    // A. First round: every load that depends on no other load, at clamped
    // indices so that all of them are unconditional and issue before the
    // first wait; their LDS writes follow. Thread tid stages field tid / TC of
    // u, v, w, VisEff (cell tid % TC), element tid of the packed gradients and
    // of the centres, CSR slot codes tid and tid + kTileBlock, and its slot's
    // column and connectivity.
    const int ci = min(tid & (TC - 1), nc - 1);
    const T x1 = (tid < TC ? a.u : tid < 2 * TC ? a.v : tid < 3 * TC ? a.w : a.visEff)[c0 + ci];
    const size_t g0 = 3 * (size_t)c0 + min(tid, 3 * nc - 1);
    const T gu = a.dudx[g0], gv = a.dvdx[g0], gw = a.dwdx[g0];
    const T xc = ((const T *)(m.cgeo + c0))[min(tid, 4 * nc - 1)];
    int eb, ee;
    if (m.rowlen) {
        eb = c0 * m.rowlen;
        ee = (c0 + nc) * m.rowlen;
    } else {
        eb = m.row[c0];
        ee = m.row[c0 + nc];
    }
    const uint16_t *esrc = NA ? ueslot : m.eslot;
    const int ne = ee - eb;
    const uint16_t q0 = esrc[eb + min(tid, ne - 1)], q1 = esrc[eb + min(tid + kTileBlock, ne - 1)];
    // phase C's fold item tid < 3 * TC: component k of cell cl
    const int k = tid / TC, cl = tid % TC;
    const bool fold = tid < 3 * TC && cl < nc;
    T *const dst = k == 0 ? a.su : (k == 1 ? a.sv : a.sw);
    const bool live = tid < ts.nslot;
    const int sl0 = min(tid, ts.nslot - 1);
    int col = NA ? (int)ucol[(size_t)t * S + sl0] : sl0;
    int2 cn;
    bool owned;
    int f = slot_conn_flat(m, ts, sl0, cn, owned);
    if ((tid & (TC - 1)) < nc) lx[(tid / TC) * W + (tid & (TC - 1))] = x1;
    if (tid < 3 * nc) {
        const int r = tid % 3, c = tid / 3;
        lx[(4 + r) * W + c] = gu;
        lx[(7 + r) * W + c] = gv;
        lx[(10 + r) * W + c] = gw;
    }
    if (tid < 4 * nc && (tid & 3) != 3) lx[(13 + (tid & 3)) * W + (tid >> 2)] = xc;
    if (tid < ne) le[tid] = q0;
    if (tid + kTileBlock < ne) le[tid + kTileBlock] = q1;
    for (int i = tid + 2 * kTileBlock; i < ne; i += kTileBlock) le[i] = esrc[eb + i];
    int out = -1;
    V4<T> s{};
    T mf = 0, rl = 0;
    v3<T> uf{m.carry_init, m.carry_init, m.carry_init};
    // second round: slot sl's face and, for an interior face, its carried
    // uFace/vFace/wFace and its cell outside the tile into table A's column
    // TC + col (cn, f and col already set)
    auto rest = [&]() {
        s = m.fgeoS[f];
        out = -1;
        if (cn.y >= 0) {
            mf = a.massFlux[f];
            rl = m.frl[f];
            if (a.cuvw) {
                const int g = m.carry_src[f];
                const T ci = m.carry_init;
                uf = g < 0 ? v3<T>{ci, ci, ci} : ld3(a.cuvw, g - m.Fi);
            }
            out = outside_cell(cn, c0, nc);
            if (out >= 0) {
                const UvwCell<T> o = uvw_cell(m, a, out);
                T *q = lx + TC + col;
                q[0] = o.u, q[W] = o.v, q[2 * W] = o.w, q[3 * W] = o.vis;
                q[4 * W] = o.du.x, q[5 * W] = o.du.y, q[6 * W] = o.du.z;
                q[7 * W] = o.dv.x, q[8 * W] = o.dv.y, q[9 * W] = o.dv.z;
                q[10 * W] = o.dw.x, q[11 * W] = o.dw.y, q[12 * W] = o.dw.z;
                q[13 * W] = o.cc.x, q[14 * W] = o.cc.y, q[15 * W] = o.cc.z;
            }
        }
    };
    // a later slot sl of this thread (tiles of more than kTileBlock slots)
    auto load = [&](int sl) {
        col = NA ? (int)ucol[(size_t)t * S + sl] : sl;
        f = slot_conn(m, ts, sl, cn, owned);
        rest();
    };
    if (live) rest();
    __syncthreads();
    // B; with WC - TC <= kTileBlock slots (every slot one thread's) a single pass
    constexpr bool kOnePass = WC != 0 && WC - TC <= kTileBlock;
    for (int sl = tid; sl < ts.nslot; sl += kTileBlock) {
        if (!kOnePass && sl != tid) load(sl);  // the slot's column belongs to this thread alone
        auto cell = [&](int c) -> UvwCell<T> {
            const T *q = lx + (c == out ? TC + col : c - c0);
            return UvwCell<T>{q[0],
                              q[W],
                              q[2 * W],
                              q[3 * W],
                              v3<T>{q[4 * W], q[5 * W], q[6 * W]},
                              v3<T>{q[7 * W], q[8 * W], q[9 * W]},
                              v3<T>{q[10 * W], q[11 * W], q[12 * W]},
                              v3<T>{q[13 * W], q[14 * W], q[15 * W]}};
        };
        T tt[10];
        bool fl = false;
        if (cn.y >= 0) {
            uvw_interior(m, a, f, s, mf, rl, cell(cn.x), cell(cn.y), owned, uf, tt);
            tt[9] = nzero<T>();
        } else {
            fl = uvw_boundary(m, a, f, s, cell(cn.x), ~cn.y, tt);
        }
        if (!NA || col < NA) {
            T *q = lx + TC + col;
            for (int j = 0; j < 10; j++) q[j * W] = tt[j];
            lf[col] = fl ? 1 : 0;
        } else {  // between two cells of the tile: never an a-updating face
            T *q = lb + (col - NA);
            for (int j = 0; j < 9; j++) q[j * WB] = tt[j];
        }
        if (kOnePass) break;
    }
    // the fold's starting value, loaded before the barrier so that the wait
    // overlaps it, and after phase B so that it holds no register there
    const T acc0 = fold ? dst[c0 + cl] : (T)0;
    __syncthreads();
    // C: su/sv/sw, then au/av/aw for cells with an a-updating face
    if (fold) {
        const int c = c0 + cl;
        T acc = acc0;
        int e0, n;
        row_range(m, c, e0, n);
        const T *la = lx + TC + (3 * k) * W;
        const T *lbk = lb + (3 * k) * WB;
        bool any = false;
        for (int e = e0; e < e0 + n; e++) {
            const int es = le[e - eb];
            const int sc = es & 0x7fff;
            T t0, t1, t2;
            if (!NA || sc < NA) {
                t0 = la[sc], t1 = la[sc + W], t2 = la[sc + 2 * W];
                any |= lf[sc] != 0;
            } else {
                t0 = lbk[sc - NA], t1 = lbk[sc - NA + WB], t2 = lbk[sc - NA + 2 * WB];
            }
            if (es & 0x8000) {
                t0 = -t0;
                t1 = -t1;
                t2 = -t2;
            }
            acc = ((acc + t0) + t1) + t2;
        }
        st_stream(dst + c, acc);
        if (any) {
            T *dsta = k == 0 ? a.au : (k == 1 ? a.av : a.aw);
            const T *l9 = lx + TC + 9 * W;
            T aa = dsta[c];
            for (int e = e0; e < e0 + n; e++) {
                const int sc = le[e - eb] & 0x7fff;
                if ((!NA || sc < NA) && lf[sc]) aa = aa + l9[sc];
            }
            dsta[c] = aa;
        }
    }
}

// a11 FluxUVWDolfynRegionLoop1 (UVWKernels.ipp:517-524)
template <class T> __global__ void k_uvw_region1(DevMesh<T> m) {
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir < m.R) st3(m.rforcet, ir, v3<T>{(T)0, (T)0, (T)0});
}

// a12 FluxUVWDolfynBndsLoop1 (UVWKernels.ipp:527-544): WALL shear summed per
// region in boundary order; lane r owns region rbase + r
// block r adds the boundary shear of WALL region r's boundaries, in ascending
// id, to its force tangent; wave c < 3 adds component c
template <class T>
__global__ __launch_bounds__(kSeqBlock) void k_uvw_bnds1(DevMesh<T> m, const int *__restrict__ rbnd,
                                                         const int *__restrict__ roff) {
    __shared__ T buf[2 * kSeqLds];
    const int r = blockIdx.x;
    if (m.rtype[r] != CUPCFD_RTYPE_WALL) return;
    const int b0 = roff[r], c = min((int)threadIdx.x >> 6, 2);
    const T s = block_sum<T, 3>(
        m.tree_sums, m.rforcet[3 * r + c], roff[r + 1] - b0,
        [&](int i, T *v) {
            const v3<T> x = ld3((const T *)m.bshear, rbnd[b0 + i]);
            v[0] = x.x, v[1] = x.y, v[2] = x.z;
        },
        buf);
    if (threadIdx.x < 192 && (threadIdx.x & 63) == 0) m.rforcet[3 * r + c] = s;
}

// ===========================================================================
// a13-a16 viscosity loops — ViscosityKernels.ipp
// ===========================================================================
template <class T>
__global__ __launch_bounds__(kBlock) void k_visc_cell1(int C, T small, T cmu, T visURF, T visLam,
                                                       const T *__restrict__ TE, const T *__restrict__ ED,
                                                       const T *__restrict__ den, T *__restrict__ visEff) {
    const int ip = blockIdx.x * kBlock + threadIdx.x;
    if (ip >= C) return;
    const T visOld = visEff[ip];
    T visNew;
    if (ED[ip] > small) visNew = visLam + cmu * den[ip] * (TE[ip] * TE[ip]) / ED[ip];
    else visNew = visLam;
    visEff[ip] = visOld + visURF * (visNew - visOld);
}

// yplustmp is carried from region to region, as in the reference
template <class T> __global__ void k_visc_region(DevMesh<T> m, T kappa) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    T yplus, yplustmp = m.carry_init, elog;
    for (int ir = 0; ir < m.R; ir++) {
        if (m.rtype[ir] == CUPCFD_RTYPE_WALL && m.rstd[ir]) {
            yplus = 11.0;
            elog = m.relog[ir];
            int i = 0;
            while ((fabs((double)(yplus - yplustmp)) > 0.001) && (i < 50)) {
                yplustmp = yplus;
                yplus = log((double)(elog * yplustmp)) / kappa;
                i = i + 1;
            }
            m.rylog[ir] = yplus;
        } else if (m.rtype[ir] == CUPCFD_RTYPE_WALL && !(m.rstd[ir])) {
            m.rylog[ir] = 0.0;
        }
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_visc_boundary(DevMesh<T> m, T cmu, T cmu25, T small, T kappa,
                                                          T visLam, const T *__restrict__ TE,
                                                          const T *__restrict__ den,
                                                          const T *__restrict__ visEffCell, T *visEffB) {
    const int ib = blockIdx.x * kBlock + threadIdx.x;
    if (ib >= m.B) return;
    const int i = m.bface[ib];
    const int ir = m.bregion[ib];
    const int it = m.rtype[ir];
    const int ip = m.fconn[i].x;
    if (it == CUPCFD_RTYPE_INLET) {
        visEffB[ib] =
            visLam + m.rdensity[ir] * cmu * (m.rturbke[ir] * m.rturbke[ir]) / (m.rturbdiss[ir] + small);
    } else if (it == CUPCFD_RTYPE_OUTLET || it == CUPCFD_RTYPE_SYMP) {
        visEffB[ib] = visEffCell[ip];
    } else if (it == CUPCFD_RTYPE_WALL) {
        const T dist = m.bdist[ib];
        T turb = TE[ip];
        const T zero = 0.0;
        if (turb < zero) turb = zero;
        const T utau = cmu25 * sqrt((double)turb);
        T yplus = dist * utau * den[ip] / (visLam + small);
        m.byplus[ib] = yplus;
        T uplus;
        if (m.rstd[ir]) {
            if (yplus < m.rylog[ir]) {
                uplus = yplus;
                if (yplus < small) {
                    yplus = 0.0;
                    uplus = small;
                }
            } else {
                T tmp = m.relog[ir] * m.byplus[ib];
                if (tmp < 1.1) tmp = 1.1;
                uplus = log((double)tmp) / kappa;
            }
            m.buplus[ib] = uplus;
            visEffB[ib] = smax((T)(1.0), yplus / uplus) * visLam;
        } else {
            const T d0 = 0.0, z0 = 0.1;
            uplus = log((double)((dist - d0) / z0)) / kappa;
            m.buplus[ib] = uplus;
            visEffB[ib] = smax((T)(1.0), yplus / uplus) * visLam;
        }
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_visc_cell2(int C, int B, T visLam, T *visEffCell, T *visEffB) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < C) visEffCell[k] = smin(visEffCell[k], (T)(1000000.0) * visLam);
    else if (k < C + B) visEffB[k - C] = smin(visEffB[k - C], (T)(1000000.0) * visLam);
}

// ===========================================================================
// layout helpers: mesh order <-> internal order
// ===========================================================================
template <class E>
__global__ __launch_bounds__(kBlock) void k_permute(int n, int ncomp, const int *__restrict__ order,
                                                    const E *__restrict__ src, E *__restrict__ dst, int to_internal) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const size_t a = (size_t)k * ncomp, b = (size_t)order[k] * ncomp;
    for (int j = 0; j < ncomp; j++) {
        if (to_internal) dst[a + j] = src[b + j];
        else dst[b + j] = src[a + j];
    }
}

// ===========================================================================
// Batched gathers: K scalars on one mesh (no counterpart in the reference).
// Scalar k's results are those of the plain kernel on scalar k's arrays, bit
// for bit. Per-scalar arrays are scalar-innermost (phi[c*K + k], vectors
// [(c*3 + j)*K + k]), so the G lanes that share a cell read one contiguous
// run per field, and the cell's connectivity and geometry are read once for
// all K. A group of G lanes (G = 4, 16 or 64; K > 64 takes blockIdx.y chunks
// of 64) owns one cell and walks its CSR row in order, each lane folding its
// own scalar exactly as the plain kernels' folds do.
// ===========================================================================
struct BatchLane {
    int c, k;
    bool live;
};
template <class T, int G>
__device__ __forceinline__ BatchLane batch_lane(const DevMesh<T> &m, int K, int per, int nblk) {
    BatchLane r;
    const int lb = xcd_block(blockIdx.x, per);
    r.c = lb * (kBlock / G) + threadIdx.x / G;
    r.k = blockIdx.y * G + threadIdx.x % G;
    r.live = lb < nblk && r.c < m.C && r.k < K;
    return r;
}

// Row of cell c as the batched kernels walk it. RL > 0: rows of exactly RL
// entries (uniform meshes), whose entries, face values and neighbour values
// are all loaded before any term is computed (no store sits between them, so
// every load of the row is in flight at once); RL = 0: any row, loaded entry
// by entry. With G = 64 the cell index is made wave-uniform, so the cell's
// connectivity and geometry are read with scalar loads.
template <int G> __device__ __forceinline__ int batch_uniform(int c) {
    return G == 64 ? __builtin_amdgcn_readfirstlane(c) : c;
}

// Arguments of the batched gradient: per-scalar inputs stored as TI, outputs
// as TO (float storage with double geometry and accumulation is T = double,
// TI = TO = float); the previous sweep's gradient is always T.
template <class T, class TI, class TO> struct BatchGradArgs {
    const TI *phi, *phiB;
    const T *grado;
    TO *out1, *out2;
};

// The fields of a device mesh the batched gradient reads (held in registers
// by the mesh-set kernel, which switches meshes between work units)
// Read-only device memory seen through the constant address space, so that
// wave-uniform loads become scalar loads even where the compiler cannot show
// that a work unit's stores do not alias the next unit's geometry. operator[]
// loads by value (4- and 8-byte words; a V4 as four values).
template <class X> struct RoArray {
    const __attribute__((address_space(4))) X *p;
    __device__ __forceinline__ X operator[](int64_t i) const { return p[i]; }
};
template <> struct RoArray<int2> {
    const __attribute__((address_space(4))) long long *p;
    __device__ __forceinline__ int2 operator[](int64_t i) const {
        const long long v = p[i];
        return make_int2((int)v, (int)(v >> 32));
    }
};
template <class T> struct RoArray<V4<T>> {
    const __attribute__((address_space(4))) T *p;
    __device__ __forceinline__ V4<T> operator[](int64_t i) const {
        return V4<T>{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
    }
};
template <class X, class Q> __device__ __forceinline__ X ro(const Q *q) {
    X a;
    a.p = (decltype(a.p))q;
    return a;
}
template <class T> struct GradGeom {
    int C, rowlen;
    RoArray<int> row;
    RoArray<int2> ent;
    RoArray<V4<T>> fgeoS, fgeoX, cgeo;
    RoArray<T> cvol;
    __device__ __forceinline__ explicit GradGeom(const DevMesh<T> &m)
        : C(m.C), rowlen(m.rowlen), row(ro<RoArray<int>>(m.row)), ent(ro<RoArray<int2>>(m.ent)),
          fgeoS(ro<RoArray<V4<T>>>(m.fgeoS)), fgeoX(ro<RoArray<V4<T>>>(m.fgeoX)), cgeo(ro<RoArray<V4<T>>>(m.cgeo)),
          cvol(ro<RoArray<T>>(m.cvol)) {}
};

// a value every lane of the wave holds, moved to scalar registers
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class P> __device__ __forceinline__ P *uni(P *p) {
    const uint64_t v = (uint64_t)p;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (P *)((hi << 32) | lo);
}
template <class A> __device__ __forceinline__ A uni_ro(A a) {
    const uint64_t v = (uint64_t)a.p;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    a.p = (decltype(a.p))((hi << 32) | lo);
    return a;
}
template <class T> __device__ __forceinline__ GradGeom<T> uni(GradGeom<T> g) {
    g.C = uni(g.C);
    g.rowlen = uni(g.rowlen);
    g.row = uni_ro(g.row);
    g.ent = uni_ro(g.ent);
    g.fgeoS = uni_ro(g.fgeoS);
    g.fgeoX = uni_ro(g.fgeoX);
    g.cgeo = uni_ro(g.cgeo);
    g.cvol = uni_ro(g.cvol);
    return g;
}
template <class T, class TI, class TO>
__device__ __forceinline__ BatchGradArgs<T, TI, TO> uni(BatchGradArgs<T, TI, TO> a) {
    a.phi = uni(a.phi);
    a.phiB = uni(a.phiB);
    a.out1 = uni(a.out1);
    a.out2 = uni(a.out2);
    return a;
}

// the gradient of cell c for scalar k of a batch of K (M: DevMesh or GradGeom)
template <class T, class TI, class TO, bool SWEEP0, int RL, class M>
__device__ __forceinline__ void gradient_batched_cell(const M &m, const BatchGradArgs<T, TI, TO> &a, int K, int c,
                                                      int kk) {
    const TI *phi = a.phi + kk, *phiB = a.phiB + kk;
    const T *go = SWEEP0 ? nullptr : a.grado + kk;
    auto cellv = [&](int x) -> GradCell<T> {
        GradCell<T> r;
        r.phi = (T)phi[(size_t)x * K];
        if (!SWEEP0) {
            r.go = ld3s(go, x, K);
            r.cc = xyz(m.cgeo[x]);
        }
        return r;
    };
    const GradCell<T> own = cellv(c);
    T acc[3] = {0, 0, 0};
    auto face = [&](int2 en, const V4<T> &s, const V4<T> &xf, const GradCell<T> &o) {
        T t[3];
        if (en.y >= 0) {
            const bool in = (en.x & 1) != 0;
            gradient_interior<T, SWEEP0>(xf, s, in ? o : own, in ? own : o, t);
        } else {
            t[0] = o.phi * s.x;  // o.phi holds phiB of the boundary
            t[1] = o.phi * s.y;
            t[2] = o.phi * s.z;
        }
        acc[0] = acc[0] + t[0];
        acc[1] = acc[1] + t[1];
        acc[2] = acc[2] + t[2];
    };
    auto other = [&](int2 en) -> GradCell<T> {
        if (en.y >= 0) return cellv(en.y);
        GradCell<T> r{};
        r.phi = (T)phiB[(size_t)(~en.y) * K];
        return r;
    };
    int e0, n;
    if (m.rowlen) {
        e0 = c * m.rowlen;
        n = m.rowlen;
    } else {
        e0 = m.row[c];
        n = m.row[c + 1] - e0;
    }
    if (RL) {
        int2 en[RL > 0 ? RL : 1];
        V4<T> sv[RL > 0 ? RL : 1], xv[RL > 0 ? RL : 1];
        GradCell<T> ov[RL > 0 ? RL : 1];
#pragma unroll
        for (int i = 0; i < RL; i++) en[i] = m.ent[e0 + i];
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int f = en[i].x >> 1;
            sv[i] = m.fgeoS[f];
            if (!SWEEP0 && en[i].y >= 0) xv[i] = m.fgeoX[f];
            ov[i] = other(en[i]);
        }
#pragma unroll
        for (int i = 0; i < RL; i++) face(en[i], sv[i], SWEEP0 ? sv[i] : xv[i], ov[i]);
    } else {
        for (int e = e0; e < e0 + n; e++) {
            const int2 en = m.ent[e];
            const int f = en.x >> 1;
            const V4<T> s = m.fgeoS[f];
            face(en, s, (SWEEP0 || en.y < 0) ? s : m.fgeoX[f], other(en));
        }
    }
    const T vol = m.cvol[c];
    const T fact = 1.0 / vol;
    for (int j = 0; j < 3; j++) {
        const T r = acc[j] * fact;
        const size_t o = ((size_t)c * 3 + j) * K + kk;
        st_stream(a.out1 + o, (TO)r);
        if (a.out2) st_stream(a.out2 + o, (TO)r);
    }
}

template <class T, class TI, class TO, bool SWEEP0, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_gradient_batched(DevMesh<T> m, BatchGradArgs<T, TI, TO> a, int K, int per,
                                                             int nblk) {
    const BatchLane L = batch_lane<T, G>(m, K, per, nblk);
    if (!L.live) return;
    gradient_batched_cell<T, TI, TO, SWEEP0, RL>(m, a, K, batch_uniform<G>(L.c), L.k);
}

// Batched gradient, first sweep, for lane groups narrower than a wave (G = 4
// or 16: a wave spans 16 or 4 cells, so its lanes cannot share scalar loads) on meshes
// whose rows all hold RL entries. The block's rows, face vectors and cell
// volumes are loaded once into LDS, one lane per entry, instead of by all G
// lanes of each cell: the rows before the first barrier, the face vectors and
// volumes while every lane's phi gathers are in flight, before the second.
// Terms and their order are gradient_batched_cell's.
template <class T, class TI, class TO, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_gradient_batched_staged(DevMesh<T> m, BatchGradArgs<T, TI, TO> a, int K,
                                                                    int per, int nblk) {
    constexpr int NC = kBlock / G, NP = NC * RL;
    __shared__ int2 sen[NP];
    __shared__ V4<T> ss[NP];
    __shared__ T svol[NC];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cb = lb * NC, tid = threadIdx.x;
    for (int p = tid; p < NP; p += kBlock) {
        const int c = cb + p / RL;
        sen[p] = c < m.C ? m.ent[(size_t)c * RL + p % RL] : make_int2(0, 0);
    }
    __syncthreads();
    const int cl = tid / G, c = cb + cl, kk = blockIdx.y * G + tid % G;
    const bool live = c < m.C && kk < K;
    TI phiCs = 0, ov[RL];  // storage type until first use
    if (live) {
        const TI *phi = a.phi + kk, *phiB = a.phiB + kk;
        phiCs = phi[(size_t)c * K];
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int2 en = sen[cl * RL + i];
            ov[i] = *(en.y >= 0 ? phi + (size_t)en.y * K : phiB + (size_t)(~en.y) * K);
        }
    }
    for (int p = tid; p < NP; p += kBlock)
        if (cb + p / RL < m.C) ss[p] = m.fgeoS[sen[p].x >> 1];
    for (int p = tid; p < NC; p += kBlock)
        if (cb + p < m.C) svol[p] = m.cvol[cb + p];
    __syncthreads();
    if (!live) return;
    const GradCell<T> own{(T)phiCs, {}, {}};
    T acc[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < RL; i++) {
        const int2 en = sen[cl * RL + i];
        const V4<T> s = ss[cl * RL + i];
        T t[3];
        if (en.y >= 0) {
            const GradCell<T> o{(T)ov[i], {}, {}};
            const bool in = (en.x & 1) != 0;
            gradient_interior<T, true>(s, s, in ? o : own, in ? own : o, t);
        } else {
            const T ob = (T)ov[i];
            t[0] = ob * s.x;
            t[1] = ob * s.y;
            t[2] = ob * s.z;
        }
        acc[0] = acc[0] + t[0];
        acc[1] = acc[1] + t[1];
        acc[2] = acc[2] + t[2];
    }
    const T fact = 1.0 / svol[cl];
    for (int j = 0; j < 3; j++) {
        const T r = acc[j] * fact;
        const size_t o = ((size_t)c * 3 + j) * K + kk;
        st_stream(a.out1 + o, (TO)r);
        if (a.out2) st_stream(a.out2 + o, (TO)r);
    }
}

// A set of meshes for one persistent launch: their device descriptors and,
// per lane-group width, the prefix of work units (blocks of kBlock / G cells)
template <class T> struct MeshSetArgs {
    const DevMesh<T> *meshes;
    const int *unit0;  // n + 1 entries
    int n, nunits;
};

// Batched gradient (nGradient = 1) over every mesh of a set in one persistent
// launch: logical block lb (XCD-aware, as xcd_block) takes the contiguous
// unit range [lb*U/P, (lb+1)*U/P) of the U units of all meshes and leaves
// after its last unit. phi[i] .. out2[i] are mesh i's arrays. The mesh
// descriptor and the walk's state want about 106 SGPRs, which would allow 7
// waves per SIMD; waves_per_eu(8) has the compiler spill the excess into VGPR
// lanes instead (no scratch), for the 8 waves the per-mesh kernel runs.
template <class T, class TI, class TO, int G, int RL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_gradient_batched_set(
    MeshSetArgs<T> s, const TI *const *phi, const TI *const *phiB, TO *const *out1, TO *const *out2, int K) {
    const int P = gridDim.x;
    const int lb = xcd_block(blockIdx.x, P / kXcd);
    const int u0 = (int)((int64_t)lb * s.nunits / P), u1 = (int)((int64_t)(lb + 1) * s.nunits / P);
    if (u0 >= u1) return;
    int i = 0;  // the mesh of unit u0, by bisection; later units advance it
    {
        int hi = s.n - 1;
        while (i < hi) {
            const int mid = (i + hi + 1) >> 1;
            if (s.unit0[mid] <= u0) i = mid;
            else hi = mid - 1;
        }
    }
    GradGeom<T> m = uni(GradGeom<T>(s.meshes[i]));
    BatchGradArgs<T, TI, TO> a = uni(BatchGradArgs<T, TI, TO>{phi[i], phiB[i], nullptr, out1[i], out2 ? out2[i] : nullptr});
    int ub = s.unit0[i], ue = s.unit0[i + 1];
    const int kk = blockIdx.y * G + threadIdx.x % G;
    for (int u = u0; u < u1; u++) {
        if (u >= ue) {  // next mesh with units
            do i++;
            while (s.unit0[i + 1] <= u);
            m = uni(GradGeom<T>(s.meshes[i]));
            a = uni(BatchGradArgs<T, TI, TO>{phi[i], phiB[i], nullptr, out1[i], out2 ? out2[i] : nullptr});
            ub = s.unit0[i];
            ue = s.unit0[i + 1];
        }
        const int c = batch_uniform<G>((u - ub) * (kBlock / G) + (int)threadIdx.x / G);
        if (c < m.C && kk < K) gradient_batched_cell<T, TI, TO, true, RL>(m, a, K, c, kk);
    }
}

// The parts of an interior face's scalar terms that do not depend on phi
// (ScalarKernels.ipp:135-217), evaluated once per face for a whole batch by
// k_scalar_coef with the expressions scalar_interior uses: the face
// diffusivity Visac and VisFace, Xpn, the upwind weights min/max(mf, 0) (in
// double, as the reference's std::min/max of a promoted mass flux) and, on
// meshes without carried PhiFace, fce = mf * PhiFace.
template <class T> struct ScalarFaceCoef {
    T visac, visface, xpx, xpy, xpz, fce;
    double fmn, fmx;
};

template <class T, class TS>
__global__ __launch_bounds__(kBlock) void k_scalar_coef(DevMesh<T> m, ScalarArgs<T, TS> a, ScalarFaceCoef<T> *coef) {
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= m.Fi) return;
    const int2 cn = m.fconn[f];
    const T facn = m.fgeoS[f].w;
    const T facp = 1.0 - facn;
    T Visac = a.visEff[cn.y] * facn + a.visEff[cn.x] * facp;
    Visac = scalar_visac(a, Visac);
    const v3<T> Xpn = sub3(xyz(m.cgeo[cn.y]), xyz(m.cgeo[cn.x]));
    const T VisFace = Visac * m.frl[f];
    const T mf = a.massFlux[f];
    ScalarFaceCoef<T> q;
    q.visac = Visac;
    q.visface = VisFace;
    q.xpx = Xpn.x;
    q.xpy = Xpn.y;
    q.xpz = Xpn.z;
    q.fce = m.typed ? (T)0 : mf * m.carry_init;
    q.fmn = fmin((double)mf, 0.0);
    q.fmx = fmax((double)mf, 0.0);
    coef[f] = q;
    a.rface[2 * (size_t)f] = -VisFace - fmax((double)mf, 0.0);
    a.rface[2 * (size_t)f + 1] = -VisFace + fmin((double)mf, 0.0);
}

// Batched scalar flux on meshes whose regions are all DEFAULT, with RL faces
// in every row: a boundary entry then adds -0.0 three times to Su and touches
// nothing else (scalar_boundary), so only interior entries are evaluated; the
// row's entries and neighbour values are all loaded before the first term (two
// memory round trips per cell instead of two per entry). Terms and their
// order are k_scalar_batched's.
template <class T, class TS, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_scalar_batched_plain(DevMesh<T> m, ScalarArgs<T, TS> a,
                                                                 const ScalarFaceCoef<T> *coef, int K, int per,
                                                                 int nblk) {
    const BatchLane L = batch_lane<T, G>(m, K, per, nblk);
    if (!L.live) return;
    const int c = batch_uniform<G>(L.c);
    const TS *phi = a.phi + L.k, *dphi = a.dPhidx + L.k;
    TS *sup = a.su + L.k;
    const TS phiCs = phi[(size_t)c * K];
    const v3<TS> gCs = ld3s(dphi, c, K);
    const TS sus = sup[(size_t)c * K];
    const int e0 = c * RL;
    int2 en[RL];
    TS ophi[RL];  // storage type until first use; a boundary entry gathers the cell's own row
    v3<TS> og[RL];
#pragma unroll
    for (int i = 0; i < RL; i++) en[i] = m.ent[e0 + i];
#pragma unroll
    for (int i = 0; i < RL; i++) {
        const int o = en[i].y >= 0 ? en[i].y : c;
        ophi[i] = phi[(size_t)o * K];
        og[i] = ld3s(dphi, o, K);
    }
    const T phiC = (T)phiCs;
    const v3<T> gC = widen3<T>(gCs);
    T su = (T)sus;
#pragma unroll
    for (int i = 0; i < RL; i++) {
        if (en[i].y < 0) continue;
        const int f = en[i].x >> 1;
        const V4<T> s = m.fgeoS[f];
        const ScalarFaceCoef<T> q = coef[f];
        const bool in = (en[i].x & 1) != 0;
        const T oph = (T)ophi[i];
        const v3<T> ogi = widen3<T>(og[i]);
        const T phiP = in ? oph : phiC, phiN = in ? phiC : oph;
        const v3<T> gP = in ? ogi : gC, gN = in ? gC : ogi;
        const v3<T> S{s.x, s.y, s.z};
        const T facn = s.w;
        const T facp = 1.0 - facn;
        const v3<T> g = lerp3(gN, facn, gP, facp);
        const T fde1 = q.visac * dot3(g, S);
        const T fci = q.fmn * (double)phiN + q.fmx * (double)phiP;
        const T fdi = q.visface * dot3(g, v3<T>{q.xpx, q.xpy, q.xpz});
        const T blend = a.gammaBlend * (q.fce - fci);
        T t0 = -blend, t1 = fde1, t2 = -fdi;
        if (in) {
            t0 = -t0;
            t1 = -t1;
            t2 = -t2;
        }
        su = ((su + t0) + t1) + t2;
    }
    sup[(size_t)c * K] = (TS)su;
}

// k_scalar_batched_plain for 4- and 16-lane groups (a wave spans 16 or 4
// cells, so its lanes cannot share scalar loads): the block's rows, and the face vectors and
// coefficients of their interior entries, are loaded once into LDS, one lane
// per entry, instead of by all G lanes of each cell. The rows come before the
// first barrier; the face data while every lane's phi and gradient gathers are
// in flight, before the second. Terms and their order are
// k_scalar_batched_plain's.
template <class T, class TS, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_scalar_batched_staged(DevMesh<T> m, ScalarArgs<T, TS> a,
                                                                  const ScalarFaceCoef<T> *coef, int K, int per,
                                                                  int nblk) {
    constexpr int NC = kBlock / G, NP = NC * RL;
    __shared__ int2 sen[NP];
    __shared__ V4<T> ss[NP];
    __shared__ ScalarFaceCoef<T> sq[NP];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cb = lb * NC, tid = threadIdx.x;
    for (int p = tid; p < NP; p += kBlock) {
        const int c = cb + p / RL;
        sen[p] = c < m.C ? m.ent[(size_t)c * RL + p % RL] : make_int2(0, -1);
    }
    __syncthreads();
    const int cl = tid / G, c = cb + cl, kk = blockIdx.y * G + tid % G;
    const bool live = c < m.C && kk < K;
    const TS *phi = a.phi + kk, *dphi = a.dPhidx + kk;
    TS *sup = a.su + kk;
    TS phiCs = 0, sus = 0, ophi[RL];  // storage type until first use
    v3<TS> gCs{}, og[RL];
    if (live) {
        phiCs = phi[(size_t)c * K];
        gCs = ld3s(dphi, c, K);
        sus = sup[(size_t)c * K];
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int2 en = sen[cl * RL + i];
            const int o = en.y >= 0 ? en.y : c;  // a boundary entry gathers the cell's own row
            ophi[i] = phi[(size_t)o * K];
            og[i] = ld3s(dphi, o, K);
        }
    }
    for (int p = tid; p < NP; p += kBlock) {
        const int2 en = sen[p];
        if (en.y >= 0) {
            const int f = en.x >> 1;
            ss[p] = m.fgeoS[f];
            sq[p] = coef[f];
        }
    }
    __syncthreads();
    if (!live) return;
    const T phiC = (T)phiCs;
    const v3<T> gC = widen3<T>(gCs);
    T su = (T)sus;
#pragma unroll
    for (int i = 0; i < RL; i++) {
        const int2 en = sen[cl * RL + i];
        if (en.y < 0) continue;
        const V4<T> s = ss[cl * RL + i];
        const ScalarFaceCoef<T> q = sq[cl * RL + i];
        const bool in = (en.x & 1) != 0;
        const T oph = (T)ophi[i];
        const v3<T> ogi = widen3<T>(og[i]);
        const T phiP = in ? oph : phiC, phiN = in ? phiC : oph;
        const v3<T> gP = in ? ogi : gC, gN = in ? gC : ogi;
        const v3<T> S{s.x, s.y, s.z};
        const T facn = s.w;
        const T facp = 1.0 - facn;
        const v3<T> g = lerp3(gN, facn, gP, facp);
        const T fde1 = q.visac * dot3(g, S);
        const T fci = q.fmn * (double)phiN + q.fmx * (double)phiP;
        const T fdi = q.visface * dot3(g, v3<T>{q.xpx, q.xpy, q.xpz});
        const T blend = a.gammaBlend * (q.fce - fci);
        T t0 = -blend, t1 = fde1, t2 = -fdi;
        if (in) {
            t0 = -t0;
            t1 = -t1;
            t2 = -t2;
        }
        su = ((su + t0) + t1) + t2;
    }
    sup[(size_t)c * K] = (TS)su;
}

template <class T, class TS, int G>
__global__ __launch_bounds__(kBlock) void k_scalar_batched(DevMesh<T> m, ScalarArgs<T, TS> a0,
                                                           const ScalarFaceCoef<T> *coef, int K, int per, int nblk) {
    const BatchLane L = batch_lane<T, G>(m, K, per, nblk);
    if (!L.live) return;
    // this lane's view: per-scalar arrays offset to scalar k, stride K
    ScalarArgs<T, TS> a = a0;
    a.phi += L.k;
    a.phiB += L.k;
    a.au += L.k;
    a.su += L.k;
    a.dPhidx += L.k;
    a.ks = K;
    a.write_state = (L.k == K - 1);  // K plain calls leave the last scalar's mesh state
    const int c = batch_uniform<G>(L.c);
    const ScalarCell<T> own{(T)a.phi[(size_t)c * K], a.visEff[c], ld3s_as<T>(a.dPhidx, c, K), xyz(m.cgeo[c])};
    T su = (T)a.su[(size_t)c * K];
    T au = 0;
    bool au_loaded = false;
    int e0, n;
    row_range(m, c, e0, n);
    for (int e = e0; e < e0 + n; e++) {
        const int2 en = m.ent[e];
        const int f = en.x >> 1;
        const V4<T> s = m.fgeoS[f];
        T t[4];
        bool ta = false;
        if (en.y >= 0) {
            // scalar_interior's terms with the face's phi-free parts from coef
            const ScalarFaceCoef<T> q = coef[f];
            const T ophi = (T)a.phi[(size_t)en.y * K];
            const v3<T> og = ld3s_as<T>(a.dPhidx, en.y, K);
            const bool in = (en.x & 1) != 0;
            const T phiP = in ? ophi : own.phi, phiN = in ? own.phi : ophi;
            const v3<T> gP = in ? og : own.g, gN = in ? own.g : og;
            const v3<T> S{s.x, s.y, s.z};
            const T facn = s.w;
            const T facp = 1.0 - facn;
            const v3<T> g = lerp3(gN, facn, gP, facp);
            const T fce = m.typed ? a.massFlux[f] * scalar_carry(m, a, f) : q.fce;
            const T fde1 = q.visac * dot3(g, S);
            const T fci = q.fmn * (double)phiN + q.fmx * (double)phiP;
            const T fdi = q.visface * dot3(g, v3<T>{q.xpx, q.xpy, q.xpz});
            const T blend = a.gammaBlend * (fce - fci);
            t[0] = -blend;
            t[1] = fde1;
            t[2] = -fdi;
            if (in) {
                t[0] = -t[0];
                t[1] = -t[1];
                t[2] = -t[2];
            }
        } else {
            ta = scalar_boundary(m, a, f, s, c, own, ~en.y, t);
        }
        su = ((su + t[0]) + t[1]) + t[2];
        if (ta) {
            if (!au_loaded) {
                au = (T)a.au[(size_t)c * K];
                au_loaded = true;
            }
            au = au + t[3];
        }
    }
    a.su[(size_t)c * K] = (TS)su;
    if (au_loaded) a.au[(size_t)c * K] = (TS)au;
}

// ===========================================================================
// Adjoint of GradientPhiGaussDolfyn (nGradient = 1) followed by
// FluxScalarDolfynFaceLoop w.r.t. (phi, MassFlux), on all-DEFAULT meshes
// (no reference counterpart; derived from GradientKernels.ipp:38-136 and
// ScalarKernels.ipp:135-217). For interior face f (cell1 P, cell2 N):
//   w_f = Visac*S - VisFace*Xpn,  D_f = Subar[P] - Subar[N]
//   pass A (per cell c):  h_c = (1/vol_c) sum_f (D_f*fac)*w_f,
//                         dir_c = sum_f (D_f*GammaBlend) * (c == P ? max(mf,0) : min(mf,0))
//   pass B (per cell c):  phibar_c = dir_c + sum_f fac * dot(h_P + h_N, S_f)
//   faces:                mfbar_f = -(D_f*GammaBlend*(PhiFace - dfci)) + rterm
//                         (written in pass A by the face's cell1 entry)
// with fac = facp for c == P, facn for c == N, and the sums over each cell's
// faces in row order (ascending mesh face id). Boundary faces add nothing.
// ===========================================================================
template <class T> struct AdjArgs {
    ScalarArgs<T> s;  // phi, visEff, massFlux and the diffusivity parameters
    const T *subar, *rbar;
    T *h, *dir, *phibar, *mfbar;
    // typed meshes: each boundary face's pass-A terms (gradient cotangent x3,
    // phi term) at bterm[4 (g - Fi)] (k_adj_bnd); null on all-DEFAULT meshes
    const T *bterm;
    // pass B: each cell's starting value (dir, or PhiBar of the sweep after),
    // and for a sweep before the last, the h it hands to that sweep
    const T *binit;
    T *hnext;
};

// PhiFace as the scalar loop carries it into face f (the carry-init value on
// all-DEFAULT meshes)
template <class T> __device__ __forceinline__ T adj_phiface(const DevMesh<T> &m, const AdjArgs<T> &a, int f) {
    if (!a.s.cval) return m.carry_init;
    const int g = a.s.src[f];
    return g < 0 ? m.carry_init : a.s.cval[g - m.Fi];
}

// Pass-A terms of CSR entry en of cell c: t[0..2] = (D*fac)*w, t[3] = the direct phi term
template <class T>
__device__ __forceinline__ void adj_a_terms(const DevMesh<T> &m, const AdjArgs<T> &a, int c, int2 en, T t[4]) {
    if (en.y < 0) {
        if (a.bterm) {  // evaluated ahead by k_adj_bnd, which also wrote the face's MassFluxBar
            const V4<T> bt = *(const V4<T> *)(a.bterm + 4 * (size_t)((en.x >> 1) - m.Fi));
            t[0] = bt.x, t[1] = bt.y, t[2] = bt.z, t[3] = bt.w;
        } else {
            t[0] = t[1] = t[2] = t[3] = nzero<T>();
            a.mfbar[en.x >> 1] = 0.0;  // a DEFAULT boundary face has no mass-flux cotangent
        }
        return;
    }
    const int f = en.x >> 1;
    const bool in = (en.x & 1) != 0;
    const int ip = in ? en.y : c, nn = in ? c : en.y;
    const V4<T> s = m.fgeoS[f];
    const T facn = s.w;
    const T facp = 1.0 - facn;
    T Visac = a.s.visEff[nn] * facn + a.s.visEff[ip] * facp;
    Visac = scalar_visac(a.s, Visac);
    const T VisFace = Visac * m.frl[f];
    const v3<T> Xpn = sub3(xyz(m.cgeo[nn]), xyz(m.cgeo[ip]));
    const T D = a.subar[ip] - a.subar[nn];
    const T Dg = D * a.s.gammaBlend;
    const T fac = in ? facn : facp;
    const T tt = D * fac;
    t[0] = tt * (Visac * s.x - VisFace * Xpn.x);
    t[1] = tt * (Visac * s.y - VisFace * Xpn.y);
    t[2] = tt * (Visac * s.z - VisFace * Xpn.z);
    const T mf = a.s.massFlux[f];
    t[3] = Dg * (in ? (T)fmin((double)mf, 0.0) : (T)fmax((double)mf, 0.0));
    if (!in) {  // the face's cell1 entry (one per face) also writes its mass-flux cotangent
        const T dfci = mf < 0 ? a.s.phi[nn] : (mf > 0 ? a.s.phi[ip] : (T)0.0);
        const T rterm = mf < 0 ? a.rbar[2 * (size_t)f + 1] : (mf > 0 ? -a.rbar[2 * (size_t)f] : (T)0.0);
        const T PhiFace = adj_phiface(m, a, f);
        a.mfbar[f] = -(Dg * (PhiFace - dfci)) + rterm;
    }
}

// Pass-B term t = fac*dot(h_P + h_N, S) of CSR entry en of cell c, and with
// CORR the previous sweep's gradient cotangent t*corr_f, corr_f = Xf -
// (Cc[N]*facn + Cc[P]*facp) (GradientKernels.ipp:88-96)
template <class T, bool CORR = false>
__device__ __forceinline__ T adj_b_term(const DevMesh<T> &m, const AdjArgs<T> &a, int c, int2 en,
                                        v3<T> *g = nullptr) {
    if (en.y < 0) {
        if (CORR) *g = v3<T>{nzero<T>(), nzero<T>(), nzero<T>()};
        return nzero<T>();
    }
    const int f = en.x >> 1;
    const bool in = (en.x & 1) != 0;
    const int ip = in ? en.y : c, nn = in ? c : en.y;
    const V4<T> s = m.fgeoS[f];
    const T facn = s.w;
    const T facp = 1.0 - facn;
    const T fac = in ? facn : facp;
    const v3<T> hp = ld3(a.h, ip), hn = ld3(a.h, nn);
    const v3<T> hs{hp.x + hn.x, hp.y + hn.y, hp.z + hn.z};
    const T pF = dot3(hs, v3<T>{s.x, s.y, s.z});
    const T t = fac * pF;
    if (CORR) {
        const v3<T> xac = lerp3(xyz(m.cgeo[nn]), facn, xyz(m.cgeo[ip]), facp);
        const v3<T> corr = sub3(xyz(m.fgeoX[f]), xac);
        *g = v3<T>{t * corr.x, t * corr.y, t * corr.z};
    }
    return t;
}

template <class T> __device__ __forceinline__ void adj_a_store(const DevMesh<T> &m, const AdjArgs<T> &a, int c, int k,
                                                               T acc) {
    if (k < 3) {
        const T fact = 1.0 / m.cvol[c];
        a.h[3 * (size_t)c + k] = acc * fact;
    } else {
        a.dir[c] = acc;
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_a(DevMesh<T> m, AdjArgs<T> a, int per, int nblk) {
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int c = lb * kBlock + threadIdx.x;
    if (c >= m.C) return;
    T acc[4] = {0, 0, 0, 0};
    int e0, n;
    row_range(m, c, e0, n);
    for (int e = e0; e < e0 + n; e++) {
        T t[4];
        adj_a_terms(m, a, c, m.ent[e], t);
        for (int k = 0; k < 4; k++) acc[k] = acc[k] + t[k];
    }
    for (int k = 0; k < 4; k++) adj_a_store(m, a, c, k, acc[k]);
}

template <class T, int LPC>
__global__ __launch_bounds__(LPC * 32) void k_adj_a_vec(DevMesh<T> m, AdjArgs<T> a, int per, int nblk) {
    constexpr int CPB = kCellsPerBlock, NT = 4, P = LPC + 1;
    __shared__ T lds[NT * CPB * P];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cl = threadIdx.x / LPC, j = threadIdx.x % LPC;
    const int c = lb * CPB + cl;
    int e0 = 0, n = 0;
    if (c < m.C) row_range(m, c, e0, n);
    if (j < n) {
        T t[NT];
        adj_a_terms(m, a, c, m.ent[e0 + j], t);
        for (int k = 0; k < NT; k++) lds[(k * CPB + cl) * P + j] = t[k];
    }
    __syncthreads();
    if (c >= m.C) return;
    for (int k = j; k < NT; k += LPC) {
        T acc = 0;
        for (int i = 0; i < n; i++) acc = acc + lds[(k * CPB + cl) * P + i];
        adj_a_store(m, a, c, k, acc);
    }
}

// Pass B: PhiBar_c = binit_c + sum of the row's terms; with CORR also
// hnext_c = (sum of the row's t*corr_f) * (1/vol_c)
template <class T, bool CORR>
__global__ __launch_bounds__(kBlock) void k_adj_b(DevMesh<T> m, AdjArgs<T> a, int per, int nblk) {
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int c = lb * kBlock + threadIdx.x;
    if (c >= m.C) return;
    T acc = a.binit[c];
    T gx = 0, gy = 0, gz = 0;
    int e0, n;
    row_range(m, c, e0, n);
    for (int e = e0; e < e0 + n; e++) {
        v3<T> g;
        acc = acc + adj_b_term<T, CORR>(m, a, c, m.ent[e], &g);
        if (CORR) {
            gx = gx + g.x;
            gy = gy + g.y;
            gz = gz + g.z;
        }
    }
    a.phibar[c] = acc;
    if (CORR) {
        const T fact = 1.0 / m.cvol[c];
        st3(a.hnext, c, v3<T>{gx * fact, gy * fact, gz * fact});
    }
}

template <class T, int LPC, bool CORR>
__global__ __launch_bounds__(LPC * 32) void k_adj_b_vec(DevMesh<T> m, AdjArgs<T> a, int per, int nblk) {
    constexpr int CPB = kCellsPerBlock, P = LPC + 1, NT = CORR ? 4 : 1;
    static_assert(!CORR || LPC >= 4, "a lane per folded row");
    __shared__ T lds[NT * CPB * P];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cl = threadIdx.x / LPC, j = threadIdx.x % LPC;
    const int c = lb * CPB + cl;
    int e0 = 0, n = 0;
    if (c < m.C) row_range(m, c, e0, n);
    if (j < n) {
        v3<T> g;
        lds[cl * P + j] = adj_b_term<T, CORR>(m, a, c, m.ent[e0 + j], &g);
        if (CORR) {
            lds[(1 * CPB + cl) * P + j] = g.x;
            lds[(2 * CPB + cl) * P + j] = g.y;
            lds[(3 * CPB + cl) * P + j] = g.z;
        }
    }
    __syncthreads();
    if (c >= m.C || j >= NT) return;
    T acc = j == 0 ? a.binit[c] : (T)0;
    for (int i = 0; i < n; i++) acc = acc + lds[(j * CPB + cl) * P + i];
    if (j == 0) a.phibar[c] = acc;
    else a.hnext[3 * (size_t)c + (j - 1)] = acc * (1.0 / m.cvol[c]);
}

// ---------------------------------------------------------------------------
// Tile schedule of the adjoint passes (the forward tiles' three phases): each
// slot's terms are evaluated once for both of its cells (the two sides'
// terms differ by fac, and in pass A by max / min(mf, 0), not only by sign),
// then each (cell, component) folds its row in face order, so the sums are
// the lane schedule's bit for bit.
// ---------------------------------------------------------------------------
// cell values pass A reads at a cell
template <class T> struct AdjCellA {
    T subar, vis, phi;
    v3<T> cc;
};

// Pass A: h_c = (sum of the row's gradient-cotangent terms) / vol_c and dir_c;
// MassFluxBar of each owned interior face (and of owned DEFAULT boundary
// faces) is written by its slot.
template <class T>
__global__ __launch_bounds__(kTileBlock) void k_adj_a_tile(DevMesh<T> m, AdjArgs<T> a, int per, int nblk) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TC = kTileCells, NF = 6, NT = 8;  // cells: SuBar, VisEff, phi, centre xyz; slots: 4 terms per side
    const int S = m.tslots;
    T *lc = (T *)smem;                         // [NF][TC]
    T *lt = lc + NF * TC;                      // [NT][S]
    uint16_t *le = (uint16_t *)(lt + NT * S);  // [tnnz]
    const int t = xcd_block(blockIdx.x, per);
    if (t >= nblk) return;
    const int tid = threadIdx.x;
    const int c0 = t * TC, nc = min(TC, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    // First round: every load that depends on no other load, at clamped
    // indices so that all of them issue before the first wait; their LDS
    // writes follow. Thread tid stages SuBar, VisEff or phi (tid / TC = 0, 1,
    // 2) of cell tid % TC, element tid of the centres, CSR slot codes tid and
    // tid + kTileBlock, its slot's connectivity and its fold item's row.
    const int ci = min(tid & (TC - 1), nc - 1);
    const T x1 = (tid < TC ? a.subar : tid < 2 * TC ? a.s.visEff : a.s.phi)[c0 + ci];
    const T xc = ((const T *)(m.cgeo + c0))[min(tid, 4 * nc - 1)];
    int eb, ee;
    if (m.rowlen) {
        eb = c0 * m.rowlen;
        ee = (c0 + nc) * m.rowlen;
    } else {
        eb = m.row[c0];
        ee = m.row[c0 + nc];
    }
    const int ne = ee - eb;  // >= 1: a tile whose cells have no faces takes no tile schedule
    const uint16_t q0 = m.eslot[eb + min(tid, ne - 1)], q1 = m.eslot[eb + min(tid + kTileBlock, ne - 1)];
    const int k = tid / TC, cl = tid % TC;  // fold item: component k (h xyz, dir) of cell cl
    const bool fold = cl < nc;
    int e0, n;
    row_range(m, c0 + ci, e0, n);
    const bool live = tid < ts.nslot;
    const int sl0 = min(tid, ts.nslot - 1);
    int2 cn;
    bool owned;
    int f = slot_conn_flat(m, ts, sl0, cn, owned);
    if (tid < 3 * TC && (tid & (TC - 1)) < nc) lc[(tid / TC) * TC + (tid & (TC - 1))] = x1;
    if (tid < 4 * nc && (tid & 3) != 3) lc[(3 + (tid & 3)) * TC + (tid >> 2)] = xc;
    if (tid < ne) le[tid] = q0;
    if (tid + kTileBlock < ne) le[tid + kTileBlock] = q1;
    for (int i = tid + 2 * kTileBlock; i < ne; i += kTileBlock) le[i] = m.eslot[eb + i];
    int out = -1;
    V4<T> s{};
    T mf = 0, rl = 0, pf = m.carry_init, rb0 = 0, rb1 = 0;
    AdjCellA<T> o{};
    // second round: slot sl's face values and its cell outside the tile (cn
    // and f already set)
    auto rest = [&]() {
        out = -1;
        if (cn.y < 0) return;
        s = m.fgeoS[f];
        mf = a.s.massFlux[f];
        rl = m.frl[f];
        if (owned) {
            pf = adj_phiface(m, a, f);
            rb0 = a.rbar[2 * (size_t)f];
            rb1 = a.rbar[2 * (size_t)f + 1];
        }
        out = outside_cell(cn, c0, nc);
        if (out >= 0) o = AdjCellA<T>{a.subar[out], a.s.visEff[out], a.s.phi[out], xyz(m.cgeo[out])};
    };
    // a later slot sl of this thread (tiles of more than kTileBlock slots)
    auto load = [&](int sl) {
        f = slot_conn(m, ts, sl, cn, owned);
        rest();
    };
    if (live) rest();
    __syncthreads();
    auto cell = [&](int c) -> AdjCellA<T> {
        const bool o_ = c == out;
        const int l = o_ ? 0 : c - c0;
        return AdjCellA<T>{o_ ? o.subar : lc[l], o_ ? o.vis : lc[TC + l], o_ ? o.phi : lc[2 * TC + l],
                           v3<T>{o_ ? o.cc.x : lc[3 * TC + l], o_ ? o.cc.y : lc[4 * TC + l],
                                 o_ ? o.cc.z : lc[5 * TC + l]}};
    };
    for (int sl = tid; sl < ts.nslot; sl += kTileBlock) {
        if (sl != tid) load(sl);
        T *q = lt + sl;
        const T z = nzero<T>();
        if (cn.y < 0) {  // P side only
            if (a.bterm) {
                const V4<T> bt = *(const V4<T> *)(a.bterm + 4 * (size_t)(f - m.Fi));
                q[0] = bt.x, q[S] = bt.y, q[2 * S] = bt.z, q[3 * S] = bt.w;
            } else {
                q[0] = q[S] = q[2 * S] = q[3 * S] = z;
                if (owned) a.mfbar[f] = 0.0;
            }
            continue;
        }
        const AdjCellA<T> P = cell(cn.x), N = cell(cn.y);
        const T facn = s.w;
        const T facp = 1.0 - facn;
        T Visac = N.vis * facn + P.vis * facp;
        Visac = scalar_visac(a.s, Visac);
        const T VisFace = Visac * rl;
        const v3<T> Xpn = sub3(N.cc, P.cc);
        const T D = P.subar - N.subar;
        const T Dg = D * a.s.gammaBlend;
        const T wx = Visac * s.x - VisFace * Xpn.x, wy = Visac * s.y - VisFace * Xpn.y,
                wz = Visac * s.z - VisFace * Xpn.z;
        const T tp = D * facp, tn = D * facn;
        q[0] = tp * wx, q[S] = tp * wy, q[2 * S] = tp * wz;
        q[3 * S] = Dg * (T)fmax((double)mf, 0.0);
        q[4 * S] = tn * wx, q[5 * S] = tn * wy, q[6 * S] = tn * wz;
        q[7 * S] = Dg * (T)fmin((double)mf, 0.0);
        if (owned) {
            const T dfci = mf < 0 ? N.phi : (mf > 0 ? P.phi : (T)0.0);
            const T rterm = mf < 0 ? rb1 : (mf > 0 ? -rb0 : (T)0.0);
            a.mfbar[f] = -(Dg * (pf - dfci)) + rterm;
        }
    }
    __syncthreads();
    if (fold) {
        T acc = 0;
        for (int e = e0; e < e0 + n; e++) {
            const int es = le[e - eb];
            const int sl = es & 0x7fff;
            acc = acc + lt[(k + ((es & 0x8000) ? 4 : 0)) * S + sl];
        }
        adj_a_store(m, a, c0 + cl, k, acc);
    }
}

// Pass B: PhiBar_c = binit_c + sum of the row's t = fac*dot(h_P + h_N, S); with
// CORR also hnext_c = (sum of t*corr_f) / vol_c.
template <class T, bool CORR>
__global__ __launch_bounds__(kTileBlock) void k_adj_b_tile(DevMesh<T> m, AdjArgs<T> a, int per, int nblk) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TC = kTileCells, NF = CORR ? 6 : 3, NT = CORR ? 8 : 2;  // slot rows: t per side (+ t*corr per side)
    const int S = m.tslots;
    T *lc = (T *)smem;                         // [NF][TC]: h xyz (+ centre xyz)
    T *lt = lc + NF * TC;                      // [NT][S]
    uint16_t *le = (uint16_t *)(lt + NT * S);  // [tnnz]
    const int t = xcd_block(blockIdx.x, per);
    if (t >= nblk) return;
    const int tid = threadIdx.x;
    const int c0 = t * TC, nc = min(TC, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    // First round: every load that depends on no other load, at clamped
    // indices so that all of them issue before the first wait; their LDS
    // writes follow. Thread tid stages element tid of the packed h (and of
    // the centres), CSR slot codes tid and tid + kTileBlock, its slot's
    // connectivity and its fold item's row and starting value.
    const int ci = min(tid & (TC - 1), nc - 1);
    const T gd = a.h[3 * (size_t)c0 + min(tid, 3 * nc - 1)];
    T xc = 0;
    if (CORR) xc = ((const T *)(m.cgeo + c0))[min(tid, 4 * nc - 1)];
    int eb, ee;
    if (m.rowlen) {
        eb = c0 * m.rowlen;
        ee = (c0 + nc) * m.rowlen;
    } else {
        eb = m.row[c0];
        ee = m.row[c0 + nc];
    }
    const int ne = ee - eb;  // >= 1: a tile whose cells have no faces takes no tile schedule
    const uint16_t q0 = m.eslot[eb + min(tid, ne - 1)], q1 = m.eslot[eb + min(tid + kTileBlock, ne - 1)];
    const int k = tid / TC, cl = tid % TC;  // fold item: PhiBar (k = 0) or component k - 1 of hnext
    const bool fold = k < (CORR ? 4 : 1) && cl < nc;
    int e0, n;
    row_range(m, c0 + ci, e0, n);
    const T b0 = a.binit[c0 + ci];
    const bool live = tid < ts.nslot;
    const int sl0 = min(tid, ts.nslot - 1);
    int2 cn;
    bool owned;
    int f = slot_conn_flat(m, ts, sl0, cn, owned);
    if (tid < 3 * nc) lc[(tid % 3) * TC + tid / 3] = gd;
    if (CORR && tid < 4 * nc && (tid & 3) != 3) lc[(3 + (tid & 3)) * TC + (tid >> 2)] = xc;
    if (tid < ne) le[tid] = q0;
    if (tid + kTileBlock < ne) le[tid + kTileBlock] = q1;
    for (int i = tid + 2 * kTileBlock; i < ne; i += kTileBlock) le[i] = m.eslot[eb + i];
    T acc = k == 0 ? b0 : (T)0;
    int out = -1;
    V4<T> s{}, xf{};
    v3<T> oh{}, occ{};
    // second round: slot sl's face values and its cell outside the tile (cn
    // and f already set)
    auto rest = [&]() {
        out = -1;
        if (cn.y < 0) return;
        s = m.fgeoS[f];
        if (CORR) xf = m.fgeoX[f];
        out = outside_cell(cn, c0, nc);
        if (out >= 0) {
            oh = ld3(a.h, out);
            if (CORR) occ = xyz(m.cgeo[out]);
        }
    };
    // a later slot sl of this thread (tiles of more than kTileBlock slots)
    auto load = [&](int sl) {
        f = slot_conn(m, ts, sl, cn, owned);
        rest();
    };
    if (live) rest();
    __syncthreads();
    auto hof = [&](int c) -> v3<T> {
        if (c == out) return oh;
        const int l = c - c0;
        return v3<T>{lc[l], lc[TC + l], lc[2 * TC + l]};
    };
    auto ccof = [&](int c) -> v3<T> {
        if (c == out) return occ;
        const int l = c - c0;
        return v3<T>{lc[3 * TC + l], lc[4 * TC + l], lc[5 * TC + l]};
    };
    for (int sl = tid; sl < ts.nslot; sl += kTileBlock) {
        if (sl != tid) load(sl);
        T *q = lt + sl;
        const T z = nzero<T>();
        if (cn.y < 0) {
            for (int j = 0; j < NT; j++) q[j * S] = z;
            continue;
        }
        const v3<T> hp = hof(cn.x), hn = hof(cn.y);
        const v3<T> hs{hp.x + hn.x, hp.y + hn.y, hp.z + hn.z};
        const T pF = dot3(hs, v3<T>{s.x, s.y, s.z});
        const T facn = s.w;
        const T facp = 1.0 - facn;
        const T tp = facp * pF, tn = facn * pF;
        q[0] = tp;
        q[S] = tn;
        if (CORR) {
            const v3<T> xac = lerp3(ccof(cn.y), facn, ccof(cn.x), facp);
            const v3<T> corr = sub3(xyz(xf), xac);
            q[2 * S] = tp * corr.x, q[3 * S] = tp * corr.y, q[4 * S] = tp * corr.z;
            q[5 * S] = tn * corr.x, q[6 * S] = tn * corr.y, q[7 * S] = tn * corr.z;
        }
    }
    __syncthreads();
    if (fold) {
        for (int e = e0; e < e0 + n; e++) {
            const int es = le[e - eb];
            const int sl = es & 0x7fff;
            const bool side2 = (es & 0x8000) != 0;
            const int row = k == 0 ? (side2 ? 1 : 0) : (side2 ? 4 : 1) + k;
            acc = acc + lt[row * S + sl];
        }
        const int c = c0 + cl;
        if (k == 0) a.phibar[c] = acc;
        else a.hnext[3 * (size_t)c + (k - 1)] = acc * (1.0 / m.cvol[c]);
    }
}

// Typed meshes: the cotangent of the PhiFace each boundary face k = g - Fi
// leaves is the sum, in ascending mesh face id, of its readers' terms
// (readers rd[rd_off[k] .. rd_off[k + 1])): an interior reader's
// -((D*GammaBlend)*mf), an INLET's that carries PhiFace -(SuBar[P]*f).
// k_adj_carry_terms evaluates every face's term in face order (coalesced;
// 0 for faces that read no carried value), and k_adj_carry_sum gives each
// source one wave: its lanes gather 64 readers' terms at a time, and the
// wave adds them in reader order (each term read from its lane with
// readlane, so the chain of dependent adds is the only serial work).
template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_carry_terms(DevMesh<T> m, AdjArgs<T> a, T *coef) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= m.F) return;
    const int2 cn = m.fconn[r];
    const T mf = a.s.massFlux[r];
    T c = 0.0;
    if (cn.y >= 0) {
        const T D = a.subar[cn.x] - a.subar[cn.y];
        const T Dg = D * a.s.gammaBlend;
        c = -(Dg * mf);
    } else if (!a.s.inletAssigns && m.rtype[m.bregion[~cn.y]] == CUPCFD_RTYPE_INLET) {
        const T Visac = scalar_visac(a.s, a.s.visEffB[~cn.y]);
        const T VisFace = Visac * m.frl[r];
        const T f = -VisFace + (T)fmin((double)mf, 0.0);
        c = -(a.subar[cn.x] * f);
    }
    coef[r] = c;
}

// lane l's value of v, as a wave-uniform value (l wave-uniform)
__device__ __forceinline__ double lane_value(double v, int l) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_carry_sum(int nb, const int *rd_off, const int *rd, const T *coef,
                                                          T *pfbar) {
    const int k = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);  // one wave per source
    if (k >= nb) return;
    const int lane = threadIdx.x & 63;
    const int e0 = uni(rd_off[k]), e1 = uni(rd_off[k + 1]);
    T acc = 0.0;
    for (int b = e0; b < e1; b += 64) {
        const int e = b + lane;
        const T c = e < e1 ? coef[rd[e]] : (T)0;
        const int cnt = min(64, e1 - b);
        for (int i = 0; i < cnt; i++) acc = acc + lane_value(c, i);
    }
    if (lane == 0) pfbar[k] = acc;
}

// Typed meshes: pass-A terms of boundary face g = Fi + k of cell P
// (ScalarKernels.ipp:239-447) into bterm[4k..4k+3], and its MassFluxBar:
//   INLET:  gradient cotangent SuBar[P]*w_b; MassFluxBar -(SuBar[P]*PhiFace) if mf < 0
//   OUTLET: (SuBar[P]*w_b) + PFbar*Xpn; phi term PFbar
//   WALL (enthalpy, flux region): phi term (Hcoef*SuBar[P]) + PFbar
// with w_b = Visac*S - VisFace*Xpn, Xpn = Xf - Cc[P]; -0.0 where a face adds nothing.
template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_bnd(DevMesh<T> m, AdjArgs<T> a, const T *pfbar, T *bterm) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m.F - m.Fi) return;
    const int g = m.Fi + k;
    const int ip = m.fconn[g].x, ib = ~m.fconn[g].y;
    const int ir = m.bregion[ib];
    const int it = m.rtype[ir];
    const T z = nzero<T>();
    V4<T> t{z, z, z, z};
    T mfb = 0.0;
    const T sb = a.subar[ip];
    if (it == CUPCFD_RTYPE_INLET || it == CUPCFD_RTYPE_OUTLET) {
        const V4<T> s = m.fgeoS[g];
        T Visac = it == CUPCFD_RTYPE_INLET ? a.s.visEffB[ib] : a.s.visEff[ip];
        Visac = scalar_visac(a.s, Visac);
        const v3<T> Xpn = sub3(xyz(m.fgeoX[g]), xyz(m.cgeo[ip]));
        const T VisFace = Visac * m.frl[g];
        const v3<T> w{Visac * s.x - VisFace * Xpn.x, Visac * s.y - VisFace * Xpn.y, Visac * s.z - VisFace * Xpn.z};
        if (it == CUPCFD_RTYPE_INLET) {
            t = V4<T>{sb * w.x, sb * w.y, sb * w.z, z};
            const T mf = a.s.massFlux[g];
            const T PhiFace = a.s.inletAssigns ? (T)0 : adj_phiface(m, a, g);
            if (mf < 0) mfb = -(sb * PhiFace);
        } else {
            const T pb = pfbar[k];
            t = V4<T>{sb * w.x + pb * Xpn.x, sb * w.y + pb * Xpn.y, sb * w.z + pb * Xpn.z, pb};
        }
    } else if (it == CUPCFD_RTYPE_WALL && a.s.wallActive && !m.radiab[ir] && m.rflux[ir]) {
        T Hcoef, VisFace;
        (void)scalar_wall_phiface(m, a.s, g, ib, ir, ip, Hcoef, VisFace);
        t.w = Hcoef * sb + pfbar[k];
    }
    *(V4<T> *)(bterm + 4 * (size_t)k) = t;
    a.mfbar[g] = mfb;
}

// ---------------------------------------------------------------------------
// Batched adjoint: K scalars sharing MassFlux, VisEff and RFace (no reference
// counterpart). PhiBar_k is, bit for bit, the single-scalar adjoint's for
// scalar k. MassFluxBar_f = rterm_f, then + (-(D_k*GammaBlend*(PhiFace -
// dfci_k))) for k = 0..K-1 in order, so that K = 1 gives the single-scalar
// result. Per-scalar arrays are scalar-innermost as for the batched forward
// kernels (SuBar and PhiBar [c*K + k]).
// ---------------------------------------------------------------------------
// The phi-free parts of an interior face's adjoint terms, once per face for a
// whole batch, with the expressions of the single-scalar pass A
template <class T> struct AdjFaceCoef {
    T wx, wy, wz, fmn, fmx;
};

// TS: storage type of the per-scalar arrays (phi, SuBar, PhiBar); the shared
// arrays, the scratch fields h and dir, and all arithmetic are T
template <class T, class TS = T> struct AdjBatchArgs {
    ScalarArgs<T, TS> s;  // phi ([c*K + k]), visEff, massFlux and the diffusivity parameters
    const TS *subar;
    const T *rbar;
    T *h, *dir;
    TS *phibar;
    T *mfbar;
    AdjFaceCoef<T> *coef;
};

template <class T, class TS>
__global__ __launch_bounds__(kBlock) void k_adj_coef(DevMesh<T> m, AdjBatchArgs<T, TS> a) {
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= m.Fi) return;
    const int2 cn = m.fconn[f];
    const int ip = cn.x, nn = cn.y;
    const V4<T> s = m.fgeoS[f];
    const T facn = s.w;
    const T facp = 1.0 - facn;
    T Visac = a.s.visEff[nn] * facn + a.s.visEff[ip] * facp;
    Visac = scalar_visac(a.s, Visac);
    const T VisFace = Visac * m.frl[f];
    const v3<T> Xpn = sub3(xyz(m.cgeo[nn]), xyz(m.cgeo[ip]));
    const T mf = a.s.massFlux[f];
    a.coef[f] = AdjFaceCoef<T>{Visac * s.x - VisFace * Xpn.x, Visac * s.y - VisFace * Xpn.y,
                               Visac * s.z - VisFace * Xpn.z, (T)fmin((double)mf, 0.0), (T)fmax((double)mf, 0.0)};
}

// RL > 0: every row has RL entries, all loaded (with the neighbours' values)
// before the first term; RL = 0: rows walked one entry at a time
template <class T, class TS, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_adj_a_batched(DevMesh<T> m, AdjBatchArgs<T, TS> a, int K, int per,
                                                          int nblk) {
    const BatchLane L = batch_lane<T, G>(m, K, per, nblk);
    if (!L.live) return;
    const int c = batch_uniform<G>(L.c), kk = L.k;
    T acc[4] = {0, 0, 0, 0};
    // boundary faces add nothing (the single-scalar fold adds -0.0)
    auto term = [&](int2 en, T sIp, T sNn) {
        const int f = en.x >> 1;
        const bool in = (en.x & 1) != 0;
        const AdjFaceCoef<T> q = a.coef[f];
        const T facn = m.fgeoS[f].w;
        const T facp = 1.0 - facn;
        const T D = sIp - sNn;
        const T Dg = D * a.s.gammaBlend;
        const T fac = in ? facn : facp;
        const T tt = D * fac;
        acc[0] = acc[0] + tt * q.wx;
        acc[1] = acc[1] + tt * q.wy;
        acc[2] = acc[2] + tt * q.wz;
        acc[3] = acc[3] + Dg * (in ? q.fmn : q.fmx);
    };
    if (RL) {
        const TS scs = a.subar[(size_t)c * K + kk];
        int2 en[RL > 0 ? RL : 1];
        TS so[RL > 0 ? RL : 1];  // storage type until first use; a boundary entry gathers the cell's own row
#pragma unroll
        for (int i = 0; i < RL; i++) en[i] = m.ent[c * RL + i];
#pragma unroll
        for (int i = 0; i < RL; i++) so[i] = a.subar[(size_t)(en[i].y >= 0 ? en[i].y : c) * K + kk];
        const T sc = (T)scs;
#pragma unroll
        for (int i = 0; i < RL; i++) {
            if (en[i].y < 0) continue;
            const bool in = (en[i].x & 1) != 0;
            const T sv = (T)so[i];
            term(en[i], in ? sv : sc, in ? sc : sv);
        }
    } else {
        int e0, n;
        row_range(m, c, e0, n);
        for (int e = e0; e < e0 + n; e++) {
            const int2 en = m.ent[e];
            if (en.y < 0) continue;
            const bool in = (en.x & 1) != 0;
            const int ip = in ? en.y : c, nn = in ? c : en.y;
            term(en, (T)a.subar[(size_t)ip * K + kk], (T)a.subar[(size_t)nn * K + kk]);
        }
    }
    const T fact = 1.0 / m.cvol[c];
    for (int j = 0; j < 3; j++) st_stream(a.h + ((size_t)c * 3 + j) * K + kk, acc[j] * fact);
    st_stream(a.dir + (size_t)c * K + kk, acc[3]);
}

template <class T, class TS, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_adj_b_batched(DevMesh<T> m, AdjBatchArgs<T, TS> a, int K, int per,
                                                          int nblk) {
    const BatchLane L = batch_lane<T, G>(m, K, per, nblk);
    if (!L.live) return;
    const int c = batch_uniform<G>(L.c), kk = L.k;
    T acc = a.dir[(size_t)c * K + kk];
    auto term = [&](int2 en, const v3<T> &hp, const v3<T> &hn) {
        const int f = en.x >> 1;
        const bool in = (en.x & 1) != 0;
        const V4<T> s = m.fgeoS[f];
        const T facn = s.w;
        const T facp = 1.0 - facn;
        const T fac = in ? facn : facp;
        const v3<T> hs{hp.x + hn.x, hp.y + hn.y, hp.z + hn.z};
        const T pF = dot3(hs, v3<T>{s.x, s.y, s.z});
        acc = acc + fac * pF;
    };
    if (RL) {
        const v3<T> hc = ld3s(a.h + kk, c, K);
        int2 en[RL > 0 ? RL : 1];
        v3<T> ho[RL > 0 ? RL : 1];
#pragma unroll
        for (int i = 0; i < RL; i++) en[i] = m.ent[c * RL + i];
#pragma unroll
        for (int i = 0; i < RL; i++)
            if (en[i].y >= 0) ho[i] = ld3s(a.h + kk, en[i].y, K);
#pragma unroll
        for (int i = 0; i < RL; i++) {
            if (en[i].y < 0) continue;
            const bool in = (en[i].x & 1) != 0;
            term(en[i], in ? ho[i] : hc, in ? hc : ho[i]);
        }
    } else {
        int e0, n;
        row_range(m, c, e0, n);
        for (int e = e0; e < e0 + n; e++) {
            const int2 en = m.ent[e];
            if (en.y < 0) continue;
            const bool in = (en.x & 1) != 0;
            const int ip = in ? en.y : c, nn = in ? c : en.y;
            term(en, ld3s(a.h + kk, ip, K), ld3s(a.h + kk, nn, K));
        }
    }
    st_stream(a.phibar + (size_t)c * K + kk, (TS)acc);
}

// Pass A for 64-lane groups on uniform rows when all K <= 64 scalars of a cell
// sit in one group, with the mass-flux cotangent folded in (no separate
// faces pass): each face's cell1 entry evaluates the face's K terms
// -(D*GammaBlend*(PhiFace - dfci)) from the SuBar difference pass A already
// forms and the upwind cell's phi, and puts them in LDS; after one barrier a
// lane per (cell, entry) adds the RFace term and then the K terms in scalar
// order, as k_adj_faces_batched does. A face whose cell1 entry is a boundary
// entry gets 0.
template <class T, class TS, int RL>
__global__ __launch_bounds__(kBlock) void k_adj_a_batched_mf(DevMesh<T> m, AdjBatchArgs<T, TS> a, int K, int per,
                                                             int nblk) {
    constexpr int G = 64, NC = kBlock / G, NS = NC * RL, LS = G + 1;  // LS: row stride, odd against bank conflicts
    __shared__ T lt[NS * LS];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int tid = threadIdx.x, cl = tid / G, kk = tid % G;
    const int c = __builtin_amdgcn_readfirstlane(lb * NC + cl);
    if (c < m.C) {
        const bool live = kk < K;
        const T PhiFace = m.carry_init;
        TS scs = 0, pcs = 0, so[RL], po[RL];  // storage type until first use
        int2 en[RL];
#pragma unroll
        for (int i = 0; i < RL; i++) en[i] = m.ent[c * RL + i];
        if (live) {
            scs = a.subar[(size_t)c * K + kk];
            pcs = a.s.phi[(size_t)c * K + kk];
#pragma unroll
            for (int i = 0; i < RL; i++) {
                const int o = en[i].y >= 0 ? en[i].y : c;  // a boundary entry gathers the cell's own row
                so[i] = a.subar[(size_t)o * K + kk];
                if (!(en[i].x & 1)) po[i] = a.s.phi[(size_t)o * K + kk];
            }
        }
        const T sc = (T)scs, pc = (T)pcs;
        T acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < RL; i++) {
            T t = 0.0;
            if (en[i].y >= 0 && live) {
                const int f = en[i].x >> 1;
                const bool in = (en[i].x & 1) != 0;
                const AdjFaceCoef<T> q = a.coef[f];
                const T facn = m.fgeoS[f].w;
                const T facp = 1.0 - facn;
                const T sv = (T)so[i];
                const T D = (in ? sv : sc) - (in ? sc : sv);
                const T Dg = D * a.s.gammaBlend;
                const T fac = in ? facn : facp;
                const T tt = D * fac;
                acc[0] = acc[0] + tt * q.wx;
                acc[1] = acc[1] + tt * q.wy;
                acc[2] = acc[2] + tt * q.wz;
                acc[3] = acc[3] + Dg * (in ? q.fmn : q.fmx);
                if (!in) {
                    const T mf = a.s.massFlux[f];
                    const bool up = mf < 0 || mf > 0;
                    const T dfci = up ? (mf < 0 ? (T)po[i] : pc) : (T)0.0;
                    t = -(Dg * (PhiFace - dfci));
                }
            }
            lt[(cl * RL + i) * LS + kk] = t;
        }
        if (live) {
            const T fact = 1.0 / m.cvol[c];
            for (int j = 0; j < 3; j++) st_stream(a.h + ((size_t)c * 3 + j) * K + kk, acc[j] * fact);
            st_stream(a.dir + (size_t)c * K + kk, acc[3]);
        }
    }
    __syncthreads();
    if (tid < NS) {
        const int c2 = lb * NC + tid / RL;
        if (c2 < m.C) {
            const int2 e = m.ent[c2 * RL + tid % RL];
            if (!(e.x & 1)) {
                const int f = e.x >> 1;
                T s = 0.0;
                if (e.y >= 0) {
                    const T mf = a.s.massFlux[f];
                    s = mf < 0 ? a.rbar[2 * (size_t)f + 1] : (mf > 0 ? -a.rbar[2 * (size_t)f] : (T)0.0);
                    const T *l = lt + tid * LS;
                    for (int k = 0; k < K; k++) s = s + l[k];
                }
                a.mfbar[f] = s;
            }
        }
    }
}

// Passes A and B for 4- and 16-lane groups on meshes whose rows all hold RL
// entries (a wave spans 16 or 4 cells, so its lanes cannot share scalar loads): the block's
// rows and the per-face values their interior entries use are loaded once
// into LDS, one lane per entry, instead of by all G lanes of each cell. The
// rows come before the first barrier; the face values while every lane's
// per-scalar gathers are in flight, before the second. Terms and their order
// are k_adj_a_batched's and k_adj_b_batched's.
// MF: K <= G, so each face's K mass-flux terms are formed here and added by a
// lane per (cell, entry) after a third barrier, as in k_adj_a_batched_mf.
template <class T, class TS, int G, int RL, bool MF>
__global__ __launch_bounds__(kBlock) void k_adj_a_batched_staged(DevMesh<T> m, AdjBatchArgs<T, TS> a, int K, int per,
                                                                 int nblk) {
    constexpr int NC = kBlock / G, NP = NC * RL, LS = G + 1;
    __shared__ int2 sen[NP];
    __shared__ AdjFaceCoef<T> sq[NP];
    __shared__ T sfn[NP];
    __shared__ T svol[NC];
    __shared__ T lt[MF ? NP * LS : 1];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cb = lb * NC, tid = threadIdx.x;
    for (int p = tid; p < NP; p += kBlock) {
        const int c = cb + p / RL;
        sen[p] = c < m.C ? m.ent[(size_t)c * RL + p % RL] : make_int2(0, -1);
    }
    __syncthreads();
    const int cl = tid / G, c = cb + cl, kk = blockIdx.y * G + tid % G;
    const bool live = c < m.C && kk < K;
    TS scs = 0, pcs = 0, so[RL], po[RL];  // storage type until first use
    if (live) {
        scs = a.subar[(size_t)c * K + kk];
        if (MF) pcs = a.s.phi[(size_t)c * K + kk];
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int2 en = sen[cl * RL + i];
            const int o = en.y >= 0 ? en.y : c;  // a boundary entry gathers the cell's own row
            so[i] = a.subar[(size_t)o * K + kk];
            if (MF && !(en.x & 1)) po[i] = a.s.phi[(size_t)o * K + kk];
        }
    }
    for (int p = tid; p < NP; p += kBlock) {
        const int2 en = sen[p];
        if (en.y >= 0) {
            const int f = en.x >> 1;
            sq[p] = a.coef[f];
            sfn[p] = m.fgeoS[f].w;
        }
    }
    for (int p = tid; p < NC; p += kBlock)
        if (cb + p < m.C) svol[p] = m.cvol[cb + p];
    __syncthreads();
    if (live) {
        const T PhiFace = m.carry_init;
        const T sc = (T)scs, pc = (T)pcs;
        T acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int2 en = sen[cl * RL + i];
            if (en.y < 0) continue;
            const bool in = (en.x & 1) != 0;
            const AdjFaceCoef<T> q = sq[cl * RL + i];
            const T facn = sfn[cl * RL + i];
            const T facp = 1.0 - facn;
            const T sv = (T)so[i];
            const T D = (in ? sv : sc) - (in ? sc : sv);
            const T Dg = D * a.s.gammaBlend;
            const T fac = in ? facn : facp;
            const T tt = D * fac;
            acc[0] = acc[0] + tt * q.wx;
            acc[1] = acc[1] + tt * q.wy;
            acc[2] = acc[2] + tt * q.wz;
            acc[3] = acc[3] + Dg * (in ? q.fmn : q.fmx);
            if (MF && !in) {
                const T mf = a.s.massFlux[en.x >> 1];
                const bool up = mf < 0 || mf > 0;
                const T dfci = up ? (mf < 0 ? (T)po[i] : pc) : (T)0.0;
                lt[(cl * RL + i) * LS + tid % G] = -(Dg * (PhiFace - dfci));
            }
        }
        const T fact = 1.0 / svol[cl];
        for (int j = 0; j < 3; j++) st_stream(a.h + ((size_t)c * 3 + j) * K + kk, acc[j] * fact);
        st_stream(a.dir + (size_t)c * K + kk, acc[3]);
    }
    if (!MF) return;
    __syncthreads();
    for (int p = tid; p < NP; p += kBlock) {
        const int2 e = sen[p];
        if (cb + p / RL >= m.C || (e.x & 1)) continue;
        const int f = e.x >> 1;
        T s = 0.0;
        if (e.y >= 0) {
            const T mf = a.s.massFlux[f];
            s = mf < 0 ? a.rbar[2 * (size_t)f + 1] : (mf > 0 ? -a.rbar[2 * (size_t)f] : (T)0.0);
            const T *l = lt + p * LS;
            for (int k = 0; k < K; k++) s = s + l[k];
        }
        a.mfbar[f] = s;
    }
}

template <class T, class TS, int G, int RL>
__global__ __launch_bounds__(kBlock) void k_adj_b_batched_staged(DevMesh<T> m, AdjBatchArgs<T, TS> a, int K, int per,
                                                                 int nblk) {
    constexpr int NC = kBlock / G, NP = NC * RL;
    __shared__ int2 sen[NP];
    __shared__ V4<T> ss[NP];
    const int lb = xcd_block(blockIdx.x, per);
    if (lb >= nblk) return;
    const int cb = lb * NC, tid = threadIdx.x;
    for (int p = tid; p < NP; p += kBlock) {
        const int c = cb + p / RL;
        sen[p] = c < m.C ? m.ent[(size_t)c * RL + p % RL] : make_int2(0, -1);
    }
    __syncthreads();
    const int cl = tid / G, c = cb + cl, kk = blockIdx.y * G + tid % G;
    const bool live = c < m.C && kk < K;
    T acc = 0;
    v3<T> hc{}, ho[RL];
    if (live) {
        acc = a.dir[(size_t)c * K + kk];
        hc = ld3s(a.h + kk, c, K);
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int2 en = sen[cl * RL + i];
            if (en.y >= 0) ho[i] = ld3s(a.h + kk, en.y, K);
        }
    }
    for (int p = tid; p < NP; p += kBlock) {
        const int2 en = sen[p];
        if (en.y >= 0) ss[p] = m.fgeoS[en.x >> 1];
    }
    __syncthreads();
    if (!live) return;
#pragma unroll
    for (int i = 0; i < RL; i++) {
        const int2 en = sen[cl * RL + i];
        if (en.y < 0) continue;
        const bool in = (en.x & 1) != 0;
        const v3<T> hp = in ? ho[i] : hc, hn = in ? hc : ho[i];
        const V4<T> s = ss[cl * RL + i];
        const T facn = s.w;
        const T facp = 1.0 - facn;
        const T fac = in ? facn : facp;
        const v3<T> hs{hp.x + hn.x, hp.y + hn.y, hp.z + hn.z};
        const T pF = dot3(hs, v3<T>{s.x, s.y, s.z});
        acc = acc + fac * pF;
    }
    st_stream(a.phibar + (size_t)c * K + kk, (TS)acc);
}

// The mass-flux cotangent of the batched adjoint, G lanes per face (G = 4,
// 16 or 64, the smallest that holds K; K > 64 takes chunks of 64): lane j of
// a face's group evaluates scalar base + j's term, reading the per-scalar rows
// coalesced, and puts it in LDS; every lane of the group then adds the terms
// in scalar order, each read at a group-uniform LDS address. So MassFluxBar
// is the RFace term followed by the K scalars' terms, added in order, as K
// single-scalar calls accumulate it.
template <class T, class TS, int G>
__global__ __launch_bounds__(kBlock) void k_adj_faces_batched(DevMesh<T> m, AdjBatchArgs<T, TS> a, int K) {
    constexpr int FPB = kBlock / G;  // faces per block
    __shared__ T lds[kBlock];
    const int grp = threadIdx.x / G, j = threadIdx.x % G;
    const int f = blockIdx.x * FPB + grp;
    const bool valid = f < m.F;
    const int2 cn = valid ? m.fconn[f] : make_int2(0, -1);
    const bool interior = valid && cn.y >= 0;
    const T mf = interior ? a.s.massFlux[f] : (T)0.0;
    const T PhiFace = m.carry_init;
    T acc = 0.0;
    if (interior) acc = mf < 0 ? a.rbar[2 * (size_t)f + 1] : (mf > 0 ? -a.rbar[2 * (size_t)f] : (T)0.0);
    const bool up = mf < 0 || mf > 0;
    const TS *sp = a.subar + (size_t)cn.x * K, *sn = a.subar + (size_t)max(cn.y, 0) * K;
    const TS *pu = a.s.phi + (size_t)(mf < 0 ? max(cn.y, 0) : cn.x) * K;  // the upwind cell's phi
    T *lg = lds + grp * G;
    for (int base = 0; base < K; base += G) {
        const int kk = base + j;
        T t = 0.0;
        if (interior && kk < K) {
            const T D = (T)sp[kk] - (T)sn[kk];
            const T Dg = D * a.s.gammaBlend;
            const T dfci = up ? (T)pu[kk] : (T)0.0;
            t = -(Dg * (PhiFace - dfci));
        }
        lg[j] = t;
        __syncthreads();
        const int cnt = min(G, K - base);
        if (interior)
            for (int i = 0; i < cnt; i++) acc = acc + lg[i];
        __syncthreads();
    }
    if (valid && j == 0) a.mfbar[f] = interior ? acc : (T)0.0;
}

// values in [0.5, 1.5) from a hash of the index: finite, non-zero inputs for timing runs
template <class T> __global__ __launch_bounds__(kBlock) void k_fill_timing(T *x, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 16;
    x[i] = (T)(0.5 + (double)(h & 0xffffu) / 65536.0);
}

// drop-in staging: permutations between the mesh's numbering and the internal
// one, run on the device so the host only makes contiguous copies
template <class T>
__global__ __launch_bounds__(kBlock) void k_stage_gather(T *__restrict__ dst, const T *__restrict__ raw,
                                                         const int *__restrict__ order, int64_t total, int ncomp,
                                                         int64_t stride) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int64_t k = i / ncomp;
    dst[i] = raw[(int64_t)order[k] * stride + (i - k * ncomp)];
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_stage_scatter(T *__restrict__ raw, const T *__restrict__ src,
                                                          const int *__restrict__ order, int64_t total, int ncomp) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int64_t k = i / ncomp;
    raw[(int64_t)order[k] * ncomp + (i - k * ncomp)] = src[i];
}
template <class T> int stage_gather(T *dst, const T *raw, const int *order, int n, int ncomp, int64_t stride) {
    const int64_t total = (int64_t)n * ncomp;
    if (total > 0)
        hipLaunchKernelGGL(k_stage_gather<T>, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           nullptr, dst, raw, order, total, ncomp, stride);
    return hipGetLastError() == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}
template <class T> int stage_scatter(T *raw, const T *src, const int *order, int n, int ncomp) {
    const int64_t total = (int64_t)n * ncomp;
    if (total > 0)
        hipLaunchKernelGGL(k_stage_scatter<T>, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           nullptr, raw, src, order, total, ncomp);
    return hipGetLastError() == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}
template int stage_gather<double>(double *, const double *, const int *, int, int, int64_t);
template int stage_gather<float>(float *, const float *, const int *, int, int, int64_t);
template int stage_scatter<double>(double *, const double *, const int *, int, int);
template int stage_scatter<float>(float *, const float *, const int *, int, int);

}  // namespace cupcfd_hip

// ===========================================================================
// Resident-mode C ABI
// ===========================================================================
using namespace cupcfd_hip;

namespace {

inline int blocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

inline int launch_ok() { return hipGetLastError() == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR; }

// Grid of `per` logical blocks per XCD covering `items` at `per_block` each
struct Grid {
    int blocks, per, nblk;
};
inline Grid xcd_grid(int64_t items, int per_block) {
    const int nblk = (int)((items + per_block - 1) / per_block);
    const int per = (nblk + kXcd - 1) / kXcd;
    return Grid{per * kXcd, per, nblk};
}

// Dynamic LDS of a tile kernel: nf cell fields, nt term arrays of tslots
// entries, the tile's CSR slots and (flags) one byte per slot; 0 if above a
// workgroup's 64 KiB
template <class T> size_t tile_lds(const DevMesh<T> &d, int nf, int nt, bool flags) {
    const size_t b = ((size_t)nf * kTileCells + (size_t)nt * d.tslots) * sizeof(T) + 2 * (size_t)d.tnnz +
                     (flags ? (size_t)d.tslots : 0);
    return b <= 65536 ? b : 0;
}

// Row width of the scalar and UVW tiles' [fields][W] LDS layout: a
// compile-time width when the mesh's tiles have at most kTileSlots1 (every
// tile of 4x4x4 hex cells has 240) or kTileSlots2 slots, else 0 (TC +
// tslots at run time).
constexpr int kTileSlots1 = 240, kTileSlots2 = 320;
static_assert(kTileSlots1 == 240, "k_uvw_tile's waves_per_eu names this width");
template <class T> int tile_width(const DevMesh<T> &d) {
    if (d.tslots <= kTileSlots1) return kTileCells + kTileSlots1;
    if (d.tslots <= kTileSlots2) return kTileCells + kTileSlots2;
    return 0;
}
// its dynamic LDS: nf rows of W columns, the tile's CSR slots, a flag byte per slot
template <class T> size_t tile_lds_cols(const DevMesh<T> &d, int nf) {
    const int w = tile_width(d);
    const size_t b = (size_t)nf * (w ? w : kTileCells + d.tslots) * sizeof(T) + 2 * (size_t)d.tnnz + d.tslots;
    return b <= 65536 ? b : 0;
}

// The UVW tile's split layout: table A of 16 rows x (TC + kSplitA) columns,
// table B of 9 rows x (kTileSlots1 - kSplitA), the CSR column codes, and a
// flag byte per table-A slot. kSplitA = 96 fits 4x4x4 hex tiles (48 owned
// faces to the three neighbours above, 48 halo faces from the three below;
// boundary faces take the place of either at the mesh's edges), which leaves
// 144 faces between cells of the tile.
constexpr int kSplitA = 96;
template <class T> size_t tile_lds_split(const DevMesh<T> &d) {
    return ((size_t)16 * (kTileCells + kSplitA) + (size_t)9 * (kTileSlots1 - kSplitA)) * sizeof(T) +
           2 * (size_t)d.tnnz + kSplitA;
}

// slot sl of tile t has a cell outside the tile, or is a boundary face
template <class T>
__device__ __forceinline__ bool slot_kind_a(const DevMesh<T> &m, const TileSlots &ts, int sl, int c0, int nc) {
    int2 cn;
    bool owned;
    (void)slot_conn(m, ts, sl, cn, owned);
    return cn.y < 0 || outside_cell(cn, c0, nc) >= 0;
}

// per tile (one workgroup): its counts of table-A and table-B slots, folded
// into mx[0], mx[1] by atomic maximum
template <class T> __global__ __launch_bounds__(kTileBlock) void k_split_count(DevMesh<T> m, int *mx) {
    __shared__ int cnt[2];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int c0 = t * kTileCells, nc = min(kTileCells, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    if (tid < 2) cnt[tid] = 0;
    __syncthreads();
    for (int sl = tid; sl < ts.nslot; sl += kTileBlock) atomicAdd(&cnt[slot_kind_a(m, ts, sl, c0, nc) ? 0 : 1], 1);
    __syncthreads();
    if (tid < 2) atomicMax(mx + tid, cnt[tid]);
}

// per tile (one workgroup; every tile has at most kTileBlock slots): the
// column of each slot, table-A slots in slot order from 0 and table-B slots
// in slot order from kSplitA, and the tile's CSR entries with the slot
// replaced by its column
template <class T>
__global__ __launch_bounds__(kTileBlock) void k_split_cols(DevMesh<T> m, uint8_t *col, uint16_t *es) {
    __shared__ int wa[kTileBlock / 64], wb[kTileBlock / 64];
    __shared__ uint8_t lc[kTileBlock];
    const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = t * kTileCells, nc = min(kTileCells, m.C - c0);
    const TileSlots ts = tile_slots(m, t);
    const bool live = tid < ts.nslot;
    const bool ka = live && slot_kind_a(m, ts, tid, c0, nc);
    const uint64_t ba = __ballot(ka), bb = __ballot(live && !ka);
    const uint64_t below = (1ull << lane) - 1;
    if (lane == 0) {
        wa[wave] = __popcll(ba);
        wb[wave] = __popcll(bb);
    }
    __syncthreads();
    int oa = 0, ob = 0;
    for (int w = 0; w < wave; w++) oa += wa[w], ob += wb[w];
    if (live) {
        const int c = ka ? oa + __popcll(ba & below) : kSplitA + ob + __popcll(bb & below);
        lc[tid] = (uint8_t)c;
        col[(size_t)t * m.tslots + tid] = (uint8_t)c;
    }
    __syncthreads();
    int e0, e1;
    if (m.rowlen) {
        e0 = c0 * m.rowlen;
        e1 = (c0 + nc) * m.rowlen;
    } else {
        e0 = m.row[c0];
        e1 = m.row[c0 + nc];
    }
    for (int i = tid; i < e1 - e0; i += kTileBlock) {
        const uint16_t v = m.eslot[e0 + i];
        es[e0 + i] = (uint16_t)(lc[v & 0x7fff] | (v & 0x8000));
    }
}


// Gather schedule of gradient, scalar and UVW: the kernel's autotuned schedule
// if any, else the mesh's. Automatic (schedule 0)
// and 64 take the cell-tile kernels when the mesh's tiles fit their slot
// index and `lds` bytes of LDS (set here); otherwise lanes per cell: the mesh's
// schedule if set (1 = one thread per cell), else the row length of uniform 4-
// or 6-face meshes (every lane busy), else 8 if every row fits, else one
// thread per cell. Returns kTile, 4, 6, 8, or 0 for one thread per cell.
constexpr int kTile = 64;   // the tile schedule's code in cupcfd_hip_mesh_set_schedule
constexpr int kColour = 2;  // colour groups (gradient; the other gathers take the automatic schedule)
enum { kGatherGradient = 0, kGatherScalar = 1, kGatherUvw = 2, kMassFaces = 3 };
template <class T> int gather_schedule(const Mesh<T> &M, int kind, int nf, int nt, bool flags, size_t &lds) {
    int l = M.schedule_k[kind] >= 0 ? M.schedule_k[kind] : M.schedule_lpc;
    lds = 0;
    if (l == kColour) {
        if (kind == kGatherGradient) return kColour;
        l = 0;
    }
    if (l == kTile || l == 0) {
        // gradient: nf cell fields and nt term rows of tslots; scalar / UVW: nf rows of W columns
        if (M.d.ntiles > 0) lds = kind == kGatherGradient ? tile_lds(M.d, nf, nt, flags) : tile_lds_cols(M.d, nf);
        if (lds) return kTile;
        l = 0;
    }
    if (l == 0) {
        if (M.d.rowlen == 4 || M.d.rowlen == 6) l = M.d.rowlen;
        else l = M.d.maxrow <= 4 ? 4 : (M.d.maxrow <= 8 ? 8 : 1);
    }
    if (l != 1 && M.d.maxrow > l) l = 1;
    return l == 4 || l == 6 || l == 8 ? l : 0;
}

template <class T> Mesh<T> *get(cupcfd_hip_mesh *h);
template <> Mesh<double> *get<double>(cupcfd_hip_mesh *h) { return h && h->precision == 64 ? h->m64 : nullptr; }
template <> Mesh<float> *get<float>(cupcfd_hip_mesh *h) { return h && h->precision == 32 ? h->m32 : nullptr; }

template <class T>
int dev_gradient(cupcfd_hip_mesh *h, int nGradient, const T *phi, const T *phiB, T *g, T *go, void *stream) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    hipStream_t s = (hipStream_t)stream;
    if (d.C == 0) return CUPCFD_E_SUCCESS;
    if (nGradient <= 0) {
        // the reference zeroes dPhidxo and leaves dPhidx untouched
        return hipMemsetAsync(go, 0, 3 * (size_t)d.C * sizeof(T), s) == hipSuccess ? CUPCFD_E_SUCCESS
                                                                                   : CUPCFD_E_ERROR;
    }
    // Sweeps before the last write alternately into two scratch fields (a
    // sweep reads every cell's previous gradient, so it cannot write in
    // place); the last sweep writes dPhidx and dPhidxo together.
    T *t[2] = {nullptr, nullptr};
    if (nGradient > 1) {
        t[0] = (T *)M->ws_get(42, 3 * (size_t)d.C * sizeof(T));
        t[1] = (T *)M->ws_get(43, 3 * (size_t)d.C * sizeof(T));
        if (!t[0] || !t[1]) return CUPCFD_E_ERROR;
    }
    size_t lds;
    // tile LDS: 7 cell fields and the 3 output components per cell, 3 term rows of tslots
    const int lpc = gather_schedule(*M, kGatherGradient, 10, 3, false, lds);
    const Grid gr = lpc == kTile ? xcd_grid(d.ntiles, 1) : xcd_grid(d.C, lpc ? kCellsPerBlock : kBlock);
    const int nthr = lpc == kTile ? kTileBlock : (lpc ? lpc * kCellsPerBlock : kBlock);
    T *acc = nullptr;
    if (lpc == kColour) {
        // the colouring is built on first use (allocates and synchronises)
        if (build_face_colours(*M) != CUPCFD_E_SUCCESS) return CUPCFD_E_ERROR;
        acc = (T *)M->ws_get(52, 3 * (size_t)d.C * sizeof(T));
        if (!acc) return CUPCFD_E_ERROR;
    }
    const T *prev = nullptr;
    for (int it = 0; it < nGradient; it++) {
        const bool last = (it == nGradient - 1);
        T *dst = last ? g : t[it & 1];
        GradArgs<T> ga{phi, phiB, prev, dst, last ? go : nullptr};
        const bool s0 = (it == 0);
        if (lpc == kColour) {
            hipLaunchKernelGGL(k_fill_zero<T>, dim3(blocks(3 * d.C)), dim3(kBlock), 0, s, 3 * (size_t)d.C, acc);
            for (int q = 0; q < M->ncolours; q++) {
                const int *fq = M->colour_faces + M->colour_off[q];
                const int nq = M->colour_off[q + 1] - M->colour_off[q];
                if (nq == 0) continue;
                if (s0) hipLaunchKernelGGL((k_gradient_colour<T, true>), dim3(blocks(nq)), dim3(kBlock), 0, s, d, ga, fq, nq, acc);
                else hipLaunchKernelGGL((k_gradient_colour<T, false>), dim3(blocks(nq)), dim3(kBlock), 0, s, d, ga, fq, nq, acc);
            }
            hipLaunchKernelGGL(k_gradient_colour_store<T>, dim3(blocks(d.C)), dim3(kBlock), 0, s, d, ga,
                               (const T *)acc);
        } else if (lpc == kTile) {
            if (s0) hipLaunchKernelGGL((k_gradient_tile<T, true>), gr.blocks, nthr, lds, s, d, ga, gr.per, gr.nblk);
            else hipLaunchKernelGGL((k_gradient_tile<T, false>), gr.blocks, nthr, lds, s, d, ga, gr.per, gr.nblk);
        } else if (lpc == 8) {
            if (s0) hipLaunchKernelGGL((k_gradient_vec<T, true, 8>), gr.blocks, nthr, 0, s, d, ga, gr.per, gr.nblk);
            else hipLaunchKernelGGL((k_gradient_vec<T, false, 8>), gr.blocks, nthr, 0, s, d, ga, gr.per, gr.nblk);
        } else if (lpc == 6) {
            if (s0) hipLaunchKernelGGL((k_gradient_vec<T, true, 6>), gr.blocks, nthr, 0, s, d, ga, gr.per, gr.nblk);
            else hipLaunchKernelGGL((k_gradient_vec<T, false, 6>), gr.blocks, nthr, 0, s, d, ga, gr.per, gr.nblk);
        } else if (lpc == 4) {
            if (s0) hipLaunchKernelGGL((k_gradient_vec<T, true, 4>), gr.blocks, nthr, 0, s, d, ga, gr.per, gr.nblk);
            else hipLaunchKernelGGL((k_gradient_vec<T, false, 4>), gr.blocks, nthr, 0, s, d, ga, gr.per, gr.nblk);
        } else {
            if (s0) hipLaunchKernelGGL((k_gradient<T, true>), gr.blocks, kBlock, 0, s, d, ga, gr.per, gr.nblk);
            else hipLaunchKernelGGL((k_gradient<T, false>), gr.blocks, kBlock, 0, s, d, ga, gr.per, gr.nblk);
        }
        if (hipGetLastError() != hipSuccess) return CUPCFD_E_ERROR;
        prev = dst;
    }
    return CUPCFD_E_SUCCESS;
}

template <class T>
int dev_mass(cupcfd_hip_mesh *h, const T *dudx, const T *dvdx, const T *dwdx, const T *dpdx, const T *den, T *denB,
             const T *u, const T *v, const T *w, T *massFlux, const T *p, const T *ar, T *su, T *rface, T small,
             int32_t *icinl, int32_t *icout, int32_t *icsym, int32_t *icwal, int sTE, int sTD, int sV, int sE,
             const T *teC, T *teB, const T *edC, T *edB, const T *vC, T *vB, const T *tC, T *tB, void *stream) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    MassArgs<T> a{dudx, dvdx, dwdx, dpdx, den, u, v, w, p, ar, denB, massFlux, su, rface, small, sTE, sTD, sV, sE,
                  teC, edC, vC, tC, teB, edB, vB, tB};
    const int nbb = blocks(d.nbc);
    // cell tiles (every interior face is owned by the tile of its cell1) for
    // the automatic and tile schedules, unless the autotuner timed one thread
    // per face faster on this mesh; one thread per face otherwise
    const int sch = M->schedule_k[kMassFaces] >= 0 ? M->schedule_k[kMassFaces] : M->schedule_lpc;
    if (d.ntiles && (sch == 0 || sch == kTile)) {
        const Grid gt = xcd_grid(d.ntiles, 1);
        static_assert(kTileBlock == kBlock, "boundary-cell blocks of k_mass_tile count kBlock cells");
        hipLaunchKernelGGL(k_mass_tile<T>, dim3(gt.blocks + nbb), dim3(kTileBlock), 0, (hipStream_t)stream, d, a,
                           gt.per, gt.nblk);
    } else {
        const Grid gi = xcd_grid(d.Fi, kBlock);
        if (gi.blocks + nbb > 0)
            hipLaunchKernelGGL(k_mass<T>, dim3(gi.blocks + nbb), dim3(kBlock), 0, (hipStream_t)stream, d, a, gi.per,
                               gi.nblk);
    }
    // the face loop's counters count boundary faces by region type
    if (icinl) *icinl += M->ntype[CUPCFD_RTYPE_INLET];
    if (icout) *icout += M->ntype[CUPCFD_RTYPE_OUTLET];
    if (icsym) *icsym += M->ntype[CUPCFD_RTYPE_SYMP];
    if (icwal) *icwal += M->ntype[CUPCFD_RTYPE_WALL];
    return launch_ok();
}

// The scalar loop's arguments for one call (ks = 1: one scalar; batches set
// their own stride in the kernel)
template <class T, class TS>
int scalar_args(Mesh<T> &M, ScalarArgs<T, TS> &a, const TS *Phi, TS *PhiB, const T *VisEff, TS *Au, TS *Su,
                const T *Den, const T *MF, const T *TE, const T *CpB, const T *VisEffB, T *RFace, const TS *dPhidx,
                int SolveTurb, int SolveEnthalpy, T vislam, int ivar, int VarT, T Sigma_T, T Prandtl, int VarTE,
                T Sigma_k, int VarED, T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T Small, T TMCmu) {
    a = ScalarArgs<T, TS>{};
    a.phi = Phi;
    a.visEff = VisEff;
    a.den = Den;
    a.massFlux = MF;
    a.te = TE;
    a.cpB = CpB;
    a.visEffB = VisEffB;
    a.dPhidx = dPhidx;
    a.phiB = PhiB;
    a.au = Au;
    a.su = Su;
    a.rface = RFace;
    a.solveTurb = SolveTurb;
    a.solveEnthalpy = SolveEnthalpy;
    a.ivar = ivar;
    a.varT = VarT;
    a.varTE = VarTE;
    a.varED = VarED;
    a.inletAssigns = !(ivar == VarT || ivar == VarTE || ivar == VarED);
    a.wallActive = (SolveEnthalpy && ivar == VarT) ? 1 : 0;
    a.visLam = vislam;
    a.sigmaT = Sigma_T;
    a.prandtl = Prandtl;
    a.sigmaK = Sigma_k;
    a.sigmaE = Sigma_e;
    a.sigmaS = Sigma_s;
    a.schmidt = Schmidt;
    if (SolveTurb) {
        a.visD1 = ivar == VarT ? Sigma_T : ivar == VarTE ? Sigma_k : ivar == VarED ? Sigma_e : Sigma_s;
        a.visD2 = ivar == VarT ? Prandtl : Schmidt;
        a.visHasD2 = !(ivar == VarTE || ivar == VarED);
    } else {
        a.visD1 = 1;
        a.visD2 = ivar == VarT ? Prandtl : Schmidt;
        a.visHasD2 = 1;
    }
    a.gammaBlend = GammaBlend;
    a.small = Small;
    // SLres and Cmu25 (ScalarKernels.ipp:394-395) are per-call constants
    a.slres = 9.24 * (pow((Prandtl / Sigma_T), 0.75) - 1.0) * (1.0 + 0.28 * exp(-0.007 * Prandtl / Sigma_T));
    a.cmu25 = pow(TMCmu, 0.25);
    a.ks = 1;
    a.write_state = 1;
    a.src = nullptr;
    a.cval = nullptr;
    a.bterm = nullptr;
    a.bflag = nullptr;
    if (M.d.typed) {
        a.src = scalar_carry_src<T>(M, a.inletAssigns != 0, a.wallActive != 0);
        if (!a.src) return CUPCFD_E_ERROR;
    }
    return CUPCFD_E_SUCCESS;
}

template <class T>
int dev_scalar(cupcfd_hip_mesh *h, const T *Phi, T *PhiB, const T *VisEff, T *Au, T *Su, const T *Den, const T *MF,
               const T *TE, const T *CpB, const T *VisEffB, T *RFace, const T *dPhidx, int SolveTurb,
               int SolveEnthalpy, T vislam, int ivar, int VarT, T Sigma_T, T Prandtl, int VarTE, T Sigma_k, int VarED,
               T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large, T TMCmu, void *stream) {
    (void)Large;
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    ScalarArgs<T> a;
    const int rc = scalar_args(*M, a, Phi, PhiB, VisEff, Au, Su, Den, MF, TE, CpB, VisEffB, RFace, dPhidx, SolveTurb,
                               SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e,
                               Sigma_s, Schmidt, GammaBlend, Small, TMCmu);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    if (d.C) {
        size_t lds;
        const int lpc = gather_schedule(*M, kGatherScalar, 8, 0, true, lds);
        const Grid gr = lpc == kTile ? xcd_grid(d.ntiles, 1) : xcd_grid(d.C, lpc ? kCellsPerBlock : kBlock);
        const int nthr = lpc == kTile ? kTileBlock : (lpc ? lpc * kCellsPerBlock : kBlock);
        hipStream_t st = (hipStream_t)stream;
        const int w = tile_width(d);
        if (lpc == kTile && d.typed && d.F > d.Fi) {
            const int nb = d.F - d.Fi;
            // the boundary faces' terms, PhiB and mesh state ahead of the tiles
            T *bterm = (T *)M->ws_get(55, 4 * (size_t)nb * sizeof(T));
            uint8_t *bflag = (uint8_t *)M->ws_get(56, (size_t)nb);
            if (!bterm || !bflag) return CUPCFD_E_ERROR;
            hipLaunchKernelGGL(k_scalar_bnd_terms<T>, dim3(blocks(nb)), dim3(kBlock), 0, st, d, a, bterm, bflag);
            a.bterm = bterm;
            a.bflag = bflag;
            if (a.src) {  // carried PhiFace per source face, read in phase A
                T *cval = (T *)M->ws_get(53, (size_t)nb * sizeof(T));
                if (!cval) return CUPCFD_E_ERROR;
                hipLaunchKernelGGL(k_scalar_carry_vals<T>, dim3(blocks(nb)), dim3(kBlock), 0, st, d, a, cval);
                a.cval = cval;
            }
        }
        if (lpc == kTile && w == kTileCells + kTileSlots1)
            hipLaunchKernelGGL((k_scalar_tile<T, kTileCells + kTileSlots1>), gr.blocks, nthr, lds, st, d, a, gr.per,
                               gr.nblk);
        else if (lpc == kTile && w == kTileCells + kTileSlots2)
            hipLaunchKernelGGL((k_scalar_tile<T, kTileCells + kTileSlots2>), gr.blocks, nthr, lds, st, d, a, gr.per,
                               gr.nblk);
        else if (lpc == kTile) hipLaunchKernelGGL((k_scalar_tile<T, 0>), gr.blocks, nthr, lds, st, d, a, gr.per, gr.nblk);
        else if (lpc == 8) hipLaunchKernelGGL((k_scalar_vec<T, 8>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
        else if (lpc == 6) hipLaunchKernelGGL((k_scalar_vec<T, 6>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
        else if (lpc == 4) hipLaunchKernelGGL((k_scalar_vec<T, 4>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
        else hipLaunchKernelGGL(k_scalar<T>, gr.blocks, kBlock, 0, st, d, a, gr.per, gr.nblk);
    }
    return launch_ok();
}

// Lanes per cell of the batched kernels: 4, 16 or 64 (the smallest that holds K)
inline int batch_group(int K) { return K <= 4 ? 4 : (K <= 16 ? 16 : 64); }
inline Grid batch_grid(int C, int G) { return xcd_grid(C, kBlock / G); }

template <class T, class TI, class TO>
void launch_gradient_batched(const DevMesh<T> &d, const BatchGradArgs<T, TI, TO> &ga, bool s0, int K, int G,
                             const Grid &gr, hipStream_t s) {
    const dim3 grid(gr.blocks, (K + G - 1) / G);
    const int rl = d.rowlen;
#define CUPCFD_GB(GG, RR)                                                                                        \
    if (s0)                                                                                                      \
        hipLaunchKernelGGL((k_gradient_batched<T, TI, TO, true, GG, RR>), grid, kBlock, 0, s, d, ga, K, gr.per,  \
                           gr.nblk);                                                                             \
    else                                                                                                         \
        hipLaunchKernelGGL((k_gradient_batched<T, TI, TO, false, GG, RR>), grid, kBlock, 0, s, d, ga, K, gr.per, \
                           gr.nblk);
    // the preloaded rows pay off when the row's connectivity is wave-uniform
    // (G = 64); 4- and 16-lane groups share each row through LDS on a first
    // sweep over uniform rows, and otherwise walk the row entry by entry
    if (G == 4 && s0 && (rl == 4 || rl == 6)) {
        if (rl == 4)
            hipLaunchKernelGGL((k_gradient_batched_staged<T, TI, TO, 4, 4>), grid, kBlock, 0, s, d, ga, K, gr.per,
                               gr.nblk);
        else
            hipLaunchKernelGGL((k_gradient_batched_staged<T, TI, TO, 4, 6>), grid, kBlock, 0, s, d, ga, K, gr.per,
                               gr.nblk);
    } else if (G == 4) {
        CUPCFD_GB(4, 0)
    } else if (G == 16 && s0 && (rl == 4 || rl == 6)) {
        if (rl == 4)
            hipLaunchKernelGGL((k_gradient_batched_staged<T, TI, TO, 16, 4>), grid, kBlock, 0, s, d, ga, K, gr.per,
                               gr.nblk);
        else
            hipLaunchKernelGGL((k_gradient_batched_staged<T, TI, TO, 16, 6>), grid, kBlock, 0, s, d, ga, K, gr.per,
                               gr.nblk);
    } else if (G == 16) {
        CUPCFD_GB(16, 0)
    } else if (rl == 6) {
        CUPCFD_GB(64, 6)
    } else if (rl == 4) {
        CUPCFD_GB(64, 4)
    } else {
        CUPCFD_GB(64, 0)
    }
#undef CUPCFD_GB
}

// Batched gradient sweeps; the sweeps before the last keep their gradient in
// T scratch, the last writes dPhidx and dPhidxo as TO
template <class T, class TI, class TO>
int dev_gradient_batched(cupcfd_hip_mesh *h, int nGradient, int K, const TI *phi, const TI *phiB, TO *g, TO *go,
                         void *stream) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    if (K < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const DevMesh<T> &d = M->d;
    hipStream_t s = (hipStream_t)stream;
    if (d.C == 0) return CUPCFD_E_SUCCESS;
    const size_t nv = 3 * (size_t)d.C * K;
    if (nGradient <= 0)
        return hipMemsetAsync(go, 0, nv * sizeof(TO), s) == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
    T *t[2] = {nullptr, nullptr};
    if (nGradient > 1) {
        t[0] = (T *)M->ws_get(44, nv * sizeof(T));
        t[1] = (T *)M->ws_get(45, nv * sizeof(T));
        if (!t[0] || !t[1]) return CUPCFD_E_ERROR;
    }
    const int G = batch_group(K);
    const Grid gr = batch_grid(d.C, G);
    const T *prev = nullptr;
    for (int it = 0; it < nGradient; it++) {
        const bool s0 = (it == 0);
        if (it == nGradient - 1) {
            launch_gradient_batched(d, BatchGradArgs<T, TI, TO>{phi, phiB, prev, g, go}, s0, K, G, gr, s);
        } else {
            launch_gradient_batched(d, BatchGradArgs<T, TI, T>{phi, phiB, prev, t[it & 1], nullptr}, s0, K, G, gr, s);
            prev = t[it & 1];
        }
        if (hipGetLastError() != hipSuccess) return CUPCFD_E_ERROR;
    }
    return CUPCFD_E_SUCCESS;
}


template <class T, class TS = T>
int dev_scalar_batched(cupcfd_hip_mesh *h, int K, const TS *Phi, TS *PhiB, const T *VisEff, TS *Au, TS *Su,
                       const T *Den, const T *MF, const T *TE, const T *CpB, const T *VisEffB, T *RFace,
                       const TS *dPhidx,
                       int SolveTurb, int SolveEnthalpy, T vislam, int ivar, int VarT, T Sigma_T, T Prandtl, int VarTE,
                       T Sigma_k, int VarED, T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large,
                       T TMCmu, void *stream) {
    (void)Large;
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    if (K < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const DevMesh<T> &d = M->d;
    ScalarArgs<T, TS> a;
    const int rc = scalar_args(*M, a, Phi, PhiB, VisEff, Au, Su, Den, MF, TE, CpB, VisEffB, RFace, dPhidx, SolveTurb,
                               SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e,
                               Sigma_s, Schmidt, GammaBlend, Small, TMCmu);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    if (d.C == 0) return CUPCFD_E_SUCCESS;
    hipStream_t st = (hipStream_t)stream;
    // the faces' phi-free coefficients, once for all K scalars (and RFace)
    ScalarFaceCoef<T> *coef = (ScalarFaceCoef<T> *)M->ws_get(46, (size_t)(d.Fi ? d.Fi : 1) * sizeof(ScalarFaceCoef<T>));
    if (!coef) return CUPCFD_E_ERROR;
    if (d.Fi) hipLaunchKernelGGL((k_scalar_coef<T, TS>), dim3(blocks(d.Fi)), dim3(kBlock), 0, st, d, a, coef);
    const int G = batch_group(K);
    const Grid gr = batch_grid(d.C, G);
    const dim3 grid(gr.blocks, (K + G - 1) / G);
    // untyped meshes with uniform rows: interior entries only, whole row preloaded
    const int RL = d.typed ? 0 : (d.rowlen == 6 ? 6 : (d.rowlen == 4 ? 4 : 0));
#define SCALAR_BATCHED_LAUNCH(GG)                                                                                     \
    do {                                                                                                             \
        if (RL == 6)                                                                                                 \
            hipLaunchKernelGGL((k_scalar_batched_plain<T, TS, GG, 6>), grid, kBlock, 0, st, d, a, coef, K, gr.per,   \
                               gr.nblk);                                                                             \
        else if (RL == 4)                                                                                            \
            hipLaunchKernelGGL((k_scalar_batched_plain<T, TS, GG, 4>), grid, kBlock, 0, st, d, a, coef, K, gr.per,   \
                               gr.nblk);                                                                             \
        else hipLaunchKernelGGL((k_scalar_batched<T, TS, GG>), grid, kBlock, 0, st, d, a, coef, K, gr.per, gr.nblk); \
    } while (0)
    if (G == 4 && RL == 6)
        hipLaunchKernelGGL((k_scalar_batched_staged<T, TS, 4, 6>), grid, kBlock, 0, st, d, a, coef, K, gr.per,
                           gr.nblk);
    else if (G == 4 && RL == 4)
        hipLaunchKernelGGL((k_scalar_batched_staged<T, TS, 4, 4>), grid, kBlock, 0, st, d, a, coef, K, gr.per,
                           gr.nblk);
    else if (G == 4) SCALAR_BATCHED_LAUNCH(4);
    else if (G == 16 && RL == 6)
        hipLaunchKernelGGL((k_scalar_batched_staged<T, TS, 16, 6>), grid, kBlock, 0, st, d, a, coef, K, gr.per,
                           gr.nblk);
    else if (G == 16 && RL == 4)
        hipLaunchKernelGGL((k_scalar_batched_staged<T, TS, 16, 4>), grid, kBlock, 0, st, d, a, coef, K, gr.per,
                           gr.nblk);
    else if (G == 16) SCALAR_BATCHED_LAUNCH(16);
    else SCALAR_BATCHED_LAUNCH(64);
#undef SCALAR_BATCHED_LAUNCH
    return launch_ok();
}

template <class T>
int dev_uvw(cupcfd_hip_mesh *h, T gammaBlend, T small, const T *u, const T *v, const T *w, T *uB, T *vB, T *wB,
            const T *visEff, const T *visEffB, T *massFlux, const T *dudx, const T *dvdx, const T *dwdx, T *rFace,
            T *su, T *sv, T *sw, T *au, T *av, T *aw, void *stream) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    UvwArgs<T> a{u, v, w, visEff, visEffB, dudx, dvdx, dwdx, uB, vB, wB, massFlux, rFace, su, sv, sw, au, av, aw,
                 gammaBlend, small};
    if (d.C) {
        size_t lds;
        const int lpc = gather_schedule(*M, kGatherUvw, 16, 0, true, lds);
        const Grid gr = lpc == kTile ? xcd_grid(d.ntiles, 1) : xcd_grid(d.C, lpc ? kCellsPerBlock : kBlock);
        const int nthr = lpc == kTile ? kTileBlock : (lpc ? lpc * kCellsPerBlock : kBlock);
        hipStream_t st = (hipStream_t)stream;
        const int w = tile_width(d);
        if (lpc == kTile && d.typed && d.F > d.Fi) {  // carried uFace/vFace/wFace per source face, read in phase A
            T *cuvw = (T *)M->ws_get(54, 3 * (size_t)(d.F - d.Fi) * sizeof(T));
            if (!cuvw) return CUPCFD_E_ERROR;
            hipLaunchKernelGGL(k_uvw_carry_vals<T>, dim3(blocks(d.F - d.Fi)), dim3(kBlock), 0, st, d, a, cuvw);
            a.cuvw = cuvw;
        }
        const uint8_t *nc8 = nullptr;
        const uint16_t *nc16 = nullptr;
        // the split layout (built for FP64 meshes only, where it lifts the
        // tile from four to five workgroups per CU; in FP32 the plain layout
        // already runs seven, and the split one ran 168 us against 160 at 128^3)
        if (lpc == kTile && w == kTileCells + kTileSlots1 && M->uvw_split)
            hipLaunchKernelGGL((k_uvw_tile<T, kTileCells + kTileSlots1, kSplitA>), gr.blocks, nthr,
                               tile_lds_split(d), st, d, a, gr.per, gr.nblk, (const uint8_t *)M->uvw_col,
                               (const uint16_t *)M->uvw_eslot);
        else if (lpc == kTile && w == kTileCells + kTileSlots1)
            hipLaunchKernelGGL((k_uvw_tile<T, kTileCells + kTileSlots1, 0>), gr.blocks, nthr, lds, st, d, a, gr.per,
                               gr.nblk, nc8, nc16);
        else if (lpc == kTile && w == kTileCells + kTileSlots2)
            hipLaunchKernelGGL((k_uvw_tile<T, kTileCells + kTileSlots2, 0>), gr.blocks, nthr, lds, st, d, a, gr.per,
                               gr.nblk, nc8, nc16);
        else if (lpc == kTile)
            hipLaunchKernelGGL((k_uvw_tile<T, 0, 0>), gr.blocks, nthr, lds, st, d, a, gr.per, gr.nblk, nc8, nc16);
        else if (lpc == 8) hipLaunchKernelGGL((k_uvw_vec<T, 8>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
        else if (lpc == 6) hipLaunchKernelGGL((k_uvw_vec<T, 6>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
        else if (lpc == 4) hipLaunchKernelGGL((k_uvw_vec<T, 4>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
        else hipLaunchKernelGGL(k_uvw<T>, gr.blocks, kBlock, 0, st, d, a, gr.per, gr.nblk);
    }
    return launch_ok();
}

// Lanes per cell of the adjoint gathers (the lane-per-face schedule; a
// tile schedule of the mesh maps to the row length)
template <class T> int adj_lanes(const Mesh<T> &M) {
    int l = M.schedule_lpc;
    if (l == 0 || l == kTile) {
        if (M.d.rowlen == 4 || M.d.rowlen == 6) l = M.d.rowlen;
        else l = M.d.maxrow <= 4 ? 4 : (M.d.maxrow <= 8 ? 8 : 1);
    }
    if (l != 1 && M.d.maxrow > l) l = 1;
    return l == 4 || l == 6 || l == 8 ? l : 0;
}

template <class T, bool CORR>
void adj_launch_b(int lpc, const Grid &gr, int nthr, hipStream_t st, const DevMesh<T> &d, const AdjArgs<T> &a) {
    if (lpc == 8) hipLaunchKernelGGL((k_adj_b_vec<T, 8, CORR>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
    else if (lpc == 6) hipLaunchKernelGGL((k_adj_b_vec<T, 6, CORR>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
    else if (lpc == 4) hipLaunchKernelGGL((k_adj_b_vec<T, 4, CORR>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
    else hipLaunchKernelGGL((k_adj_b<T, CORR>), gr.blocks, kBlock, 0, st, d, a, gr.per, gr.nblk);
}

// Pass A, then pass B once per gradient sweep, last sweep first: on the cell
// tiles when the mesh has them and its schedule is automatic or tiles, else
// on the mesh's lane schedule. a.h receives the last sweep's h; hb is a
// second buffer of 3C for nGradient > 1 (the sweeps alternate between the two).
template <class T> int adj_passes(const Mesh<T> &M, AdjArgs<T> a, int nGradient, T *hb, hipStream_t st) {
    const DevMesh<T> &d = M.d;
    const bool tiles = (M.schedule_lpc == 0 || M.schedule_lpc == kTile) && d.ntiles > 0 &&
                       tile_lds(d, 6, 8, false) > 0;
    const int lpc = adj_lanes(M);
    const Grid gr = xcd_grid(d.C, lpc ? kCellsPerBlock : kBlock);
    const int nthr = lpc ? lpc * kCellsPerBlock : kBlock;
    const Grid tg = xcd_grid(d.ntiles, 1);
    if (tiles)
        hipLaunchKernelGGL(k_adj_a_tile<T>, tg.blocks, kTileBlock, tile_lds(d, 6, 8, false), st, d, a, tg.per,
                           tg.nblk);
    else if (lpc == 8) hipLaunchKernelGGL((k_adj_a_vec<T, 8>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
    else if (lpc == 6) hipLaunchKernelGGL((k_adj_a_vec<T, 6>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
    else if (lpc == 4) hipLaunchKernelGGL((k_adj_a_vec<T, 4>), gr.blocks, nthr, 0, st, d, a, gr.per, gr.nblk);
    else hipLaunchKernelGGL(k_adj_a<T>, gr.blocks, kBlock, 0, st, d, a, gr.per, gr.nblk);
    a.binit = a.dir;
    T *const hbuf[2] = {a.h, hb};
    for (int s = nGradient; s >= 1; s--) {
        const int cur = (nGradient - s) & 1;
        a.h = hbuf[cur];
        a.hnext = hbuf[cur ^ 1];
        if (tiles && s > 1)
            hipLaunchKernelGGL((k_adj_b_tile<T, true>), tg.blocks, kTileBlock, tile_lds(d, 6, 8, false), st, d, a,
                               tg.per, tg.nblk);
        else if (tiles)
            hipLaunchKernelGGL((k_adj_b_tile<T, false>), tg.blocks, kTileBlock, tile_lds(d, 3, 2, false), st, d, a,
                               tg.per, tg.nblk);
        else if (s > 1) adj_launch_b<T, true>(lpc, gr, nthr, st, d, a);
        else adj_launch_b<T, false>(lpc, gr, nthr, st, d, a);
        a.binit = a.phibar;
    }
    return launch_ok();
}

template <class T>
int dev_scalar_adjoint(cupcfd_hip_mesh *h, const T *Phi, const T *VisEff, const T *MF, const T *SuBar,
                       const T *RFaceBar, int SolveTurb, T vislam, int ivar, int VarT, T Sigma_T, T Prandtl, int VarTE,
                       T Sigma_k, int VarED, T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T *PhiBar, T *MFBar,
                       void *stream, T *scratch_h = nullptr, T *scratch_dir = nullptr) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    if (d.typed) return CUPCFD_E_NOT_IMPLEMENTED;
    AdjArgs<T> a{};
    const int rc = scalar_args<T, T>(*M, a.s, Phi, nullptr, VisEff, nullptr, nullptr, nullptr, MF, nullptr, nullptr,
                               nullptr, nullptr, nullptr, SolveTurb, 0, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE,
                               Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, (T)0, (T)1);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    a.subar = SuBar;
    a.rbar = RFaceBar;
    a.phibar = PhiBar;
    a.mfbar = MFBar;
    if (d.C) {
        // h and dir are written for every cell by pass A before pass B reads them
        a.h = scratch_h ? scratch_h : (T *)M->ws_get(47, 3 * (size_t)d.C * sizeof(T));
        a.dir = scratch_dir ? scratch_dir : (T *)M->ws_get(48, (size_t)d.C * sizeof(T));
        if (!a.h || !a.dir) return CUPCFD_E_ERROR;
        return adj_passes(*M, a, 1, (T *)nullptr, (hipStream_t)stream);
    }
    return launch_ok();
}

// The same chain with nGradient gradient sweeps, on any mesh (typed
// boundaries, the carried PhiFace and the enthalpy wall included); dPhidx
// is the forward gradient the scalar loop read.
template <class T>
int dev_scalar_adjoint_full(cupcfd_hip_mesh *h, int nGradient, const T *Phi, const T *VisEff, const T *MF,
                            const T *dPhidx, const T *TE, const T *Den, const T *CpB, const T *VisEffB,
                            const T *SuBar, const T *RFaceBar, int SolveTurb, int SolveEnthalpy, T vislam, int ivar,
                            int VarT, T Sigma_T, T Prandtl, int VarTE, T Sigma_k, int VarED, T Sigma_e, T Sigma_s,
                            T Schmidt, T GammaBlend, T Small, T TMCmu, T *PhiBar, T *MFBar, void *stream) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    if (nGradient < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    if (!Phi || !VisEff || !MF || !SuBar || !RFaceBar || !PhiBar || !MFBar) return CUPCFD_E_NULL_PTR;
    if (d.typed && (!dPhidx || !TE || !Den || !CpB || !VisEffB)) return CUPCFD_E_NULL_PTR;
    AdjArgs<T> a{};
    const int rc = scalar_args<T, T>(*M, a.s, Phi, nullptr, VisEff, nullptr, nullptr, Den, MF, TE, CpB, VisEffB,
                                     nullptr, dPhidx, SolveTurb, SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl,
                                     VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, Small, TMCmu);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    a.subar = SuBar;
    a.rbar = RFaceBar;
    a.phibar = PhiBar;
    a.mfbar = MFBar;
    hipStream_t st = (hipStream_t)stream;
    if (!d.C) return launch_ok();
    a.h = (T *)M->ws_get(47, 3 * (size_t)d.C * sizeof(T));
    a.dir = (T *)M->ws_get(48, (size_t)d.C * sizeof(T));
    T *hb = nGradient > 1 ? (T *)M->ws_get(57, 3 * (size_t)d.C * sizeof(T)) : nullptr;
    if (!a.h || !a.dir || (nGradient > 1 && !hb)) return CUPCFD_E_ERROR;
    const int nb = d.F - d.Fi;
    if (d.typed && nb > 0) {
        // the carried PhiFace each source face leaves, the readers' cotangent
        // of it per source, then each boundary face's terms
        T *cval = (T *)M->ws_get(61, (size_t)nb * sizeof(T));
        T *pfbar = (T *)M->ws_get(58, (size_t)nb * sizeof(T));
        T *bterm = (T *)M->ws_get(59, 4 * (size_t)nb * sizeof(T));
        const int *rd_off = nullptr, *rd = nullptr;
        int nrd = 0;
        if (!cval || !pfbar || !bterm ||
            scalar_carry_readers<T>(*M, a.s.inletAssigns != 0, a.s.wallActive != 0, &rd_off, &rd, &nrd) !=
                CUPCFD_E_SUCCESS)
            return CUPCFD_E_ERROR;
        T *coef = (T *)M->ws_get(62, (size_t)d.F * sizeof(T));
        if (!coef) return CUPCFD_E_ERROR;
        hipLaunchKernelGGL(k_scalar_carry_vals<T>, dim3(blocks(nb)), dim3(kBlock), 0, st, d, a.s, cval);
        a.s.cval = cval;
        if (nrd) hipLaunchKernelGGL(k_adj_carry_terms<T>, dim3(blocks(d.F)), dim3(kBlock), 0, st, d, a, coef);
        hipLaunchKernelGGL(k_adj_carry_sum<T>, dim3((nb + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, st, nb,
                           rd_off, rd, (const T *)coef, pfbar);
        hipLaunchKernelGGL(k_adj_bnd<T>, dim3(blocks(nb)), dim3(kBlock), 0, st, d, a, (const T *)pfbar, bterm);
        a.bterm = bterm;
    }
    return adj_passes(*M, a, nGradient, hb, st);
}

template <class T, class TS = T>
int dev_scalar_adjoint_batched(cupcfd_hip_mesh *h, int K, const TS *Phi, const T *VisEff, const T *MF,
                               const TS *SuBar, const T *RFaceBar, int SolveTurb, T vislam, int ivar, int VarT,
                               T Sigma_T, T Prandtl, int VarTE, T Sigma_k, int VarED, T Sigma_e, T Sigma_s, T Schmidt,
                               T GammaBlend, TS *PhiBar, T *MFBar, void *stream) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    if (K < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const DevMesh<T> &d = M->d;
    if (d.typed) return CUPCFD_E_NOT_IMPLEMENTED;
    AdjBatchArgs<T, TS> a{};
    const int rc = scalar_args<T, TS>(*M, a.s, Phi, nullptr, VisEff, nullptr, nullptr, nullptr, MF, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, SolveTurb, 0, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE,
                                  Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, (T)0, (T)1);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    a.subar = SuBar;
    a.rbar = RFaceBar;
    a.phibar = PhiBar;
    a.mfbar = MFBar;
    hipStream_t st = (hipStream_t)stream;
    if (d.C) {
        a.h = (T *)M->ws_get(49, 3 * (size_t)d.C * K * sizeof(T));
        a.dir = (T *)M->ws_get(50, (size_t)d.C * K * sizeof(T));
        a.coef = (AdjFaceCoef<T> *)M->ws_get(51, (size_t)(d.Fi ? d.Fi : 1) * sizeof(AdjFaceCoef<T>));
        if (!a.h || !a.dir || !a.coef) return CUPCFD_E_ERROR;
        if (d.Fi) hipLaunchKernelGGL((k_adj_coef<T, TS>), dim3(blocks(d.Fi)), dim3(kBlock), 0, st, d, a);
        const int G = batch_group(K);
        const Grid gr = batch_grid(d.C, G);
        const dim3 grid(gr.blocks, (K + G - 1) / G);
        // uniform 6- or 4-face rows: each row loaded whole before its first term
        const int RL = d.rowlen == 6 ? 6 : (d.rowlen == 4 ? 4 : 0);
#define ADJ_BATCHED_LAUNCH(GG, RR)                                                                               \
    do {                                                                                                          \
        hipLaunchKernelGGL((k_adj_a_batched<T, TS, GG, RR>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);         \
        hipLaunchKernelGGL((k_adj_b_batched<T, TS, GG, RR>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);         \
    } while (0)
#define ADJ_BATCHED_ROWS(GG)                                                                                     \
    do {                                                                                                          \
        if (RL == 6) ADJ_BATCHED_LAUNCH(GG, 6);                                                                   \
        else if (RL == 4) ADJ_BATCHED_LAUNCH(GG, 4);                                                              \
        else ADJ_BATCHED_LAUNCH(GG, 0);                                                                           \
    } while (0)
        if (G == 64 && K <= 64 && RL == 6) {
            hipLaunchKernelGGL((k_adj_a_batched_mf<T, TS, 6>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
            hipLaunchKernelGGL((k_adj_b_batched<T, TS, 64, 6>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
        } else if (G == 64 && K <= 64 && RL == 4) {
            hipLaunchKernelGGL((k_adj_a_batched_mf<T, TS, 4>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
            hipLaunchKernelGGL((k_adj_b_batched<T, TS, 64, 4>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
        } else if (G == 4 && RL == 6) {
            hipLaunchKernelGGL((k_adj_a_batched_staged<T, TS, 4, 6, true>), grid, kBlock, 0, st, d, a, K, gr.per,
                               gr.nblk);
            hipLaunchKernelGGL((k_adj_b_batched_staged<T, TS, 4, 6>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
        } else if (G == 4 && RL == 4) {
            hipLaunchKernelGGL((k_adj_a_batched_staged<T, TS, 4, 4, true>), grid, kBlock, 0, st, d, a, K, gr.per,
                               gr.nblk);
            hipLaunchKernelGGL((k_adj_b_batched_staged<T, TS, 4, 4>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
        } else if (G == 4) ADJ_BATCHED_ROWS(4);
        else if (G == 16 && RL == 6) {
            // G == 16 means 4 < K <= 16 (batch_group): one chunk of scalars, so pass A
            // writes the mass-flux cotangent (mf_in_a below)
            hipLaunchKernelGGL((k_adj_a_batched_staged<T, TS, 16, 6, true>), grid, kBlock, 0, st, d, a, K, gr.per,
                               gr.nblk);
            hipLaunchKernelGGL((k_adj_b_batched_staged<T, TS, 16, 6>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
        } else if (G == 16 && RL == 4) {
            // G == 16 means 4 < K <= 16 (batch_group): one chunk of scalars, so pass A
            // writes the mass-flux cotangent (mf_in_a below)
            hipLaunchKernelGGL((k_adj_a_batched_staged<T, TS, 16, 4, true>), grid, kBlock, 0, st, d, a, K, gr.per,
                               gr.nblk);
            hipLaunchKernelGGL((k_adj_b_batched_staged<T, TS, 16, 4>), grid, kBlock, 0, st, d, a, K, gr.per, gr.nblk);
        } else if (G == 16) ADJ_BATCHED_ROWS(16);
        else ADJ_BATCHED_ROWS(64);
#undef ADJ_BATCHED_ROWS
#undef ADJ_BATCHED_LAUNCH
    }
    // the mass-flux cotangent, unless pass A has written it (one chunk of scalars, uniform rows)
    const bool mf_in_a = d.C && K <= 64 && (d.rowlen == 6 || d.rowlen == 4);
    if (d.F && !mf_in_a) {
        const int G = batch_group(K);
        const dim3 grid((unsigned)((d.F + kBlock / G - 1) / (kBlock / G)));
        if (G == 4) hipLaunchKernelGGL((k_adj_faces_batched<T, TS, 4>), grid, dim3(kBlock), 0, st, d, a, K);
        else if (G == 16) hipLaunchKernelGGL((k_adj_faces_batched<T, TS, 16>), grid, dim3(kBlock), 0, st, d, a, K);
        else hipLaunchKernelGGL((k_adj_faces_batched<T, TS, 64>), grid, dim3(kBlock), 0, st, d, a, K);
    }
    return launch_ok();
}

// Per-kernel schedule choice by timing: each candidate schedule (cell tiles,
// the lane-per-face width the mesh allows, one thread per cell) of the
// gradient, scalar and UVW gathers, and cell tiles or one thread per face for
// the mass flux, runs on scratch inputs on a private stream (one warm-up, then
// the best of three timed calls) and the fastest is kept.
// The mesh's mutable boundary and region state is saved and restored around
// the runs. Every schedule gives the same bits, so only speed is chosen.
template <class T> int autotune(cupcfd_hip_mesh *h, int32_t *chosen) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const DevMesh<T> &d = M->d;
    const size_t C = d.C, F = d.F, B = d.B, R = d.R;
    int cand[4];
    int nc = 0;
    cand[nc++] = kTile;
    if (d.rowlen == 4 || d.rowlen == 6) cand[nc++] = d.rowlen;
    else if (d.maxrow <= 8) cand[nc++] = d.maxrow <= 4 ? 4 : 8;
    cand[nc++] = 1;
    cand[nc++] = kColour;  // tried for the gradient only
    if (C == 0) {
        for (int k = 0; k < 4; k++) {
            M->schedule_k[k] = -1;
            if (chosen && k < 3) chosen[k] = 0;
        }
        return CUPCFD_E_SUCCESS;
    }
    // scratch: 24 cell arrays, 8 boundary arrays, 3 face arrays, and the saved state
    const size_t nstate = 5 * B + 3 * B + R + 3 * R;
    const size_t n = 24 * C + 8 * B + 3 * F + nstate;
    T *pool = nullptr;
    if (hipMalloc(&pool, (n ? n : 1) * sizeof(T)) != hipSuccess) return CUPCFD_E_ERROR;
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess) {
        (void)hipFree(pool);
        return CUPCFD_E_ERROR;
    }
    T *state = pool + 24 * C + 8 * B + 3 * F;
    hipLaunchKernelGGL(k_fill_timing<T>, dim3((unsigned)((n - nstate + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                       pool, n - nstate);
    T *sv[8] = {d.byplus, d.buplus, d.bq, d.bh, d.bt, d.bshear, d.rylog, d.rforcet};
    const size_t sn[8] = {B, B, B, B, B, 3 * B, R, 3 * R};
    {
        T *q = state;
        for (int k = 0; k < 8; k++) {
            if (sn[k]) (void)hipMemcpyAsync(q, sv[k], sn[k] * sizeof(T), hipMemcpyDeviceToDevice, st);
            q += sn[k];
        }
    }
    T *c = pool, *b = pool + 24 * C, *f = pool + 24 * C + 8 * B;
    auto cell = [&](int i) { return c + (size_t)i * C; };  // i < 24 (a 3-vector takes 3)
    auto bnd = [&](int i) { return b + (size_t)i * B; };
    T *mf = f, *rface = f + F;
    // scalar loop parameters of BenchmarkKernels (BenchmarkKernels.cpp:677-695)
    const T sp[] = {(T)0.001, (T)1.4, (T)1.2e-3, (T)1.5, (T)1.6, (T)1.7, (T)1.4e-3, (T)1.1, (T)1e-18, (T)1e18, (T)1.8};
    auto run = [&](int kind) -> int {
        if (kind == kGatherGradient)
            return dev_gradient<T>(h, 1, cell(0), bnd(0), cell(1), cell(4), st);
        if (kind == kGatherScalar)
            return dev_scalar<T>(h, cell(0), bnd(1), cell(7), cell(8), cell(9), cell(10), mf, cell(11), bnd(2),
                                 bnd(3), rface, cell(12), 1, 0, sp[0], 1, 2, sp[1], sp[2], 3, sp[3], 4, sp[4], sp[5],
                                 sp[6], sp[7], sp[8], sp[9], sp[10], st);
        if (kind == kMassFaces)
            return dev_mass<T>(h, cell(1), cell(4), cell(12), cell(21), cell(7), bnd(0), cell(0), cell(8), cell(9), mf,
                               cell(10), cell(11), cell(15), rface, (T)1e-18, nullptr, nullptr, nullptr, nullptr, 0, 0,
                               0, 0, cell(16), bnd(1), cell(17), bnd(2), cell(18), bnd(3), cell(19), bnd(4), st);
        return dev_uvw<T>(h, (T)0.9, (T)1e-18, cell(0), cell(7), cell(10), bnd(4), bnd(5), bnd(6), cell(11), bnd(7),
                          mf, cell(1), cell(4), cell(12), rface, cell(15), cell(16), cell(17), cell(18), cell(19),
                          cell(20), st);
    };
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int rc = CUPCFD_E_SUCCESS;
    for (int kind = 0; kind < 4 && rc == CUPCFD_E_SUCCESS; kind++) {
        float best = 0;
        int pick = cand[0];
        for (int j = 0; j < nc && rc == CUPCFD_E_SUCCESS; j++) {
            if (cand[j] == kColour && kind != kGatherGradient) continue;
            if (kind == kMassFaces && cand[j] != kTile && cand[j] != 1) continue;  // tiles or one thread per face
            M->schedule_k[kind] = cand[j];
            rc = run(kind);  // warm-up (and lazily built tables)
            float tmin = 0;
            for (int r = 0; r < 3 && rc == CUPCFD_E_SUCCESS; r++) {
                (void)hipEventRecord(e0, st);
                rc = run(kind);
                (void)hipEventRecord(e1, st);
                (void)hipEventSynchronize(e1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (r == 0 || ms < tmin) tmin = ms;
            }
            if (j == 0 || tmin < best) {
                best = tmin;
                pick = cand[j];
            }
        }
        M->schedule_k[kind] = pick;
        if (chosen && kind < 3) chosen[kind] = pick;
    }
    {
        T *q = state;
        for (int k = 0; k < 8; k++) {
            if (sn[k]) (void)hipMemcpyAsync(sv[k], q, sn[k] * sizeof(T), hipMemcpyDeviceToDevice, st);
            q += sn[k];
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) rc = CUPCFD_E_ERROR;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(pool);
    if (rc != CUPCFD_E_SUCCESS)
        for (int &k : M->schedule_k) k = -1;
    return rc;
}

}  // namespace

namespace cupcfd_hip {

template <class T> void tile_split_build(Mesh<T> &M) {
    const DevMesh<T> &d = M.d;
    M.uvw_split = 0;
    if (sizeof(T) != 8 || d.ntiles <= 0 || tile_width(d) != kTileCells + kTileSlots1) return;  // FP64 only
    int *mx = nullptr;
    if (hipMalloc(&mx, 2 * sizeof(int)) != hipSuccess) return;
    int h[2] = {0, 0};
    bool ok = hipMemcpy(mx, h, sizeof h, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_split_count<T>, dim3(d.ntiles), dim3(kTileBlock), 0, 0, d, mx);
        ok = hipGetLastError() == hipSuccess && hipMemcpy(h, mx, sizeof h, hipMemcpyDeviceToHost) == hipSuccess;
    }
    (void)hipFree(mx);
    M.split_na = h[0];
    M.split_nb = h[1];
    if (!ok || h[0] > kSplitA || h[1] > kTileSlots1 - kSplitA || tile_lds_split(d) > 65536) return;
    if (!M.uvw_col && hipMalloc(&M.uvw_col, (size_t)d.ntiles * d.tslots) != hipSuccess) return;
    if (!M.uvw_eslot && hipMalloc(&M.uvw_eslot, (size_t)d.nnz * sizeof(uint16_t)) != hipSuccess) return;
    hipLaunchKernelGGL(k_split_cols<T>, dim3(d.ntiles), dim3(kTileBlock), 0, 0, d, M.uvw_col, M.uvw_eslot);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return;
    M.uvw_split = 1;
}
template void tile_split_build<double>(Mesh<double> &);
template void tile_split_build<float>(Mesh<float> &);

}  // namespace cupcfd_hip

// A set of device meshes of one precision for persistent multi-mesh launches
struct cupcfd_hip_meshset {
    int precision;
    int n;
    void *meshes;      // DevMesh<T>[n] on the device
    int *unit0[3];     // per lane-group width G = 4, 16, 64: prefix of work units, n + 1 entries, device
    int nunits[3];
    int rowlen;        // common row length of every member, else 0
};

namespace {

inline int group_index(int G) { return G == 4 ? 0 : (G == 16 ? 1 : 2); }

template <class T> int meshset_build(cupcfd_hip_mesh *const *meshes, int n, cupcfd_hip_meshset *set) {
    std::vector<DevMesh<T>> d(n);
    int rl = -1;
    for (int i = 0; i < n; i++) {
        Mesh<T> *M = get<T>(meshes[i]);
        if (!M) return CUPCFD_E_NULL_PTR;
        d[i] = M->d;
        rl = (rl == -1 || rl == M->d.rowlen) ? M->d.rowlen : 0;
    }
    set->rowlen = rl < 0 ? 0 : rl;
    if (hipMalloc(&set->meshes, (size_t)n * sizeof(DevMesh<T>)) != hipSuccess) return CUPCFD_E_ERROR;
    if (hipMemcpy(set->meshes, d.data(), (size_t)n * sizeof(DevMesh<T>), hipMemcpyHostToDevice) != hipSuccess)
        return CUPCFD_E_ERROR;
    const int Gs[3] = {4, 16, 64};
    for (int g = 0; g < 3; g++) {
        std::vector<int> u(n + 1, 0);
        const int cpb = kBlock / Gs[g];
        for (int i = 0; i < n; i++) {
            const int64_t v = (int64_t)u[i] + (d[i].C + cpb - 1) / cpb;
            if (v >= INT32_MAX) return CUPCFD_E_CONFIG_INVALID_VALUE;
            u[i + 1] = (int)v;
        }
        set->nunits[g] = u[n];
        if (hipMalloc(&set->unit0[g], (size_t)(n + 1) * sizeof(int)) != hipSuccess) return CUPCFD_E_ERROR;
        if (hipMemcpy(set->unit0[g], u.data(), (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return CUPCFD_E_ERROR;
    }
    return CUPCFD_E_SUCCESS;
}

template <class T>
int dev_gradient_batched_set(cupcfd_hip_meshset *set, int K, const T *const *phi, const T *const *phiB,
                             T *const *g, T *const *go, void *stream) {
    if (!set || !phi || !phiB || !g) return CUPCFD_E_NULL_PTR;
    if (set->precision != (int)(8 * sizeof(T))) return CUPCFD_E_NULL_PTR;
    if (K < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const int G = batch_group(K), gi = group_index(G);
    const int U = set->nunits[gi];
    if (U == 0) return CUPCFD_E_SUCCESS;
    MeshSetArgs<T> a{(const DevMesh<T> *)set->meshes, set->unit0[gi], set->n, U};
    hipStream_t s = (hipStream_t)stream;
    const int rl = set->rowlen;
    // persistent grid: as many workgroups as can be resident at once (every
    // range starts at once; a larger grid would leave ranges waiting for a
    // slot), a multiple of the 8 XCDs, and no more than the units
    auto grid_for = [&](const void *kernel) {
        int dev = 0, cus = 256, per_cu = 1;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0);
        const int cap = std::max(1, cus * std::max(1, per_cu)) / kXcd * kXcd;
        const int P = std::max(kXcd, std::min(cap, ((U + kXcd - 1) / kXcd) * kXcd));
        return dim3(P, (K + G - 1) / G);
    };
#define CUPCFD_GS(GG, RR)                                                                                      \
    {                                                                                                          \
        const void *kfn = (const void *)k_gradient_batched_set<T, T, T, GG, RR>;                               \
        hipLaunchKernelGGL((k_gradient_batched_set<T, T, T, GG, RR>), grid_for(kfn), kBlock, 0, s, a, phi, phiB, \
                           g, go, K);                                                                          \
    }
    if (G == 4) {
        CUPCFD_GS(4, 0)
    } else if (G == 16) {
        CUPCFD_GS(16, 0)
    } else if (rl == 6) {
        CUPCFD_GS(64, 6)
    } else if (rl == 4) {
        CUPCFD_GS(64, 4)
    } else {
        CUPCFD_GS(64, 0)
    }
#undef CUPCFD_GS
    return launch_ok();
}

int permute(cupcfd_hip_mesh *h, bool cells, const void *src, void *dst, int ncomp, int eb, void *stream, int dir) {
    if (!h || !src || !dst) return CUPCFD_E_NULL_PTR;
    if (ncomp < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    int n;
    const std::vector<int32_t> *ord;
    if (h->precision == 64) {
        n = cells ? h->m64->d.C : h->m64->d.F;
        ord = cells ? &h->m64->cell_order : &h->m64->face_order;
    } else {
        n = cells ? h->m32->d.C : h->m32->d.F;
        ord = cells ? &h->m32->cell_order : &h->m32->face_order;
    }
    if (n == 0) return CUPCFD_E_SUCCESS;
    // device copy of the order, cached in a workspace slot
    void *dord;
    int slot = cells ? 40 : 41;
    if (h->precision == 64) dord = h->m64->ws_get(slot, n * sizeof(int));
    else dord = h->m32->ws_get(slot, n * sizeof(int));
    if (!dord) return CUPCFD_E_ERROR;
    if (hipMemcpyAsync(dord, ord->data(), n * sizeof(int), hipMemcpyHostToDevice, (hipStream_t)stream) !=
        hipSuccess)
        return CUPCFD_E_ERROR;
    hipStream_t s = (hipStream_t)stream;
    if (eb == 8)
        hipLaunchKernelGGL(k_permute<uint64_t>, dim3(blocks(n)), dim3(kBlock), 0, s, n, ncomp, (const int *)dord,
                           (const uint64_t *)src, (uint64_t *)dst, dir);
    else if (eb == 4)
        hipLaunchKernelGGL(k_permute<uint32_t>, dim3(blocks(n)), dim3(kBlock), 0, s, n, ncomp, (const int *)dord,
                           (const uint32_t *)src, (uint32_t *)dst, dir);
    else
        return CUPCFD_E_CONFIG_INVALID_VALUE;
    return launch_ok();
}

}  // namespace

// ---------------------------------------------------------------------------
// Fixed device scratch of the allocation-free adjoint entry point
// (cupcfd_dev_scalar_adjoint_noalloc_f64): the pass-A fields h (3 per cell)
// and dir (1 per cell), 32 bytes per cell, for meshes of up to 2,097,152
// cells (a 128^3 grid). Each GPU the library's code object is loaded on holds
// its own copy until the process ends. Host code uses only the per-device
// address hipGetSymbolAddress returns, never the array's name.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// scratch's address on the current device, looked up once per device
int scratch_on_current_device(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return CUPCFD_E_ERROR;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CUPCFD_E_ERROR;
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return CUPCFD_E_SUCCESS;
}
}  // namespace

extern "C" {

int cupcfd_dev_cells_to_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                 int32_t elem_bytes, void *stream) {
    return permute(mesh, true, src, dst, ncomp, elem_bytes, stream, 1);
}
int cupcfd_dev_cells_from_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                   int32_t elem_bytes, void *stream) {
    return permute(mesh, true, src, dst, ncomp, elem_bytes, stream, 0);
}
int cupcfd_dev_faces_to_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                 int32_t elem_bytes, void *stream) {
    return permute(mesh, false, src, dst, ncomp, elem_bytes, stream, 1);
}
int cupcfd_dev_faces_from_internal(cupcfd_hip_mesh *mesh, const void *src, void *dst, int32_t ncomp,
                                   int32_t elem_bytes, void *stream) {
    return permute(mesh, false, src, dst, ncomp, elem_bytes, stream, 0);
}

#define DEV_API(T, S)                                                                                              \
    int cupcfd_dev_gradient_phi_gauss_dolfyn_##S(cupcfd_hip_mesh *mesh, int32_t nGradient, const T *phiCell,       \
                                                 const T *phiBoundary, T *dPhidxCell, T *dPhidxoCell,               \
                                                 void *stream) {                                                   \
        return dev_gradient<T>(mesh, nGradient, phiCell, phiBoundary, dPhidxCell, dPhidxoCell, stream);       \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_face_loop_##S(                                                                        \
        cupcfd_hip_mesh *mesh, const T *dudx, const T *dvdx, const T *dwdx, const T *dpdx, const T *denCell,       \
        T *denBoundary, const T *uCell, const T *vCell, const T *wCell, T *massFlux, const T *p, const T *ar,      \
        T *su, T *rface, T small, int32_t *icinl, int32_t *icout, int32_t *icsym, int32_t *icwal,                  \
        int32_t solveTurbEnergy, int32_t solveTurbDiss, int32_t solveVisc, int32_t solveEnthalpy,                  \
        const T *teCell, T *teBoundary, const T *edCell, T *edBoundary, const T *viseffCell, T *viseffBoundary,    \
        const T *tCell, T *tBoundary, void *stream) {                                                              \
        return dev_mass<T>(mesh, dudx, dvdx, dwdx, dpdx, denCell, denBoundary, uCell, vCell, wCell, massFlux, p,   \
                           ar, su, rface, small, icinl, icout, icsym, icwal, solveTurbEnergy, solveTurbDiss,       \
                           solveVisc, solveEnthalpy, teCell, teBoundary, edCell, edBoundary, viseffCell,           \
                           viseffBoundary, tCell, tBoundary, stream);                                              \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_boundary_loop1_##S(cupcfd_hip_mesh *mesh, const T *massFlux, T *flowin,               \
                                                void *stream) {                                                    \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        hipLaunchKernelGGL(k_mass_bl1<T>, dim3(1), dim3(kSeqBlock), 0, (hipStream_t)stream, M->d, massFlux,        \
                           flowin);                                                                                \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_boundary_loop2_##S(cupcfd_hip_mesh *mesh, const T *massFlux, T *flowRegion,           \
                                                T *flowout, void *stream) {                                        \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!region_boundaries(*M)) return CUPCFD_E_ERROR;                                                         \
        hipLaunchKernelGGL(k_mass_bl2<T>, dim3(M->d.R + 1), dim3(kSeqBlock), 0, (hipStream_t)stream, M->d,         \
                           massFlux, flowRegion, flowout, (const int *)M->rbnd, (const int *)M->rbnd_off);         \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_boundary_loop3_##S(cupcfd_hip_mesh *mesh, T flowin, T *ratearea, void *stream) {      \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        hipLaunchKernelGGL(k_mass_bl3<T>, dim3(1), dim3(kSeqBlock), 0, (hipStream_t)stream, M->d, flowin,          \
                           ratearea);                                                                              \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_boundary_loop4_##S(cupcfd_hip_mesh *mesh, T *massFlux, T *uBoundary, T *vBoundary,    \
                                                T *wBoundary, const T *denBoundary, T ratearea, T *flowout,        \
                                                void *stream) {                                                    \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        const int n = M->d.nblist[CUPCFD_RTYPE_OUTLET];                                                            \
        if (n)                                                                                                     \
            hipLaunchKernelGGL(k_mass_bl4_faces<T>, dim3(blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, M->d,   \
                               massFlux, uBoundary, vBoundary, wBoundary, denBoundary, ratearea);                  \
        hipLaunchKernelGGL(k_outlet_sum<T>, dim3(1), dim3(kSeqBlock), 0, (hipStream_t)stream, M->d,                \
                           (const T *)massFlux, flowout, 1);                                                       \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_boundary_loop5_##S(cupcfd_hip_mesh *mesh, T *massFlux, T *su, T *uBoundary,           \
                                                T *vBoundary, T *wBoundary, T fact, int32_t solveU,                \
                                                int32_t solveV, int32_t solveW, const T *flowFact, T *flowout2,    \
                                                void *stream) {                                                    \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (M->d.noc)                                                                                              \
            hipLaunchKernelGGL(k_mass_bl5_cells<T>, dim3(blocks(M->d.noc)), dim3(kBlock), 0, (hipStream_t)stream,  \
                               M->d, massFlux, su, uBoundary, vBoundary, wBoundary, fact, solveU, solveV, solveW,  \
                               flowFact);                                                                          \
        hipLaunchKernelGGL(k_outlet_sum<T>, dim3(1), dim3(kSeqBlock), 0, (hipStream_t)stream, M->d,                \
                           (const T *)massFlux, flowout2, 0);                                                      \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_mass_region_loop_##S(cupcfd_hip_mesh *mesh, T *flowFact, const T *flowRegion, T flowIn,    \
                                             void *stream) {                                                       \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (M->d.R)                                                                                                \
            hipLaunchKernelGGL(k_mass_region<T>, dim3(blocks(M->d.R)), dim3(kBlock), 0, (hipStream_t)stream,       \
                               M->d, flowFact, flowRegion, flowIn);                                                \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_scalar_face_loop_##S(                                                                      \
        cupcfd_hip_mesh *mesh, const T *PhiCell, T *PhiBoundary, const T *VisEff, T *Au, T *Su, const T *Den,      \
        const T *MassFlux, const T *TE, const T *CpBoundary, const T *visEffBoundary, T *RFace, const T *dPhidx,   \
        int32_t SolveTurb, int32_t SolveEnthalpy, T vislam, int32_t ivar, int32_t VarT, T Sigma_T, T Prandtl,      \
        int32_t VarTE, T Sigma_k, int32_t VarED, T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large,  \
        T TMCmu, void *stream) {                                                                                   \
        return dev_scalar<T>(mesh, PhiCell, PhiBoundary, VisEff, Au, Su, Den, MassFlux, TE, CpBoundary,            \
                             visEffBoundary, RFace, dPhidx, SolveTurb, SolveEnthalpy, vislam, ivar, VarT, Sigma_T, \
                             Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, Small, Large,  \
                             TMCmu, stream);                                                                       \
    }                                                                                                              \
    int cupcfd_dev_flux_uvw_face_loop1_##S(cupcfd_hip_mesh *mesh, T gammaBlend, T small, const T *uCell,           \
                                           const T *vCell, const T *wCell, T *uBoundary, T *vBoundary,             \
                                           T *wBoundary, const T *visEffCell, const T *visEffBoundary,             \
                                           T *massFlux, const T *dudx, const T *dvdx, const T *dwdx, T *rFace,     \
                                           T *su, T *sv, T *sw, T *au, T *av, T *aw, void *stream) {               \
        return dev_uvw<T>(mesh, gammaBlend, small, uCell, vCell, wCell, uBoundary, vBoundary, wBoundary,           \
                          visEffCell, visEffBoundary, massFlux, dudx, dvdx, dwdx, rFace, su, sv, sw, au, av, aw,   \
                          stream);                                                                                 \
    }                                                                                                              \
    int cupcfd_dev_flux_uvw_region_loop1_##S(cupcfd_hip_mesh *mesh, void *stream) {                                \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (M->d.R)                                                                                                \
            hipLaunchKernelGGL(k_uvw_region1<T>, dim3(blocks(M->d.R)), dim3(kBlock), 0, (hipStream_t)stream,       \
                               M->d);                                                                              \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_flux_uvw_bnds_loop1_##S(cupcfd_hip_mesh *mesh, void *stream) {                                  \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!region_boundaries(*M)) return CUPCFD_E_ERROR;                                                         \
        if (M->d.R)                                                                                                \
            hipLaunchKernelGGL(k_uvw_bnds1<T>, dim3(M->d.R), dim3(kSeqBlock), 0, (hipStream_t)stream, M->d,        \
                               (const int *)M->rbnd, (const int *)M->rbnd_off);                                    \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_viscosity_cell_loop1_##S(cupcfd_hip_mesh *mesh, T small, T cmu, T visURF, T visLam,             \
                                            const T *TE, const T *ED, const T *den, T *visEff, void *stream) {     \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (M->d.C)                                                                                                \
            hipLaunchKernelGGL(k_visc_cell1<T>, dim3(blocks(M->d.C)), dim3(kBlock), 0, (hipStream_t)stream,        \
                               M->d.C, small, cmu, visURF, visLam, TE, ED, den, visEff);                           \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_viscosity_region_loop_##S(T kappa, cupcfd_hip_mesh *mesh, void *stream) {                       \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        hipLaunchKernelGGL(k_visc_region<T>, dim3(1), dim3(64), 0, (hipStream_t)stream, M->d, kappa);              \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_viscosity_boundary_loop_##S(cupcfd_hip_mesh *mesh, T tmcmu, T large, T small, T kappa,          \
                                               T visLam, const T *TE, const T *den, const T *visEffCell,           \
                                               T *visEffBoundary, void *stream) {                                  \
        (void)large;                                                                                               \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        const T cmu = tmcmu;                                                                                       \
        const T cmu25 = pow(cmu, 0.25);                                                                            \
        if (M->d.B)                                                                                                \
            hipLaunchKernelGGL(k_visc_boundary<T>, dim3(blocks(M->d.B)), dim3(kBlock), 0, (hipStream_t)stream,     \
                               M->d, cmu, cmu25, small, kappa, visLam, TE, den, visEffCell, visEffBoundary);       \
        return launch_ok();                                                                                        \
    }                                                                                                              \
    int cupcfd_dev_viscosity_cell_loop2_##S(cupcfd_hip_mesh *mesh, T visLam, T *visEffCell, T *visEffBoundary,     \
                                            void *stream) {                                                        \
        Mesh<T> *M = get<T>(mesh);                                                                                 \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        const int n = M->d.C + M->d.B;                                                                             \
        if (n)                                                                                                     \
            hipLaunchKernelGGL(k_visc_cell2<T>, dim3(blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, M->d.C,     \
                               M->d.B, visLam, visEffCell, visEffBoundary);                                        \
        return launch_ok();                                                                                        \
    }

DEV_API(double, f64)
DEV_API(float, f32)

#undef DEV_API

#define BATCH_API(T, S)                                                                                            \
    int cupcfd_dev_gradient_batched_##S(cupcfd_hip_mesh *mesh, int32_t nGradient, int32_t K, const T *phiCell,      \
                                        const T *phiBoundary, T *dPhidxCell, T *dPhidxoCell, void *stream) {       \
        return dev_gradient_batched<T, T, T>(mesh, nGradient, K, phiCell, phiBoundary, dPhidxCell, dPhidxoCell,    \
                                             stream);                                                          \
    }                                                                                                              \
    int cupcfd_dev_flux_scalar_batched_##S(                                                                        \
        cupcfd_hip_mesh *mesh, int32_t K, const T *PhiCell, T *PhiBoundary, const T *VisEff, T *Au, T *Su,         \
        const T *Den, const T *MassFlux, const T *TE, const T *CpBoundary, const T *visEffBoundary, T *RFace,       \
        const T *dPhidx, int32_t SolveTurb, int32_t SolveEnthalpy, T vislam, int32_t ivar, int32_t VarT,           \
        T Sigma_T, T Prandtl, int32_t VarTE, T Sigma_k, int32_t VarED, T Sigma_e, T Sigma_s, T Schmidt,            \
        T GammaBlend, T Small, T Large, T TMCmu, void *stream) {                                                   \
        return dev_scalar_batched<T>(mesh, K, PhiCell, PhiBoundary, VisEff, Au, Su, Den, MassFlux, TE, CpBoundary, \
                                     visEffBoundary, RFace, dPhidx, SolveTurb, SolveEnthalpy, vislam, ivar, VarT,  \
                                     Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt,           \
                                     GammaBlend, Small, Large, TMCmu, stream);                                     \
    }

BATCH_API(double, f64)
BATCH_API(float, f32)

#undef BATCH_API

int cupcfd_dev_gradient_batched_f32s(cupcfd_hip_mesh *mesh, int32_t nGradient, int32_t K, const float *phiCell,
                                     const float *phiBoundary, float *dPhidxCell, float *dPhidxoCell, void *stream) {
    return dev_gradient_batched<double, float, float>(mesh, nGradient, K, phiCell, phiBoundary, dPhidxCell,
                                                      dPhidxoCell, stream);
}

int cupcfd_dev_flux_scalar_batched_f32s(cupcfd_hip_mesh *mesh, int32_t K, const float *PhiCell, float *PhiBoundary,
                                         const double *VisEff, float *Au, float *Su, const double *Den,
                                         const double *MassFlux, const double *TE, const double *CpBoundary,
                                         const double *visEffBoundary, double *RFace, const float *dPhidx,
                                         int32_t SolveTurb, int32_t SolveEnthalpy, double vislam, int32_t ivar,
                                         int32_t VarT, double Sigma_T, double Prandtl, int32_t VarTE, double Sigma_k,
                                         int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt,
                                         double GammaBlend, double Small, double Large, double TMCmu, void *stream) {
    return dev_scalar_batched<double, float>(mesh, K, PhiCell, PhiBoundary, VisEff, Au, Su, Den, MassFlux, TE,
                                             CpBoundary, visEffBoundary, RFace, dPhidx, SolveTurb, SolveEnthalpy,
                                             vislam, ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e,
                                             Sigma_s, Schmidt, GammaBlend, Small, Large, TMCmu, stream);
}

int cupcfd_dev_scalar_adjoint_f64(cupcfd_hip_mesh *mesh, const double *PhiCell, const double *VisEff,
                                  const double *MassFlux, const double *SuBar, const double *RFaceBar,
                                  int32_t SolveTurb, double vislam, int32_t ivar, int32_t VarT, double Sigma_T,
                                  double Prandtl, int32_t VarTE, double Sigma_k, int32_t VarED, double Sigma_e,
                                  double Sigma_s, double Schmidt, double GammaBlend, double *PhiBar,
                                  double *MassFluxBar, void *stream) {
    return dev_scalar_adjoint<double>(mesh, PhiCell, VisEff, MassFlux, SuBar, RFaceBar, SolveTurb, vislam, ivar, VarT,
                                      Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend,
                                      PhiBar, MassFluxBar, stream);
}

int cupcfd_dev_scalar_adjoint_full_f64(cupcfd_hip_mesh *mesh, int32_t nGradient, const double *PhiCell,
                                       const double *VisEff, const double *MassFlux, const double *dPhidx,
                                       const double *TE, const double *Den, const double *CpBoundary,
                                       const double *visEffBoundary, const double *SuBar, const double *RFaceBar,
                                       int32_t SolveTurb, int32_t SolveEnthalpy, double vislam, int32_t ivar,
                                       int32_t VarT, double Sigma_T, double Prandtl, int32_t VarTE, double Sigma_k,
                                       int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt,
                                       double GammaBlend, double Small, double TMCmu, double *PhiBar,
                                       double *MassFluxBar, void *stream) {
    return dev_scalar_adjoint_full<double>(mesh, nGradient, PhiCell, VisEff, MassFlux, dPhidx, TE, Den, CpBoundary,
                                           visEffBoundary, SuBar, RFaceBar, SolveTurb, SolveEnthalpy, vislam, ivar,
                                           VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt,
                                           GammaBlend, Small, TMCmu, PhiBar, MassFluxBar, stream);
}

int cupcfd_dev_scalar_adjoint_noalloc_f64(cupcfd_hip_mesh *mesh, const double *PhiCell, const double *VisEff,
                                          const double *MassFlux, const double *SuBar, const double *RFaceBar,
                                          int32_t SolveTurb, double vislam, int32_t ivar, int32_t VarT,
                                          double Sigma_T, double Prandtl, int32_t VarTE, double Sigma_k, int32_t VarED,
                                          double Sigma_e, double Sigma_s, double Schmidt, double GammaBlend,
                                          double *PhiBar, double *MassFluxBar, void *stream) {
    Mesh<double> *M = get<double>(mesh);
    if (!M) return CUPCFD_E_NULL_PTR;
    const size_t C = (size_t)M->d.C;
    if (4 * C * sizeof(double) > kNoAllocScratchBytes) return CUPCFD_E_CONFIG_INVALID_VALUE;
    double *sp = nullptr;
    const int rc = scratch_on_current_device(&sp);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    return dev_scalar_adjoint<double>(mesh, PhiCell, VisEff, MassFlux, SuBar, RFaceBar, SolveTurb, vislam, ivar, VarT,
                                      Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend,
                                      PhiBar, MassFluxBar, stream, sp, sp + 3 * C);
}

int cupcfd_dev_scalar_adjoint_batched_f64(cupcfd_hip_mesh *mesh, int32_t K, const double *PhiCell,
                                          const double *VisEff, const double *MassFlux, const double *SuBar,
                                          const double *RFaceBar, int32_t SolveTurb, double vislam, int32_t ivar,
                                          int32_t VarT, double Sigma_T, double Prandtl, int32_t VarTE, double Sigma_k,
                                          int32_t VarED, double Sigma_e, double Sigma_s, double Schmidt,
                                          double GammaBlend, double *PhiBar, double *MassFluxBar, void *stream) {
    return dev_scalar_adjoint_batched<double>(mesh, K, PhiCell, VisEff, MassFlux, SuBar, RFaceBar, SolveTurb, vislam,
                                              ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s,
                                              Schmidt, GammaBlend, PhiBar, MassFluxBar, stream);
}

int cupcfd_hip_mesh_face_colours(cupcfd_hip_mesh *mesh, int32_t *colour, int32_t *ncolours) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    auto fetch = [&](auto *M) -> int {
        if (build_face_colours(*M) != CUPCFD_E_SUCCESS) return CUPCFD_E_ERROR;
        if (ncolours) *ncolours = M->ncolours;
        if (!colour || M->d.F == 0) return CUPCFD_E_SUCCESS;
        std::vector<int32_t> c(M->d.F);
        if (hipMemcpy(c.data(), M->colour_of, c.size() * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
            return CUPCFD_E_ERROR;
        for (int k = 0; k < M->d.F; k++) colour[M->face_order[k]] = c[k];  // mesh numbering
        return CUPCFD_E_SUCCESS;
    };
    return mesh->precision == 64 ? fetch(mesh->m64) : fetch(mesh->m32);
}

int cupcfd_dev_scalar_adjoint_batched_f32s(cupcfd_hip_mesh *mesh, int32_t K, const float *PhiCell,
                                            const double *VisEff, const double *MassFlux, const float *SuBar,
                                            const double *RFaceBar, int32_t SolveTurb, double vislam, int32_t ivar,
                                            int32_t VarT, double Sigma_T, double Prandtl, int32_t VarTE,
                                            double Sigma_k, int32_t VarED, double Sigma_e, double Sigma_s,
                                            double Schmidt, double GammaBlend, float *PhiBar, double *MassFluxBar,
                                            void *stream) {
    return dev_scalar_adjoint_batched<double, float>(mesh, K, PhiCell, VisEff, MassFlux, SuBar, RFaceBar, SolveTurb,
                                                     vislam, ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED,
                                                     Sigma_e, Sigma_s, Schmidt, GammaBlend, PhiBar, MassFluxBar,
                                                     stream);
}

int cupcfd_hip_meshset_create(cupcfd_hip_mesh *const *meshes, int32_t n, cupcfd_hip_meshset **out) {
    if (!meshes || !out) return CUPCFD_E_NULL_PTR;
    *out = nullptr;
    if (n < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    for (int i = 0; i < n; i++)
        if (!meshes[i] || meshes[i]->precision != meshes[0]->precision) return CUPCFD_E_NULL_PTR;
    cupcfd_hip_meshset *set = new (std::nothrow) cupcfd_hip_meshset();
    if (!set) return CUPCFD_E_ERROR;
    set->precision = meshes[0]->precision;
    set->n = n;
    const int rc = set->precision == 64 ? meshset_build<double>(meshes, n, set) : meshset_build<float>(meshes, n, set);
    if (rc != CUPCFD_E_SUCCESS) {
        cupcfd_hip_meshset_destroy(set);
        return rc;
    }
    *out = set;
    return CUPCFD_E_SUCCESS;
}

int cupcfd_hip_meshset_destroy(cupcfd_hip_meshset *set) {
    if (!set) return CUPCFD_E_NULL_PTR;
    if (set->meshes) (void)hipFree(set->meshes);
    for (int *u : set->unit0)
        if (u) (void)hipFree(u);
    delete set;
    return CUPCFD_E_SUCCESS;
}

int cupcfd_dev_gradient_batched_set_f64(cupcfd_hip_meshset *set, int32_t K, const double *const *phiCell,
                                        const double *const *phiBoundary, double *const *dPhidxCell,
                                        double *const *dPhidxoCell, void *stream) {
    return dev_gradient_batched_set<double>(set, K, phiCell, phiBoundary, dPhidxCell, dPhidxoCell, stream);
}
int cupcfd_dev_gradient_batched_set_f32(cupcfd_hip_meshset *set, int32_t K, const float *const *phiCell,
                                        const float *const *phiBoundary, float *const *dPhidxCell,
                                        float *const *dPhidxoCell, void *stream) {
    return dev_gradient_batched_set<float>(set, K, phiCell, phiBoundary, dPhidxCell, dPhidxoCell, stream);
}

}  // extern "C"

namespace {

// Autotuned schedules, cached per (GPU architecture, precision, mesh class)
// for the process and, when CUPCFD_HIP_TUNE_CACHE names a file, across
// processes: one "key gradient scalar uvw mass" line per entry (a line
// without the mass entry leaves the mass flux on its automatic schedule). The mesh class is
// what the choice depends on: row lengths, whether tiles apply and at which
// LDS width, and the cell count to the nearest power of two.
std::mutex tune_mutex;
std::map<std::string, std::vector<int>> tune_cache;

template <class T> std::string tune_key(const Mesh<T> &M) {
    hipDeviceProp_t prop{};
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    std::string arch(prop.gcnArchName);
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    const DevMesh<T> &d = M.d;
    int lg = 0;
    while (lg < 31 && ((int64_t)1 << (lg + 1)) <= (int64_t)d.C) lg++;
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s|p%d|rl%d|mr%d|tiles%d|tw%d|typed%d|c2^%d", arch.c_str(),
                  (int)(8 * sizeof(T)), d.rowlen, d.maxrow, d.ntiles > 0 ? 1 : 0, tile_width(d), d.typed, lg);
    return buf;
}

void tune_load(const char *path) {
    FILE *fp = std::fopen(path, "r");
    if (!fp) return;
    char line[256], key[160];
    while (std::fgets(line, sizeof line, fp)) {
        int g, sc, u, ms = -1;
        const int n = std::sscanf(line, "%159s %d %d %d %d", key, &g, &sc, &u, &ms);
        if (n >= 4) tune_cache[key] = {g, sc, u, n == 5 ? ms : -1};
    }
    std::fclose(fp);
}

void tune_save(const char *path) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE *fp = std::fopen(tmp.c_str(), "w");
    if (!fp) return;
    for (const auto &e : tune_cache)
        std::fprintf(fp, "%s %d %d %d %d\n", e.first.c_str(), e.second[0], e.second[1], e.second[2], e.second[3]);
    if (std::fclose(fp) == 0) (void)std::rename(tmp.c_str(), path);
}

template <class T> int autotune_cached(cupcfd_hip_mesh *h, int32_t *chosen) {
    Mesh<T> *M = get<T>(h);
    if (!M) return CUPCFD_E_NULL_PTR;
    const std::string key = tune_key(*M);
    const char *path = std::getenv("CUPCFD_HIP_TUNE_CACHE");
    {
        std::lock_guard<std::mutex> lock(tune_mutex);
        if (path && *path) tune_load(path);
        auto it = tune_cache.find(key);
        if (it != tune_cache.end()) {
            for (int k = 0; k < 4; k++) {
                M->schedule_k[k] = it->second[k];
                if (chosen && k < 3) chosen[k] = it->second[k];
            }
            return CUPCFD_E_SUCCESS;
        }
    }
    int32_t pick[3];
    const int rc = autotune<T>(h, pick);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    if (chosen)
        for (int k = 0; k < 3; k++) chosen[k] = pick[k];
    std::lock_guard<std::mutex> lock(tune_mutex);
    tune_cache[key] = {pick[0], pick[1], pick[2], M->schedule_k[kMassFaces]};
    if (path && *path) tune_save(path);
    return CUPCFD_E_SUCCESS;
}

}  // namespace

extern "C" {

int cupcfd_hip_mesh_autotune(cupcfd_hip_mesh *mesh, int32_t *schedules) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    return mesh->precision == 64 ? autotune_cached<double>(mesh, schedules) : autotune_cached<float>(mesh, schedules);
}

// Drop the process's autotune cache (the next cupcfd_hip_mesh_autotune of
// each mesh class times again, or reads CUPCFD_HIP_TUNE_CACHE).
int cupcfd_hip_autotune_forget(void) {
    std::lock_guard<std::mutex> lock(tune_mutex);
    tune_cache.clear();
    return CUPCFD_E_SUCCESS;
}

const char *cupcfd_hip_version(void) { return "cupcfd_hip 0.1 gfx950"; }

}  // extern "C"

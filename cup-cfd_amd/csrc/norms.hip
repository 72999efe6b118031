// Residual norms of cell fields (the source terms the face loops accumulate:
// Su of the scalar loop, su / sv / sw of the UVW loop), for the convergence
// test a solver makes after each sweep. The reference's kernels return their
// residuals in these arrays and leave the norm to the caller; this is the
// device-side reduction, so no field crosses PCIe.
//
// For each field x of n values: max |x| and sqrt(sum x^2), accumulated in
// double whatever the storage type. The order of every addition is fixed by
// the launch shape alone (kNormGrid workgroups of kNormBlock threads, a
// grid-stride walk per thread four elements at a time, a fixed xor tree per
// wavefront, the waves of a workgroup in order, then one workgroup over the
// partials in order), so two calls on the same data give the same bits on
// any MI355X. NaN propagates into
// both norms. The partials live in a mesh workspace slot sized on the first
// call; later calls allocate nothing and can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "devmesh.hpp"

namespace cupcfd_hip {
namespace {

constexpr int kNormBlock = 256;
constexpr int kNormGrid = 1024;  // 4 workgroups per CU on 256 CUs
constexpr int kNormWaves = kNormBlock / 64;
constexpr int kMaxNormFields = 8;

struct NormArgs {
    const void *field[kMaxNormFields];
    int nfields;
    int n;
    double *part;  // [field][workgroup][2]
    double *out;   // [field][2]: max |x|, sqrt(sum x^2)
};

// max that keeps a NaN once it has seen one
__device__ inline double nanmax(double m, double v) { return (v > m || v != v) ? v : m; }

// (max, sum) of one wavefront, in a fixed xor-tree order
__device__ inline void wave_reduce(double &mx, double &ss) {
    for (int o = 32; o >= 1; o >>= 1) {
        const double m2 = __shfl_xor(mx, o, 64), s2 = __shfl_xor(ss, o, 64);
        mx = nanmax(mx, m2);
        ss = ss + s2;
    }
}

// (max, sum) of NF fields over a workgroup: waves in order 0..kNormWaves-1;
// valid in thread 0
template <int NF> __device__ inline void block_reduce(double *mx, double *ss, double *lds) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int i = 0; i < NF; i++) {
        wave_reduce(mx[i], ss[i]);
        if (l == 0) {
            lds[2 * (i * kNormWaves + w)] = mx[i];
            lds[2 * (i * kNormWaves + w) + 1] = ss[i];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NF; i++) {
            mx[i] = lds[2 * (i * kNormWaves)];
            ss[i] = lds[2 * (i * kNormWaves) + 1];
            for (int q = 1; q < kNormWaves; q++) {
                mx[i] = nanmax(mx[i], lds[2 * (i * kNormWaves + q)]);
                ss[i] = ss[i] + lds[2 * (i * kNormWaves + q) + 1];
            }
        }
    }
}

// Each thread walks its grid-stride sequence four elements at a time, every
// field's four loads issued before the first accumulation.
template <class TS, int NF>
__global__ __launch_bounds__(kNormBlock) void k_norm_partial(NormArgs a) {
    __shared__ double lds[2 * NF * kNormWaves];
    double mx[NF], ss[NF];
#pragma unroll
    for (int i = 0; i < NF; i++) mx[i] = ss[i] = 0.0;
    const int stride = kNormGrid * kNormBlock;
    int j = blockIdx.x * kNormBlock + threadIdx.x;
    for (; j + 3 * stride < a.n; j += 4 * stride) {
        double v[NF][4];
#pragma unroll
        for (int i = 0; i < NF; i++) {
            const TS *x = (const TS *)a.field[i];
#pragma unroll
            for (int u = 0; u < 4; u++) v[i][u] = (double)x[j + u * stride];
        }
#pragma unroll
        for (int i = 0; i < NF; i++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                mx[i] = nanmax(mx[i], fabs(v[i][u]));
                ss[i] = ss[i] + v[i][u] * v[i][u];
            }
    }
    for (; j < a.n; j += stride) {
#pragma unroll
        for (int i = 0; i < NF; i++) {
            const double v = (double)((const TS *)a.field[i])[j];
            mx[i] = nanmax(mx[i], fabs(v));
            ss[i] = ss[i] + v * v;
        }
    }
    block_reduce<NF>(mx, ss, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NF; i++) {
            a.part[2 * ((size_t)i * kNormGrid + blockIdx.x)] = mx[i];
            a.part[2 * ((size_t)i * kNormGrid + blockIdx.x) + 1] = ss[i];
        }
    }
}

// one workgroup per field: the field's partials in workgroup order
__global__ __launch_bounds__(kNormBlock) void k_norm_final(NormArgs a) {
    __shared__ double lds[2 * kNormWaves];
    const int i = blockIdx.x;
    double mx = 0.0, ss = 0.0;
    for (int b = threadIdx.x; b < kNormGrid; b += kNormBlock) {
        mx = nanmax(mx, a.part[2 * ((size_t)i * kNormGrid + b)]);
        ss = ss + a.part[2 * ((size_t)i * kNormGrid + b) + 1];
    }
    block_reduce<1>(&mx, &ss, lds);
    if (threadIdx.x == 0) {
        a.out[2 * i] = mx;
        a.out[2 * i + 1] = sqrt(ss);
    }
}

constexpr int kNormSlot = 60;  // workspace slot of the partials

template <class T, class TS>
int residual_norms(cupcfd_hip_mesh *h, int nfields, const TS *const *fields, double *out, void *stream) {
    if (!h || !fields || !out) return CUPCFD_E_NULL_PTR;
    if (nfields < 1 || nfields > kMaxNormFields) return CUPCFD_E_CONFIG_INVALID_VALUE;
    Mesh<T> *M = (h->precision == 64) ? (Mesh<T> *)(void *)h->m64 : (Mesh<T> *)(void *)h->m32;
    if (!M) return CUPCFD_E_NULL_PTR;
    NormArgs a{};
    for (int i = 0; i < nfields; i++) {
        if (!fields[i]) return CUPCFD_E_NULL_PTR;
        a.field[i] = fields[i];
    }
    a.nfields = nfields;
    a.n = M->d.C;
    a.out = out;
    a.part = (double *)M->ws_get(kNormSlot, 2 * sizeof(double) * kNormGrid * kMaxNormFields);
    if (!a.part) return CUPCFD_E_ERROR;
    hipStream_t s = (hipStream_t)stream;
    switch (nfields) {
#define CUPCFD_NORM_CASE(NF)                                                                 \
    case NF:                                                                                 \
        hipLaunchKernelGGL((k_norm_partial<TS, NF>), dim3(kNormGrid), dim3(kNormBlock), 0, s, a); \
        break;
        CUPCFD_NORM_CASE(1)
        CUPCFD_NORM_CASE(2)
        CUPCFD_NORM_CASE(3)
        CUPCFD_NORM_CASE(4)
        CUPCFD_NORM_CASE(5)
        CUPCFD_NORM_CASE(6)
        CUPCFD_NORM_CASE(7)
        CUPCFD_NORM_CASE(8)
#undef CUPCFD_NORM_CASE
    }
    hipLaunchKernelGGL(k_norm_final, dim3(nfields), dim3(kNormBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

}  // namespace
}  // namespace cupcfd_hip

extern "C" {

int cupcfd_dev_residual_norms_f64(cupcfd_hip_mesh *mesh, int32_t nfields, const double *const *fields, double *norms,
                                  void *stream) {
    if (mesh && mesh->precision != 64) return CUPCFD_E_CONFIG_INVALID_VALUE;
    return cupcfd_hip::residual_norms<double, double>(mesh, nfields, fields, norms, stream);
}

int cupcfd_dev_residual_norms_f32(cupcfd_hip_mesh *mesh, int32_t nfields, const float *const *fields, double *norms,
                                  void *stream) {
    if (mesh && mesh->precision != 32 && mesh->precision != 64) return CUPCFD_E_CONFIG_INVALID_VALUE;
    if (mesh && mesh->precision == 64)
        return cupcfd_hip::residual_norms<double, float>(mesh, nfields, fields, norms, stream);
    return cupcfd_hip::residual_norms<float, float>(mesh, nfields, fields, norms, stream);
}

}  // extern "C"

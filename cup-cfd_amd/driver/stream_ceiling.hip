// Streaming ceilings of one GPU: the rates a plain stream of HBM traffic reaches,
// against which the face loops' traffic rates are read. Kernels over 1 GiB
// buffers (16-byte lanes, grid-stride loops): read-only (plain and
// non-temporal loads), copy (plain and non-temporal), write-only
// (non-temporal stores), and the mass-flux face loop's face streams alone
// (three 32-byte and one 16-byte record, the 8-byte connectivity read and
// 24 bytes written per face, one thread per face, at the 128^3 mesh's
// interior face count). Each row: the best over four grid sizes of the mean
// of 20 timed launches (HIP events), after 3 untimed ones.
//
//   ./cup-cfd_amd/stream_ceiling
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

#define CHECK(x)                                                                             \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(1);                                                                    \
        }                                                                                    \
    } while (0)

constexpr int kThreads = 256;
constexpr int kUnroll = 4;

template <bool NT> __device__ __forceinline__ d2 load(const d2 *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT> __device__ __forceinline__ void store(d2 *p, d2 v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// every element i < n is read once; the sum is stored only if it takes a value
// no input gives, so the loads cannot be dropped
template <bool NT> __global__ __launch_bounds__(kThreads) void k_read(const d2 *__restrict__ a, size_t n, double *out) {
    double s = 0;
    const size_t stride = (size_t)gridDim.x * kThreads * kUnroll;
    for (size_t i = (size_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; i + (kUnroll - 1) * kThreads < n;
         i += stride) {
        d2 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = load<NT>(a + i + u * kThreads);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) s += v[u].x + v[u].y;
    }
    if (s == -1.0) out[0] = s;
}

template <bool NT> __global__ __launch_bounds__(kThreads) void k_copy(const d2 *__restrict__ a, d2 *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * kThreads * kUnroll;
    for (size_t i = (size_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; i + (kUnroll - 1) * kThreads < n;
         i += stride) {
        d2 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = load<NT>(a + i + u * kThreads);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) store<NT>(b + i + u * kThreads, v[u]);
    }
}

__global__ __launch_bounds__(kThreads) void k_fill(d2 *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) store<true>(b + i, d2{1.0, 2.0});
}

struct alignas(32) R4 {
    double x, y, z, w;
};
struct alignas(16) R2 {
    double x, y;
};

// the mass-flux face loop's face streams without its cell gathers
__global__ __launch_bounds__(kThreads) void k_face_streams(const R4 *__restrict__ S, const R4 *__restrict__ X,
                                                           const R4 *__restrict__ P, const R2 *__restrict__ N,
                                                           const int2 *__restrict__ cn, double *__restrict__ mf,
                                                           double *__restrict__ rf, int F) {
    const int f = blockIdx.x * kThreads + threadIdx.x;
    if (f >= F) return;
    const R4 s = S[f], x = X[f], p = P[f];
    const R2 q = N[f];
    const int2 c = cn[f];
    const double r = s.x * x.y + p.z * q.x + s.w * x.w + p.x * q.y + (double)(c.x - c.y);
    mf[f] = r;
    rf[2 * (size_t)f] = -r;
    rf[2 * (size_t)f + 1] = 0.5 * r;
}

hipEvent_t e0, e1;

template <class L> double best_us(L launch, const int *grids, int ng) {
    double best = 1e30;
    for (int g = 0; g < ng; g++) {
        for (int w = 0; w < 3; w++) launch(grids[g]);
        CHECK(hipDeviceSynchronize());
        const int reps = 20;
        CHECK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) launch(grids[g]);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / reps;
        if (us < best) best = us;
    }
    return best;
}

void row(const char *name, double bytes, double us) {
    std::printf("| %s | %.1f | %.1f | %.2f |\n", name, bytes / 1e6, us, bytes / (us * 1e-6) / 1e12);
    std::fflush(stdout);
}

}  // namespace

int main() {
    const size_t bytes = (size_t)1 << 30, n = bytes / sizeof(d2);
    d2 *a, *b;
    double *o;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMalloc(&o, sizeof(double)));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int grids[4] = {1024, 2048, 4096, 8192};
    std::printf("| stream | MB per launch | µs | TB/s |\n|---|---|---|---|\n");
    row("read, plain loads", (double)bytes,
        best_us([&](int g) { hipLaunchKernelGGL(k_read<false>, dim3(g), dim3(kThreads), 0, 0, a, n, o); }, grids, 4));
    row("read, non-temporal loads", (double)bytes,
        best_us([&](int g) { hipLaunchKernelGGL(k_read<true>, dim3(g), dim3(kThreads), 0, 0, a, n, o); }, grids, 4));
    row("copy, plain", 2.0 * bytes,
        best_us([&](int g) { hipLaunchKernelGGL(k_copy<false>, dim3(g), dim3(kThreads), 0, 0, a, b, n); }, grids, 4));
    row("copy, non-temporal", 2.0 * bytes,
        best_us([&](int g) { hipLaunchKernelGGL(k_copy<true>, dim3(g), dim3(kThreads), 0, 0, a, b, n); }, grids, 4));
    row("write, non-temporal stores", (double)bytes,
        best_us([&](int g) { hipLaunchKernelGGL(k_fill, dim3(g), dim3(kThreads), 0, 0, b, n); }, grids, 4));
    CHECK(hipGetLastError());
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    const int F = 6242304;  // interior faces of the 128^3 StructGen mesh
    R4 *S, *X, *P;
    R2 *N;
    int2 *cn;
    double *mf, *rf;
    CHECK(hipMalloc(&S, F * sizeof(R4)));
    CHECK(hipMalloc(&X, F * sizeof(R4)));
    CHECK(hipMalloc(&P, F * sizeof(R4)));
    CHECK(hipMalloc(&N, F * sizeof(R2)));
    CHECK(hipMalloc(&cn, F * sizeof(int2)));
    CHECK(hipMalloc(&mf, F * sizeof(double)));
    CHECK(hipMalloc(&rf, 2 * (size_t)F * sizeof(double)));
    CHECK(hipMemset(S, 0, F * sizeof(R4)));
    CHECK(hipMemset(X, 0, F * sizeof(R4)));
    CHECK(hipMemset(P, 0, F * sizeof(R4)));
    CHECK(hipMemset(N, 0, F * sizeof(R2)));
    CHECK(hipMemset(cn, 0, F * sizeof(int2)));
    const int g1[1] = {(F + kThreads - 1) / kThreads};
    row("mass-flux face streams (120 B read, 24 B written per face)", (double)F * (3 * 32 + 16 + 8 + 24),
        best_us([&](int g) { hipLaunchKernelGGL(k_face_streams, dim3(g), dim3(kThreads), 0, 0, S, X, P, N, cn, mf, rf, F); },
                g1, 1));
    CHECK(hipGetLastError());
    return 0;
}

// cupcfd_bench — host C++ driver with the reference's BenchmarkKernels entry.
//
// Reads a cupcfd.json in CUP-CFD's schema (README.md:100-160;
// examples/benchmarks.json): "DataTypes".FloatDataType selects double or
// float, "Mesh"."MeshSource" is either "MeshSourceStructGen" (the grid) or
// "MeshSourceFile" {"FileFormat": "HDF5", "FilePath"} (MeshSourceFileConfigJSON.cpp:
// 60-95; read by cupcfd_hdf5_mesh_create), and every
// "Benchmarks"[i]."BenchmarkKernels" {"BenchmarkName", "Repetitions"} entry
// runs the 16 fvm kernels per repetition in BenchmarkKernels::runBenchmark's
// order (src/benchmarks/implementation/component/BenchmarkKernels.cpp:55-115)
// with its parameter values, through the cupcfd::fvm-shaped shim
// (cupcfd_hip_fvm.hpp) on a MiniSoA-layout mesh whose 3-vectors carry a
// vptr as EuclideanVector does. Other benchmark types are reported and
// skipped. Optional keys under "BenchmarkKernels" that the reference ignores:
// "Seed" (input stream), "Permute" (relabel cell/face ids), "Typed" (six
// typed boundary regions), "Graph" (also run the 16 kernels on device-resident
// inputs, eagerly and as one captured hipGraph, see GraphBench). Each kernel
// call is timed on the host around the (synchronous) call, as TreeTimer times
// the reference's; the report is JSON on stdout.
//
// Usage: cupcfd_bench [path/to/cupcfd.json]   (default ./cupcfd.json)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "cupcfd_hip.h"
#include "cupcfd_hip_fvm.hpp"

// ----------------------------------------------------------------------------
// Minimal JSON reader (objects, arrays, strings, numbers, booleans, null)
// ----------------------------------------------------------------------------
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
    const Json *get(const std::string &k) const {
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const std::string &s;
    size_t i = 0;
    bool ok = true;
    explicit Parser(const std::string &t) : s(t) {}
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    Json value() {
        ws();
        Json j;
        if (i >= s.size()) {
            ok = false;
            return j;
        }
        char c = s[i];
        if (c == '{') {
            j.kind = Json::OBJ;
            i++;
            ws();
            if (s[i] == '}') {
                i++;
                return j;
            }
            while (ok) {
                ws();
                Json k = value();
                ws();
                if (k.kind != Json::STR || s[i] != ':') {
                    ok = false;
                    break;
                }
                i++;
                j.obj.emplace_back(k.str, value());
                ws();
                if (s[i] == ',') {
                    i++;
                    continue;
                }
                if (s[i] == '}') {
                    i++;
                    break;
                }
                ok = false;
            }
        } else if (c == '[') {
            j.kind = Json::ARR;
            i++;
            ws();
            if (s[i] == ']') {
                i++;
                return j;
            }
            while (ok) {
                j.arr.push_back(value());
                ws();
                if (s[i] == ',') {
                    i++;
                    continue;
                }
                if (s[i] == ']') {
                    i++;
                    break;
                }
                ok = false;
            }
        } else if (c == '"') {
            j.kind = Json::STR;
            i++;
            while (i < s.size() && s[i] != '"') {
                if (s[i] == '\\' && i + 1 < s.size()) i++;
                j.str += s[i++];
            }
            i++;
        } else if (s.compare(i, 4, "true") == 0) {
            j.kind = Json::BOOL;
            j.b = true;
            i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            j.kind = Json::BOOL;
            i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            j.kind = Json::NUM;
            char *end = nullptr;
            j.num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) ok = false;
            i = end - s.c_str();
        }
        return j;
    }
};

// ----------------------------------------------------------------------------
// MiniSoA-layout mesh: the members CupCfdSoAMesh exposes, with vptr-bearing
// 3-vectors (CupCfdSoAMesh.h:71-154, EuclideanVector.h:40,57)
// ----------------------------------------------------------------------------
template <class T> struct Vec3 {
    virtual ~Vec3() {}
    T cmp[3] = {0, 0, 0};
};
enum RType { RTYPE_DEFAULT, RTYPE_INLET, RTYPE_OUTLET, RTYPE_SYMP, RTYPE_WALL };

template <class T> struct SoAMesh {
    struct {
        int lTCells = 0, lFaces = 0, lBoundaries = 0, lRegions = 0;
    } properties;
    std::vector<int> faceCell1ID, faceCell2ID, faceBoundaryID, boundaryFaceID, boundaryRegionID;
    std::vector<T> faceLambda, faceRLencos, faceArea, cellVolume, boundaryDistance, boundaryYPlus, boundaryUPlus,
        boundaryQ, boundaryH, boundaryT, regionYLog, regionELog, regionDensity, regionTurbKE, regionTurbDiss,
        regionSplvl, regionDen, regionR, regionT;
    std::vector<Vec3<T>> faceNorm, faceCenter, faceXpac, faceXnac, cellCenter, boundaryShear, regionForceTangent,
        regionUVW;
    std::vector<RType> regionType;
    std::vector<bool> regionStd, regionFlux, regionAdiab;
};

template <class T> void copy3(std::vector<Vec3<T>> &dst, const cupcfd_v3_f64 &v, int n) {
    dst.resize(n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) dst[i].cmp[k] = (T)v.base[3 * (size_t)i + k];
}
template <class T, class S> void copy1(std::vector<T> &dst, const S *src, int n) {
    dst.resize(n);
    for (int i = 0; i < n; i++) dst[i] = (T)src[i];
}

// where the mesh comes from: a StructGen grid, or an HDF5 mesh file
struct MeshSpec {
    cupcfd_structgen_params sg{};
    std::string file;  // non-empty: read this HDF5 mesh file instead
};

template <class T> int load_mesh(SoAMesh<T> &m, const MeshSpec &spec) {
    cupcfd_host_mesh *h = nullptr;
    int rc = spec.file.empty() ? cupcfd_structgen_create(&spec.sg, &h) : cupcfd_hdf5_mesh_create(spec.file.c_str(), &h);
    if (rc != CUPCFD_E_SUCCESS) return rc;
    cupcfd_mesh_view_f64 v;
    cupcfd_host_mesh_view_f64(h, &v);
    const int C = v.nCells, F = v.nFaces, B = v.nBoundaries, R = v.nRegions;
    m.properties = {C, F, B, R};
    copy1(m.faceCell1ID, v.faceCell1ID, F);
    copy1(m.faceCell2ID, v.faceCell2ID, F);
    copy1(m.faceBoundaryID, v.faceBoundaryID, F);
    copy1(m.faceLambda, v.faceLambda, F);
    copy1(m.faceRLencos, v.faceRLencos, F);
    copy1(m.faceArea, v.faceArea, F);
    copy3(m.faceNorm, v.faceNorm, F);
    copy3(m.faceCenter, v.faceCenter, F);
    copy3(m.faceXpac, v.faceXpac, F);
    copy3(m.faceXnac, v.faceXnac, F);
    copy3(m.cellCenter, v.cellCenter, C);
    copy1(m.cellVolume, v.cellVolume, C);
    copy1(m.boundaryFaceID, v.boundaryFaceID, B);
    copy1(m.boundaryRegionID, v.boundaryRegionID, B);
    copy1(m.boundaryDistance, v.boundaryDistance, B);
    copy1(m.boundaryYPlus, v.boundaryYPlus, B);
    copy1(m.boundaryUPlus, v.boundaryUPlus, B);
    copy1(m.boundaryQ, v.boundaryQ, B);
    copy1(m.boundaryH, v.boundaryH, B);
    copy1(m.boundaryT, v.boundaryT, B);
    copy3(m.boundaryShear, v.boundaryShear, B);
    m.regionType.resize(R);
    m.regionStd.resize(R);
    m.regionFlux.resize(R);
    m.regionAdiab.resize(R);
    for (int r = 0; r < R; r++) {
        m.regionType[r] = (RType)v.regionType[r];
        m.regionStd[r] = v.regionStd[r] != 0;
        m.regionFlux[r] = v.regionFlux[r] != 0;
        m.regionAdiab[r] = v.regionAdiab[r] != 0;
    }
    copy1(m.regionYLog, v.regionYLog, R);
    copy1(m.regionELog, v.regionELog, R);
    copy1(m.regionDensity, v.regionDensity, R);
    copy1(m.regionTurbKE, v.regionTurbKE, R);
    copy1(m.regionTurbDiss, v.regionTurbDiss, R);
    copy1(m.regionSplvl, v.regionSplvl, R);
    copy1(m.regionDen, v.regionDen, R);
    copy1(m.regionR, v.regionR, R);
    copy1(m.regionT, v.regionT, R);
    copy3(m.regionForceTangent, v.regionForceTangent, R);
    copy3(m.regionUVW, v.regionUVW, R);
    cupcfd_host_mesh_destroy(h);
    return CUPCFD_E_SUCCESS;
}

// ----------------------------------------------------------------------------
// BenchmarkKernels: 16 wrappers, each timing one call
// ----------------------------------------------------------------------------
template <class T> struct Bench {
    SoAMesh<T> &mesh;
    std::mt19937_64 rng;
    std::map<std::string, std::vector<double>> times;
    int failures = 0;
    Bench(SoAMesh<T> &m, uint64_t seed) : mesh(m), rng(seed) {}

    // randomUniform(1e-6, 1e-2) as the reference's wrappers fill their inputs
    // (include/utility/implementation/component/ArrayKernels.ipp:307-315), seeded
    std::vector<T> U(int n, double lo = 1e-6, double hi = 1e-2) {
        std::uniform_real_distribution<double> d(lo, hi);
        std::vector<T> v(n);
        for (auto &x : v) x = (T)d(rng);
        return v;
    }
    std::vector<Vec3<T>> U3(int n) {
        std::uniform_real_distribution<double> d(1e-6, 1e-2);
        std::vector<Vec3<T>> v(n);
        for (auto &x : v)
            for (int k = 0; k < 3; k++) x.cmp[k] = (T)d(rng);
        return v;
    }
    template <class F> void timed(const char *name, F &&f) {
        auto t0 = std::chrono::steady_clock::now();
        int rc = f();
        auto t1 = std::chrono::steady_clock::now();
        times[name].push_back(std::chrono::duration<double>(t1 - t0).count());
        if (rc != CUPCFD_E_SUCCESS) {
            failures++;
            std::fprintf(stderr, "%s returned %d\n", name, rc);
        }
    }

    void repetition() {
        namespace fvm = cupcfd_hip::fvm;
        const int C = mesh.properties.lTCells, F = mesh.properties.lFaces, B = mesh.properties.lBoundaries,
                  R = mesh.properties.lRegions;
        {  // BenchmarkKernels.cpp:117-194
            auto phi = U(C), phiB = U(B);
            auto g = U3(C), go = U3(C);
            timed("GradientPhiGaussDolfynBenchmark", [&] {
                return fvm::GradientPhiGaussDolfyn(mesh, 1, phi.data(), C, phiB.data(), B, g.data(), C, go.data(), C);
            });
        }
        std::vector<T> massFlux = U(F);
        {  // :196-326
            auto dudx = U3(C), dvdx = U3(C), dwdx = U3(C), dpdx = U3(C);
            auto den = U(C), denB = U(B), u = U(C), v = U(C), w = U(C), p = U(C), ar = U(C), su = U(C),
                 rface = U(2 * F), te = U(C), teB = U(B), ed = U(C), edB = U(B), vis = U(C), visB = U(B), t = U(C),
                 tB = U(B);
            int icinl = 0, icout = 0, icsym = 0, icwal = 0;
            timed("FluxMassDolfynFaceLoopBenchmark", [&] {
                return fvm::FluxMassDolfynFaceLoop(
                    mesh, dudx.data(), C, dvdx.data(), C, dwdx.data(), C, dpdx.data(), C, den.data(), C, denB.data(),
                    B, u.data(), C, v.data(), C, w.data(), C, massFlux.data(), F, p.data(), C, ar.data(), C,
                    su.data(), C, rface.data(), 2 * F, (T)1e-18, &icinl, &icout, &icsym, &icwal, false, false, true,
                    false, te.data(), C, teB.data(), B, ed.data(), C, edB.data(), B, vis.data(), C, visB.data(), B,
                    t.data(), C, tB.data(), B);
            });
        }
        std::vector<T> flowRegion = U(R), flowFact = U(R);
        {
            T flowin = 0, flowout = 0, ratearea = 0, flowout2 = 0;
            timed("FluxMassDolfynBoundaryLoop1Benchmark",
                  [&] { return fvm::FluxMassDolfynBoundaryLoop1(mesh, massFlux.data(), F, &flowin); });
            timed("FluxMassDolfynBoundaryLoop2Benchmark", [&] {
                return fvm::FluxMassDolfynBoundaryLoop2(mesh, massFlux.data(), F, flowRegion.data(), R, &flowout);
            });
            timed("FluxMassDolfynBoundaryLoop3Benchmark",
                  [&] { return fvm::FluxMassDolfynBoundaryLoop3(mesh, (T)1.1, &ratearea); });
            auto uB = U(B), vB = U(B), wB = U(B), denB = U(B);
            timed("FluxMassDolfynBoundaryLoop4Benchmark", [&] {
                return fvm::FluxMassDolfynBoundaryLoop4(mesh, massFlux.data(), F, uB.data(), B, vB.data(), B,
                                                        wB.data(), B, denB.data(), B, (T)1.0, &flowout);
            });
            auto su = U(C);
            timed("FluxMassDolfynBoundaryLoop5Benchmark", [&] {
                return fvm::FluxMassDolfynBoundaryLoop5(mesh, massFlux.data(), F, su.data(), C, uB.data(), B,
                                                        vB.data(), B, wB.data(), B, (T)1.1, true, true, true,
                                                        flowFact.data(), R, &flowout2);
            });
            timed("FluxMassDolfynRegionLoopBenchmark", [&] {
                return fvm::FluxMassDolfynRegionLoop(mesh, flowFact.data(), R, flowRegion.data(), R, (T)0.0);
            });
        }
        {  // :634-738, parameter values :677-695
            auto phi = U(C), phiB = U(B), vis = U(C), au = U(C), su = U(C), den = U(C), te = U(C), cpB = U(B),
                 visB = U(B), rf = U(2 * F);
            auto g = U3(C);
            timed("FluxScalarDolfynFaceLoopBenchmark", [&] {
                return fvm::FluxScalarDolfynFaceLoop(
                    mesh, phi.data(), C, phiB.data(), B, vis.data(), C, au.data(), C, su.data(), C, den.data(), C,
                    massFlux.data(), F, te.data(), C, cpB.data(), B, visB.data(), B, rf.data(), 2 * F, g.data(), C,
                    true, false, (T)0.001, 1, 2, (T)1.4, (T)1.2e-3, 3, (T)1.5, 4, (T)1.6, (T)1.7, (T)1.4e-3, (T)1.1,
                    (T)1e-18, (T)1e18, (T)1.8);
            });
        }
        {  // :740-880, :761-762
            auto u = U(C), v = U(C), w = U(C), uB = U(B), vB = U(B), wB = U(B), vis = U(C), visB = U(B),
                 rf = U(2 * F), su = U(C), sv = U(C), sw = U(C), au = U(C), av = U(C), aw = U(C);
            auto du = U3(C), dv = U3(C), dw = U3(C);
            timed("FluxUVWDolfynFaceLoop1Benchmark", [&] {
                return fvm::FluxUVWDolfynFaceLoop1(mesh, (T)0.9, (T)1e-18, u.data(), C, v.data(), C, w.data(), C,
                                                   uB.data(), B, vB.data(), B, wB.data(), B, vis.data(), C,
                                                   visB.data(), B, massFlux.data(), F, du.data(), C, dv.data(), C,
                                                   dw.data(), C, rf.data(), 2 * F, su.data(), C, sv.data(), C,
                                                   sw.data(), C, au.data(), C, av.data(), C, aw.data(), C);
            });
        }
        timed("FluxUVWDolfynRegionLoop1Benchmark", [&] { return fvm::FluxUVWDolfynRegionLoop1<T>(mesh); });
        timed("FluxUVWDolfynBndsLoop1Benchmark", [&] { return fvm::FluxUVWDolfynBndsLoop1<T>(mesh); });
        {  // :938-1133
            auto te = U(C), ed = U(C), den = U(C), vis = U(C), visB = U(B);
            timed("calculateViscosityDolfynCellLoop1Benchmark", [&] {
                return fvm::calculateViscosityDolfynCellLoop1(mesh, (T)1e-18, (T)1e-4, (T)0.232, (T)0.00547,
                                                              te.data(), C, ed.data(), C, den.data(), C, vis.data(),
                                                              C);
            });
            timed("calculateViscosityDolfynRegionLoopBenchmark",
                  [&] { return fvm::calculateViscosityDolfynRegionLoop((T)1.1, mesh); });
            timed("calculateViscosityDolfynBoundaryLoopBenchmark", [&] {
                return fvm::calculateViscosityDolfynBoundaryLoop(mesh, (T)1.0, (T)1e18, (T)1e-18, (T)1.1, (T)1.2,
                                                                 te.data(), C, den.data(), C, vis.data(), C,
                                                                 visB.data(), B);
            });
            timed("calculateViscosityDolfynCellLoop2Benchmark", [&] {
                return fvm::calculateViscosityDolfynCellLoop2(mesh, (T)1.1, vis.data(), C, visB.data(), B);
            });
        }
    }
};


// ----------------------------------------------------------------------------
// GraphBench: the same 16 kernels, with BenchmarkKernels' parameter values, on
// device-resident inputs through the resident C ABI (cupcfd_dev_*), first
// launched eagerly on a stream and then captured once into a hipGraph and
// replayed. One eager repetition before the capture makes every lazily
// allocated workspace exist, so the captured calls allocate nothing. Before
// timing, one eager repetition and one graph launch from the same device
// state are compared byte for byte.
// ----------------------------------------------------------------------------
template <class T> struct DevFns;
#define CUPCFD_DEVFNS(T, S)                                                       \
    template <> struct DevFns<T> {                                                \
        static constexpr auto gradient = cupcfd_dev_gradient_phi_gauss_dolfyn_##S; \
        static constexpr auto mass = cupcfd_dev_flux_mass_face_loop_##S;          \
        static constexpr auto bl1 = cupcfd_dev_flux_mass_boundary_loop1_##S;      \
        static constexpr auto bl2 = cupcfd_dev_flux_mass_boundary_loop2_##S;      \
        static constexpr auto bl3 = cupcfd_dev_flux_mass_boundary_loop3_##S;      \
        static constexpr auto bl4 = cupcfd_dev_flux_mass_boundary_loop4_##S;      \
        static constexpr auto bl5 = cupcfd_dev_flux_mass_boundary_loop5_##S;      \
        static constexpr auto mreg = cupcfd_dev_flux_mass_region_loop_##S;        \
        static constexpr auto scalar = cupcfd_dev_flux_scalar_face_loop_##S;      \
        static constexpr auto uvw = cupcfd_dev_flux_uvw_face_loop1_##S;           \
        static constexpr auto ureg = cupcfd_dev_flux_uvw_region_loop1_##S;        \
        static constexpr auto ubnd = cupcfd_dev_flux_uvw_bnds_loop1_##S;          \
        static constexpr auto vc1 = cupcfd_dev_viscosity_cell_loop1_##S;          \
        static constexpr auto vreg = cupcfd_dev_viscosity_region_loop_##S;        \
        static constexpr auto vbnd = cupcfd_dev_viscosity_boundary_loop_##S;      \
        static constexpr auto vc2 = cupcfd_dev_viscosity_cell_loop2_##S;          \
    };
CUPCFD_DEVFNS(double, f64)
CUPCFD_DEVFNS(float, f32)
#undef CUPCFD_DEVFNS

template <class T> struct GraphBench {
    cupcfd_hip_mesh *dev;
    int C, F, B, R;
    hipStream_t s = nullptr;
    std::mt19937_64 rng;
    struct Buf {
        T *d;
        size_t n;
    };
    std::vector<Buf> bufs;
    int icinl = 0, icout = 0, icsym = 0, icwal = 0;
    int failures = 0;
    GraphBench(cupcfd_hip_mesh *d, int C_, int F_, int B_, int R_, uint64_t seed)
        : dev(d), C(C_), F(F_), B(B_), R(R_), rng(seed) {}
    ~GraphBench() {
        for (auto &b : bufs) (void)hipFree(b.d);
        if (s) (void)hipStreamDestroy(s);
    }
    // a device array of n values drawn from U(1e-6, 1e-2), as the reference's wrappers fill theirs
    T *U(size_t n) {
        std::uniform_real_distribution<double> dist(1e-6, 1e-2);
        std::vector<T> h(n ? n : 1);
        for (auto &x : h) x = (T)dist(rng);
        T *d = nullptr;
        if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        bufs.push_back({d, h.size()});
        return d;
    }
    void ok(int rc, const char *what) {
        if (rc != CUPCFD_E_SUCCESS) {
            failures++;
            std::fprintf(stderr, "graph mode: %s returned %d\n", what, rc);
        }
    }

    T *phi, *phiB, *g, *go, *massFlux, *dudx, *dvdx, *dwdx, *dpdx, *den, *denB, *u, *v, *w, *p, *ar, *su, *rface,
        *te, *teB, *ed, *edB, *vis, *visB, *t, *tB, *flows, *flowRegion, *flowFact, *uB, *vB, *wB, *denB2, *su5,
        *sphi, *sphiB, *svis, *sau, *ssu, *sden, *ste, *scpB, *svisB, *srf, *sg, *uu, *uv, *uw, *uuB, *uvB, *uwB,
        *uvis, *uvisB, *urf, *usu, *usv, *usw, *uau, *uav, *uaw, *udu, *udv, *udw, *vte, *ved, *vden, *vvis, *vvisB;

    bool alloc() {
        T **all[] = {&phi, &phiB, &g, &go, &massFlux, &dudx, &dvdx, &dwdx, &dpdx, &den, &denB, &u, &v, &w, &p,
                     &ar, &su, &rface, &te, &teB, &ed, &edB, &vis, &visB, &t, &tB, &flows, &flowRegion, &flowFact,
                     &uB, &vB, &wB, &denB2, &su5, &sphi, &sphiB, &svis, &sau, &ssu, &sden, &ste, &scpB, &svisB,
                     &srf, &sg, &uu, &uv, &uw, &uuB, &uvB, &uwB, &uvis, &uvisB, &urf, &usu, &usv, &usw, &uau, &uav,
                     &uaw, &udu, &udv, &udw, &vte, &ved, &vden, &vvis, &vvisB};
        const size_t c = C, f = F, b = B, r = R;
        const size_t n[] = {c, b, 3 * c, 3 * c, f, 3 * c, 3 * c, 3 * c, 3 * c, c, b, c, c, c, c,
                            c, c, 2 * f, c, b, c, b, c, b, c, b, 4, r, r,
                            b, b, b, b, c, c, b, c, c, c, c, c, b, b,
                            2 * f, 3 * c, c, c, c, b, b, b, c, b, 2 * f, c, c, c, c, c,
                            c, 3 * c, 3 * c, 3 * c, c, c, c, c, b};
        static_assert(sizeof(all) / sizeof(all[0]) == sizeof(n) / sizeof(n[0]), "one size per buffer");
        for (size_t i = 0; i < sizeof(n) / sizeof(n[0]); i++)
            if (!(*all[i] = U(n[i]))) return false;
        return hipStreamCreate(&s) == hipSuccess;
    }

    // BenchmarkKernels::runBenchmark's 16 calls (BenchmarkKernels.cpp:55-115), resident
    void repetition() {
        using D = DevFns<T>;
        ok(D::gradient(dev, 1, phi, phiB, g, go, s), "gradient");
        ok(D::mass(dev, dudx, dvdx, dwdx, dpdx, den, denB, u, v, w, massFlux, p, ar, su, rface, (T)1e-18, &icinl,
                   &icout, &icsym, &icwal, 0, 0, 1, 0, te, teB, ed, edB, vis, visB, t, tB, s),
           "mass");
        ok(D::bl1(dev, massFlux, flows + 0, s), "mass_bl1");
        ok(D::bl2(dev, massFlux, flowRegion, flows + 1, s), "mass_bl2");
        ok(D::bl3(dev, (T)1.1, flows + 2, s), "mass_bl3");
        ok(D::bl4(dev, massFlux, uB, vB, wB, denB2, (T)1.0, flows + 1, s), "mass_bl4");
        ok(D::bl5(dev, massFlux, su5, uB, vB, wB, (T)1.1, 1, 1, 1, flowFact, flows + 3, s), "mass_bl5");
        ok(D::mreg(dev, flowFact, flowRegion, (T)0.0, s), "mass_region");
        ok(D::scalar(dev, sphi, sphiB, svis, sau, ssu, sden, massFlux, ste, scpB, svisB, srf, sg, 1, 0, (T)0.001, 1,
                     2, (T)1.4, (T)1.2e-3, 3, (T)1.5, 4, (T)1.6, (T)1.7, (T)1.4e-3, (T)1.1, (T)1e-18, (T)1e18,
                     (T)1.8, s),
           "scalar");
        ok(D::uvw(dev, (T)0.9, (T)1e-18, uu, uv, uw, uuB, uvB, uwB, uvis, uvisB, massFlux, udu, udv, udw, urf, usu,
                  usv, usw, uau, uav, uaw, s),
           "uvw");
        ok(D::ureg(dev, s), "uvw_region1");
        ok(D::ubnd(dev, s), "uvw_bnds1");
        ok(D::vc1(dev, (T)1e-18, (T)1e-4, (T)0.232, (T)0.00547, vte, ved, vden, vvis, s), "visc_cell1");
        ok(D::vreg((T)1.1, dev, s), "visc_region");
        ok(D::vbnd(dev, (T)1.0, (T)1e18, (T)1e-18, (T)1.1, (T)1.2, vte, vden, vvis, vvisB, s), "visc_boundary");
        ok(D::vc2(dev, (T)1.1, vvis, vvisB, s), "visc_cell2");
    }

    std::vector<std::vector<T>> snapshot() {
        std::vector<std::vector<T>> out;
        for (auto &b : bufs) {
            out.emplace_back(b.n);
            (void)hipMemcpy(out.back().data(), b.d, b.n * sizeof(T), hipMemcpyDeviceToHost);
        }
        return out;
    }
    void restore(const std::vector<std::vector<T>> &st) {
        for (size_t i = 0; i < bufs.size(); i++)
            (void)hipMemcpy(bufs[i].d, st[i].data(), bufs[i].n * sizeof(T), hipMemcpyHostToDevice);
        (void)cupcfd_hip_mesh_push_state(dev);
    }

    // eager and graph repetitions per second; returns false on any failure
    bool run(int reps, double &eager_ms, double &graph_ms, bool &same) {
        if (!alloc()) return false;
        (void)cupcfd_hip_mesh_push_state(dev);
        repetition();  // workspaces come into existence here
        if (hipStreamSynchronize(s) != hipSuccess || failures) return false;
        (void)cupcfd_hip_mesh_push_state(dev);  // the eager and the graph repetition start from the same state
        const auto st0 = snapshot();
        repetition();
        if (hipStreamSynchronize(s) != hipSuccess || failures) return false;
        const auto eager = snapshot();
        restore(st0);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        repetition();
        if (hipStreamEndCapture(s, &graph) != hipSuccess || failures) return false;
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) return false;
        if (hipGraphLaunch(exec, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return false;
        const auto replay = snapshot();
        same = true;
        for (size_t i = 0; i < eager.size(); i++)
            if (std::memcmp(eager[i].data(), replay[i].data(), eager[i].size() * sizeof(T)) != 0) same = false;
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        float ms = 0;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; r++) repetition();
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        eager_ms = ms / reps;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; r++) (void)hipGraphLaunch(exec, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        graph_ms = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return failures == 0;
    }
};

static double num(const Json *o, const char *k, double dflt) {
    const Json *v = o ? o->get(k) : nullptr;
    return v && v->kind == Json::NUM ? v->num : (v && v->kind == Json::BOOL ? (v->b ? 1.0 : 0.0) : dflt);
}

template <class T> int run_all(const Json &cfg, const MeshSpec &base) {
    const Json *benches = cfg.get("Benchmarks");
    if (!benches || benches->kind != Json::ARR) {
        std::fprintf(stderr, "cupcfd.json: no Benchmarks array\n");
        return CUPCFD_E_CONFIG_INVALID_VALUE;  // E_CONFIG_OPT_NOT_FOUND in the reference
    }
    std::printf("{\"benchmarks\": [");
    bool first = true;
    int status = CUPCFD_E_SUCCESS;
    for (const Json &b : benches->arr) {
        const Json *bk = b.get("BenchmarkKernels");
        if (!bk) {
            std::fprintf(stderr, "skipping %s (not part of this build)\n",
                         b.obj.empty() ? "?" : b.obj[0].first.c_str());
            continue;
        }
        const Json *nm = bk->get("BenchmarkName");
        const Json *reps = bk->get("Repetitions");
        if (!nm || nm->kind != Json::STR || !reps || reps->kind != Json::NUM) {
            std::fprintf(stderr, "BenchmarkKernels needs BenchmarkName and Repetitions\n");
            return CUPCFD_E_CONFIG_INVALID_VALUE;
        }
        MeshSpec spec = base;
        cupcfd_structgen_params &p = spec.sg;
        p.seed = (uint64_t)num(bk, "Seed", (double)p.seed);
        p.permute_cells = p.permute_faces = (int)num(bk, "Permute", 0);  // StructGen grids only
        p.regions = (int)num(bk, "Typed", 0);
        SoAMesh<T> mesh;
        int rc = load_mesh(mesh, spec);
        if (rc != CUPCFD_E_SUCCESS) {
            std::fprintf(stderr, "mesh source returned %d\n", rc);
            return rc;
        }
        // "MeshBuild": "host" repacks the device mesh on the host, anything
        // else on the GPU; the repack runs here, timed on its own, so the
        // first kernel's time is the kernel's alone
        const Json *mb = bk->get("MeshBuild");
        const bool on_device = !(mb && mb->kind == Json::STR && mb->str == "host");
        cupcfd_hip::shim::device_repack() = on_device;
        auto ts = std::chrono::steady_clock::now();
        cupcfd_hip_mesh *built = cupcfd_hip::shim::device<T>(mesh, &rc);
        if (!built) {
            std::fprintf(stderr, "device mesh build returned %d\n", rc);
            return rc;
        }
        const double setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        // one hash over the per-array digests: equal for the host and the GPU repack
        uint64_t dig[64], mesh_hash = 1469598103934665603ull;
        int32_t nd = 0;
        if ((rc = cupcfd_hip_mesh_digest(built, dig, 64, &nd)) != CUPCFD_E_SUCCESS) return rc;
        for (int i = 0; i < nd; i++) mesh_hash = (mesh_hash ^ dig[i]) * 1099511628211ull;
        Bench<T> bench(mesh, p.seed ^ 0x5eedull);
        const int R = (int)reps->num;
        auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < R; r++) bench.repetition();
        double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("%s{\"BenchmarkName\": \"%s\", \"Repetitions\": %d, \"LocalCells\": %d, \"LocalFaces\": %d, "
                    "\"LocalBounds\": %d, \"mesh_build\": \"%s\", \"mesh_setup_s\": %.6f, "
                    "\"mesh_digest\": \"%016llx\", \"total_s\": %.6f, \"failures\": %d, \"kernels\": {",
                    first ? "" : ", ", nm->str.c_str(), R, mesh.properties.lTCells, mesh.properties.lFaces,
                    mesh.properties.lBoundaries, on_device ? "device" : "host", setup,
                    (unsigned long long)mesh_hash, total, bench.failures);
        bool fk = true;
        for (auto &kv : bench.times) {
            double s = 0;
            for (double t : kv.second) s += t;
            std::printf("%s\"%s\": {\"calls\": %zu, \"mean_s\": %.9f}", fk ? "" : ", ", kv.first.c_str(),
                        kv.second.size(), s / kv.second.size());
            fk = false;
        }
        std::printf("}");
        if (num(bk, "Graph", 0)) {
            int grc = CUPCFD_E_SUCCESS;
            cupcfd_hip_mesh *dev = cupcfd_hip::shim::device<T>(mesh, &grc);
            double em = 0, gm = 0;
            bool same = false;
            bool good = dev != nullptr;
            if (good) {
                GraphBench<T> gb(dev, mesh.properties.lTCells, mesh.properties.lFaces, mesh.properties.lBoundaries,
                                 mesh.properties.lRegions, p.seed ^ 0x6a5eedull);
                good = gb.run(R, em, gm, same);
            }
            std::printf(", \"graph\": {\"ok\": %s, \"replay_matches_eager\": %s, \"eager_ms_per_repetition\": %.6f, "
                        "\"graph_ms_per_repetition\": %.6f}",
                        good ? "true" : "false", same ? "true" : "false", em, gm);
            if (!good || !same) status = CUPCFD_E_ERROR;
        }
        std::printf("}");
        first = false;
        cupcfd_hip::shim::forget(&mesh);
        if (bench.failures) status = CUPCFD_E_ERROR;
    }
    std::printf("], \"library\": \"%s\"}\n", cupcfd_hip_version());
    return status;
}

int main(int argc, char **argv) {
    const char *path = argc > 1 ? argv[1] : "./cupcfd.json";
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", path);
        return 1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    std::string text = ss.str();
    Parser ps(text);
    Json cfg = ps.value();
    if (!ps.ok || cfg.kind != Json::OBJ) {
        std::fprintf(stderr, "%s: malformed JSON\n", path);
        return 1;
    }
    const Json *dt = cfg.get("DataTypes");
    const Json *fdt = dt ? dt->get("FloatDataType") : nullptr;
    const Json *mdt = dt ? dt->get("MeshDataType") : nullptr;
    if (mdt && mdt->str == "MiniAoS")
        std::fprintf(stderr, "MeshDataType MiniAoS: running the MiniSoA layout this build provides\n");
    const Json *mesh = cfg.get("Mesh");
    const Json *src = mesh ? mesh->get("MeshSource") : nullptr;
    const Json *sg = src ? src->get("MeshSourceStructGen") : nullptr;
    const Json *sf = src ? src->get("MeshSourceFile") : nullptr;
    MeshSpec spec;
    if (sf) {
        const Json *fmt = sf->get("FileFormat");
        const Json *fp = sf->get("FilePath");
        if (!fmt || fmt->kind != Json::STR || fmt->str != "HDF5" || !fp || fp->kind != Json::STR) {
            std::fprintf(stderr, "MeshSourceFile needs FileFormat \"HDF5\" and a FilePath\n");
            return 1;
        }
        spec.file = fp->str;
    } else if (!sg) {
        std::fprintf(stderr, "Mesh.MeshSource needs MeshSourceStructGen or MeshSourceFile\n");
        return 1;
    }
    cupcfd_structgen_params &p = spec.sg;
    p.nx = (int)num(sg, "CellX", 0);
    p.ny = (int)num(sg, "CellY", 0);
    p.nz = (int)num(sg, "CellZ", 0);
    p.xmin = num(sg, "SpatialXMin", 0);
    p.xmax = num(sg, "SpatialXMax", 1);
    p.ymin = num(sg, "SpatialYMin", 0);
    p.ymax = num(sg, "SpatialYMax", 1);
    p.zmin = num(sg, "SpatialZMin", 0);
    p.zmax = num(sg, "SpatialZMax", 1);
    p.seed = 1;
    int rc = (fdt && fdt->str == "float") ? run_all<float>(cfg, spec) : run_all<double>(cfg, spec);
    return rc == CUPCFD_E_SUCCESS ? 0 : 2;
}

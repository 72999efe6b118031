// ORACLE — test infrastructure only.
//
// Runs CUP-CFD's own fvm kernels (include/fvm/implementation/component/*.ipp),
// compiled from the reference sources by oracle/ref.mk, on a
// mesh and inputs read from a case file, and writes every output back. It is
// the strongest check of the C restatement in oracle_kernels.inc: both run on
// the same mesh arrays and inputs, and the restatement must give the same bits.
//
// The mesh is built through CupCfdSoAMesh's public build API
// (addRegion / addBoundary / addCell / addFace / finalize,
// src/geometry/mesh/implementation/component/CupCfdSoAMesh.cpp:52-463, 842-888)
// in id order, so the reference's local ids equal the case file's ids; the
// harness checks that after finalize. Geometry (lambda, S, Xf, xpac, xnac,
// rlencos, area, Cc, vol) is taken from the case file as given. Faces and
// boundaries name four placeholder vertices: the fvm kernels read no vertex.
//
// Two ranks. The reference's mesh finalize cannot complete on one rank: the
// distributed graph broadcasts each rank's ghost list
// (src/data_structures/implementation/component/DistributedAdjacencyList.cpp:355)
// and BroadcastMPI returns E_NO_DATA for an empty one
// (include/comm/implementation/component/BroadcastMPI.ipp:28-30), so a
// partition without ghost cells fails. The harness therefore runs under
// `mpiexec -n 2`. In kernel mode rank 0 holds the case mesh plus one detached
// cell pair: a local cell D (id C), a ghost G (id C+1) owned by rank 1, and
// the face D-G (id F, the last face). D and G touch no case cell or face, the
// face comes after every case face (so no carried value changes), and every
// cell and face array is padded by those two cells and one face; the padding
// is dropped from the results.
//
// A second mode builds the reference's own StructGen mesh
// (MeshConfig::buildUnstructuredMesh over MeshSourceStructGenConfig and the
// naive partitioner, include/geometry/mesh/implementation/config/MeshConfig.ipp:66-121)
// on both ranks, and each rank writes its partition's arrays with their build
// labels (result file suffixed .0 / .1), which pins this repository's
// StructGen generator element by element.
//
// Case file format (little endian): "CUPREF1\n", then records of
//   u32 name length, name, u8 kind (0 f64, 1 f32, 2 i32, 3 u8), u64 count, data.
// Names: "mode" (u8 text), "precision" (i32: 64 or 32), "kernel" (u8 text),
// "mesh.<field>" arrays, "a.<arg>" arrays, "s.<arg>" scalars (f64 or i32).
// The result file holds "rc" (i32), every "a.<arg>" after the call and the
// mesh state "mesh.<state field>".

#include <mpi.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "Communicator.h"
#include "CupCfdSoAMesh.h"
#include "EuclideanPoint.h"
#include "EuclideanVector.h"
#include "GradientKernels.h"
#include "MassKernels.h"
#include "MeshConfig.h"
#include "MeshSourceStructGenConfig.h"
#include "PartitionerNaiveConfig.h"
#include "ScalarKernels.h"
#include "UVWKernels.h"
#include "ViscosityKernels.h"

namespace euc = cupcfd::geometry::euclidean;
namespace meshgeo = cupcfd::geometry::mesh;

enum Kind : uint8_t { K_F64 = 0, K_F32 = 1, K_I32 = 2, K_U8 = 3 };

struct Rec {
    Kind kind;
    uint64_t n;
    std::vector<unsigned char> bytes;
};

static size_t kind_size(Kind k) { return k == K_F64 ? 8 : (k == K_U8 ? 1 : 4); }

using Records = std::map<std::string, Rec>;

static Records read_case(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "CUPREF1\n", 8) != 0)
        throw std::runtime_error("bad case file magic");
    Records out;
    for (;;) {
        uint32_t len;
        if (std::fread(&len, 4, 1, f) != 1) break;
        std::string name(len, '\0');
        uint8_t kind;
        uint64_t n;
        if (std::fread(&name[0], 1, len, f) != len || std::fread(&kind, 1, 1, f) != 1 || std::fread(&n, 8, 1, f) != 1)
            throw std::runtime_error("truncated case file");
        Rec r{Kind(kind), n, std::vector<unsigned char>(n * kind_size(Kind(kind)))};
        if (!r.bytes.empty() && std::fread(r.bytes.data(), 1, r.bytes.size(), f) != r.bytes.size())
            throw std::runtime_error("truncated record " + name);
        out[name] = std::move(r);
    }
    std::fclose(f);
    return out;
}

static void write_case(const char* path, const Records& recs) {
    FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    std::fwrite("CUPREF1\n", 1, 8, f);
    for (const auto& kv : recs) {
        uint32_t len = uint32_t(kv.first.size());
        uint8_t kind = kv.second.kind;
        std::fwrite(&len, 4, 1, f);
        std::fwrite(kv.first.data(), 1, len, f);
        std::fwrite(&kind, 1, 1, f);
        std::fwrite(&kv.second.n, 8, 1, f);
        if (!kv.second.bytes.empty()) std::fwrite(kv.second.bytes.data(), 1, kv.second.bytes.size(), f);
    }
    std::fclose(f);
}

template <class T> struct KindOf;
template <> struct KindOf<double> { static constexpr Kind k = K_F64; };
template <> struct KindOf<float> { static constexpr Kind k = K_F32; };
template <> struct KindOf<int> { static constexpr Kind k = K_I32; };
template <> struct KindOf<unsigned char> { static constexpr Kind k = K_U8; };

template <class X>
static X* arr(Records& r, const std::string& name, uint64_t* n = nullptr) {
    auto it = r.find(name);
    if (it == r.end()) throw std::runtime_error("missing record " + name);
    if (it->second.kind != KindOf<X>::k) throw std::runtime_error("wrong kind for " + name);
    if (n) *n = it->second.n;
    return it->second.n ? reinterpret_cast<X*>(it->second.bytes.data()) : nullptr;
}

template <class X>
static void put(Records& r, const std::string& name, const X* p, uint64_t n) {
    Rec rec{KindOf<X>::k, n, std::vector<unsigned char>(n * sizeof(X))};
    if (n) std::memcpy(rec.bytes.data(), p, n * sizeof(X));
    r[name] = std::move(rec);
}

static std::string text(Records& r, const std::string& name) {
    uint64_t n;
    unsigned char* p = arr<unsigned char>(r, name, &n);
    return std::string(reinterpret_cast<char*>(p), n);
}

static double sval(Records& r, const std::string& name) {
    auto it = r.find(name);
    if (it == r.end()) throw std::runtime_error("missing scalar " + name);
    if (it->second.kind == K_F64) return *reinterpret_cast<double*>(it->second.bytes.data());
    if (it->second.kind == K_I32) return *reinterpret_cast<int*>(it->second.bytes.data());
    throw std::runtime_error("scalar " + name + " must be f64 or i32");
}

static int ival(Records& r, const std::string& name) { return int(sval(r, name)); }

// packed xyz <-> the reference's vector objects
template <class T>
static std::vector<euc::EuclideanVector<T, 3>> to_v3(const T* p, uint64_t n3) {
    std::vector<euc::EuclideanVector<T, 3>> v;
    v.reserve(n3 / 3);
    for (uint64_t i = 0; i < n3 / 3; i++) v.emplace_back(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    return v;
}

template <class T>
static void from_v3(const std::vector<euc::EuclideanVector<T, 3>>& v, T* p) {
    for (size_t i = 0; i < v.size(); i++)
        for (int k = 0; k < 3; k++) p[3 * i + k] = v[i].cmp[k];
}

// ---------------------------------------------------------------------------
// Mesh built from case-file arrays
// ---------------------------------------------------------------------------
static void ok(cupcfd::error::eCodes s, const char* what, int i) {
    if (s != cupcfd::error::E_SUCCESS)
        throw std::runtime_error(std::string(what) + " failed at " + std::to_string(i) + " with eCode " +
                                 std::to_string(int(s)));
}

// four placeholder vertices shared by every face and boundary
static int vlab[4] = {0, 1, 2, 3};

template <class T>
static void add_vertices(meshgeo::CupCfdSoAMesh<int, T, int>& m) {
    for (int v = 0; v < 4; v++) {
        euc::EuclideanPoint<T, 3> p(T(v & 1), T((v >> 1) & 1), T(0));
        ok(m.addVertex(v, p), "addVertex", v);
    }
}

// The detached pair D (label C, on rank 0) - G (label C+1, on rank 1) and
// their face (label F), built identically on both ranks.
template <class T>
static void add_pair(meshgeo::CupCfdSoAMesh<int, T, int>& m, int C, int F, bool rank0) {
    euc::EuclideanPoint<T, 3> pd(T(0), T(0), T(0)), pg(T(1), T(0), T(0));
    ok(m.addCell(C, pd, T(1), rank0), "addCell", C);
    ok(m.addCell(C + 1, pg, T(1), !rank0), "addCell", C + 1);
    euc::EuclideanVector<T, 3> s(T(1), T(0), T(0));
    euc::EuclideanPoint<T, 3> x(T(0.5), T(0), T(0)), z(T(0), T(0), T(0));
    ok(m.addFace(F, C, C + 1, false, T(0.5), s, vlab, 4, x, z, z, T(0), T(1)), "addFace", F);
}

template <class T>
static meshgeo::CupCfdSoAMesh<int, T, int>* build_mesh(Records& r, cupcfd::comm::Communicator& comm) {
    using Mesh = meshgeo::CupCfdSoAMesh<int, T, int>;
    const int C = ival(r, "s.nCells"), F = ival(r, "s.nFaces"), B = ival(r, "s.nBoundaries"), R = ival(r, "s.nRegions");
    Mesh* m = new Mesh(comm);
    add_vertices<T>(*m);
    if (comm.rank != 0) {  // rank 1: G and its ghost neighbour D
        add_pair<T>(*m, C, F, false);
        ok(m->finalize(), "finalize", 1);
        return m;
    }
    int* c1 = arr<int>(r, "mesh.faceCell1");
    int* c2 = arr<int>(r, "mesh.faceCell2");
    int* fb = arr<int>(r, "mesh.faceBnd");
    T* lam = arr<T>(r, "mesh.faceLambda");
    T* nrm = arr<T>(r, "mesh.faceNorm");
    T* fc = arr<T>(r, "mesh.faceCenter");
    T* xp = arr<T>(r, "mesh.faceXpac");
    T* xn = arr<T>(r, "mesh.faceXnac");
    T* rl = arr<T>(r, "mesh.faceRLencos");
    T* ar = arr<T>(r, "mesh.faceArea");
    T* cc = arr<T>(r, "mesh.cellCenter");
    T* vol = arr<T>(r, "mesh.cellVolume");
    int* breg = arr<int>(r, "mesh.bndRegion");
    T* bdist = arr<T>(r, "mesh.bndDistance");
    int* rtype = arr<int>(r, "mesh.regType");
    unsigned char* rstd = arr<unsigned char>(r, "mesh.regStd");
    unsigned char* rflux = arr<unsigned char>(r, "mesh.regFlux");
    unsigned char* radiab = arr<unsigned char>(r, "mesh.regAdiab");
    T* ryl = arr<T>(r, "mesh.regYLog");
    T* rel = arr<T>(r, "mesh.regELog");
    T* rdens = arr<T>(r, "mesh.regDensity");
    T* rte = arr<T>(r, "mesh.regTurbKE");
    T* rtd = arr<T>(r, "mesh.regTurbDiss");
    T* rspl = arr<T>(r, "mesh.regSplvl");
    T* rden = arr<T>(r, "mesh.regDen");
    T* rr = arr<T>(r, "mesh.regR");
    T* rt = arr<T>(r, "mesh.regT");
    T* rft = arr<T>(r, "mesh.regForceTangent");
    T* ruvw = arr<T>(r, "mesh.regUVW");

    for (int i = 0; i < R; i++) {
        euc::EuclideanVector<T, 3> ft(rft[3 * i], rft[3 * i + 1], rft[3 * i + 2]);
        euc::EuclideanVector<T, 3> uvw(ruvw[3 * i], ruvw[3 * i + 1], ruvw[3 * i + 2]);
        std::string name = "region" + std::to_string(i);
        ok(m->addRegion(i, meshgeo::RType(rtype[i]), rstd[i] != 0, rflux[i] != 0, radiab[i] != 0, ryl[i], rel[i],
                        rdens[i], rte[i], rtd[i], rspl[i], rden[i], rr[i], rt[i], ft, uvw, name),
           "addRegion", i);
    }
    for (int b = 0; b < B; b++) ok(m->addBoundary(b, breg[b], vlab, 4, bdist[b]), "addBoundary", b);
    for (int c = 0; c < C; c++) {
        euc::EuclideanPoint<T, 3> p(cc[3 * c], cc[3 * c + 1], cc[3 * c + 2]);
        ok(m->addCell(c, p, vol[c], true), "addCell", c);
    }
    for (int f = 0; f < F; f++) {
        bool bnd = fb[f] > -1;
        euc::EuclideanVector<T, 3> s(nrm[3 * f], nrm[3 * f + 1], nrm[3 * f + 2]);
        euc::EuclideanPoint<T, 3> x(fc[3 * f], fc[3 * f + 1], fc[3 * f + 2]);
        euc::EuclideanPoint<T, 3> pa(xp[3 * f], xp[3 * f + 1], xp[3 * f + 2]);
        euc::EuclideanPoint<T, 3> na(xn[3 * f], xn[3 * f + 1], xn[3 * f + 2]);
        ok(m->addFace(f, c1[f], bnd ? fb[f] : c2[f], bnd, lam[f], s, vlab, 4, x, pa, na, rl[f], ar[f]), "addFace", f);
    }
    add_pair<T>(*m, C, F, true);
    ok(m->finalize(), "finalize", 0);
    // the reference's local ids must be the case file's ids (D and G last)
    if (m->properties.lTCells != C + 2 || m->properties.lOCells != C + 1 || m->properties.lFaces != F + 1 ||
        m->properties.lBoundaries != B || m->properties.lRegions != R)
        throw std::runtime_error("mesh counts changed in finalize");
    for (int c = 0; c < C + 2; c++)
        if (m->getCellID(c) != c) throw std::runtime_error("cell " + std::to_string(c) + " renumbered");
    for (int f = 0; f < F; f++)
        if (m->getFaceID(f) != f || m->getFaceCell1ID(f) != c1[f] || m->getFaceCell2ID(f) != c2[f] ||
            m->getFaceBoundaryID(f) != fb[f])
            throw std::runtime_error("face " + std::to_string(f) + " renumbered");
    if (m->getFaceID(F) != F || m->getFaceCell1ID(F) != C || m->getFaceCell2ID(F) != C + 1)
        throw std::runtime_error("detached face renumbered");
    return m;
}

template <class T>
static void set_state(Records& r, meshgeo::CupCfdSoAMesh<int, T, int>& m) {
    const int B = m.properties.lBoundaries, R = m.properties.lRegions;
    T* yp = arr<T>(r, "mesh.bndYPlus");
    T* up = arr<T>(r, "mesh.bndUPlus");
    T* q = arr<T>(r, "mesh.bndQ");
    T* h = arr<T>(r, "mesh.bndH");
    T* t = arr<T>(r, "mesh.bndT");
    T* sh = arr<T>(r, "mesh.bndShear");
    T* yl = arr<T>(r, "mesh.regYLog");
    T* ft = arr<T>(r, "mesh.regForceTangent");
    for (int b = 0; b < B; b++) {
        m.setBoundaryYPlus(b, yp[b]);
        m.setBoundaryUPlus(b, up[b]);
        m.setBoundaryQ(b, q[b]);
        m.setBoundaryH(b, h[b]);
        m.setBoundaryT(b, t[b]);
        euc::EuclideanVector<T, 3> s(sh[3 * b], sh[3 * b + 1], sh[3 * b + 2]);
        m.setBoundaryShear(b, s);
    }
    for (int i = 0; i < R; i++) {
        m.setRegionYLog(i, yl[i]);
        euc::EuclideanVector<T, 3> v(ft[3 * i], ft[3 * i + 1], ft[3 * i + 2]);
        m.setRegionForceTangent(i, v);
    }
}

template <class T>
static void get_state(Records& out, meshgeo::CupCfdSoAMesh<int, T, int>& m) {
    const int B = m.properties.lBoundaries, R = m.properties.lRegions;
    std::vector<T> yp(B), up(B), q(B), h(B), t(B), sh(3 * B), yl(R), ft(3 * R);
    for (int b = 0; b < B; b++) {
        yp[b] = m.getBoundaryYPlus(b);
        up[b] = m.getBoundaryUPlus(b);
        q[b] = m.getBoundaryQ(b);
        h[b] = m.getBoundaryH(b);
        t[b] = m.getBoundaryT(b);
        euc::EuclideanVector<T, 3> s = m.getBoundaryShear(b);
        for (int k = 0; k < 3; k++) sh[3 * b + k] = s.cmp[k];
    }
    for (int i = 0; i < R; i++) {
        yl[i] = m.getRegionYLog(i);
        euc::EuclideanVector<T, 3> v = m.getRegionForceTangent(i);
        for (int k = 0; k < 3; k++) ft[3 * i + k] = v.cmp[k];
    }
    put(out, "mesh.bndYPlus", yp.data(), B);
    put(out, "mesh.bndUPlus", up.data(), B);
    put(out, "mesh.bndQ", q.data(), B);
    put(out, "mesh.bndH", h.data(), B);
    put(out, "mesh.bndT", t.data(), B);
    put(out, "mesh.bndShear", sh.data(), 3 * B);
    put(out, "mesh.regYLog", yl.data(), R);
    put(out, "mesh.regForceTangent", ft.data(), 3 * R);
}

// ---------------------------------------------------------------------------
// One kernel call with the reference's signature (include/fvm/implementation/component/*.h)
// ---------------------------------------------------------------------------
template <class T>
struct Args {
    Records& r;
    std::map<std::string, std::vector<euc::EuclideanVector<T, 3>>> v3;
    explicit Args(Records& rr) : r(rr) {}
    T* a(const std::string& n) { return arr<T>(r, "a." + n); }
    int n(const std::string& nm) {
        uint64_t k;
        arr<T>(r, "a." + nm, &k);
        return int(k);
    }
    euc::EuclideanVector<T, 3>* v(const std::string& nm) {
        uint64_t k;
        T* p = arr<T>(r, "a." + nm, &k);
        v3[nm] = to_v3<T>(p, k);
        return v3[nm].data();
    }
    int nv(const std::string& nm) { return int(v3.at(nm).size()); }
    T s(const std::string& nm) { return T(sval(r, "s." + nm)); }
    int i(const std::string& nm) { return ival(r, "s." + nm); }
    void writeback() {
        for (auto& kv : v3) from_v3<T>(kv.second, arr<T>(r, "a." + kv.first));
    }
};

template <class T>
static int run_kernel(const std::string& k, meshgeo::CupCfdSoAMesh<int, T, int>& m, Args<T>& A) {
    using namespace cupcfd::fvm;
    if (k == "gradient") {
        euc::EuclideanVector<T, 3>* g = A.v("dPhidxCell");
        euc::EuclideanVector<T, 3>* go = A.v("dPhidxoCell");
        return GradientPhiGaussDolfyn(m, A.i("nGradient"), A.a("phiCell"), A.n("phiCell"), A.a("phiBoundary"),
                                      A.n("phiBoundary"), g, A.nv("dPhidxCell"), go, A.nv("dPhidxoCell"));
    }
    if (k == "mass") {
        euc::EuclideanVector<T, 3>* du = A.v("dudx");
        euc::EuclideanVector<T, 3>* dv = A.v("dvdx");
        euc::EuclideanVector<T, 3>* dw = A.v("dwdx");
        euc::EuclideanVector<T, 3>* dp = A.v("dpdx");
        int* cnt = arr<int>(A.r, "a.counters");
        return FluxMassDolfynFaceLoop(
            m, du, A.nv("dudx"), dv, A.nv("dvdx"), dw, A.nv("dwdx"), dp, A.nv("dpdx"), A.a("denCell"), A.n("denCell"),
            A.a("denBoundary"), A.n("denBoundary"), A.a("uCell"), A.n("uCell"), A.a("vCell"), A.n("vCell"),
            A.a("wCell"), A.n("wCell"), A.a("massFlux"), A.n("massFlux"), A.a("p"), A.n("p"), A.a("ar"), A.n("ar"),
            A.a("su"), A.n("su"), A.a("rface"), A.n("rface"), A.s("small"), cnt + 0, cnt + 1, cnt + 2, cnt + 3,
            A.i("solveTurbEnergy") != 0, A.i("solveTurbDiss") != 0, A.i("solveVisc") != 0, A.i("solveEnthalpy") != 0,
            A.a("teCell"), A.n("teCell"), A.a("teBoundary"), A.n("teBoundary"), A.a("edCell"), A.n("edCell"),
            A.a("edBoundary"), A.n("edBoundary"), A.a("viseffCell"), A.n("viseffCell"), A.a("viseffBoundary"),
            A.n("viseffBoundary"), A.a("tCell"), A.n("tCell"), A.a("tBoundary"), A.n("tBoundary"));
    }
    if (k == "mass_bl1") return FluxMassDolfynBoundaryLoop1(m, A.a("massFlux"), A.n("massFlux"), A.a("flowin"));
    if (k == "mass_bl2")
        return FluxMassDolfynBoundaryLoop2(m, A.a("massFlux"), A.n("massFlux"), A.a("flowRegion"), A.n("flowRegion"),
                                           A.a("flowout"));
    if (k == "mass_bl3") return FluxMassDolfynBoundaryLoop3(m, A.s("flowin"), A.a("ratearea"));
    if (k == "mass_bl4")
        return FluxMassDolfynBoundaryLoop4(m, A.a("massFlux"), A.n("massFlux"), A.a("uBoundary"), A.n("uBoundary"),
                                           A.a("vBoundary"), A.n("vBoundary"), A.a("wBoundary"), A.n("wBoundary"),
                                           A.a("denBoundary"), A.n("denBoundary"), A.s("ratearea"), A.a("flowout"));
    if (k == "mass_bl5")
        return FluxMassDolfynBoundaryLoop5(m, A.a("massFlux"), A.n("massFlux"), A.a("su"), A.n("su"),
                                           A.a("uBoundary"), A.n("uBoundary"), A.a("vBoundary"), A.n("vBoundary"),
                                           A.a("wBoundary"), A.n("wBoundary"), A.s("fact"), A.i("solveU") != 0,
                                           A.i("solveV") != 0, A.i("solveW") != 0, A.a("flowFact"), A.n("flowFact"),
                                           A.a("flowout2"));
    if (k == "mass_region")
        return FluxMassDolfynRegionLoop(m, A.a("flowFact"), A.n("flowFact"), A.a("flowRegion"), A.n("flowRegion"),
                                        A.s("flowIn"));
    if (k == "scalar") {
        euc::EuclideanVector<T, 3>* g = A.v("dPhidx");
        return FluxScalarDolfynFaceLoop(
            m, A.a("PhiCell"), A.n("PhiCell"), A.a("PhiBoundary"), A.n("PhiBoundary"), A.a("VisEff"), A.n("VisEff"),
            A.a("Au"), A.n("Au"), A.a("Su"), A.n("Su"), A.a("Den"), A.n("Den"), A.a("MassFlux"), A.n("MassFlux"),
            A.a("TE"), A.n("TE"), A.a("CpBoundary"), A.n("CpBoundary"), A.a("visEffBoundary"),
            A.n("visEffBoundary"), A.a("RFace"), A.n("RFace"), g, A.nv("dPhidx"), A.i("SolveTurb") != 0,
            A.i("SolveEnthalpy") != 0, A.s("vislam"), A.i("ivar"), A.i("VarT"), A.s("Sigma_T"), A.s("Prandtl"),
            A.i("VarTE"), A.s("Sigma_k"), A.i("VarED"), A.s("Sigma_e"), A.s("Sigma_s"), A.s("Schmidt"),
            A.s("GammaBlend"), A.s("Small"), A.s("Large"), A.s("TMCmu"));
    }
    if (k == "uvw") {
        euc::EuclideanVector<T, 3>* du = A.v("dudx");
        euc::EuclideanVector<T, 3>* dv = A.v("dvdx");
        euc::EuclideanVector<T, 3>* dw = A.v("dwdx");
        return FluxUVWDolfynFaceLoop1(
            m, A.s("gammaBlend"), A.s("small"), A.a("uCell"), A.n("uCell"), A.a("vCell"), A.n("vCell"), A.a("wCell"),
            A.n("wCell"), A.a("uBoundary"), A.n("uBoundary"), A.a("vBoundary"), A.n("vBoundary"), A.a("wBoundary"),
            A.n("wBoundary"), A.a("visEffCell"), A.n("visEffCell"), A.a("visEffBoundary"), A.n("visEffBoundary"),
            A.a("massFlux"), A.n("massFlux"), du, A.nv("dudx"), dv, A.nv("dvdx"), dw, A.nv("dwdx"), A.a("rFace"),
            A.n("rFace"), A.a("su"), A.n("su"), A.a("sv"), A.n("sv"), A.a("sw"), A.n("sw"), A.a("au"), A.n("au"),
            A.a("av"), A.n("av"), A.a("aw"), A.n("aw"));
    }
    if (k == "uvw_region1") {
        FluxUVWDolfynRegionLoop1(m);
        return cupcfd::error::E_SUCCESS;
    }
    if (k == "uvw_bnds1") {
        FluxUVWDolfynBndsLoop1(m);
        return cupcfd::error::E_SUCCESS;
    }
    if (k == "visc_cell1")
        return calculateViscosityDolfynCellLoop1(m, A.s("small"), A.s("cmu"), A.s("visURF"), A.s("visLam"), A.a("TE"),
                                                 A.n("TE"), A.a("ED"), A.n("ED"), A.a("den"), A.n("den"),
                                                 A.a("visEff"), A.n("visEff"));
    if (k == "visc_region") {
        calculateViscosityDolfynRegionLoop(A.s("kappa"), m);
        return cupcfd::error::E_SUCCESS;
    }
    if (k == "visc_boundary")
        return calculateViscosityDolfynBoundaryLoop(m, A.s("tmcmu"), A.s("large"), A.s("small"), A.s("kappa"),
                                                    A.s("visLam"), A.a("TE"), A.n("TE"), A.a("den"), A.n("den"),
                                                    A.a("visEffCell"), A.n("visEffCell"), A.a("visEffBoundary"),
                                                    A.n("visEffBoundary"));
    if (k == "visc_cell2")
        return calculateViscosityDolfynCellLoop2(m, A.s("visLam"), A.a("visEffCell"), A.n("visEffCell"),
                                                 A.a("visEffBoundary"), A.n("visEffBoundary"));
    throw std::runtime_error("unknown kernel " + k);
}

template <class T>
static void kernel_mode(Records& r, Records& out, cupcfd::comm::Communicator& comm) {
    meshgeo::CupCfdSoAMesh<int, T, int>* m = build_mesh<T>(r, comm);
    if (comm.rank != 0) {
        delete m;
        return;
    }
    set_state<T>(r, *m);
    Args<T> A(r);
    int rc = run_kernel<T>(text(r, "kernel"), *m, A);
    A.writeback();
    put(out, "rc", &rc, 1);
    for (auto& kv : r)
        if (kv.first.compare(0, 2, "a.") == 0) out[kv.first] = kv.second;
    get_state<T>(out, *m);
    delete m;
}

// ---------------------------------------------------------------------------
// The reference's own StructGen mesh
// ---------------------------------------------------------------------------
template <class T>
static void structgen_mode(Records& r, Records& out, cupcfd::comm::Communicator& comm) {
    cupcfd::partitioner::PartitionerNaiveConfig<int, int> part;
    meshgeo::MeshSourceStructGenConfig<int, T> src(ival(r, "s.nx"), ival(r, "s.ny"), ival(r, "s.nz"),
                                                   T(sval(r, "s.xmin")), T(sval(r, "s.xmax")), T(sval(r, "s.ymin")),
                                                   T(sval(r, "s.ymax")), T(sval(r, "s.zmin")), T(sval(r, "s.zmax")));
    meshgeo::MeshConfig<int, T, int> cfg(part, src);
    meshgeo::CupCfdSoAMesh<int, T, int>* m = nullptr;
    int rc = cfg.buildUnstructuredMesh(&m, comm);
    put(out, "rc", &rc, 1);
    if (rc != cupcfd::error::E_SUCCESS) return;
    const int C = m->properties.lTCells, F = m->properties.lFaces, B = m->properties.lBoundaries,
              R = m->properties.lRegions;
    int counts[5] = {C, F, B, R, m->properties.lOCells};
    put(out, "counts", counts, 5);
    // build labels of the local ids
    std::vector<int> clab(C, -1), flab(F, -1), blab(B, -1);
    for (const auto& kv : m->cellBuildIDToLocalID) clab.at(kv.second) = kv.first;
    for (const auto& kv : m->faceBuildIDToLocalID) flab.at(kv.second) = kv.first;
    for (const auto& kv : m->boundaryBuildIDToLocalID) blab.at(kv.second) = kv.first;
    put(out, "label.cell", clab.data(), C);
    put(out, "label.face", flab.data(), F);
    put(out, "label.bnd", blab.data(), B);
    std::vector<int> c1(F), c2(F), fb(F), bf(B), br(B), rtype(R), xadj(C + 1), adj;
    std::vector<T> lam(F), nrm(3 * F), fc(3 * F), xp(3 * F), xn(3 * F), rl(F), ar(F), cc(3 * C), vol(C), bd(B);
    for (int f = 0; f < F; f++) {
        c1[f] = m->getFaceCell1ID(f);
        c2[f] = m->getFaceCell2ID(f);
        fb[f] = m->getFaceBoundaryID(f);
        lam[f] = m->getFaceLambda(f);
        rl[f] = m->getFaceRLencos(f);
        ar[f] = m->getFaceArea(f);
        euc::EuclideanVector<T, 3> s = m->getFaceNorm(f);
        euc::EuclideanPoint<T, 3> x = m->getFaceCenter(f), pa = m->getFaceXpac(f), na = m->getFaceXnac(f);
        for (int k = 0; k < 3; k++) {
            nrm[3 * f + k] = s.cmp[k];
            fc[3 * f + k] = x.cmp[k];
            xp[3 * f + k] = pa.cmp[k];
            xn[3 * f + k] = na.cmp[k];
        }
    }
    for (int c = 0; c < C; c++) {
        euc::EuclideanPoint<T, 3> x = m->getCellCenter(c);
        for (int k = 0; k < 3; k++) cc[3 * c + k] = x.cmp[k];
        vol[c] = m->getCellVolume(c);
        xadj[c] = int(adj.size());
        const int nf = m->getCellStoredNFaces(c);
        for (int j = 0; j < nf; j++) adj.push_back(m->getCellFaceID(c, j));
    }
    xadj[C] = int(adj.size());
    for (int b = 0; b < B; b++) {
        bf[b] = m->getBoundaryFaceID(b);
        br[b] = m->getBoundaryRegionID(b);
        bd[b] = m->getBoundaryDistance(b);
    }
    for (int i = 0; i < R; i++) rtype[i] = int(m->getRegionType(i));
    put(out, "mesh.faceCell1", c1.data(), F);
    put(out, "mesh.faceCell2", c2.data(), F);
    put(out, "mesh.faceBnd", fb.data(), F);
    put(out, "mesh.faceLambda", lam.data(), F);
    put(out, "mesh.faceNorm", nrm.data(), 3 * F);
    put(out, "mesh.faceCenter", fc.data(), 3 * F);
    put(out, "mesh.faceXpac", xp.data(), 3 * F);
    put(out, "mesh.faceXnac", xn.data(), 3 * F);
    put(out, "mesh.faceRLencos", rl.data(), F);
    put(out, "mesh.faceArea", ar.data(), F);
    put(out, "mesh.cellCenter", cc.data(), 3 * C);
    put(out, "mesh.cellVolume", vol.data(), C);
    put(out, "mesh.cellFaceXAdj", xadj.data(), C + 1);
    put(out, "mesh.cellFaceAdj", adj.data(), adj.size());
    put(out, "mesh.bndFace", bf.data(), B);
    put(out, "mesh.bndRegion", br.data(), B);
    put(out, "mesh.bndDistance", bd.data(), B);
    put(out, "mesh.regType", rtype.data(), R);
    delete m;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: mpiexec -n 2 %s case.bin result.bin\n", argv[0]);
        return 2;
    }
    MPI_Init(&argc, &argv);
    int status = 0;
    try {
        cupcfd::comm::Communicator comm(MPI_COMM_WORLD);
        if (comm.size != 2) throw std::runtime_error("run under mpiexec -n 2");
        Records r = read_case(argv[1]);
        Records out;
        const std::string mode = text(r, "mode");
        const int prec = ival(r, "precision");
        std::string path = argv[2];
        if (mode == "kernel") {
            prec == 64 ? kernel_mode<double>(r, out, comm) : kernel_mode<float>(r, out, comm);
        } else if (mode == "structgen") {
            prec == 64 ? structgen_mode<double>(r, out, comm) : structgen_mode<float>(r, out, comm);
            path += "." + std::to_string(comm.rank);
        } else {
            throw std::runtime_error("unknown mode " + mode);
        }
        if (comm.rank == 0 || mode == "structgen") write_case(path.c_str(), out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_harness: %s\n", e.what());
        status = 1;
    }
    MPI_Finalize();
    return status;
}

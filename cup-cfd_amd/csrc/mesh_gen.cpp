// Host mesh sources: the StructGen-numbered hex grid, seeded relabelling and
// typed boundary regions. Connectivity and numbering follow
// src/geometry/mesh/implementation/source/MeshStructGenSource.cpp as
// assembled on one rank by UnstructuredMeshInterface::addData
// (include/geometry/mesh/interface/component/UnstructuredMeshInterface.ipp:1464-1862):
// local cell / face / boundary ids equal their StructGen labels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "host_mesh.hpp"
#include "sg_geom.hpp"

namespace cupcfd_hip {

namespace {

struct Grid {
    int nx, ny, nz;
    double xmin, ymin, zmin, dx, dy, dz;
    // MeshStructGenSource::calculateLabel (MeshStructGenSource.cpp:58-70)
    static int32_t label(int x, int y, int z, int xs, int ys) { return x + y * xs + z * xs * ys; }
    int32_t cell(int x, int y, int z) const { return label(x, y, z, nx, ny); }
    // low / high spatial bounds of a cell along one axis, as getCellCenter
    // computes them (MeshStructGenSource.cpp:330-340)
    double lo(double mn, double d, int k) const { return mn + (k * d); }
    double hi(double mn, double d, int k) const { return mn + (k + 1) * d; }
};

}  // namespace

void add_regions(cupcfd_host_mesh &m, bool typed) {
    auto push = [&](int type, bool std_, bool flux, bool adiab, double ylog, double elog, double density,
                    double ke, double diss, double splvl, double den, double r, double t, double u,
                    double v, double w) {
        m.regType.push_back(type);
        m.regStd.push_back(std_);
        m.regFlux.push_back(flux);
        m.regAdiab.push_back(adiab);
        m.regYLog.push_back(ylog);
        m.regELog.push_back(elog);
        m.regDensity.push_back(density);
        m.regTurbKE.push_back(ke);
        m.regTurbDiss.push_back(diss);
        m.regSplvl.push_back(splvl);
        m.regDen.push_back(den);
        m.regR.push_back(r);
        m.regT.push_back(t);
        m.regForceTangent.insert(m.regForceTangent.end(), {0.0, 0.0, 0.0});
        m.regUVW.insert(m.regUVW.end(), {u, v, w});
    };
    if (!typed) {
        // CupCfdSoAMesh::addRegion(label, name): RTYPE_DEFAULT, all-zero
        // properties (CupCfdSoAMesh.cpp:211-222)
        push(CUPCFD_RTYPE_DEFAULT, false, false, false, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0);
    } else {
        push(CUPCFD_RTYPE_INLET, false, false, false, 0.0, 0.0, 1.0, 1e-3, 1e-4, 0.0, 1.0, 0.0, 0.0, 1.0,
             0.0, 0.0);
        push(CUPCFD_RTYPE_OUTLET, false, false, false, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0,
             0.0, 0.0);
        push(CUPCFD_RTYPE_WALL, true, false, false, 11.0, 9.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0,
             0.0, 0.0);
        push(CUPCFD_RTYPE_WALL, false, true, false, 11.0, 9.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.01, 0.5, 0.5,
             0.0, 0.0);
        push(CUPCFD_RTYPE_SYMP, false, false, false, 0, 0, 1.0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0);
        push(CUPCFD_RTYPE_SYMP, false, false, false, 0, 0, 1.0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0);
    }
    m.nRegions = (int32_t)m.regType.size();
}

void resize_mesh(cupcfd_host_mesh &m, int C, int F, int B) {
    m.nCells = C;
    m.nFaces = F;
    m.nBoundaries = B;
    m.faceCell1.assign(F, -1);
    m.faceCell2.assign(F, -1);
    m.faceBnd.assign(F, -1);
    m.faceLambda.assign(F, 0.0);
    m.faceRLencos.assign(F, 0.0);
    m.faceArea.assign(F, 0.0);
    m.faceNorm.assign(3 * (size_t)F, 0.0);
    m.faceCenter.assign(3 * (size_t)F, 0.0);
    m.faceXpac.assign(3 * (size_t)F, 0.0);
    m.faceXnac.assign(3 * (size_t)F, 0.0);
    m.cellCenter.assign(3 * (size_t)C, 0.0);
    m.cellVolume.assign(C, 0.0);
    m.bndFace.assign(B, -1);
    m.bndRegion.assign(B, 0);
    m.bndDistance.assign(B, 0.0);
    m.bndYPlus.assign(B, 0.0);
    m.bndUPlus.assign(B, 0.0);
    m.bndQ.assign(B, 0.0);
    m.bndH.assign(B, 0.0);
    m.bndT.assign(B, 0.0);
    m.bndShear.assign(3 * (size_t)B, 0.0);
}

// Seeded non-zero xpac / xnac / rlencos so the reference's xpn, dpx..dpz and
// VisFace terms are exercised (the reference's own meshes leave them 0).
void fill_face_extras(cupcfd_host_mesh &m, uint64_t seed, double h) {
    SplitMix64 rng(seed ^ 0xFACEE47A5ull);
    for (int f = 0; f < m.nFaces; f++) {
        for (int k = 0; k < 3; k++) {
            m.faceXpac[3 * f + k] = m.faceCenter[3 * f + k] + 0.05 * h * (2.0 * rng.uniform() - 1.0);
            m.faceXnac[3 * f + k] = m.faceCenter[3 * f + k] + 0.05 * h * (2.0 * rng.uniform() - 1.0);
        }
        int ip = m.faceCell1[f], in = m.faceCell2[f];
        const double *a = &m.cellCenter[3 * (size_t)ip];
        const double *b = in >= 0 ? &m.cellCenter[3 * (size_t)in] : &m.faceCenter[3 * (size_t)f];
        double d2 = 0;
        for (int k = 0; k < 3; k++) d2 += (b[k] - a[k]) * (b[k] - a[k]);
        m.faceRLencos[f] = m.faceArea[f] / std::sqrt(d2);
    }
}

int build_structgen(const cupcfd_structgen_params &p, cupcfd_host_mesh &m) {
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    Grid g;
    g.nx = p.nx;
    g.ny = p.ny;
    g.nz = p.nz;
    g.xmin = p.xmin;
    g.ymin = p.ymin;
    g.zmin = p.zmin;
    // spatial division (MeshStructGenSource.cpp:44-47)
    g.dx = (p.xmax - p.xmin) / double(p.nx);
    g.dy = (p.ymax - p.ymin) / double(p.ny);
    g.dz = (p.zmax - p.zmin) / double(p.nz);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t C64 = (int64_t)nx * ny * nz;
    const int64_t nYZ = (int64_t)ny * nz * (nx + 1), nXZ = (int64_t)nx * nz * (ny + 1),
                  nXY = (int64_t)nx * ny * (nz + 1);
    const int64_t F64 = nYZ + nXZ + nXY;
    const int64_t nBYZ = (int64_t)ny * nz * 2, nBXZ = (int64_t)nx * nz * 2, nBXY = (int64_t)nx * ny * 2;
    const int64_t B64 = nBYZ + nBXZ + nBXY;
    if (F64 >= (int64_t)1 << 30) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const int C = (int)C64, F = (int)F64, B = (int)B64;
    resize_mesh(m, C, F, B);
    add_regions(m, p.regions != 0);

    // cells: centre as getCellCenter computes it, volume from the reference's
    // five-tetrahedron split of the hexahedron around that centre
    const double half[3] = {g.dx / 2, g.dy / 2, g.dz / 2};
    const sg::HexPlan plan = sg::hex_plan(half);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int c = g.cell(x, y, z);
                double xl = g.lo(g.xmin, g.dx, x), xh = g.hi(g.xmin, g.dx, x);
                double yl = g.lo(g.ymin, g.dy, y), yh = g.hi(g.ymin, g.dy, y);
                double zl = g.lo(g.zmin, g.dz, z), zh = g.hi(g.zmin, g.dz, z);
                sg::P3 ctr{{xl + ((xh - xl) / 2.0), yl + ((yh - yl) / 2.0), zl + ((zh - zl) / 2.0)}};
                for (int k = 0; k < 3; k++) m.cellCenter[3 * (size_t)c + k] = ctr.c[k];
                m.cellVolume[c] = sg::hex_volume(plan, ctr, half);
            }

    // faces, in label order: YZ (x-normal), XZ (y-normal), XY (z-normal)
    // (getFaceCell1Labels / getFaceCell2Labels / getFaceBoundaryLabels,
    // MeshStructGenSource.cpp:518-821)
    for (int64_t L = 0; L < F64; L++) {
        int axis, n, a, b, c;  // n: node coordinate along axis; (a,b,c) cell coords
        int64_t r;
        if (L < nYZ) {
            axis = 0;
            r = L;
            n = (int)(r % (nx + 1));
            b = (int)((r / (nx + 1)) % ny);
            c = (int)(r / ((int64_t)(nx + 1) * ny));
            a = n;
        } else if (L < nYZ + nXZ) {
            axis = 1;
            r = L - nYZ;
            a = (int)(r % nx);
            n = (int)((r / nx) % (ny + 1));
            c = (int)(r / ((int64_t)nx * (ny + 1)));
            b = n;
        } else {
            axis = 2;
            r = L - nYZ - nXZ;
            a = (int)(r % nx);
            b = (int)((r / nx) % ny);
            n = (int)(r / ((int64_t)nx * ny));
            c = n;
        }
        const int nmax = axis == 0 ? nx : (axis == 1 ? ny : nz);
        const bool bnd = (n == 0 || n == nmax);
        int lo[3] = {a, b, c};  // cell coords of cell1
        if (n > 0) lo[axis] = n - 1;
        const int f = (int)L;
        m.faceCell1[f] = g.cell(lo[0], lo[1], lo[2]);
        if (!bnd) {
            int hi3[3] = {lo[0], lo[1], lo[2]};
            hi3[axis] = n;
            m.faceCell2[f] = g.cell(hi3[0], hi3[1], hi3[2]);
            m.faceLambda[f] = 0.5;
        } else {
            m.faceCell2[f] = -1;
            m.faceLambda[f] = -1.0;
            int side = (n == 0) ? 0 : 1;
            int32_t bl;
            int reg;
            if (axis == 0) {
                bl = Grid::label(side, b, c, 2, ny);
                reg = side ? REG_OUTLET : REG_INLET;
            } else if (axis == 1) {
                bl = (int32_t)nBYZ + Grid::label(a, side, c, nx, 2);
                reg = side ? REG_WALL_HI : REG_WALL_LO;
            } else {
                bl = (int32_t)(nBYZ + nBXZ) + Grid::label(a, b, side, nx, ny);
                reg = side ? REG_SYMP_HI : REG_SYMP_LO;
            }
            m.faceBnd[f] = bl;
            m.bndFace[bl] = f;
            m.bndRegion[bl] = p.regions ? reg : 0;
            double d = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
            m.bndDistance[bl] = p.regions ? d / 2.0 : 0.0;
        }
        // geometry from the face's four vertices in getFaceVerticesLabelsCSR's
        // order (MeshStructGenSource.cpp:1010-1085): node (a, b, c) with the
        // face's own axis at its node coordinate n
        const double mn[3] = {g.xmin, g.ymin, g.zmin}, dd[3] = {g.dx, g.dy, g.dz};
        int node[3] = {a, b, c};
        node[axis] = n;
        const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;  // the two in-plane axes, low then high
        const int step[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        sg::P3 v[4];
        for (int i = 0; i < 4; i++) {
            int q[3] = {node[0], node[1], node[2]};
            q[u] += step[i][0];
            q[w] += step[i][1];
            for (int k = 0; k < 3; k++) v[i].c[k] = mn[k] + (dd[k] * q[k]);
        }
        double area;
        sg::P3 ctr;
        sg::quad_area_centroid(v, &area, &ctr);
        const sg::P3 c1c{{m.cellCenter[3 * (size_t)m.faceCell1[f]], m.cellCenter[3 * (size_t)m.faceCell1[f] + 1],
                          m.cellCenter[3 * (size_t)m.faceCell1[f] + 2]}};
        const sg::P3 S = sg::quad_normal(v, c1c);
        m.faceArea[f] = area;
        for (int k = 0; k < 3; k++) {
            m.faceCenter[3 * (size_t)f + k] = ctr.c[k];
            m.faceNorm[3 * (size_t)f + k] = S.c[k];
        }
    }
    if (p.face_extras) fill_face_extras(m, p.seed, std::min(g.dx, std::min(g.dy, g.dz)));
    return CUPCFD_E_SUCCESS;
}

void relabel(cupcfd_host_mesh &m, uint64_t seed, bool cells, bool faces) {
    SplitMix64 rng(seed);
    auto fy = [&](int n) {
        std::vector<int32_t> perm(n);
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int)rng.below((uint64_t)i + 1)]);
        return perm;  // new id of old id i
    };
    if (cells) {
        std::vector<int32_t> pc = fy(m.nCells);
        std::vector<double> cc(m.cellCenter.size()), cv(m.cellVolume.size());
        for (int c = 0; c < m.nCells; c++) {
            for (int k = 0; k < 3; k++) cc[3 * (size_t)pc[c] + k] = m.cellCenter[3 * (size_t)c + k];
            cv[pc[c]] = m.cellVolume[c];
        }
        m.cellCenter.swap(cc);
        m.cellVolume.swap(cv);
        for (int f = 0; f < m.nFaces; f++) {
            m.faceCell1[f] = pc[m.faceCell1[f]];
            if (m.faceCell2[f] >= 0) m.faceCell2[f] = pc[m.faceCell2[f]];
        }
    }
    if (faces) {
        std::vector<int32_t> pf = fy(m.nFaces);
        auto perm_i = [&](std::vector<int32_t> &v) {
            std::vector<int32_t> o(v.size());
            for (int f = 0; f < m.nFaces; f++) o[pf[f]] = v[f];
            v.swap(o);
        };
        auto perm_d = [&](std::vector<double> &v, int k) {
            std::vector<double> o(v.size());
            for (int f = 0; f < m.nFaces; f++)
                for (int j = 0; j < k; j++) o[(size_t)k * pf[f] + j] = v[(size_t)k * f + j];
            v.swap(o);
        };
        perm_i(m.faceCell1);
        perm_i(m.faceCell2);
        perm_i(m.faceBnd);
        perm_d(m.faceLambda, 1);
        perm_d(m.faceRLencos, 1);
        perm_d(m.faceArea, 1);
        perm_d(m.faceNorm, 3);
        perm_d(m.faceCenter, 3);
        perm_d(m.faceXpac, 3);
        perm_d(m.faceXnac, 3);
        for (int b = 0; b < m.nBoundaries; b++) m.bndFace[b] = pf[m.bndFace[b]];
    }
}

}  // namespace cupcfd_hip

using namespace cupcfd_hip;

extern "C" int cupcfd_structgen_create(const cupcfd_structgen_params *params, cupcfd_host_mesh **out) {
    if (!params || !out) return CUPCFD_E_NULL_PTR;
    *out = nullptr;
    cupcfd_host_mesh *m = new (std::nothrow) cupcfd_host_mesh();
    if (!m) return CUPCFD_E_ERROR;
    int rc;
    try {
        rc = params->tets ? build_tetgen(*params, *m) : build_structgen(*params, *m);
        if (rc == CUPCFD_E_SUCCESS && (params->permute_cells || params->permute_faces))
            relabel(*m, params->seed, params->permute_cells != 0, params->permute_faces != 0);
    } catch (const std::bad_alloc &) {
        rc = CUPCFD_E_ERROR;
    }
    if (rc != CUPCFD_E_SUCCESS) {
        delete m;
        return rc;
    }
    *out = m;
    return CUPCFD_E_SUCCESS;
}

extern "C" void cupcfd_host_mesh_destroy(cupcfd_host_mesh *mesh) { delete mesh; }

namespace {
template <class T, class V3, class View, class Vec>
void fill_view(cupcfd_host_mesh *m, View *v, Vec &faceLambda, Vec &faceRLencos, Vec &faceArea, Vec &faceNorm,
               Vec &faceCenter, Vec &faceXpac, Vec &faceXnac, Vec &cellCenter, Vec &cellVolume,
               Vec &bndDistance, Vec &bndYPlus, Vec &bndUPlus, Vec &bndQ, Vec &bndH, Vec &bndT, Vec &bndShear,
               Vec &regYLog, Vec &regELog, Vec &regDensity, Vec &regTurbKE, Vec &regTurbDiss, Vec &regSplvl,
               Vec &regDen, Vec &regR, Vec &regT, Vec &regForceTangent, Vec &regUVW) {
    std::memset(v, 0, sizeof(*v));
    v->nCells = m->nCells;
    v->nFaces = m->nFaces;
    v->nBoundaries = m->nBoundaries;
    v->nRegions = m->nRegions;
    v->faceCell1ID = m->faceCell1.data();
    v->faceCell2ID = m->faceCell2.data();
    v->faceBoundaryID = m->faceBnd.data();
    v->faceLambda = faceLambda.data();
    v->faceRLencos = faceRLencos.data();
    v->faceArea = faceArea.data();
    v->faceNorm = V3{faceNorm.data(), 0};
    v->faceCenter = V3{faceCenter.data(), 0};
    v->faceXpac = V3{faceXpac.data(), 0};
    v->faceXnac = V3{faceXnac.data(), 0};
    v->cellCenter = V3{cellCenter.data(), 0};
    v->cellVolume = cellVolume.data();
    v->boundaryFaceID = m->bndFace.data();
    v->boundaryRegionID = m->bndRegion.data();
    v->boundaryDistance = bndDistance.data();
    v->boundaryYPlus = bndYPlus.data();
    v->boundaryUPlus = bndUPlus.data();
    v->boundaryQ = bndQ.data();
    v->boundaryH = bndH.data();
    v->boundaryT = bndT.data();
    v->boundaryShear = V3{bndShear.data(), 0};
    v->regionType = m->regType.data();
    v->regionStd = m->regStd.data();
    v->regionFlux = m->regFlux.data();
    v->regionAdiab = m->regAdiab.data();
    v->regionYLog = regYLog.data();
    v->regionELog = regELog.data();
    v->regionDensity = regDensity.data();
    v->regionTurbKE = regTurbKE.data();
    v->regionTurbDiss = regTurbDiss.data();
    v->regionSplvl = regSplvl.data();
    v->regionDen = regDen.data();
    v->regionR = regR.data();
    v->regionT = regT.data();
    v->regionForceTangent = V3{regForceTangent.data(), 0};
    v->regionUVW = V3{regUVW.data(), 0};
}

void to_f32(const std::vector<double> &s, std::vector<float> &d) {
    d.resize(s.size());
    for (size_t i = 0; i < s.size(); i++) d[i] = (float)s[i];
}
}  // namespace

extern "C" int cupcfd_host_mesh_view_f64(cupcfd_host_mesh *m, cupcfd_mesh_view_f64 *v) {
    if (!m || !v) return CUPCFD_E_NULL_PTR;
    fill_view<double, cupcfd_v3_f64>(m, v, m->faceLambda, m->faceRLencos, m->faceArea, m->faceNorm, m->faceCenter,
                                     m->faceXpac, m->faceXnac, m->cellCenter, m->cellVolume, m->bndDistance,
                                     m->bndYPlus, m->bndUPlus, m->bndQ, m->bndH, m->bndT, m->bndShear, m->regYLog,
                                     m->regELog, m->regDensity, m->regTurbKE, m->regTurbDiss, m->regSplvl,
                                     m->regDen, m->regR, m->regT, m->regForceTangent, m->regUVW);
    return CUPCFD_E_SUCCESS;
}

extern "C" int cupcfd_host_mesh_view_f32(cupcfd_host_mesh *m, cupcfd_mesh_view_f32 *v) {
    if (!m || !v) return CUPCFD_E_NULL_PTR;
    auto &s = m->f32;
    to_f32(m->faceLambda, s.faceLambda);
    to_f32(m->faceRLencos, s.faceRLencos);
    to_f32(m->faceArea, s.faceArea);
    to_f32(m->faceNorm, s.faceNorm);
    to_f32(m->faceCenter, s.faceCenter);
    to_f32(m->faceXpac, s.faceXpac);
    to_f32(m->faceXnac, s.faceXnac);
    to_f32(m->cellCenter, s.cellCenter);
    to_f32(m->cellVolume, s.cellVolume);
    to_f32(m->bndDistance, s.bndDistance);
    to_f32(m->bndYPlus, s.bndYPlus);
    to_f32(m->bndUPlus, s.bndUPlus);
    to_f32(m->bndQ, s.bndQ);
    to_f32(m->bndH, s.bndH);
    to_f32(m->bndT, s.bndT);
    to_f32(m->bndShear, s.bndShear);
    to_f32(m->regYLog, s.regYLog);
    to_f32(m->regELog, s.regELog);
    to_f32(m->regDensity, s.regDensity);
    to_f32(m->regTurbKE, s.regTurbKE);
    to_f32(m->regTurbDiss, s.regTurbDiss);
    to_f32(m->regSplvl, s.regSplvl);
    to_f32(m->regDen, s.regDen);
    to_f32(m->regR, s.regR);
    to_f32(m->regT, s.regT);
    to_f32(m->regForceTangent, s.regForceTangent);
    to_f32(m->regUVW, s.regUVW);
    fill_view<float, cupcfd_v3_f32>(m, v, s.faceLambda, s.faceRLencos, s.faceArea, s.faceNorm, s.faceCenter,
                                    s.faceXpac, s.faceXnac, s.cellCenter, s.cellVolume, s.bndDistance,
                                    s.bndYPlus, s.bndUPlus, s.bndQ, s.bndH, s.bndT, s.bndShear, s.regYLog,
                                    s.regELog, s.regDensity, s.regTurbKE, s.regTurbDiss, s.regSplvl, s.regDen,
                                    s.regR, s.regT, s.regForceTangent, s.regUVW);
    return CUPCFD_E_SUCCESS;
}

// Tetrahedral stress mesh: every hex of an nx x ny x nz grid over the box is
// split into six tetrahedra around its main diagonal (the Kuhn split: tet t
// walks from corner (0,0,0) to (1,1,1) along the axes in permutation t, so
// neighbouring hexes agree on every shared triangle), and interior vertices
// are jittered by U(-jitter, jitter) grid spacings per component. Vertices on
// the box faces keep the coordinate that puts them on their plane.
//
// The reference has no tetrahedral generator; the geometry is this build's:
//   cell centre   = mean of the 4 vertices,  volume = |det(v1-v0, v2-v0, v3-v0)| / 6
//   face normal   = (b - a) x (c - a) / 2, oriented outward from cell1
//   face area     = Heron's formula on the edge lengths
//   face centre   = mean of the 3 vertices
//   lambda        = |Xf - Cc1| / (|Xf - Cc1| + |Xf - Cc2|) (the weight of cell2), -1 on boundaries
//   boundary distance (typed regions) = |(Xf - Cc1) . n|
// Numbering: tet 6h + t for hex h = x + nx*(y + ny*z); a triangle's id is
// given when the lowest (tet, local face) that holds it is reached, and its
// cell1 is that tet; boundary ids follow face ids.
#include <algorithm>
#include <cmath>

#include "host_mesh.hpp"

namespace cupcfd_hip {

namespace {

struct FaceKey {
    uint64_t ab;     // (lowest vertex << 32) | middle vertex
    uint32_t c;      // highest vertex
    uint32_t owner;  // 4 * tet + local face
    bool operator<(const FaceKey &o) const {
        if (ab != o.ab) return ab < o.ab;
        if (c != o.c) return c < o.c;
        return owner < o.owner;
    }
    bool same(const FaceKey &o) const { return ab == o.ab && c == o.c; }
};

inline void sub(double *o, const double *a, const double *b) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
inline double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross(double *o, const double *a, const double *b) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dist(const double *a, const double *b) {
    double d[3];
    sub(d, a, b);
    return std::sqrt(dot(d, d));
}

}  // namespace

int build_tetgen(const cupcfd_structgen_params &p, cupcfd_host_mesh &m) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    if (nx < 1 || ny < 1 || nz < 1) return CUPCFD_E_CONFIG_INVALID_VALUE;
    if (p.jitter < 0.0 || p.jitter >= 0.5) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const int64_t H = (int64_t)nx * ny * nz;
    const int64_t C64 = 6 * H;
    const int64_t Vn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    // 4 C face slots and the face count must fit 32-bit ids
    if (4 * C64 >= ((int64_t)1 << 31) || Vn >= ((int64_t)1 << 31)) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const double dx = (p.xmax - p.xmin) / nx, dy = (p.ymax - p.ymin) / ny, dz = (p.zmax - p.zmin) / nz;

    // vertices, jittered in vertex-id order
    std::vector<double> X(3 * (size_t)Vn);
    SplitMix64 rng(p.seed ^ 0x7E7A7E7A5EEDull);
    const int n3[3] = {nx, ny, nz};
    const double mn[3] = {p.xmin, p.ymin, p.zmin}, dd[3] = {dx, dy, dz};
    for (int64_t v = 0; v < Vn; v++) {
        const int ijk[3] = {(int)(v % (nx + 1)), (int)((v / (nx + 1)) % (ny + 1)),
                            (int)(v / ((int64_t)(nx + 1) * (ny + 1)))};
        for (int k = 0; k < 3; k++) {
            double x = mn[k] + ijk[k] * dd[k];
            const double u = 2.0 * rng.uniform() - 1.0;
            if (ijk[k] > 0 && ijk[k] < n3[k]) x += p.jitter * dd[k] * u;
            X[3 * (size_t)v + k] = x;
        }
    }
    auto vid = [&](int i, int j, int k) -> uint32_t {
        return (uint32_t)(i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k));
    };
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

    const int C = (int)C64;
    std::vector<uint32_t> tv(4 * (size_t)C);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int64_t h = x + (int64_t)nx * (y + (int64_t)ny * z);
                for (int t = 0; t < 6; t++) {
                    int q[3] = {x, y, z};
                    uint32_t *w = &tv[4 * (size_t)(6 * h + t)];
                    w[0] = vid(q[0], q[1], q[2]);
                    q[perm[t][0]]++;
                    w[1] = vid(q[0], q[1], q[2]);
                    q[perm[t][1]]++;
                    w[2] = vid(q[0], q[1], q[2]);
                    w[3] = vid(x + 1, y + 1, z + 1);
                }
            }

    // match triangles: local face l of a tet is the one opposite vertex l
    std::vector<FaceKey> keys(4 * (size_t)C);
    for (int c = 0; c < C; c++)
        for (int l = 0; l < 4; l++) {
            uint32_t f3[3], n = 0;
            for (int q = 0; q < 4; q++)
                if (q != l) f3[n++] = tv[4 * (size_t)c + q];
            std::sort(f3, f3 + 3);
            keys[4 * (size_t)c + l] = FaceKey{((uint64_t)f3[0] << 32) | f3[1], f3[2], (uint32_t)(4 * c + l)};
        }
    std::sort(keys.begin(), keys.end());
    // per slot (4 tet + l): the other slot holding the same triangle, or -1
    std::vector<int32_t> twin(4 * (size_t)C, -1);
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].same(keys[i])) j++;
        if (j - i > 2) return CUPCFD_E_ERROR;  // a triangle in three tets: not a valid mesh
        if (j - i == 2) {
            twin[keys[i].owner] = (int32_t)keys[i + 1].owner;
            twin[keys[i + 1].owner] = (int32_t)keys[i].owner;
        }
        i = j;
    }
    keys.clear();
    keys.shrink_to_fit();
    int64_t F64 = 0, B64 = 0;
    for (size_t s = 0; s < twin.size(); s++) {
        if (twin[s] < 0) {
            F64++;
            B64++;
        } else if (twin[s] > (int32_t)s) {
            F64++;
        }
    }
    if (F64 >= ((int64_t)1 << 30)) return CUPCFD_E_CONFIG_INVALID_VALUE;
    const int F = (int)F64, B = (int)B64;
    resize_mesh(m, C, F, B);
    add_regions(m, p.regions != 0);

    for (int c = 0; c < C; c++) {
        const double *v[4];
        for (int q = 0; q < 4; q++) v[q] = &X[3 * (size_t)tv[4 * (size_t)c + q]];
        for (int k = 0; k < 3; k++) m.cellCenter[3 * (size_t)c + k] = (v[0][k] + v[1][k] + v[2][k] + v[3][k]) / 4.0;
        double e1[3], e2[3], e3[3], cr[3];
        sub(e1, v[1], v[0]);
        sub(e2, v[2], v[0]);
        sub(e3, v[3], v[0]);
        cross(cr, e2, e3);
        m.cellVolume[c] = std::fabs(dot(e1, cr)) / 6.0;
    }

    int f = 0, b = 0;
    for (int s = 0; s < 4 * C; s++) {
        const int t2 = twin[s];
        if (t2 >= 0 && t2 < s) continue;  // numbered at its lower slot
        const int c1 = s / 4, l = s % 4;
        const double *v3[3];
        uint32_t vv[3];
        int n = 0;
        for (int q = 0; q < 4; q++)
            if (q != l) {
                vv[n] = tv[4 * (size_t)c1 + q];
                v3[n] = &X[3 * (size_t)vv[n]];
                n++;
            }
        double xf[3], e1[3], e2[3], S[3], d1[3];
        for (int k = 0; k < 3; k++) xf[k] = (v3[0][k] + v3[1][k] + v3[2][k]) / 3.0;
        sub(e1, v3[1], v3[0]);
        sub(e2, v3[2], v3[0]);
        cross(S, e1, e2);
        for (int k = 0; k < 3; k++) S[k] *= 0.5;
        const double *cc1 = &m.cellCenter[3 * (size_t)c1];
        sub(d1, xf, cc1);
        if (dot(S, d1) < 0)
            for (int k = 0; k < 3; k++) S[k] = -S[k];
        const double la = dist(v3[0], v3[1]), lb = dist(v3[1], v3[2]), lc = dist(v3[2], v3[0]);
        const double hs = (la + lb + lc) / 2.0;
        const double area = std::sqrt(std::max(0.0, hs * (hs - la) * (hs - lb) * (hs - lc)));
        m.faceCell1[f] = c1;
        m.faceArea[f] = area;
        for (int k = 0; k < 3; k++) {
            m.faceNorm[3 * (size_t)f + k] = S[k];
            m.faceCenter[3 * (size_t)f + k] = xf[k];
        }
        if (t2 >= 0) {
            const int c2 = t2 / 4;
            m.faceCell2[f] = c2;
            const double dP = dist(xf, cc1), dN = dist(xf, &m.cellCenter[3 * (size_t)c2]);
            m.faceLambda[f] = dP / (dP + dN);
        } else {
            m.faceCell2[f] = -1;
            m.faceLambda[f] = -1.0;
            m.faceBnd[f] = b;
            m.bndFace[b] = f;
            // the box side the triangle lies on: the axis whose index is shared by its 3 vertices
            int reg = 0;
            for (int k = 0; k < 3; k++) {
                int idx[3];
                for (int q = 0; q < 3; q++) {
                    const uint32_t w = vv[q];
                    idx[q] = k == 0 ? (int)(w % (nx + 1))
                                    : k == 1 ? (int)((w / (nx + 1)) % (ny + 1))
                                             : (int)(w / ((uint32_t)(nx + 1) * (ny + 1)));
                }
                if (idx[0] == idx[1] && idx[1] == idx[2] && (idx[0] == 0 || idx[0] == n3[k])) {
                    const bool hi = idx[0] == n3[k];
                    reg = k == 0 ? (hi ? REG_OUTLET : REG_INLET)
                                 : k == 1 ? (hi ? REG_WALL_HI : REG_WALL_LO) : (hi ? REG_SYMP_HI : REG_SYMP_LO);
                }
            }
            m.bndRegion[b] = p.regions ? reg : 0;
            const double an = std::sqrt(dot(S, S));
            m.bndDistance[b] = p.regions ? std::fabs(dot(d1, S)) / an : 0.0;
            b++;
        }
        f++;
    }
    if (p.face_extras) fill_face_extras(m, p.seed, std::min(dx, std::min(dy, dz)));
    return CUPCFD_E_SUCCESS;
}

}  // namespace cupcfd_hip

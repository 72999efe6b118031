// StructGen geometry as the reference's MeshStructGenSource computes it, in
// the same operation order, so the generated face areas, face centres, face
// normals and cell volumes equal the reference mesh's bit for bit. Shared by
// the host generator (mesh_gen.cpp) and the device generator (devgen.hip).
//
//   vertex      sMin + (dS * node)                    MeshStructGenSource.cpp:1100-1107
//   face quad   vertex order of getFaceVerticesLabelsCSR  MeshStructGenSource.cpp:1010-1085
//   face area   Polygon3D::computeArea: Heron triangles (v0,v1,v2), (v0,v2,v3)
//               (src/geometry/shapes/interface/component/Polygon3D.cpp:31-40,
//                include/geometry/shapes/implementation/component/Triangle.ipp:54-99)
//   face centre Polygon3D::computeCentroid: area-weighted triangle centroids
//               (Polygon3D.cpp:42-55)
//   face normal cross product oriented by isVertexOrderClockwise from cell1's centre
//               (MeshStructGenSource.cpp:887-944, Polygon3D.cpp:94-119,
//                src/geometry/euclidean/implementation/component/EuclideanPlane3D.cpp:203-210)
//   cell volume Hexahedron of six Quadrilateral3D faces around the centre,
//               split into five tetrahedra (MeshStructGenSource.cpp:278-320,
//               include/geometry/shapes/implementation/component/Hexahedron.ipp:30-68,
//               src/geometry/shapes/implementation/component/Hexahedron.cpp:26-165),
//               each tetrahedron's volume by Pyramid::computeVolume
//               (include/geometry/shapes/implementation/component/Pyramid.ipp:30-49, 204-219)
//
// Vector helpers keep EuclideanVector's order: dot = ((0 + a0 b0) + a1 b1) + a2 b2
// (EuclideanVector.cpp:83-90), length = sqrt(dot), normalise = multiply by
// 1 / length (EuclideanVector.cpp:93-110). Nothing here may be contracted into
// fused multiply-adds (the library is built with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace cupcfd_hip {
namespace sg {

struct P3 {
    double c[3];
};

__host__ __device__ inline P3 sub(const P3 &a, const P3 &b) { return P3{{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}}; }
__host__ __device__ inline P3 add(const P3 &a, const P3 &b) { return P3{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
__host__ __device__ inline P3 scale(const P3 &a, double s) { return P3{{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
__host__ __device__ inline P3 divide(const P3 &a, double s) { return P3{{a.c[0] / s, a.c[1] / s, a.c[2] / s}}; }
__host__ __device__ inline double dot(const P3 &a, const P3 &b) {
    double d = 0.0;
    d += a.c[0] * b.c[0];
    d += a.c[1] * b.c[1];
    d += a.c[2] * b.c[2];
    return d;
}
__host__ __device__ inline double length(const P3 &a) { return sqrt(dot(a, a)); }
__host__ __device__ inline P3 cross(const P3 &u, const P3 &v) {
    return P3{{(u.c[1] * v.c[2]) - (u.c[2] * v.c[1]), (u.c[2] * v.c[0]) - (u.c[0] * v.c[2]),
               (u.c[0] * v.c[1]) - (u.c[1] * v.c[0])}};
}
__host__ __device__ inline P3 normalised(P3 v) {
    const double len = length(v);
    if (len != 0.0) {
        const double r = 1.0 / len;
        for (int k = 0; k < 3; k++) v.c[k] *= r;
    }
    return v;
}
// calculateNormal(p1, p2, p3) = (p2 - p1) x (p3 - p1)
__host__ __device__ inline P3 plane_normal(const P3 &p1, const P3 &p2, const P3 &p3) {
    return cross(sub(p2, p1), sub(p3, p1));
}

// Heron's formula from the three edge lengths |b-a|, |c-a|, |c-b|
__host__ __device__ inline double tri_area(const P3 &a, const P3 &b, const P3 &c) {
    const double ab = length(sub(b, a)), ac = length(sub(c, a)), bc = length(sub(c, b));
    const double s = ((ab + ac) + bc) / 2.0;
    return sqrt(((s * (s - ab)) * (s - bc)) * (s - ac));
}
__host__ __device__ inline P3 tri_centroid(const P3 &a, const P3 &b, const P3 &c) {
    return divide(add(add(a, b), c), 3.0);
}

// quad area and area-weighted centroid over the fan (v0,v1,v2), (v0,v2,v3)
__host__ __device__ inline void quad_area_centroid(const P3 v[4], double *area, P3 *ctr) {
    const double a1 = tri_area(v[0], v[1], v[2]);
    const double a2 = tri_area(v[0], v[2], v[3]);
    double a = 0.0;
    a += a1;
    a += a2;
    P3 c{{0.0, 0.0, 0.0}};
    c = add(c, scale(tri_centroid(v[0], v[1], v[2]), a1));
    c = add(c, scale(tri_centroid(v[0], v[2], v[3]), a2));
    *area = a;
    *ctr = divide(c, a);
}

// outward normal of a face seen from cell1's centre
__host__ __device__ inline P3 quad_normal(const P3 v[4], const P3 &cell1_centre) {
    const P3 obs = sub(cell1_centre, v[0]);
    const bool clockwise = dot(plane_normal(v[0], v[1], v[2]), obs) < 0.0;
    return clockwise ? plane_normal(v[0], v[1], v[2]) : plane_normal(v[2], v[1], v[0]);
}

// Corners of a cell around its centre, indexed (x > 0) | (y > 0) << 1 | (z > 0) << 2,
// as centre + (+-dX, +-dY, +-dZ) with dX = dSx / 2.
__host__ __device__ inline P3 corner(const P3 &centre, const double half[3], int k) {
    P3 o{{(k & 1) ? half[0] : -half[0], (k & 2) ? half[1] : -half[1], (k & 4) ? half[2] : -half[2]}};
    return add(centre, o);
}

// The five tetrahedra of the reference's hexahedron split, as corner indices
// (apex, base0, base1, base2) with the base already in the order the Pyramid
// constructor leaves it. The same for every cell of one grid; built once on
// the host by hex_plan().
struct HexPlan {
    int8_t tet[5][4];
};

// Volume of one tetrahedron: (1/3) * Heron(base) * |(apex - b0) . unit normal(base)|
__host__ __device__ inline double tet_volume(const P3 &apex, const P3 &b0, const P3 &b1, const P3 &b2) {
    const double area = tri_area(b0, b1, b2);
    const P3 n = normalised(plane_normal(b0, b1, b2));
    const double height = fabs(dot(sub(apex, b0), n));
    return ((1.0 / 3.0) * area) * height;
}

__host__ __device__ inline double hex_volume(const HexPlan &plan, const P3 &centre, const double half[3]) {
    P3 p[8];
    for (int k = 0; k < 8; k++) p[k] = corner(centre, half, k);
    double v = 0.0;
    for (int t = 0; t < 5; t++) {
        const int8_t *q = plan.tet[t];
        const double tv = tet_volume(p[q[0]], p[q[1]], p[q[2]], p[q[3]]);
        v = (t == 0) ? tv : v + tv;
    }
    return v;
}

// The reference's hexahedron construction for one cell shape, on corner
// indices: faces top, bottom, front, back, right, left (MeshStructGenSource.cpp:299-313);
// orientFaceNormals reverses faces whose normal points at the centroid of the
// face centroids; detectOpposingFaces keeps face 0 as top, finds the first face
// sharing two vertices including top vertex 0 (adjacent) and the first face
// sharing none with top and two with adjacent (bottom), then rotates the bottom
// face so that its vertex next to top vertex 0 comes first; the tetrahedra are
// (h1; h5 h6 h8), (h1; h3 h8 h4), (h3; h6 h8 h7), (h1; h2 h3 h6), (h1; h3 h8 h6)
// with top = (h6 h7 h8 h5) and bottom = (h2 h1 h4 h3); each Pyramid reverses
// its base when the base normal points towards the apex. The orientation tests
// depend only on the signs of the spacings, so they are evaluated here once on
// a cell centred at the origin.
inline HexPlan hex_plan(const double half[3]) {
    const P3 zero{{0.0, 0.0, 0.0}};
    P3 p[8];
    for (int k = 0; k < 8; k++) p[k] = corner(zero, half, k);
    // tlf trf tlb trb blf brf blb brb
    const int tlf = 6, trf = 7, tlb = 4, trb = 5, blf = 2, brf = 3, blb = 0, brb = 1;
    int faces[6][4] = {{tlf, trf, trb, tlb}, {blf, brf, brb, blb}, {tlf, trf, brf, blf},
                       {tlb, trb, brb, blb}, {trf, trb, brb, brf}, {tlf, tlb, blb, blf}};
    auto qv = [&](int f, P3 v[4]) {
        for (int i = 0; i < 4; i++) v[i] = p[faces[f][i]];
    };
    // orientFaceNormals
    P3 cen[6];
    for (int f = 0; f < 6; f++) {
        P3 v[4];
        qv(f, v);
        double a;
        quad_area_centroid(v, &a, &cen[f]);
    }
    P3 oc = cen[0];
    for (int f = 1; f < 6; f++) oc = add(oc, cen[f]);
    oc = divide(oc, 6.0);
    for (int f = 0; f < 6; f++) {
        P3 v[4];
        qv(f, v);
        if (dot(plane_normal(v[0], v[1], v[2]), sub(cen[f], oc)) < 0.0) {
            int r[4] = {faces[f][3], faces[f][2], faces[f][1], faces[f][0]};
            for (int i = 0; i < 4; i++) faces[f][i] = r[i];
        }
    }
    // detectOpposingFaces
    const int top = 0;
    int adj = -1;
    for (int i = 1; i < 6 && adj < 0; i++) {
        int shared = 0;
        bool v0 = false;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                if (faces[top][a] == faces[i][b]) {
                    shared++;
                    if (a == 0) v0 = true;
                    break;
                }
        if (shared == 2 && v0) adj = i;
    }
    int bot = -1;
    for (int i = 1; i < 6 && bot < 0 && adj >= 0; i++) {
        int st = 0, sa = 0;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                if (faces[top][a] == faces[i][b]) st++;
                if (faces[adj][a] == faces[i][b]) sa++;
            }
        if (st == 0 && sa == 2) bot = i;
    }
    HexPlan plan{};
    if (adj < 0 || bot < 0) return plan;  // not reached for a box
    for (int s = 0; s < 4; s++) {
        if (faces[adj][s] != faces[top][0]) continue;
        const int va = faces[adj][(s + 1) % 4], vb = faces[adj][(s + 3) % 4];
        for (int b = 0; b < 4; b++) {
            if (faces[bot][b] == va || faces[bot][b] == vb) {
                int rot[4];
                for (int i = 0; i < 4; i++) rot[(i + 4 - b) % 4] = faces[bot][i];
                for (int i = 0; i < 4; i++) faces[bot][i] = rot[i];
                break;
            }
        }
        break;
    }
    const int h6 = faces[top][0], h7 = faces[top][1], h8 = faces[top][2], h5 = faces[top][3];
    const int h2 = faces[bot][0], h1 = faces[bot][1], h4 = faces[bot][2], h3 = faces[bot][3];
    const int tets[5][4] = {{h1, h5, h6, h8}, {h1, h3, h8, h4}, {h3, h6, h8, h7}, {h1, h2, h3, h6}, {h1, h3, h8, h6}};
    for (int t = 0; t < 5; t++) {
        int q[4] = {tets[t][0], tets[t][1], tets[t][2], tets[t][3]};
        // Pyramid constructor: reverse the base if its normal points at the apex
        const P3 n = plane_normal(p[q[1]], p[q[2]], p[q[3]]);
        if (dot(n, sub(tri_centroid(p[q[1]], p[q[2]], p[q[3]]), p[q[0]])) < 0.0) {
            const int x = q[1];
            q[1] = q[3];
            q[3] = x;
        }
        for (int k = 0; k < 4; k++) plan.tet[t][k] = (int8_t)q[k];
    }
    return plan;
}

}  // namespace sg
}  // namespace cupcfd_hip

// Host-side MiniSoA mesh container: one std::vector per CupCfdSoAMesh member
// the fvm kernels touch (include/geometry/mesh/implementation/component/CupCfdSoAMesh.h:71-154),
// 3-vectors packed xyz. Geometry is held in double; float copies are made on
// request for the single-precision path.
#pragma once

#include <cstdint>
#include <vector>

#include "cupcfd_hip.h"

struct cupcfd_host_mesh {
    int32_t nCells = 0, nFaces = 0, nBoundaries = 0, nRegions = 0;
    std::vector<int32_t> faceCell1, faceCell2, faceBnd;
    std::vector<double> faceLambda, faceRLencos, faceArea;
    std::vector<double> faceNorm, faceCenter, faceXpac, faceXnac;  // 3F
    std::vector<double> cellCenter;                                // 3C
    std::vector<double> cellVolume;
    std::vector<int32_t> bndFace, bndRegion;
    std::vector<double> bndDistance, bndYPlus, bndUPlus, bndQ, bndH, bndT;
    std::vector<double> bndShear;  // 3B
    std::vector<int32_t> regType;
    std::vector<uint8_t> regStd, regFlux, regAdiab;
    std::vector<double> regYLog, regELog, regDensity, regTurbKE, regTurbDiss, regSplvl, regDen, regR,
        regT;
    std::vector<double> regForceTangent, regUVW;  // 3R

    // float mirror, filled by cupcfd_host_mesh_view_f32
    struct F32 {
        std::vector<float> faceLambda, faceRLencos, faceArea, faceNorm, faceCenter, faceXpac, faceXnac,
            cellCenter, cellVolume, bndDistance, bndYPlus, bndUPlus, bndQ, bndH, bndT, bndShear, regYLog,
            regELog, regDensity, regTurbKE, regTurbDiss, regSplvl, regDen, regR, regT, regForceTangent,
            regUVW;
    } f32;
};

namespace cupcfd_hip {
// splitmix64: the seeded stream behind every permutation and jitter here
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0,1): top 53 bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) { return next() % n; }
};

// typed regions, one per side of the box: x-low INLET, x-high OUTLET, y-low
// standard WALL, y-high moving non-standard WALL, z-low / z-high SYMP
enum { REG_INLET = 0, REG_OUTLET, REG_WALL_LO, REG_WALL_HI, REG_SYMP_LO, REG_SYMP_HI };
// one DEFAULT region, or the six typed regions
void add_regions(cupcfd_host_mesh &m, bool typed);
// size every per-cell, per-face and per-boundary array (zero / -1 filled)
void resize_mesh(cupcfd_host_mesh &m, int C, int F, int B);
// seeded non-zero xpac / xnac and rlencos = area / |Cc2 - Cc1|
void fill_face_extras(cupcfd_host_mesh &m, uint64_t seed, double h);

int build_structgen(const cupcfd_structgen_params &p, cupcfd_host_mesh &m);
int build_tetgen(const cupcfd_structgen_params &p, cupcfd_host_mesh &m);
void relabel(cupcfd_host_mesh &m, uint64_t seed, bool cells, bool faces);
}  // namespace cupcfd_hip

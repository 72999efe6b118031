// Tetrahedral stress mesh: every hex of a StructGen grid split into six
// tetrahedra around its main diagonal, vertices jittered.
#include "host_mesh.hpp"

namespace cupcfd_hip {
int build_tetgen(const cupcfd_structgen_params &p, cupcfd_host_mesh &m) {
    (void)p;
    (void)m;
    return CUPCFD_E_NOT_IMPLEMENTED;
}
}  // namespace cupcfd_hip

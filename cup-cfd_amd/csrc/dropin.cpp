// Drop-in C ABI (cupcfd_fvm_*): the reference's calling convention — host
// arrays in the mesh's numbering with (pointer, length) pairs, synchronous,
// mesh state mutated in place — built on the resident kernels. Each call
// stages its arrays into the device mesh's internal order, runs the resident
// kernel, copies results back and mirrors the mesh state the kernel set.
#include <cstring>
#include <vector>

#include "devmesh.hpp"

namespace cupcfd_hip {
namespace {

template <class T> struct V3Of;
template <> struct V3Of<double> { using type = cupcfd_v3_f64; };
template <> struct V3Of<float> { using type = cupcfd_v3_f32; };

template <class T> Mesh<T> *mesh_of(cupcfd_hip_mesh *h);
template <> Mesh<double> *mesh_of<double>(cupcfd_hip_mesh *h) { return h && h->precision == 64 ? h->m64 : nullptr; }
template <> Mesh<float> *mesh_of<float>(cupcfd_hip_mesh *h) { return h && h->precision == 32 ? h->m32 : nullptr; }

// Staging of one drop-in call. Device buffers come from the mesh's workspace
// pool (slots 0..31 for the kernels' arrays, 64..127 for the raw copies of
// the caller's arrays). The caller's arrays are copied as they are, in the
// mesh's numbering (staged_h2d / staged_d2h); the permutation to and from the
// internal numbering runs on the device (stage_gather / stage_scatter).
template <class T> struct Stage {
    using V3 = typename V3Of<T>::type;
    Mesh<T> &M;
    int slot = 0, raw_slot = 0;
    int rc = CUPCFD_E_SUCCESS;
    std::vector<T> h;
    explicit Stage(Mesh<T> &m) : M(m) {}

    T *dev(size_t n) {
        if (rc != CUPCFD_E_SUCCESS) return nullptr;
        if (slot > 31) {
            rc = CUPCFD_E_ERROR;
            return nullptr;
        }
        T *p = (T *)M.ws_get(slot++, (n ? n : 1) * sizeof(T));
        if (!p) rc = CUPCFD_E_ERROR;
        return p;
    }
    bool check(const void *p, int64_t have, int64_t need) {
        if (rc != CUPCFD_E_SUCCESS) return false;
        if (need > 0 && !p) {
            rc = CUPCFD_E_NULL_PTR;
            return false;
        }
        if (have < need) {
            rc = CUPCFD_E_INVALID_INDEX;
            return false;
        }
        return true;
    }
    T *raw(size_t n) {
        if (rc != CUPCFD_E_SUCCESS) return nullptr;
        if (raw_slot > 63) {
            rc = CUPCFD_E_ERROR;
            return nullptr;
        }
        T *p = (T *)M.ws_get(64 + raw_slot++, (n ? n : 1) * sizeof(T));
        if (!p) rc = CUPCFD_E_ERROR;
        return p;
    }
    const int *order(bool faces) {
        if (rc != CUPCFD_E_SUCCESS) return nullptr;
        const int *o = device_order(M, faces);
        if (!o) rc = CUPCFD_E_ERROR;
        return o;
    }
    void copy_in(void *d, const void *src, size_t bytes) {
        if (rc == CUPCFD_E_SUCCESS) ok(staged_h2d(d, src, bytes));
    }
    void copy_out(void *dst, const void *d, size_t bytes) {
        if (rc == CUPCFD_E_SUCCESS) ok(staged_d2h(dst, d, bytes));
    }
    void ok(int r) {
        if (rc == CUPCFD_E_SUCCESS && r != CUPCFD_E_SUCCESS) rc = r;
    }
    // n items of ncomp values, mesh order (item i at src + i * stride values) -> internal order on the device
    T *permuted(const T *src, int n, int ncomp, int64_t stride, bool faces) {
        const int *o = order(faces);
        const size_t span = n ? (size_t)(n - 1) * stride + ncomp : 0;  // values from src that are read
        T *r = raw(span), *d = dev((size_t)n * ncomp);
        copy_in(r, src, span * sizeof(T));
        if (rc == CUPCFD_E_SUCCESS) ok(stage_gather<T>(d, r, o, n, ncomp, stride));
        return d;
    }
    // n items of ncomp values, internal order on the device -> packed mesh order at dst
    void unpermuted(const T *d, T *dst, int n, int ncomp, bool faces) {
        const int *o = order(faces);
        T *r = raw((size_t)n * ncomp);
        if (rc == CUPCFD_E_SUCCESS) ok(stage_scatter<T>(r, d, o, n, ncomp));
        copy_out(dst, r, (size_t)n * ncomp * sizeof(T));
    }
    T *upload(size_t n) {
        T *d = dev(n);
        if (d) copy_in(d, h.data(), n * sizeof(T));
        return d;
    }
    void download(const T *d, size_t n) {
        h.resize(n);
        copy_out(h.data(), d, n * sizeof(T));
    }
    static T *at3(const V3 &v, int64_t i, int k) {
        int64_t stride = v.stride ? v.stride : (int64_t)(3 * sizeof(T));
        return (T *)((char *)v.base + i * stride) + k;
    }
    // --- inputs -------------------------------------------------------------
    T *cells(const T *src, int n) {  // ncomp 1, mesh order -> internal
        const int C = M.d.C;
        if (!check(src, n, C)) return nullptr;
        return permuted(src, C, 1, 1, false);
    }
    T *cells3(const V3 &v, int n) {
        const int C = M.d.C;
        if (!check(v.base, n, C)) return nullptr;
        const int64_t stride = v.stride ? v.stride : (int64_t)(3 * sizeof(T));
        if (stride % (int64_t)sizeof(T) == 0)
            return permuted((const T *)v.base, C, 3, stride / (int64_t)sizeof(T), false);
        h.resize(3 * (size_t)C);  // a stride that is not a whole number of values: packed on the host
        for (int k = 0; k < C; k++)
            for (int j = 0; j < 3; j++) h[3 * (size_t)k + j] = *at3(v, M.cell_order[k], j);
        return upload(3 * (size_t)C);
    }
    T *faces(const T *src, int n, int ncomp) {
        const int F = M.d.F;
        if (!check(src, n, (int64_t)F * ncomp)) return nullptr;
        return permuted(src, F, ncomp, ncomp, true);
    }
    T *plain(const T *src, int n, int need) {  // boundary / region arrays
        if (!check(src, n, need)) return nullptr;
        h.assign(src, src + need);
        return upload(need);
    }
    T *scalar(T v) {
        h.assign(1, v);
        return upload(1);
    }
    // --- outputs ------------------------------------------------------------
    void cells_out(const T *d, T *dst) { unpermuted(d, dst, M.d.C, 1, false); }
    void cells3_out(const T *d, const V3 &v) {
        const int C = M.d.C;
        const int64_t stride = v.stride ? v.stride : (int64_t)(3 * sizeof(T));
        if (stride == (int64_t)(3 * sizeof(T))) {  // packed: straight into the caller's array
            unpermuted(d, (T *)v.base, C, 3, false);
            return;
        }
        // strided (e.g. vector objects with a vptr): packed on the host, then
        // only the three components of each element are written
        h.resize(3 * (size_t)C);
        unpermuted(d, h.data(), C, 3, false);
        if (rc == CUPCFD_E_SUCCESS)
            for (int k = 0; k < C; k++)
                for (int j = 0; j < 3; j++) *at3(v, k, j) = h[3 * (size_t)k + j];
    }
    void faces_out(const T *d, T *dst, int ncomp) { unpermuted(d, dst, M.d.F, ncomp, true); }
    void plain_out(const T *d, T *dst, int n) {
        download(d, n);
        if (rc == CUPCFD_E_SUCCESS && n) std::memcpy(dst, h.data(), n * sizeof(T));
    }
    int finish() {
        if (rc == CUPCFD_E_SUCCESS && hipDeviceSynchronize() != hipSuccess) rc = CUPCFD_E_ERROR;
        return rc;
    }
};

inline int sync_or(int rc) {
    if (rc != CUPCFD_E_SUCCESS) return rc;
    return hipDeviceSynchronize() == hipSuccess ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

}  // namespace
}  // namespace cupcfd_hip

using namespace cupcfd_hip;

// The resident entry points, declared in cupcfd_hip.h, do the device work.
#define DROPIN_API(T, V3, S)                                                                                       \
    int cupcfd_fvm_gradient_phi_gauss_dolfyn_##S(cupcfd_hip_mesh *mesh, int32_t nGradient, const T *phiCell,       \
                                                 int32_t nPhiCell, const T *phiBoundary, int32_t nPhiBoundary,     \
                                                 V3 dPhidxCell, int32_t nDPhidxCell, V3 dPhidxoCell,                \
                                                 int32_t nDPhidxoCell) {                                           \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        Stage<T> st(*M);                                                                                           \
        const int C = M->d.C;                                                                                      \
        T *phi = nGradient > 0 ? st.cells(phiCell, nPhiCell) : nullptr;                                            \
        T *phiB = nGradient > 0 ? st.plain(phiBoundary, nPhiBoundary, M->d.B) : nullptr;                           \
        st.check(dPhidxCell.base, nDPhidxCell, nGradient > 0 ? C : 0);                                             \
        st.check(dPhidxoCell.base, nDPhidxoCell, C);                                                               \
        T *g = st.dev(3 * (size_t)C), *go = st.dev(3 * (size_t)C);                                                 \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = cupcfd_dev_gradient_phi_gauss_dolfyn_##S(mesh, nGradient, phi, phiB, g, go, nullptr);             \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        if (nGradient > 0) st.cells3_out(g, dPhidxCell);                                                           \
        st.cells3_out(go, dPhidxoCell);                                                                            \
        /* entries past the mesh's cells: zeroed by the reference's reset loops */                                 \
        for (int i = C; i < nDPhidxoCell; i++)                                                                     \
            for (int j = 0; j < 3; j++) *Stage<T>::at3(dPhidxoCell, i, j) = (T)0;                                  \
        if (nGradient > 0)                                                                                         \
            for (int i = C; i < nDPhidxCell; i++)                                                                  \
                for (int j = 0; j < 3; j++) *Stage<T>::at3(dPhidxCell, i, j) = (T)0;                               \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_face_loop_##S(                                                                        \
        cupcfd_hip_mesh *mesh, V3 dudx, int32_t nDudx, V3 dvdx, int32_t nDvdx, V3 dwdx, int32_t nDwdx, V3 dpdx,    \
        int32_t nDpdx, const T *denCell, int32_t nDenCell, T *denBoundary, int32_t nDenBoundary, const T *uCell,   \
        int32_t nUCell, const T *vCell, int32_t nVCell, const T *wCell, int32_t nWCell, T *massFlux,               \
        int32_t nMassFlux, const T *p, int32_t nP, const T *ar, int32_t nAr, T *su, int32_t nSu, T *rface,         \
        int32_t nRFace, T small, int32_t *icinl, int32_t *icout, int32_t *icsym, int32_t *icwal,                   \
        int32_t solveTurbEnergy, int32_t solveTurbDiss, int32_t solveVisc, int32_t solveEnthalpy,                  \
        const T *teCell, int32_t nTeCell, T *teBoundary, int32_t nTeBoundary, const T *edCell, int32_t nEdCell,    \
        T *edBoundary, int32_t nEdBoundary, const T *viseffCell, int32_t nViseffCell, T *viseffBoundary,           \
        int32_t nViseffBoundary, const T *tCell, int32_t nTCell, T *tBoundary, int32_t nTBoundary) {               \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!icinl || !icout || !icsym || !icwal) return CUPCFD_E_NULL_PTR;                                        \
        Stage<T> st(*M);                                                                                           \
        const int B = M->d.B;                                                                                      \
        T *d_dudx = st.cells3(dudx, nDudx), *d_dvdx = st.cells3(dvdx, nDvdx), *d_dwdx = st.cells3(dwdx, nDwdx);   \
        T *d_dpdx = st.cells3(dpdx, nDpdx);                                                                        \
        T *d_den = st.cells(denCell, nDenCell), *d_u = st.cells(uCell, nUCell), *d_v = st.cells(vCell, nVCell);   \
        T *d_w = st.cells(wCell, nWCell), *d_p = st.cells(p, nP), *d_ar = st.cells(ar, nAr);                       \
        T *d_su = st.cells(su, nSu);                                                                               \
        T *d_mf = st.faces(massFlux, nMassFlux, 1), *d_rf = st.faces(rface, nRFace, 2);                            \
        T *d_denB = st.plain(denBoundary, nDenBoundary, B);                                                        \
        /* per-type boundary copies: only arrays the OUTLET branch can touch */                                    \
        const bool out = M->ntype[CUPCFD_RTYPE_OUTLET] > 0;                                                        \
        T *d_teC = (out && solveTurbEnergy) ? st.cells(teCell, nTeCell) : nullptr;                                 \
        T *d_teB = (out && solveTurbEnergy) ? st.plain(teBoundary, nTeBoundary, B) : nullptr;                      \
        T *d_edC = (out && solveTurbDiss) ? st.cells(edCell, nEdCell) : nullptr;                                   \
        T *d_edB = (out && solveTurbDiss) ? st.plain(edBoundary, nEdBoundary, B) : nullptr;                        \
        T *d_vC = (out && solveVisc) ? st.cells(viseffCell, nViseffCell) : nullptr;                                \
        T *d_vB = (out && solveVisc) ? st.plain(viseffBoundary, nViseffBoundary, B) : nullptr;                     \
        T *d_tC = (out && solveEnthalpy) ? st.cells(tCell, nTCell) : nullptr;                                      \
        T *d_tB = (out && solveEnthalpy) ? st.plain(tBoundary, nTBoundary, B) : nullptr;                           \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = cupcfd_dev_flux_mass_face_loop_##S(                                                               \
            mesh, d_dudx, d_dvdx, d_dwdx, d_dpdx, d_den, d_denB, d_u, d_v, d_w, d_mf, d_p, d_ar, d_su, d_rf,       \
            small, icinl, icout, icsym, icwal, solveTurbEnergy, solveTurbDiss, solveVisc, solveEnthalpy, d_teC,    \
            d_teB, d_edC, d_edB, d_vC, d_vB, d_tC, d_tB, nullptr);                                                 \
        rc = sync_or(rc);                                                                                          \
        if (rc == CUPCFD_E_SUCCESS) {                                                                              \
            st.cells_out(d_su, su);                                                                                \
            st.faces_out(d_mf, massFlux, 1);                                                                       \
            st.faces_out(d_rf, rface, 2);                                                                          \
            st.plain_out(d_denB, denBoundary, B);                                                                  \
            if (d_teB) st.plain_out(d_teB, teBoundary, B);                                                         \
            if (d_edB) st.plain_out(d_edB, edBoundary, B);                                                         \
            if (d_vB) st.plain_out(d_vB, viseffBoundary, B);                                                       \
            if (d_tB) st.plain_out(d_tB, tBoundary, B);                                                            \
            rc = st.finish();                                                                                      \
        }                                                                                                          \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_boundary_loop1_##S(cupcfd_hip_mesh *mesh, const T *massFlux, int32_t nMassFlux,       \
                                                T *flowin) {                                                       \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!flowin) return CUPCFD_E_NULL_PTR;                                                                     \
        Stage<T> st(*M);                                                                                           \
        T *d_mf = st.faces(massFlux, nMassFlux, 1), *d_out = st.scalar((T)0);                                      \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_mass_boundary_loop1_##S(mesh, d_mf, d_out, nullptr));                     \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.plain_out(d_out, flowin, 1);                                                                            \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_boundary_loop2_##S(cupcfd_hip_mesh *mesh, const T *massFlux, int32_t nMassFlux,       \
                                                T *flowRegion, int32_t nFlowRegion, T *flowout) {                  \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!flowout) return CUPCFD_E_NULL_PTR;                                                                    \
        Stage<T> st(*M);                                                                                           \
        T *d_mf = st.faces(massFlux, nMassFlux, 1);                                                                \
        T *d_fr = st.plain(flowRegion, nFlowRegion, M->d.R), *d_out = st.scalar((T)0);                             \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_mass_boundary_loop2_##S(mesh, d_mf, d_fr, d_out, nullptr));               \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.plain_out(d_out, flowout, 1);                                                                           \
        st.plain_out(d_fr, flowRegion, M->d.R);                                                                    \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_boundary_loop3_##S(cupcfd_hip_mesh *mesh, T flowin, T *ratearea) {                    \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!ratearea) return CUPCFD_E_NULL_PTR;                                                                   \
        Stage<T> st(*M);                                                                                           \
        T *d_out = st.scalar((T)0);                                                                                \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_mass_boundary_loop3_##S(mesh, flowin, d_out, nullptr));                   \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.plain_out(d_out, ratearea, 1);                                                                          \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_boundary_loop4_##S(cupcfd_hip_mesh *mesh, T *massFlux, int32_t nMassFlux,             \
                                                T *uBoundary, int32_t nUBoundary, T *vBoundary,                    \
                                                int32_t nVBoundary, T *wBoundary, int32_t nWBoundary,              \
                                                const T *denBoundary, int32_t nDenBoundary, T ratearea,            \
                                                T *flowout) {                                                      \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!flowout) return CUPCFD_E_NULL_PTR;                                                                    \
        Stage<T> st(*M);                                                                                           \
        const int B = M->d.B;                                                                                      \
        T *d_mf = st.faces(massFlux, nMassFlux, 1);                                                                \
        T *d_u = st.plain(uBoundary, nUBoundary, B), *d_v = st.plain(vBoundary, nVBoundary, B);                    \
        T *d_w = st.plain(wBoundary, nWBoundary, B), *d_d = st.plain(denBoundary, nDenBoundary, B);                \
        T *d_out = st.scalar((T)0);                                                                                \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(                                                                                          \
            cupcfd_dev_flux_mass_boundary_loop4_##S(mesh, d_mf, d_u, d_v, d_w, d_d, ratearea, d_out, nullptr));    \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.faces_out(d_mf, massFlux, 1);                                                                           \
        st.plain_out(d_u, uBoundary, B);                                                                           \
        st.plain_out(d_v, vBoundary, B);                                                                           \
        st.plain_out(d_w, wBoundary, B);                                                                           \
        st.plain_out(d_out, flowout, 1);                                                                           \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_boundary_loop5_##S(cupcfd_hip_mesh *mesh, T *massFlux, int32_t nMassFlux, T *su,      \
                                                int32_t nSu, T *uBoundary, int32_t nUBoundary, T *vBoundary,       \
                                                int32_t nVBoundary, T *wBoundary, int32_t nWBoundary, T fact,      \
                                                int32_t solveU, int32_t solveV, int32_t solveW,                    \
                                                const T *flowFact, int32_t nFlowFact, T *flowout2) {               \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        if (!flowout2) return CUPCFD_E_NULL_PTR;                                                                   \
        Stage<T> st(*M);                                                                                           \
        const int B = M->d.B;                                                                                      \
        T *d_mf = st.faces(massFlux, nMassFlux, 1), *d_su = st.cells(su, nSu);                                     \
        T *d_u = st.plain(uBoundary, nUBoundary, B), *d_v = st.plain(vBoundary, nVBoundary, B);                    \
        T *d_w = st.plain(wBoundary, nWBoundary, B), *d_ff = st.plain(flowFact, nFlowFact, M->d.R);                \
        T *d_out = st.scalar(*flowout2);                                                                           \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_mass_boundary_loop5_##S(mesh, d_mf, d_su, d_u, d_v, d_w, fact, solveU,    \
                                                                 solveV, solveW, d_ff, d_out, nullptr));           \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.faces_out(d_mf, massFlux, 1);                                                                           \
        st.cells_out(d_su, su);                                                                                    \
        st.plain_out(d_u, uBoundary, B);                                                                           \
        st.plain_out(d_v, vBoundary, B);                                                                           \
        st.plain_out(d_w, wBoundary, B);                                                                           \
        st.plain_out(d_out, flowout2, 1);                                                                          \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_mass_region_loop_##S(cupcfd_hip_mesh *mesh, T *flowFact, int32_t nFlowFact,                \
                                             const T *flowRegion, int32_t nFlowRegion, T flowIn) {                 \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        Stage<T> st(*M);                                                                                           \
        const int R = M->d.R;                                                                                      \
        T *d_ff = st.plain(flowFact, nFlowFact, R), *d_fr = st.plain(flowRegion, nFlowRegion, R);                  \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_mass_region_loop_##S(mesh, d_ff, d_fr, flowIn, nullptr));                 \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.plain_out(d_ff, flowFact, R);                                                                           \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_flux_scalar_face_loop_##S(                                                                      \
        cupcfd_hip_mesh *mesh, const T *PhiCell, int32_t nPhiCell, T *PhiBoundary, int32_t nPhiBoundary,           \
        const T *VisEff, int32_t nVisEff, T *Au, int32_t nAu, T *Su, int32_t nSu, const T *Den, int32_t nDen,      \
        const T *MassFlux, int32_t nMassFlux, const T *TE, int32_t nTE, const T *CpBoundary, int32_t nCpBoundary,  \
        const T *visEffBoundary, int32_t nVisEffBoundary, T *RFace, int32_t nRFace, V3 dPhidx, int32_t ndPhidx,    \
        int32_t SolveTurb, int32_t SolveEnthalpy, T vislam, int32_t ivar, int32_t VarT, T Sigma_T, T Prandtl,      \
        int32_t VarTE, T Sigma_k, int32_t VarED, T Sigma_e, T Sigma_s, T Schmidt, T GammaBlend, T Small, T Large,  \
        T TMCmu) {                                                                                                 \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        const int B = M->d.B;                                                                                      \
        const bool wall = SolveEnthalpy && ivar == VarT && M->ntype[CUPCFD_RTYPE_WALL] > 0;                        \
        if (wall && cupcfd_hip_mesh_push_state(mesh) != CUPCFD_E_SUCCESS) return CUPCFD_E_ERROR;                   \
        Stage<T> st(*M);                                                                                           \
        T *d_phi = st.cells(PhiCell, nPhiCell), *d_phiB = st.plain(PhiBoundary, nPhiBoundary, B);                  \
        T *d_vis = st.cells(VisEff, nVisEff), *d_au = st.cells(Au, nAu), *d_su = st.cells(Su, nSu);               \
        T *d_den = st.cells(Den, nDen), *d_mf = st.faces(MassFlux, nMassFlux, 1);                                  \
        T *d_visB = st.plain(visEffBoundary, nVisEffBoundary, B), *d_rf = st.faces(RFace, nRFace, 2);              \
        T *d_g = st.cells3(dPhidx, ndPhidx);                                                                       \
        T *d_te = wall ? st.cells(TE, nTE) : nullptr;                                                              \
        T *d_cp = wall ? st.plain(CpBoundary, nCpBoundary, B) : nullptr;                                           \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_scalar_face_loop_##S(                                                     \
            mesh, d_phi, d_phiB, d_vis, d_au, d_su, d_den, d_mf, d_te, d_cp, d_visB, d_rf, d_g, SolveTurb,         \
            SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE, Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, \
            GammaBlend, Small, Large, TMCmu, nullptr));                                                            \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.plain_out(d_phiB, PhiBoundary, B);                                                                      \
        st.cells_out(d_au, Au);                                                                                    \
        st.cells_out(d_su, Su);                                                                                    \
        st.faces_out(d_rf, RFace, 2);                                                                              \
        rc = st.finish();                                                                                          \
        if (rc == CUPCFD_E_SUCCESS && wall) rc = cupcfd_hip_mesh_pull_state(mesh);                                 \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_flux_uvw_face_loop1_##S(                                                                        \
        cupcfd_hip_mesh *mesh, T gammaBlend, T small, const T *uCell, int32_t nUCell, const T *vCell,              \
        int32_t nVCell, const T *wCell, int32_t nWCell, T *uBoundary, int32_t nUBoundary, T *vBoundary,            \
        int32_t nVBoundary, T *wBoundary, int32_t nWBoundary, const T *visEffCell, int32_t nVisEffCell,            \
        const T *visEffBoundary, int32_t nVisEffBoundary, T *massFlux, int32_t nMassFlux, V3 dudx, int32_t nDudx,  \
        V3 dvdx, int32_t nDvdx, V3 dwdx, int32_t nDwdx, T *rFace, int32_t nRFace, T *su, int32_t nSu, T *sv,       \
        int32_t nSv, T *sw, int32_t nSw, T *au, int32_t nAu, T *av, int32_t nAv, T *aw, int32_t nAw) {             \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        const int B = M->d.B;                                                                                      \
        const bool wall = M->ntype[CUPCFD_RTYPE_WALL] > 0;                                                         \
        if (wall && cupcfd_hip_mesh_push_state(mesh) != CUPCFD_E_SUCCESS) return CUPCFD_E_ERROR;                   \
        Stage<T> st(*M);                                                                                           \
        T *d_u = st.cells(uCell, nUCell), *d_v = st.cells(vCell, nVCell), *d_w = st.cells(wCell, nWCell);         \
        T *d_vis = st.cells(visEffCell, nVisEffCell), *d_mf = st.faces(massFlux, nMassFlux, 1);                    \
        T *d_gu = st.cells3(dudx, nDudx), *d_gv = st.cells3(dvdx, nDvdx), *d_gw = st.cells3(dwdx, nDwdx);         \
        T *d_rf = st.faces(rFace, nRFace, 2), *d_su = st.cells(su, nSu), *d_sv = st.cells(sv, nSv);                \
        T *d_sw = st.cells(sw, nSw), *d_au = st.cells(au, nAu), *d_av = st.cells(av, nAv);                         \
        T *d_aw = st.cells(aw, nAw);                                                                               \
        T *d_uB = st.plain(uBoundary, nUBoundary, B), *d_vB = st.plain(vBoundary, nVBoundary, B);                  \
        T *d_wB = st.plain(wBoundary, nWBoundary, B), *d_visB = st.plain(visEffBoundary, nVisEffBoundary, B);      \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_flux_uvw_face_loop1_##S(mesh, gammaBlend, small, d_u, d_v, d_w, d_uB, d_vB,    \
                                                            d_wB, d_vis, d_visB, d_mf, d_gu, d_gv, d_gw, d_rf,     \
                                                            d_su, d_sv, d_sw, d_au, d_av, d_aw, nullptr));         \
        if (rc == CUPCFD_E_SUCCESS) {                                                                              \
            st.faces_out(d_mf, massFlux, 1);                                                                       \
            st.faces_out(d_rf, rFace, 2);                                                                          \
            st.cells_out(d_su, su);                                                                                \
            st.cells_out(d_sv, sv);                                                                                \
            st.cells_out(d_sw, sw);                                                                                \
            st.cells_out(d_au, au);                                                                                \
            st.cells_out(d_av, av);                                                                                \
            st.cells_out(d_aw, aw);                                                                                \
            st.plain_out(d_uB, uBoundary, B);                                                                      \
            st.plain_out(d_vB, vBoundary, B);                                                                      \
            st.plain_out(d_wB, wBoundary, B);                                                                      \
            rc = st.finish();                                                                                      \
        }                                                                                                          \
        if (rc == CUPCFD_E_SUCCESS && wall) rc = cupcfd_hip_mesh_pull_state(mesh);                                 \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_flux_uvw_region_loop1_##S(cupcfd_hip_mesh *mesh) {                                              \
        if (!mesh_of<T>(mesh)) return CUPCFD_E_NULL_PTR;                                                           \
        int rc = cupcfd_hip_mesh_push_state(mesh);                                                                 \
        if (rc == CUPCFD_E_SUCCESS) rc = sync_or(cupcfd_dev_flux_uvw_region_loop1_##S(mesh, nullptr));             \
        if (rc == CUPCFD_E_SUCCESS) rc = cupcfd_hip_mesh_pull_state(mesh);                                         \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_flux_uvw_bnds_loop1_##S(cupcfd_hip_mesh *mesh) {                                                \
        if (!mesh_of<T>(mesh)) return CUPCFD_E_NULL_PTR;                                                           \
        int rc = cupcfd_hip_mesh_push_state(mesh);                                                                 \
        if (rc == CUPCFD_E_SUCCESS) rc = sync_or(cupcfd_dev_flux_uvw_bnds_loop1_##S(mesh, nullptr));               \
        if (rc == CUPCFD_E_SUCCESS) rc = cupcfd_hip_mesh_pull_state(mesh);                                         \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_viscosity_cell_loop1_##S(cupcfd_hip_mesh *mesh, T small, T cmu, T visURF, T visLam,             \
                                            const T *TE, int32_t nTE, const T *ED, int32_t nED, const T *den,      \
                                            int32_t nDen, T *visEff, int32_t nVisEff) {                            \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        Stage<T> st(*M);                                                                                           \
        T *d_te = st.cells(TE, nTE), *d_ed = st.cells(ED, nED), *d_den = st.cells(den, nDen);                      \
        T *d_vis = st.cells(visEff, nVisEff);                                                                      \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_viscosity_cell_loop1_##S(mesh, small, cmu, visURF, visLam, d_te, d_ed, d_den,  \
                                                             d_vis, nullptr));                                     \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.cells_out(d_vis, visEff);                                                                               \
        return st.finish();                                                                                        \
    }                                                                                                              \
    int cupcfd_fvm_viscosity_region_loop_##S(T kappa, cupcfd_hip_mesh *mesh) {                                     \
        if (!mesh_of<T>(mesh)) return CUPCFD_E_NULL_PTR;                                                           \
        int rc = cupcfd_hip_mesh_push_state(mesh);                                                                 \
        if (rc == CUPCFD_E_SUCCESS) rc = sync_or(cupcfd_dev_viscosity_region_loop_##S(kappa, mesh, nullptr));      \
        if (rc == CUPCFD_E_SUCCESS) rc = cupcfd_hip_mesh_pull_state(mesh);                                         \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_viscosity_boundary_loop_##S(cupcfd_hip_mesh *mesh, T tmcmu, T large, T small, T kappa,          \
                                               T visLam, const T *TE, int32_t nTE, const T *den, int32_t nDen,     \
                                               const T *visEffCell, int32_t nVisEffCell, T *visEffBoundary,        \
                                               int32_t nVisEffBoundary) {                                          \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        int rc = cupcfd_hip_mesh_push_state(mesh);                                                                 \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        Stage<T> st(*M);                                                                                           \
        const bool wall = M->ntype[CUPCFD_RTYPE_WALL] > 0;                                                         \
        const bool cellv = M->ntype[CUPCFD_RTYPE_OUTLET] + M->ntype[CUPCFD_RTYPE_SYMP] > 0;                        \
        T *d_te = wall ? st.cells(TE, nTE) : nullptr, *d_den = wall ? st.cells(den, nDen) : nullptr;               \
        T *d_vc = cellv ? st.cells(visEffCell, nVisEffCell) : nullptr;                                             \
        T *d_vb = st.plain(visEffBoundary, nVisEffBoundary, M->d.B);                                               \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        rc = sync_or(cupcfd_dev_viscosity_boundary_loop_##S(mesh, tmcmu, large, small, kappa, visLam, d_te,        \
                                                            d_den, d_vc, d_vb, nullptr));                          \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.plain_out(d_vb, visEffBoundary, M->d.B);                                                                \
        rc = st.finish();                                                                                          \
        if (rc == CUPCFD_E_SUCCESS) rc = cupcfd_hip_mesh_pull_state(mesh);                                         \
        return rc;                                                                                                 \
    }                                                                                                              \
    int cupcfd_fvm_viscosity_cell_loop2_##S(cupcfd_hip_mesh *mesh, T visLam, T *visEffCell, int32_t nVisEffCell,   \
                                            T *visEffBoundary, int32_t nVisEffBoundary) {                          \
        Mesh<T> *M = mesh_of<T>(mesh);                                                                             \
        if (!M) return CUPCFD_E_NULL_PTR;                                                                          \
        Stage<T> st(*M);                                                                                           \
        T *d_vc = st.cells(visEffCell, nVisEffCell), *d_vb = st.plain(visEffBoundary, nVisEffBoundary, M->d.B);    \
        if (st.rc != CUPCFD_E_SUCCESS) return st.rc;                                                               \
        int rc = sync_or(cupcfd_dev_viscosity_cell_loop2_##S(mesh, visLam, d_vc, d_vb, nullptr));                  \
        if (rc != CUPCFD_E_SUCCESS) return rc;                                                                     \
        st.cells_out(d_vc, visEffCell);                                                                            \
        st.plain_out(d_vb, visEffBoundary, M->d.B);                                                                \
        return st.finish();                                                                                        \
    }

extern "C" {
DROPIN_API(double, cupcfd_v3_f64, f64)
DROPIN_API(float, cupcfd_v3_f32, f32)
}

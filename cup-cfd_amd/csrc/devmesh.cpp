// Device-mesh repack: validates a MiniSoA view and builds the HBM layout
// described in devmesh.hpp. Runs once per mesh, on the host, before any
// kernel (the reference builds its own cell→face CSR at the same point,
// CupCfdSoAMesh::updateCellFaceMap, src/geometry/mesh/implementation/component/CupCfdSoAMesh.cpp:549-668).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "devmesh.hpp"

namespace cupcfd_hip {

template <class T> Mesh<T>::~Mesh() {
    for (void *p : allocs) (void)hipFree(p);
    for (void *p : ws) (void)hipFree(p);
    for (void *p : ws_retired) (void)hipFree(p);
    for (int k = 0; k < 4; k++) {
        if (sc_src[k]) (void)hipFree(sc_src[k]);
        if (sc_rd_off[k]) (void)hipFree(sc_rd_off[k]);
        if (sc_rd[k]) (void)hipFree(sc_rd[k]);
    }
    if (rbnd) (void)hipFree(rbnd);
    if (rbnd_off) (void)hipFree(rbnd_off);
    if (cell_order_d) (void)hipFree(cell_order_d);
    if (face_order_d) (void)hipFree(face_order_d);
    if (uvw_col) (void)hipFree(uvw_col);
    if (uvw_eslot) (void)hipFree(uvw_eslot);
}

template <class T> void *Mesh<T>::ws_get(int slot, size_t bytes) {
    if ((int)ws.size() <= slot) {
        ws.resize(slot + 1, nullptr);
        ws_size.resize(slot + 1, 0);
    }
    if (ws_size[slot] < bytes) {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        if (ws[slot]) ws_retired.push_back(ws[slot]);
        ws[slot] = p;
        ws_size[slot] = bytes;
    }
    return ws[slot];
}

template struct Mesh<double>;
template struct Mesh<float>;

namespace {

template <class V3> const auto *v3ptr(const V3 &v, int64_t i, int k) {
    using S = std::remove_pointer_t<decltype(v.base)>;
    int64_t stride = v.stride ? v.stride : (int64_t)(3 * sizeof(S));
    return (const S *)((const char *)v.base + i * stride) + k;
}

struct Uploader {
    std::vector<void *> &allocs;
    bool ok = true;
    template <class X> X *put(const X *src, size_t n) {
        if (!ok) return nullptr;
        if (n == 0) n = 1;  // keep every pointer valid
        void *p = nullptr;
        if (hipMalloc(&p, n * sizeof(X)) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        allocs.push_back(p);
        if (src && hipMemcpy(p, src, n * sizeof(X), hipMemcpyHostToDevice) != hipSuccess) ok = false;
        return (X *)p;
    }
};

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

template <class T, class View> int build(Mesh<T> &M, const View &v, int flags) {
    const int C = v.nCells, F = v.nFaces, B = v.nBoundaries, R = v.nRegions;
    if (int rc = view_check(v); rc != CUPCFD_E_SUCCESS) return rc;

    // ---- validation -------------------------------------------------------
    int Fi = 0;
    for (int f = 0; f < F; f++) {
        int c1 = v.faceCell1ID[f], c2 = v.faceCell2ID[f], b = v.faceBoundaryID[f];
        if (c1 < 0 || c1 >= C) return CUPCFD_E_INVALID_INDEX;
        if (b > -1) {
            if (b >= B) return CUPCFD_E_INVALID_INDEX;
        } else {
            if (c2 < 0 || c2 >= C) return CUPCFD_E_INVALID_INDEX;
            Fi++;
        }
    }
    for (int b = 0; b < B; b++) {
        int f = v.boundaryFaceID[b], r = v.boundaryRegionID[b];
        if (f < 0 || f >= F || r < 0 || r >= R) return CUPCFD_E_INVALID_INDEX;
    }
    for (int r = 0; r < R; r++)
        if (v.regionType[r] < 0 || v.regionType[r] > 4) return CUPCFD_E_CONFIG_INVALID_VALUE;

    // ---- cell order ---------------------------------------------------------
    M.cell_order.resize(C);
    for (int c = 0; c < C; c++) M.cell_order[c] = c;
    if ((flags & CUPCFD_MESH_REORDER) && C > 1) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < C; c++)
            for (int k = 0; k < 3; k++) {
                double x = (double)*v3ptr(v.cellCenter, c, k);
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
        std::vector<uint64_t> key(C);
        for (int c = 0; c < C; c++) {
            uint64_t q[3];
            for (int k = 0; k < 3; k++) {
                double span = hi[k] - lo[k];
                double t = span > 0 ? ((double)*v3ptr(v.cellCenter, c, k) - lo[k]) / span : 0.0;
                q[k] = (uint64_t)std::min(2097151.0, std::max(0.0, std::floor(t * 2097151.0 + 0.5)));
            }
            key[c] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
        }
        std::stable_sort(M.cell_order.begin(), M.cell_order.end(),
                         [&](int a, int b) { return key[a] < key[b]; });
    }
    M.cell_rank.assign(C, 0);
    for (int k = 0; k < C; k++) M.cell_rank[M.cell_order[k]] = k;

    // ---- face order: (boundary?, owner) buckets, mesh order within, then ----
    // ---- interior faces grouped by neighbour tile within each owner tile ---
    std::vector<int> cnt(2 * (size_t)C + 1, 0);
    auto bucket = [&](int f) {
        int own = M.cell_rank[v.faceCell1ID[f]];
        return (v.faceBoundaryID[f] > -1 ? C : 0) + own;
    };
    for (int f = 0; f < F; f++) cnt[bucket(f) + 1]++;
    for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
    M.face_order.assign(F, 0);
    for (int f = 0; f < F; f++) M.face_order[cnt[bucket(f)]++] = f;
    // Interior faces, stably regrouped within each owner tile (kTileCells
    // consecutive internal cells): first the faces whose cell2 is in the same
    // tile, then those shared with each other tile, by that tile. A tile that
    // reads its halo faces from a neighbour then reads one contiguous run of
    // the neighbour's face records instead of records spread over all of its
    // range. Owner order (then mesh order) is kept within each group.
    {
        // the bucket order already runs tile by tile: sort each tile's run
        std::vector<int> grp(F, 0);
        for (int f = 0; f < F; f++)
            if (v.faceBoundaryID[f] <= -1) {
                const int t1 = M.cell_rank[v.faceCell1ID[f]] / kTileCells;
                const int t2 = M.cell_rank[v.faceCell2ID[f]] / kTileCells;
                grp[f] = t2 == t1 ? 0 : t2 + 1;
            }
        for (int k0 = 0; k0 < Fi;) {
            const int t = M.cell_rank[v.faceCell1ID[M.face_order[k0]]] / kTileCells;
            int k1 = k0 + 1;
            while (k1 < Fi && M.cell_rank[v.faceCell1ID[M.face_order[k1]]] / kTileCells == t) k1++;
            std::stable_sort(M.face_order.begin() + k0, M.face_order.begin() + k1,
                             [&](int a, int b) { return grp[a] < grp[b]; });
            k0 = k1;
        }
    }
    M.face_rank.assign(F, 0);
    for (int k = 0; k < F; k++) M.face_rank[M.face_order[k]] = k;

    // ---- face streams -------------------------------------------------------
    std::vector<int2> fconn(F);
    std::vector<V4<T>> fS(F), fX(F), fP(F);
    std::vector<V2<T>> fN(F);
    std::vector<T> frl(F);
    M.fconn_h.assign(2 * (size_t)F, 0);
    for (int k = 0; k < F; k++) {
        int f = M.face_order[k];
        int b = v.faceBoundaryID[f];
        int nb = b > -1 ? ~b : M.cell_rank[v.faceCell2ID[f]];
        fconn[k] = make_int2(M.cell_rank[v.faceCell1ID[f]], nb);
        M.fconn_h[2 * k] = fconn[k].x;
        M.fconn_h[2 * k + 1] = nb;
        fS[k] = {*v3ptr(v.faceNorm, f, 0), *v3ptr(v.faceNorm, f, 1), *v3ptr(v.faceNorm, f, 2), v.faceLambda[f]};
        fX[k] = {*v3ptr(v.faceCenter, f, 0), *v3ptr(v.faceCenter, f, 1), *v3ptr(v.faceCenter, f, 2), v.faceArea[f]};
        fP[k] = {*v3ptr(v.faceXpac, f, 0), *v3ptr(v.faceXpac, f, 1), *v3ptr(v.faceXpac, f, 2),
                 *v3ptr(v.faceXnac, f, 0)};
        fN[k] = {*v3ptr(v.faceXnac, f, 1), *v3ptr(v.faceXnac, f, 2)};
        frl[k] = v.faceRLencos[f];
    }
    std::vector<V4<T>> cg(C);
    std::vector<T> cv(C);
    for (int k = 0; k < C; k++) {
        int c = M.cell_order[k];
        cg[k] = {*v3ptr(v.cellCenter, c, 0), *v3ptr(v.cellCenter, c, 1), *v3ptr(v.cellCenter, c, 2),
                 v.cellVolume[c]};
        cv[k] = v.cellVolume[c];
    }

    // ---- cell→face CSR, rows in mesh face order ----------------------------
    std::vector<int> row(C + 1, 0);
    for (int f = 0; f < F; f++) {
        row[M.cell_rank[v.faceCell1ID[f]] + 1]++;
        if (v.faceBoundaryID[f] <= -1) row[M.cell_rank[v.faceCell2ID[f]] + 1]++;
    }
    for (int c = 0; c < C; c++) row[c + 1] += row[c];
    const int nnz = row[C];
    std::vector<int2> ent(nnz);
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (int f = 0; f < F; f++) {
        const int k = M.face_rank[f];
        const int c1 = M.cell_rank[v.faceCell1ID[f]];
        const int b = v.faceBoundaryID[f];
        if (b > -1) {
            ent[fill[c1]++] = make_int2(k << 1, ~b);
        } else {
            const int c2 = M.cell_rank[v.faceCell2ID[f]];
            ent[fill[c1]++] = make_int2(k << 1, c2);
            ent[fill[c2]++] = make_int2((k << 1) | 1, c1);
        }
    }

    // ---- boundaries / regions ----------------------------------------------
    std::vector<int> bface(B), bcell(B);
    M.btype_h.assign(B, 0);
    for (int t = 0; t < 5; t++) M.ntype[t] = 0;
    std::vector<int> blist[5];
    for (int b = 0; b < B; b++) {
        int f = v.boundaryFaceID[b];
        bface[b] = M.face_rank[f];
        bcell[b] = M.cell_rank[v.faceCell1ID[f]];
        int t = v.regionType[v.boundaryRegionID[b]];
        M.btype_h[b] = t;
        blist[t].push_back(b);
    }
    // face-loop counters count boundary faces by type (MassKernels.ipp:248-304)
    for (int f = 0; f < F; f++) {
        int b = v.faceBoundaryID[f];
        if (b > -1) M.ntype[v.regionType[v.boundaryRegionID[b]]]++;
    }
    std::vector<int> bcells;
    {
        std::vector<char> has(C, 0);
        for (int k = Fi; k < F; k++) has[fconn[k].x] = 1;
        for (int c = 0; c < C; c++)
            if (has[c]) bcells.push_back(c);
    }
    // OUTLET-owning cells with their OUTLET boundaries in ascending id
    std::vector<int> ocells, orow(1, 0), oent;
    {
        std::vector<int> ocnt(C, 0);
        for (int b : blist[CUPCFD_RTYPE_OUTLET]) ocnt[bcell[b]]++;
        std::vector<int> start(C, -1);
        for (int c = 0; c < C; c++)
            if (ocnt[c]) {
                start[c] = orow.back();
                ocells.push_back(c);
                orow.push_back(orow.back() + ocnt[c]);
            }
        oent.assign(orow.back(), 0);
        for (int b : blist[CUPCFD_RTYPE_OUTLET]) oent[start[bcell[b]]++] = b;
    }
    bool typed = false;
    for (int r = 0; r < R; r++)
        if (v.regionType[r] != CUPCFD_RTYPE_DEFAULT) typed = true;
    std::vector<int> carry;
    if (typed) {
        carry.assign(F, -1);
        int last_io = -1, last_vis = -1;
        for (int f = 0; f < F; f++) {
            int k = M.face_rank[f];
            int b = v.faceBoundaryID[f];
            if (b <= -1) {
                carry[k] = last_io;
                last_vis = k;
            } else {
                int t = v.regionType[v.boundaryRegionID[b]];
                if (t == CUPCFD_RTYPE_SYMP) carry[k] = last_vis;
                if (t == CUPCFD_RTYPE_INLET || t == CUPCFD_RTYPE_OUTLET) {
                    last_io = k;
                    last_vis = k;
                }
            }
        }
    }

    std::vector<T> bdist(B), rbuf;
    std::vector<uint8_t> rstd(R), rflux(R), radiab(R);
    for (int b = 0; b < B; b++) bdist[b] = v.boundaryDistance[b];
    M.bregion_h.assign(v.boundaryRegionID, v.boundaryRegionID + B);
    for (int r = 0; r < R; r++) {
        rstd[r] = v.regionStd ? v.regionStd[r] : 0;
        rflux[r] = v.regionFlux ? v.regionFlux[r] : 0;
        radiab[r] = v.regionAdiab ? v.regionAdiab[r] : 0;
    }

    // ---- upload -------------------------------------------------------------
    Uploader up{M.allocs};
    DevMesh<T> &d = M.d;
    d.C = C;
    d.F = F;
    d.Fi = Fi;
    d.B = B;
    d.R = R;
    d.nnz = nnz;
    d.maxrow = 0;
    d.rowlen = C ? row[1] - row[0] : 0;
    for (int c = 0; c < C; c++) {
        d.maxrow = std::max(d.maxrow, row[c + 1] - row[c]);
        if (row[c + 1] - row[c] != d.rowlen) d.rowlen = 0;
    }
    d.fconn = up.put(fconn.data(), F);
    d.fgeoS = up.put(fS.data(), F);
    d.fgeoX = up.put(fX.data(), F);
    d.fgeoP = up.put(fP.data(), F);
    d.fgeoN = up.put(fN.data(), F);
    d.frl = up.put(frl.data(), F);
    d.cgeo = up.put(cg.data(), C);
    d.cvol = up.put(cv.data(), C);
    d.row = up.put(row.data(), C + 1);
    d.ent = up.put(ent.data(), nnz);
    d.bface = up.put(bface.data(), B);
    d.bregion = up.put(v.boundaryRegionID, B);
    d.bcell = up.put(bcell.data(), B);
    d.bdist = up.put(bdist.data(), B);
    d.byplus = up.put<T>(nullptr, B);
    d.buplus = up.put<T>(nullptr, B);
    d.bq = up.put<T>(nullptr, B);
    d.bh = up.put<T>(nullptr, B);
    d.bt = up.put<T>(nullptr, B);
    d.bshear = up.put<T>(nullptr, 3 * (size_t)B);
    d.rtype = up.put(v.regionType, R);
    d.rstd = up.put(rstd.data(), R);
    d.rflux = up.put(rflux.data(), R);
    d.radiab = up.put(radiab.data(), R);
    d.rylog = up.put<T>(nullptr, R);
    d.rforcet = up.put<T>(nullptr, 3 * (size_t)R);
    auto put_r = [&](const T *src) {
        rbuf.assign(R, (T)0);
        if (src)
            for (int r = 0; r < R; r++) rbuf[r] = src[r];
        return (const T *)up.put(rbuf.data(), R);
    };
    d.relog = put_r(v.regionELog);
    d.rdensity = put_r(v.regionDensity);
    d.rturbke = put_r(v.regionTurbKE);
    d.rturbdiss = put_r(v.regionTurbDiss);
    d.rsplvl = put_r(v.regionSplvl);
    d.rden = put_r(v.regionDen);
    d.rr = put_r(v.regionR);
    d.rt = put_r(v.regionT);
    {
        std::vector<T> u(3 * (size_t)R, (T)0);
        if (v.regionUVW.base)
            for (int r = 0; r < R; r++)
                for (int k = 0; k < 3; k++) u[3 * r + k] = *v3ptr(v.regionUVW, r, k);
        d.ruvw = up.put(u.data(), 3 * (size_t)R);
    }
    d.carry_src = typed ? up.put(carry.data(), F) : nullptr;
    d.typed = typed ? 1 : 0;
    d.carry_init = (T)0;
    d.bcells = up.put(bcells.data(), bcells.size());
    d.nbc = (int)bcells.size();
    for (int t = 0; t < 5; t++) {
        d.blist[t] = up.put(blist[t].data(), blist[t].size());
        d.nblist[t] = (int)blist[t].size();
    }
    d.ocells = up.put(ocells.data(), ocells.size());
    d.orow = up.put(orow.data(), orow.size());
    d.oent = up.put(oent.data(), oent.size());
    d.noc = (int)ocells.size();

    // ---- cell tiles ---------------------------------------------------------
    {
        const int TC = kTileCells;
        const int NT = (C + TC - 1) / TC;
        std::vector<int> istart(C + 1, 0), bstart(C + 1, 0);  // first owned interior / boundary face per cell
        for (int k = 0; k < Fi; k++) istart[fconn[k].x + 1]++;
        for (int k = Fi; k < F; k++) bstart[fconn[k].x + 1]++;
        for (int c = 0; c < C; c++) {
            istart[c + 1] += istart[c];
            bstart[c + 1] += bstart[c];
        }
        std::vector<int4> tinfo(NT);
        for (int t = 0; t < NT; t++) {
            const int c0 = t * TC, c1 = std::min(C, c0 + TC);
            tinfo[t] = make_int4(istart[c0], istart[c1] - istart[c0], Fi + bstart[c0], bstart[c1] - bstart[c0]);
        }
        std::vector<int> thp(NT + 1, 0), thf;
        for (int k = 0; k < Fi; k++)
            if (fconn[k].y / TC != fconn[k].x / TC) thp[fconn[k].y / TC + 1]++;
        for (int t = 0; t < NT; t++) thp[t + 1] += thp[t];
        thf.assign(thp[NT], 0);
        {
            std::vector<int> hf(thp.begin(), thp.end() - 1);
            for (int k = 0; k < Fi; k++)  // ascending k within each tile
                if (fconn[k].y / TC != fconn[k].x / TC) thf[hf[fconn[k].y / TC]++] = k;
        }
        std::vector<int2> thc(thf.size());
        for (size_t h = 0; h < thf.size(); h++) thc[h] = fconn[thf[h]];
        int tslots = 0, tnnz = 0;
        bool fits = true;
        std::vector<uint16_t> eslot(nnz, 0);
        for (int t = 0; t < NT && fits; t++) {
            const int4 ti = tinfo[t];
            const int nslot = ti.y + ti.w + (thp[t + 1] - thp[t]);
            tslots = std::max(tslots, nslot);
            // slot codes are 15 bits; a tile whose cells have no faces has no slots to stage
            if (nslot > 32767 || nslot == 0) fits = false;
            const int c0 = t * TC, c1 = std::min(C, c0 + TC);
            tnnz = std::max(tnnz, row[c1] - row[c0]);
            // owned faces (whose per-face outputs this tile writes) have their owner in the tile
            for (int k = ti.x; k < ti.x + ti.y; k++)
                if (fconn[k].x < c0 || fconn[k].x >= c1) fits = false;
            for (int k = ti.z; k < ti.z + ti.w; k++)
                if (fconn[k].x < c0 || fconn[k].x >= c1) fits = false;
            for (int c = c0; c < c1 && fits; c++)
                for (int e = row[c]; e < row[c + 1]; e++) {
                    const int k = ent[e].x >> 1, role = ent[e].x & 1;
                    int slot;
                    if (k >= Fi) slot = ti.y + (k - ti.z);
                    else if (fconn[k].x / TC == t) slot = k - ti.x;
                    else {
                        const int *h0 = thf.data() + thp[t], *h1 = thf.data() + thp[t + 1];
                        slot = ti.y + ti.w + (int)(std::lower_bound(h0, h1, k) - h0);
                    }
                    // the slot must name this entry's face the way the kernels decode it
                    int back;
                    if (slot < ti.y) back = ti.x + slot;
                    else if (slot < ti.y + ti.w) back = ti.z + (slot - ti.y);
                    else if (slot < nslot) back = thf[thp[t] + (slot - ti.y - ti.w)];
                    else back = -1;
                    if (back != k) fits = false;
                    eslot[e] = (uint16_t)(slot | (role << 15));
                }
        }
        d.tcells = TC;
        d.ntiles = fits ? NT : 0;
        d.tslots = tslots;
        d.tnnz = tnnz;
        d.thc = up.put(thc.data(), thc.size());
        d.tinfo = up.put(tinfo.data(), tinfo.size());
        d.thp = up.put(thp.data(), thp.size());
        d.thf = up.put(thf.data(), thf.size());
        d.eslot = up.put(eslot.data(), eslot.size());
    }
    if (!up.ok) return CUPCFD_E_ERROR;
    M.radiab_h = radiab;
    return CUPCFD_E_SUCCESS;
}

}  // namespace

// Copy the view's mutable mesh state to the device (push) or back (pull).
template <class T, class View> int sync_state(Mesh<T> &M, const View &v, bool push) {
    const DevMesh<T> &d = M.d;
    const int B = d.B, R = d.R;
    std::vector<T> buf;
    auto xfer = [&](T *dev, T *host, int n) -> bool {
        if (!host || n == 0) return true;
        if (push) return hipMemcpy(dev, host, n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
        return hipMemcpy(host, dev, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess;
    };
    auto xfer3 = [&](T *dev, const decltype(v.boundaryShear) &hv, int n) -> bool {
        if (!hv.base || n == 0) return true;
        int64_t stride = hv.stride ? hv.stride : (int64_t)(3 * sizeof(T));
        buf.resize(3 * (size_t)n);
        if (push) {
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) buf[3 * i + k] = ((T *)((char *)hv.base + i * stride))[k];
            return hipMemcpy(dev, buf.data(), 3 * n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
        }
        if (hipMemcpy(buf.data(), dev, 3 * n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return false;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) ((T *)((char *)hv.base + i * stride))[k] = buf[3 * i + k];
        return true;
    };
    bool ok = xfer(d.byplus, v.boundaryYPlus, B) && xfer(d.buplus, v.boundaryUPlus, B) &&
              xfer(d.bq, v.boundaryQ, B) && xfer(d.bh, v.boundaryH, B) && xfer(d.bt, v.boundaryT, B) &&
              xfer3(d.bshear, v.boundaryShear, B) && xfer(d.rylog, v.regionYLog, R) &&
              xfer3(d.rforcet, v.regionForceTangent, R);
    return ok ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

int sync_state_any(cupcfd_hip_mesh *h, bool push) {
    if (h->precision == 64) return sync_state(*h->m64, h->m64->view64, push);
    return sync_state(*h->m32, h->m32->view32, push);
}

// Scalar-loop PhiFace source per face for one assignment rule
// (ScalarKernels.ipp:240-447): OUTLET faces always assign PhiFace; INLET faces
// assign 0 unless ivar is VarT/VarTE/VarED; WALL faces assign it when the
// enthalpy wall branch runs on a non-adiabatic region. Entry k is the last
// assigning face before face k in mesh face order, or -1.
template <class T> int *scalar_carry_src(Mesh<T> &M, bool inlet_assigns, bool wall_assigns) {
    int key = (inlet_assigns ? 1 : 0) | (wall_assigns ? 2 : 0);
    if (M.sc_src[key]) return M.sc_src[key];
    const int F = M.d.F;
    std::vector<int> src(F, -1);
    int last = -1;
    for (int f = 0; f < F; f++) {
        int k = M.face_rank[f];
        int nb = M.fconn_h[2 * k + 1];
        if (nb >= 0) {
            src[k] = last;
            continue;
        }
        int b = ~nb;
        int t = M.btype_h[b];
        if (t == CUPCFD_RTYPE_INLET) {
            src[k] = last;
            if (inlet_assigns) last = k;
        } else if (t == CUPCFD_RTYPE_OUTLET) {
            last = k;
        } else if (t == CUPCFD_RTYPE_WALL) {
            if (wall_assigns && !M.radiab_h[M.bregion_h[b]]) last = k;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, (F ? F : 1) * sizeof(int)) != hipSuccess) return nullptr;
    if (F && hipMemcpy(p, src.data(), F * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    M.sc_src[key] = (int *)p;
    return M.sc_src[key];
}

template int *scalar_carry_src<double>(Mesh<double> &, bool, bool);
template int *scalar_carry_src<float>(Mesh<float> &, bool, bool);

template <class T>
int scalar_carry_readers(Mesh<T> &M, bool inlet_assigns, bool wall_assigns, const int **off, const int **rd,
                         int *nrd) {
    const int key = (inlet_assigns ? 1 : 0) | (wall_assigns ? 2 : 0);
    if (!M.sc_rd_off[key]) {
        const int F = M.d.F, Fi = M.d.Fi, nb = F - Fi;
        // the sources in mesh face order, as scalar_carry_src walks them
        std::vector<int> src(F, -1);
        int last = -1;
        for (int f = 0; f < F; f++) {
            const int k = M.face_rank[f];
            const int nb_ = M.fconn_h[2 * k + 1];
            if (nb_ >= 0) {
                src[k] = last;
                continue;
            }
            const int b = ~nb_;
            const int t = M.btype_h[b];
            if (t == CUPCFD_RTYPE_INLET) {
                if (!inlet_assigns) src[k] = last;  // it reads the carried value
                else last = k;
            } else if (t == CUPCFD_RTYPE_OUTLET) {
                last = k;
            } else if (t == CUPCFD_RTYPE_WALL) {
                if (wall_assigns && !M.radiab_h[M.bregion_h[b]]) last = k;
            }
        }
        std::vector<int> cnt(nb + 1, 0), list;
        for (int k = 0; k < F; k++)
            if (src[k] >= 0) cnt[src[k] - Fi + 1]++;
        for (int k = 0; k < nb; k++) cnt[k + 1] += cnt[k];
        list.assign(cnt[nb], 0);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int f = 0; f < F; f++) {  // ascending mesh face id within each source
            const int k = M.face_rank[f];
            if (src[k] >= 0) list[fill[src[k] - Fi]++] = k;
        }
        void *po = nullptr, *pr = nullptr;
        if (hipMalloc(&po, (nb + 1) * sizeof(int)) != hipSuccess) return CUPCFD_E_ERROR;
        if (hipMalloc(&pr, (list.empty() ? 1 : list.size()) * sizeof(int)) != hipSuccess) {
            (void)hipFree(po);
            return CUPCFD_E_ERROR;
        }
        bool ok = hipMemcpy(po, cnt.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        if (ok && !list.empty())
            ok = hipMemcpy(pr, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) {
            (void)hipFree(po);
            (void)hipFree(pr);
            return CUPCFD_E_ERROR;
        }
        M.sc_rd_off[key] = (int *)po;
        M.sc_rd[key] = (int *)pr;
        M.sc_nrd[key] = (int)list.size();
    }
    *off = M.sc_rd_off[key];
    *rd = M.sc_rd[key];
    *nrd = M.sc_nrd[key];
    return CUPCFD_E_SUCCESS;
}

template int scalar_carry_readers<double>(Mesh<double> &, bool, bool, const int **, const int **, int *);
template int scalar_carry_readers<float>(Mesh<float> &, bool, bool, const int **, const int **, int *);

// A stable counting sort of the boundary ids by region: each region's
// boundaries in ascending id, the order the reference's boundary loops add
// them in (one pass over b = 0 .. B-1).
template <class T> bool region_boundaries(Mesh<T> &M) {
    if (M.rbnd && M.rbnd_off) return true;
    const int B = M.d.B, R = M.d.R;
    std::vector<int> off(R + 1, 0), lst(B ? B : 1);
    for (int b = 0; b < B; b++) off[M.bregion_h[b] + 1]++;
    for (int r = 0; r < R; r++) off[r + 1] += off[r];
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int b = 0; b < B; b++) lst[pos[M.bregion_h[b]]++] = b;
    void *pl = nullptr, *po = nullptr;
    if (hipMalloc(&pl, lst.size() * sizeof(int)) != hipSuccess) return false;
    if (hipMalloc(&po, off.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(pl, lst.data(), lst.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(po, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(pl);
        if (po) (void)hipFree(po);
        return false;
    }
    M.rbnd = (int *)pl;
    M.rbnd_off = (int *)po;
    return true;
}

template bool region_boundaries<double>(Mesh<double> &);
template bool region_boundaries<float>(Mesh<float> &);

template <class T> const int *device_order(Mesh<T> &M, bool faces) {
    int *&d = faces ? M.face_order_d : M.cell_order_d;
    if (d) return d;
    const std::vector<int32_t> &h = faces ? M.face_order : M.cell_order;
    void *p = nullptr;
    if (hipMalloc(&p, (h.empty() ? 1 : h.size()) * sizeof(int)) != hipSuccess) return nullptr;
    if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    d = (int *)p;
    return d;
}

template const int *device_order<double>(Mesh<double> &, bool);
template const int *device_order<float>(Mesh<float> &, bool);

}  // namespace cupcfd_hip

using namespace cupcfd_hip;

namespace {

// Build a device mesh from a view, on the host or on the device; the handle
// keeps the view for drop-in state mirroring.
template <class T, class View>
int create_mesh(const View *view, int32_t flags, cupcfd_hip_mesh **out, bool on_device, View Mesh<T>::*keep) {
    if (!view || !out) return CUPCFD_E_NULL_PTR;
    *out = nullptr;
    auto *M = new (std::nothrow) Mesh<T>();
    if (!M) return CUPCFD_E_ERROR;
    int rc;
    try {
        rc = on_device ? build_device(*M, *view, flags) : build(*M, *view, flags);
    } catch (const std::bad_alloc &) {
        rc = CUPCFD_E_ERROR;
    }
    if (rc == CUPCFD_E_SUCCESS) {
        M->*keep = *view;
        rc = sync_state(*M, *view, true);
    }
    if (rc == CUPCFD_E_SUCCESS) tile_split_build(*M);
    auto *h = rc == CUPCFD_E_SUCCESS ? new (std::nothrow) cupcfd_hip_mesh{} : nullptr;
    if (!h) {
        delete M;
        return rc == CUPCFD_E_SUCCESS ? CUPCFD_E_ERROR : rc;
    }
    if constexpr (sizeof(T) == 8) *h = cupcfd_hip_mesh{64, M, nullptr};
    else *h = cupcfd_hip_mesh{32, nullptr, M};
    *out = h;
    return CUPCFD_E_SUCCESS;
}

}  // namespace

extern "C" int cupcfd_hip_mesh_create_f64(const cupcfd_mesh_view_f64 *view, int32_t flags, cupcfd_hip_mesh **out) {
    return create_mesh<double>(view, flags, out, false, &Mesh<double>::view64);
}

extern "C" int cupcfd_hip_mesh_create_f32(const cupcfd_mesh_view_f32 *view, int32_t flags, cupcfd_hip_mesh **out) {
    return create_mesh<float>(view, flags, out, false, &Mesh<float>::view32);
}

extern "C" int cupcfd_hip_mesh_create_device_f64(const cupcfd_mesh_view_f64 *view, int32_t flags,
                                                 cupcfd_hip_mesh **out) {
    return create_mesh<double>(view, flags, out, true, &Mesh<double>::view64);
}

extern "C" int cupcfd_hip_mesh_create_device_f32(const cupcfd_mesh_view_f32 *view, int32_t flags,
                                                 cupcfd_hip_mesh **out) {
    return create_mesh<float>(view, flags, out, true, &Mesh<float>::view32);
}

extern "C" int cupcfd_hip_mesh_destroy(cupcfd_hip_mesh *mesh) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    (void)hipDeviceSynchronize();
    delete mesh->m64;
    delete mesh->m32;
    delete mesh;
    return CUPCFD_E_SUCCESS;
}

extern "C" int cupcfd_hip_mesh_counts(const cupcfd_hip_mesh *mesh, int32_t *nCells, int32_t *nFaces,
                                      int32_t *nBoundaries, int32_t *nRegions) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    int C, F, B, R;
    if (mesh->precision == 64) {
        C = mesh->m64->d.C, F = mesh->m64->d.F, B = mesh->m64->d.B, R = mesh->m64->d.R;
    } else {
        C = mesh->m32->d.C, F = mesh->m32->d.F, B = mesh->m32->d.B, R = mesh->m32->d.R;
    }
    if (nCells) *nCells = C;
    if (nFaces) *nFaces = F;
    if (nBoundaries) *nBoundaries = B;
    if (nRegions) *nRegions = R;
    return CUPCFD_E_SUCCESS;
}

extern "C" int cupcfd_hip_mesh_cell_order(const cupcfd_hip_mesh *mesh, int32_t *order) {
    if (!mesh || !order) return CUPCFD_E_NULL_PTR;
    const auto &v = mesh->precision == 64 ? mesh->m64->cell_order : mesh->m32->cell_order;
    std::copy(v.begin(), v.end(), order);
    return CUPCFD_E_SUCCESS;
}

extern "C" int cupcfd_hip_mesh_face_order(const cupcfd_hip_mesh *mesh, int32_t *order) {
    if (!mesh || !order) return CUPCFD_E_NULL_PTR;
    const auto &v = mesh->precision == 64 ? mesh->m64->face_order : mesh->m32->face_order;
    std::copy(v.begin(), v.end(), order);
    return CUPCFD_E_SUCCESS;
}

namespace {

const char *const kDigestFields[] = {
    "counts",  "fconn",   "fgeoS",   "fgeoX",   "fgeoP",    "fgeoN",   "frl",      "cgeo",    "cvol",
    "row",     "ent",     "bface",   "bregion", "bcell",    "bdist",   "rtype",    "rstd",    "rflux",
    "radiab",  "relog",   "rdensity", "rturbke", "rturbdiss", "rsplvl", "rden",     "rr",      "rt",
    "ruvw",    "carry_src", "bcells", "blist",  "ocells",   "orow",    "oent",     "halo_count", "tinfo",
    "thp",     "thf",     "thc",     "eslot",   "cell_order", "face_order", "cell_rank", "face_rank", "fconn_h",
    "btype_h", "bregion_h", "radiab_h", "ntype"};
constexpr int kNDigest = (int)(sizeof(kDigestFields) / sizeof(kDigestFields[0]));

// 64-bit hash of a byte string, 8 bytes per step (FNV-1a constants)
struct Hash64 {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            h = (h ^ w) * 0x100000001b3ull;
            h ^= h >> 29;
        }
        for (; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
        h ^= n;
    }
};

template <class T> int mesh_digest(const Mesh<T> &M, uint64_t *out) {
    const DevMesh<T> &d = M.d;
    std::vector<unsigned char> buf;
    bool ok = true;
    int i = 0;
    auto dev = [&](const void *p, size_t bytes) {
        Hash64 h;
        if (ok && bytes) {
            buf.resize(bytes);
            ok = hipMemcpy(buf.data(), p, bytes, hipMemcpyDeviceToHost) == hipSuccess;
            h.add(buf.data(), bytes);
        }
        out[i++] = h.h;
    };
    auto host = [&](const void *p, size_t bytes) {
        Hash64 h;
        h.add(p, bytes);
        out[i++] = h.h;
    };
    const size_t C = d.C, F = d.F, B = d.B, R = d.R, nnz = d.nnz;
    const int32_t counts[] = {d.C,     d.F,     d.Fi,    d.B,          d.R,          d.nnz,        d.maxrow,
                              d.rowlen, d.typed, d.nbc,   d.nblist[0],  d.nblist[1],  d.nblist[2],  d.nblist[3],
                              d.nblist[4], d.noc, d.tcells, d.ntiles,   d.tslots,     d.tnnz};
    host(counts, sizeof(counts));
    dev(d.fconn, F * sizeof(int2));
    dev(d.fgeoS, F * sizeof(V4<T>));
    dev(d.fgeoX, F * sizeof(V4<T>));
    dev(d.fgeoP, F * sizeof(V4<T>));
    dev(d.fgeoN, F * sizeof(V2<T>));
    dev(d.frl, F * sizeof(T));
    dev(d.cgeo, C * sizeof(V4<T>));
    dev(d.cvol, C * sizeof(T));
    dev(d.row, (C + 1) * sizeof(int));
    dev(d.ent, nnz * sizeof(int2));
    dev(d.bface, B * sizeof(int));
    dev(d.bregion, B * sizeof(int));
    dev(d.bcell, B * sizeof(int));
    dev(d.bdist, B * sizeof(T));
    dev(d.rtype, R * sizeof(int));
    dev(d.rstd, R);
    dev(d.rflux, R);
    dev(d.radiab, R);
    for (const T *p : {d.relog, d.rdensity, d.rturbke, d.rturbdiss, d.rsplvl, d.rden, d.rr, d.rt}) dev(p, R * sizeof(T));
    dev(d.ruvw, 3 * R * sizeof(T));
    dev(d.carry_src, d.typed ? F * sizeof(int) : 0);
    dev(d.bcells, d.nbc * sizeof(int));
    {  // the five by-type lists as one sequence
        Hash64 h;
        for (int t = 0; t < 5 && ok; t++) {
            buf.resize(d.nblist[t] * sizeof(int));
            if (d.nblist[t])
                ok = hipMemcpy(buf.data(), d.blist[t], buf.size(), hipMemcpyDeviceToHost) == hipSuccess;
            h.add(buf.data(), buf.size());
        }
        out[i++] = h.h;
    }
    dev(d.ocells, d.noc * sizeof(int));
    dev(d.orow, (d.noc + 1) * sizeof(int));
    dev(d.oent, d.nblist[CUPCFD_RTYPE_OUTLET] * sizeof(int));
    // tile arrays are read only when every tile fits its 15-bit slots
    const size_t NT = d.ntiles;
    int H = 0;
    if (NT && ok) ok = hipMemcpy(&H, d.thp + NT, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    host(&H, sizeof(H));
    dev(d.tinfo, NT * sizeof(int4));
    dev(d.thp, NT ? (NT + 1) * sizeof(int) : 0);
    dev(d.thf, H * sizeof(int));
    dev(d.thc, H * sizeof(int2));
    dev(d.eslot, NT ? nnz * sizeof(uint16_t) : 0);
    host(M.cell_order.data(), M.cell_order.size() * sizeof(int32_t));
    host(M.face_order.data(), M.face_order.size() * sizeof(int32_t));
    host(M.cell_rank.data(), M.cell_rank.size() * sizeof(int32_t));
    host(M.face_rank.data(), M.face_rank.size() * sizeof(int32_t));
    host(M.fconn_h.data(), M.fconn_h.size() * sizeof(int32_t));
    host(M.btype_h.data(), M.btype_h.size() * sizeof(int32_t));
    host(M.bregion_h.data(), M.bregion_h.size() * sizeof(int32_t));
    host(M.radiab_h.data(), M.radiab_h.size());
    host(M.ntype, sizeof(M.ntype));
    return ok && i == kNDigest ? CUPCFD_E_SUCCESS : CUPCFD_E_ERROR;
}

}  // namespace

extern "C" int cupcfd_hip_mesh_digest(const cupcfd_hip_mesh *mesh, uint64_t *digest, int32_t capacity,
                                      int32_t *count) {
    if (!mesh || !digest || !count) return CUPCFD_E_NULL_PTR;
    *count = kNDigest;
    if (capacity < kNDigest) return CUPCFD_E_ARRAY_SIZE_MISMATCH;
    return mesh->precision == 64 ? mesh_digest(*mesh->m64, digest) : mesh_digest(*mesh->m32, digest);
}

extern "C" const char *cupcfd_hip_mesh_digest_field(int32_t i) {
    return i >= 0 && i < kNDigest ? kDigestFields[i] : nullptr;
}

extern "C" int cupcfd_hip_mesh_set_carry_init(cupcfd_hip_mesh *mesh, double value) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    if (mesh->precision == 64) mesh->m64->d.carry_init = value;
    else mesh->m32->d.carry_init = (float)value;
    return CUPCFD_E_SUCCESS;
}

extern "C" int cupcfd_hip_mesh_set_ordered_sums(cupcfd_hip_mesh *mesh, int32_t sequential) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    if (mesh->precision == 64) mesh->m64->d.tree_sums = sequential ? 0 : 1;
    else mesh->m32->d.tree_sums = sequential ? 0 : 1;
    return CUPCFD_E_SUCCESS;
}

namespace {
template <class T> void tile_layout(const Mesh<T> &M, int32_t *info) {
    const int32_t v[6] = {M.d.ntiles, M.d.tslots, M.d.tnnz, M.split_na, M.split_nb, M.uvw_split};
    for (int i = 0; i < 6; i++) info[i] = v[i];
}
}  // namespace

extern "C" int cupcfd_hip_mesh_tile_layout(const cupcfd_hip_mesh *mesh, int32_t info[6]) {
    if (!mesh || !info) return CUPCFD_E_NULL_PTR;
    if (mesh->precision == 64) tile_layout(*mesh->m64, info);
    else tile_layout(*mesh->m32, info);
    return CUPCFD_E_SUCCESS;
}

extern "C" int cupcfd_hip_mesh_push_state(cupcfd_hip_mesh *mesh) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    return sync_state_any(mesh, true);
}

extern "C" int cupcfd_hip_mesh_pull_state(cupcfd_hip_mesh *mesh) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    return sync_state_any(mesh, false);
}

extern "C" int cupcfd_hip_mesh_set_schedule(cupcfd_hip_mesh *mesh, int32_t lanes_per_cell) {
    if (!mesh) return CUPCFD_E_NULL_PTR;
    if (lanes_per_cell != 0 && lanes_per_cell != 1 && lanes_per_cell != 2 && lanes_per_cell != 4 &&
        lanes_per_cell != 6 && lanes_per_cell != 8 && lanes_per_cell != 64)
        return CUPCFD_E_CONFIG_INVALID_VALUE;
    auto set = [&](auto *M) {
        M->schedule_lpc = lanes_per_cell;
        for (int &k : M->schedule_k) k = -1;  // an explicit schedule replaces an autotuned one
    };
    if (mesh->precision == 64) set(mesh->m64);
    else set(mesh->m32);
    return CUPCFD_E_SUCCESS;
}

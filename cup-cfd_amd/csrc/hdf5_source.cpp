// Mesh source for the reference's HDF5 mesh files (MeshHDF5Source,
// src/geometry/mesh/implementation/source/MeshHDF5Source.cpp): a MiniSoA
// mesh assembled as UnstructuredMeshInterface::addData
// (include/geometry/mesh/interface/component/UnstructuredMeshInterface.ipp:1464-1862)
// assembles it from that source on one rank.
//
// File layout (1-based labels; 0 means "none"):
//   root attributes ncel, nfac, nbnd, nreg, nvrt, maxfaces (int)
//   /cell/nfaces[ncel], /cell/cface[ncel][maxfaces], /cell/x[ncel][3], /cell/vol[ncel]
//   /face/cell1[nfac], /face/cell2[nfac] (0: boundary face), /face/bnd[nfac],
//   /face/lambda[nfac], /face/area[nfac], /face/n[nfac][3], /face/x[nfac][3]
//   /bnd/face[nbnd], /bnd/rid[nbnd], /bnd/distance[nbnd]
// Real datasets are read as double (HDF5 converts the file's float exactly).
//
// Assembly (all cells local, no ghosts):
//   cells      ascending label; id = position
//   faces      the distinct labels of the cells' face lists, ascending
//   boundaries the distinct boundary labels of those faces, ascending
//   regions    every region label 1..nreg, each an RTYPE_DEFAULT region with
//              all-zero properties (addRegion(label, name), CupCfdSoAMesh.cpp:211-222)
//   xpac = xnac = 0, rlencos = 0 (addData passes zeros)
//
// libhdf5 is loaded at run time (dlopen), so the library does not depend on
// it: CUPCFD_HDF5_LIBRARY names it, else libhdf5.so / libhdf5.so.103 on the
// loader path, else /opt/conda/lib/libhdf5.so.103. Without it every call
// returns CUPCFD_E_NOT_IMPLEMENTED.
#include <dlfcn.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "host_mesh.hpp"

namespace cupcfd_hip {
namespace {

// the HDF5 1.10 C API this reader uses (hid_t is 64-bit from 1.10 on)
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

struct H5 {
    bool ok = false;
    herr_t (*open)();
    hid_t (*Fopen)(const char *, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Dopen2)(hid_t, const char *, hid_t);
    herr_t (*Dclose)(hid_t);
    hid_t (*Dget_space)(hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    herr_t (*Sclose)(hid_t);
    hid_t (*Aopen)(hid_t, const char *, hid_t);
    herr_t (*Aread)(hid_t, hid_t, void *);
    herr_t (*Aclose)(hid_t);
    hid_t (*Aget_space)(hid_t);
    herr_t (*Eset_auto2)(hid_t, void *, void *);
    hid_t *native_int = nullptr, *native_double = nullptr;
};

std::once_flag g_h5_once;
H5 g_h5;

template <class F> bool sym(void *lib, const char *name, F &f) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

void load_h5() {
    std::vector<std::string> names;
    if (const char *e = std::getenv("CUPCFD_HDF5_LIBRARY")) names.push_back(e);
    names.push_back("libhdf5.so");
    names.push_back("libhdf5.so.103");
    names.push_back("/opt/conda/lib/libhdf5.so.103");
    void *lib = nullptr;
    for (const auto &n : names)
        if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!lib) return;
    H5 &h = g_h5;
    bool ok = sym(lib, "H5open", h.open) && sym(lib, "H5Fopen", h.Fopen) && sym(lib, "H5Fclose", h.Fclose) &&
              sym(lib, "H5Dopen2", h.Dopen2) && sym(lib, "H5Dclose", h.Dclose) &&
              sym(lib, "H5Dget_space", h.Dget_space) && sym(lib, "H5Dread", h.Dread) &&
              sym(lib, "H5Sget_simple_extent_ndims", h.Sget_simple_extent_ndims) &&
              sym(lib, "H5Sget_simple_extent_dims", h.Sget_simple_extent_dims) && sym(lib, "H5Sclose", h.Sclose) &&
              sym(lib, "H5Aopen", h.Aopen) && sym(lib, "H5Aread", h.Aread) && sym(lib, "H5Aclose", h.Aclose) &&
              sym(lib, "H5Aget_space", h.Aget_space) && sym(lib, "H5Eset_auto2", h.Eset_auto2);
    if (!ok) return;
    h.native_int = reinterpret_cast<hid_t *>(dlsym(lib, "H5T_NATIVE_INT_g"));
    h.native_double = reinterpret_cast<hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
    if (!h.native_int || !h.native_double || h.open() < 0) return;
    h.Eset_auto2(0 /* H5E_DEFAULT */, nullptr, nullptr);  // report errors by return code only
    h.ok = true;
}

const H5 *h5() {
    std::call_once(g_h5_once, load_h5);
    return g_h5.ok ? &g_h5 : nullptr;
}

constexpr hid_t kDefault = 0;     // H5P_DEFAULT
constexpr hid_t kAll = 0;         // H5S_ALL
constexpr unsigned kReadOnly = 0; // H5F_ACC_RDONLY

struct File {
    const H5 &h;
    hid_t id = -1;
    File(const H5 &hh, const char *path) : h(hh) { id = h.Fopen(path, kReadOnly, kDefault); }
    ~File() {
        if (id >= 0) h.Fclose(id);
    }

    bool attr_int(const char *name, int *out) const {
        hid_t a = h.Aopen(id, name, kDefault);
        if (a < 0) return false;
        hid_t sp = h.Aget_space(a);
        hsize_t dims[4] = {0, 0, 0, 0};
        const int nd = sp >= 0 ? h.Sget_simple_extent_ndims(sp) : -1;
        bool ok = nd >= 0 && nd <= 1;
        if (ok && nd == 1) ok = h.Sget_simple_extent_dims(sp, dims, nullptr) == 1 && dims[0] == 1;
        if (sp >= 0) h.Sclose(sp);
        ok = ok && h.Aread(a, *h.native_int, out) >= 0;
        h.Aclose(a);
        return ok;
    }

    // a whole dataset of `rows` x `cols` values (cols = 1: a 1-d dataset)
    template <class X> bool read(const char *path, hid_t type, size_t rows, size_t cols, std::vector<X> &out) const {
        hid_t d = h.Dopen2(id, path, kDefault);
        if (d < 0) return false;
        hid_t sp = h.Dget_space(d);
        hsize_t dims[4] = {0, 0, 0, 0};
        const int nd = sp >= 0 ? h.Sget_simple_extent_ndims(sp) : -1;
        bool ok = (nd == 1 && cols == 1) || nd == 2;
        if (ok) ok = h.Sget_simple_extent_dims(sp, dims, nullptr) == nd;
        if (ok) ok = dims[0] == rows && (nd == 1 || dims[1] == cols);
        if (sp >= 0) h.Sclose(sp);
        if (ok) {
            out.assign(rows * cols, X());
            ok = rows * cols == 0 || h.Dread(d, type, kAll, kAll, kDefault, out.data()) >= 0;
        }
        h.Dclose(d);
        return ok;
    }
};

int build_hdf5(const char *path, cupcfd_host_mesh &m) {
    const H5 *hp = h5();
    if (!hp) return CUPCFD_E_NOT_IMPLEMENTED;
    const H5 &h = *hp;
    struct stat st;
    if (stat(path, &st) != 0) return CUPCFD_E_FILE_MISSING;
    File f(h, path);
    if (f.id < 0) return CUPCFD_E_INVALID_FILE_FORMAT;  // there, but not an HDF5 file libhdf5 can open
    int ncel, nfac, nbnd, nreg, maxfaces;
    if (!f.attr_int("ncel", &ncel) || !f.attr_int("nfac", &nfac) || !f.attr_int("nbnd", &nbnd) ||
        !f.attr_int("nreg", &nreg) || !f.attr_int("maxfaces", &maxfaces))
        return CUPCFD_E_INVALID_FILE_FORMAT;
    if (ncel < 1 || nfac < 1 || nbnd < 0 || nreg < 1 || maxfaces < 1) return CUPCFD_E_INVALID_FILE_FORMAT;
    const hid_t I = *h.native_int, D = *h.native_double;
    std::vector<int> nfaces, cface, c1, c2, fbnd, bface, brid;
    std::vector<double> cx, cvol, lam, area, fn, fx, bdist;
    bool ok = f.read("/cell/nfaces", I, ncel, 1, nfaces) && f.read("/cell/cface", I, ncel, maxfaces, cface) &&
              f.read("/cell/x", D, ncel, 3, cx) && f.read("/cell/vol", D, ncel, 1, cvol) &&
              f.read("/face/cell1", I, nfac, 1, c1) && f.read("/face/cell2", I, nfac, 1, c2) &&
              f.read("/face/bnd", I, nfac, 1, fbnd) && f.read("/face/lambda", D, nfac, 1, lam) &&
              f.read("/face/area", D, nfac, 1, area) && f.read("/face/n", D, nfac, 3, fn) &&
              f.read("/face/x", D, nfac, 3, fx);
    if (ok && nbnd > 0)
        ok = f.read("/bnd/face", I, nbnd, 1, bface) && f.read("/bnd/rid", I, nbnd, 1, brid) &&
             f.read("/bnd/distance", D, nbnd, 1, bdist);
    if (!ok) return CUPCFD_E_INVALID_FILE_FORMAT;

    // faces: the distinct labels of every cell's face list, ascending
    std::vector<int> flab;
    for (int c = 0; c < ncel; c++) {
        if (nfaces[c] < 0 || nfaces[c] > maxfaces) return CUPCFD_E_INVALID_FILE_FORMAT;
        for (int j = 0; j < nfaces[c]; j++) {
            const int l = cface[(size_t)c * maxfaces + j];
            if (l < 1 || l > nfac) return CUPCFD_E_INVALID_INDEX;
            flab.push_back(l);
        }
    }
    std::sort(flab.begin(), flab.end());
    flab.erase(std::unique(flab.begin(), flab.end()), flab.end());
    const int F = (int)flab.size();
    // boundaries: the distinct boundary labels of those faces, ascending
    std::vector<int> blab;
    for (int l : flab)
        if (c2[l - 1] == 0) {
            const int b = fbnd[l - 1];
            if (b < 1 || b > nbnd) return CUPCFD_E_INVALID_INDEX;
            blab.push_back(b);
        }
    std::sort(blab.begin(), blab.end());
    if (std::adjacent_find(blab.begin(), blab.end()) != blab.end()) return CUPCFD_E_INVALID_FILE_FORMAT;
    const int B = (int)blab.size();
    std::vector<int> bid(nbnd + 1, -1);
    for (int i = 0; i < B; i++) bid[blab[i]] = i;

    resize_mesh(m, ncel, F, B);
    for (int r = 0; r < nreg; r++) {
        m.regType.push_back(CUPCFD_RTYPE_DEFAULT);
        m.regStd.push_back(0);
        m.regFlux.push_back(0);
        m.regAdiab.push_back(0);
        for (auto *v : {&m.regYLog, &m.regELog, &m.regDensity, &m.regTurbKE, &m.regTurbDiss, &m.regSplvl, &m.regDen,
                        &m.regR, &m.regT})
            v->push_back(0.0);
        m.regForceTangent.insert(m.regForceTangent.end(), {0.0, 0.0, 0.0});
        m.regUVW.insert(m.regUVW.end(), {0.0, 0.0, 0.0});
    }
    m.nRegions = nreg;
    for (int c = 0; c < ncel; c++) {
        for (int k = 0; k < 3; k++) m.cellCenter[3 * (size_t)c + k] = cx[3 * (size_t)c + k];
        m.cellVolume[c] = cvol[c];
    }
    for (int i = 0; i < F; i++) {
        const int l = flab[i] - 1;
        const int a = c1[l], b = c2[l];
        if (a < 1 || a > ncel || b < 0 || b > ncel) return CUPCFD_E_INVALID_INDEX;
        m.faceCell1[i] = a - 1;
        m.faceCell2[i] = b > 0 ? b - 1 : -1;
        m.faceBnd[i] = b > 0 ? -1 : bid[fbnd[l]];
        m.faceLambda[i] = lam[l];
        m.faceArea[i] = area[l];
        for (int k = 0; k < 3; k++) {
            m.faceNorm[3 * (size_t)i + k] = fn[3 * (size_t)l + k];
            m.faceCenter[3 * (size_t)i + k] = fx[3 * (size_t)l + k];
        }
        if (b == 0) m.bndFace[m.faceBnd[i]] = i;
    }
    for (int i = 0; i < B; i++) {
        const int l = blab[i] - 1;
        const int r = brid[l];
        if (r < 1 || r > nreg) return CUPCFD_E_INVALID_INDEX;
        m.bndRegion[i] = r - 1;
        m.bndDistance[i] = bdist[l];
    }
    return CUPCFD_E_SUCCESS;
}

}  // namespace
}  // namespace cupcfd_hip

extern "C" int cupcfd_hdf5_mesh_create(const char *path, cupcfd_host_mesh **out) {
    if (!path || !out) return CUPCFD_E_NULL_PTR;
    *out = nullptr;
    cupcfd_host_mesh *m = new (std::nothrow) cupcfd_host_mesh();
    if (!m) return CUPCFD_E_ERROR;
    int rc;
    try {
        rc = cupcfd_hip::build_hdf5(path, *m);
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

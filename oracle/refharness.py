"""ORACLE — test infrastructure only.

Python side of oracle/ref_harness.cpp: CUP-CFD's own fvm kernels, mesh class
and StructGen source, compiled from the reference sources by
``make -f oracle/ref.mk`` into oracle/_ref/cupcfd_ref_harness. Exists only in
a tree that holds the reference sources; tests skip what needs it when the
binary is absent, and tests/golden/make_ref_goldens.py stores its outputs as
fixtures so that the oracle stays pinned without it.

The kernel functions take the same arguments as pyoracle's (and
cupcfd_hip.fvm's), so tests/kernel_runs.py drives all three the same way. Each
call writes the mesh, the mesh state and the arguments to a case file, runs
the harness, and copies every in/out array and the new mesh state back.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_ref", "cupcfd_ref_harness")
REF_TREE = os.environ.get("CUPCFD_REFERENCE", "/root/reference")
MPIEXEC = os.environ.get("CUPCFD_MPIEXEC", "/opt/conda/bin/mpiexec")

# Kernel arguments by extent. The harness's mesh carries a detached cell pair
# and one extra face after the case mesh (the reference's mesh finalize needs a
# ghost cell on every rank), so cell arrays get two padding entries, face
# arrays one, and the two-per-face RFace arrays two; results drop them.
CELL_ARGS = {"phiCell", "dPhidxCell", "dPhidxoCell", "dudx", "dvdx", "dwdx", "dpdx", "denCell", "uCell", "vCell",
             "wCell", "p", "ar", "su", "teCell", "edCell", "viseffCell", "tCell", "PhiCell", "VisEff", "Au", "Su",
             "Den", "TE", "dPhidx", "visEffCell", "sv", "sw", "au", "av", "aw", "ED", "den", "visEff"}
FACE_ARGS = {"massFlux", "MassFlux"}
FACE2_ARGS = {"rface", "RFace", "rFace"}

_KIND = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int32): 2, np.dtype(np.uint8): 3}
_DT = {0: np.float64, 1: np.float32, 2: np.int32, 3: np.uint8}

MESH_ARRAYS = ["faceCell1", "faceCell2", "faceBnd", "faceLambda", "faceNorm", "faceCenter", "faceXpac", "faceXnac",
               "faceRLencos", "faceArea", "cellCenter", "cellVolume", "bndFace", "bndRegion", "bndDistance",
               "regType", "regStd", "regFlux", "regAdiab", "regYLog", "regELog", "regDensity", "regTurbKE",
               "regTurbDiss", "regSplvl", "regDen", "regR", "regT", "regForceTangent", "regUVW"]
STATE_KEYS = ("bndYPlus", "bndUPlus", "bndQ", "bndH", "bndT", "bndShear", "regYLog", "regForceTangent")
INT_KEYS = {"faceCell1", "faceCell2", "faceBnd", "bndFace", "bndRegion", "regType"}
BOOL_KEYS = {"regStd", "regFlux", "regAdiab"}


def available():
    return os.path.exists(BIN)


def build():
    """Build the harness from the reference tree (only where that tree exists)."""
    if not os.path.isdir(REF_TREE):
        raise FileNotFoundError(f"no reference tree at {REF_TREE}")
    subprocess.run(["make", "-s", "-j8", "-f", os.path.join(HERE, "ref.mk"), f"REF={REF_TREE}"], check=True)


def write_case(path, recs):
    with open(path, "wb") as f:
        f.write(b"CUPREF1\n")
        for name, a in recs.items():
            a = np.ascontiguousarray(a)
            nb = name.encode()
            f.write(struct.pack("<I", len(nb)) + nb + struct.pack("<BQ", _KIND[a.dtype], a.size))
            f.write(a.tobytes())


def read_case(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    assert data[:8] == b"CUPREF1\n", "bad result file"
    i = 8
    while i < len(data):
        (n,) = struct.unpack_from("<I", data, i)
        i += 4
        name = data[i:i + n].decode()
        i += n
        kind, cnt = struct.unpack_from("<BQ", data, i)
        i += 9
        dt = np.dtype(_DT[kind])
        out[name] = np.frombuffer(data, dtype=dt, count=cnt, offset=i).copy()
        i += cnt * dt.itemsize
    return out


def run_case(recs):
    """Run the harness on a case (dict name -> array) and return its result records."""
    if not available():
        raise FileNotFoundError(f"{BIN} is not built (make -f oracle/ref.mk)")
    with tempfile.TemporaryDirectory(prefix="cupref") as d:
        cin, cout = os.path.join(d, "case.bin"), os.path.join(d, "result.bin")
        write_case(cin, recs)
        p = subprocess.run([MPIEXEC, "-n", "2", BIN, cin, cout], capture_output=True, text=True, cwd=d,
                           timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"reference harness failed ({p.returncode}): {p.stderr.strip()[-2000:]}")
        if os.path.exists(cout):
            return read_case(cout)
        return [read_case(cout + ".0"), read_case(cout + ".1")]


def _text(s):
    return np.frombuffer(s.encode(), dtype=np.uint8).copy()


VEC_FIELDS = ("faceNorm", "faceCenter", "faceXpac", "faceXnac", "cellCenter")


def structgen(nx, ny, nz, bounds=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), precision=64):
    """The reference's own StructGen mesh (MeshConfig + MeshSourceStructGenConfig + naive partitioner) built
    on two ranks, merged back into build-label order: {field: array}, 3-vectors as (n, 3), cell and
    boundary references as labels. Each element comes from the rank that owns its cell (a face from the
    rank owning its cell1); elements both ranks hold must agree bit for bit, which is checked."""
    recs = {"mode": _text("structgen"), "precision": np.array([precision], np.int32),
            "s.nx": np.array([nx], np.int32), "s.ny": np.array([ny], np.int32), "s.nz": np.array([nz], np.int32)}
    for k, v in zip(("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"), bounds):
        recs["s." + k] = np.array([float(v)])
    return _merge(run_case(recs), precision, base=0)


def _merge(parts, precision, base):
    for r in parts:
        rc = int(r["rc"][0])
        if rc != 95:
            raise RuntimeError(f"reference mesh build returned eCodes {rc}")
    dt = np.float64 if precision == 64 else np.float32
    for r in parts:  # build labels -> 0-based ids
        for k in ("label.cell", "label.face", "label.bnd"):
            r[k] = r[k] - base
    C = max(int(r["label.cell"].max()) + 1 for r in parts)
    F = max(int(r["label.face"].max()) + 1 for r in parts)
    B = max(int(r["label.bnd"].max()) + 1 if r["label.bnd"].size else 0 for r in parts)
    out = {"faceCell1": np.full(F, -2, np.int32), "faceCell2": np.full(F, -2, np.int32),
           "faceBnd": np.full(F, -2, np.int32), "faceLambda": np.zeros(F, dt), "faceRLencos": np.zeros(F, dt),
           "faceArea": np.zeros(F, dt), "cellVolume": np.zeros(C, dt), "bndFace": np.full(B, -2, np.int32),
           "bndRegion": np.full(B, -2, np.int32), "bndDistance": np.zeros(B, dt)}
    for k in VEC_FIELDS:
        out[k] = np.zeros((C if k == "cellCenter" else F, 3), dt)
    seen_c, seen_f, seen_b = np.zeros(C, bool), np.zeros(F, bool), np.zeros(B, bool)
    copies = []  # faces a rank holds through a ghost cell1: checked against the owner's copy
    for r in parts:
        nown = int(r["counts"][4])
        cl, fl, bl = r["label.cell"], r["label.face"], r["label.bnd"]
        own = np.arange(cl.size) < nown
        cc = r["mesh.cellCenter"].reshape(-1, 3)
        out["cellCenter"][cl[own]] = cc[own]
        out["cellVolume"][cl[own]] = r["mesh.cellVolume"][own]
        seen_c[cl[own]] = True
        c1 = r["mesh.faceCell1"]
        fown = own[c1]
        lab = fl[fown]
        c2 = r["mesh.faceCell2"]
        fb = r["mesh.faceBnd"]
        vals = {"faceCell1": cl[c1], "faceCell2": np.where(c2 >= 0, cl[np.maximum(c2, 0)], -1),
                "faceBnd": np.where(fb >= 0, bl[np.maximum(fb, 0)], -1)}
        for k in ("faceLambda", "faceRLencos", "faceArea"):
            vals[k] = r["mesh." + k]
        for k in VEC_FIELDS[:4]:
            vals[k] = r["mesh." + k].reshape(-1, 3)
        for k, v in vals.items():
            prev = seen_f[lab]
            if prev.any():  # a face both ranks hold as owned-cell1 cannot happen; keep the check anyway
                assert np.array_equal(out[k][lab[prev]], v[fown][prev]), k
            out[k][lab] = v[fown]
        seen_f[lab] = True
        bown = fb[fown] >= 0
        bl_own = bl[fb[fown][bown]]
        out["bndFace"][bl_own] = lab[bown]
        out["bndRegion"][bl_own] = r["mesh.bndRegion"][fb[fown][bown]]
        out["bndDistance"][bl_own] = r["mesh.bndDistance"][fb[fown][bown]]
        seen_b[bl_own] = True
        other = ~fown
        copies.append((fl[other], {k: v[other] for k, v in vals.items()}))
    assert seen_c.all() and seen_f.all() and seen_b.all(), "a rank did not report every element"
    for lab, vals in copies:
        for k, v in vals.items():
            assert np.array_equal(out[k][lab], v), f"ranks disagree on {k}"
    out["regType"] = parts[0]["mesh.regType"]
    out["counts"] = np.array([C, F, B, int(parts[0]["counts"][3])])
    return out


class RefMesh:
    """A mesh given as arrays (``host.a`` of a HostMesh, or any object with ``a`` and counts), in precision 64
    or 32, with its own copy of the mutable state, like pyoracle.OracleMesh."""

    def __init__(self, host, precision=64):
        self.precision = precision
        self.dtype = np.float64 if precision == 64 else np.float32
        self.nCells, self.nFaces, self.nBoundaries, self.nRegions = (host.nCells, host.nFaces, host.nBoundaries,
                                                                    host.nRegions)
        self.arr = {}
        for k in MESH_ARRAYS + list(STATE_KEYS):
            a = np.asarray(host.a[k])
            if k in INT_KEYS:
                a = np.ascontiguousarray(a, dtype=np.int32).copy()
            elif k in BOOL_KEYS:
                a = np.ascontiguousarray(a, dtype=np.uint8).copy()
            else:
                a = np.ascontiguousarray(a, dtype=self.dtype).copy()
            self.arr[k] = a

    def state(self):
        return {k: self.arr[k].copy() for k in STATE_KEYS}

    def set_state(self, st):
        for k, v in st.items():
            self.arr[k][...] = v

    def call(self, kernel, arrays, scalars):
        """Run ``kernel`` on this mesh. ``arrays``: name -> numpy array (updated in place from the result);
        ``scalars``: name -> float or int. Returns the reference's eCode."""
        recs = {"mode": _text("kernel"), "precision": np.array([self.precision], np.int32),
                "kernel": _text(kernel),
                "s.nCells": np.array([self.nCells], np.int32), "s.nFaces": np.array([self.nFaces], np.int32),
                "s.nBoundaries": np.array([self.nBoundaries], np.int32),
                "s.nRegions": np.array([self.nRegions], np.int32)}
        for k, a in self.arr.items():
            recs["mesh." + k] = a.reshape(-1)
        pads = {}
        for k, a in arrays.items():
            if a is None:
                continue
            want = np.int32 if k == "counters" else self.dtype
            assert a.dtype == want, (k, a.dtype)
            flat = a.reshape(-1)
            per = 3 if a.ndim == 2 else 1
            pad = 2 * per if k in CELL_ARGS else (1 if k in FACE_ARGS else (2 if k in FACE2_ARGS else 0))
            if pad:
                fill = np.zeros(pad, self.dtype) if k in FACE_ARGS or k in FACE2_ARGS else np.ones(pad, self.dtype)
                flat = np.concatenate([flat, fill])
            pads[k] = pad
            recs["a." + k] = flat
        for k, v in scalars.items():
            recs["s." + k] = np.array([v], np.int32) if isinstance(v, (bool, int, np.integer)) else \
                np.array([float(v)])
        res = run_case(recs)
        for k, a in arrays.items():
            if a is not None:
                r = res["a." + k]
                a[...] = (r[:r.size - pads[k]] if pads[k] else r).reshape(a.shape)
        for k in STATE_KEYS:
            self.arr[k][...] = res["mesh." + k].reshape(self.arr[k].shape)
        return int(res["rc"][0])


# ---------------------------------------------------------------------------
# The reference's kernel signatures (include/fvm/implementation/component/*.h),
# as pyoracle and cupcfd_hip.fvm spell them.
# ---------------------------------------------------------------------------
def GradientPhiGaussDolfyn(m, nGradient, phiCell, phiBoundary, dPhidxCell, dPhidxoCell):
    return m.call("gradient", dict(phiCell=phiCell, phiBoundary=phiBoundary, dPhidxCell=dPhidxCell,
                                   dPhidxoCell=dPhidxoCell), dict(nGradient=int(nGradient)))


def FluxMassDolfynFaceLoop(m, dudx, dvdx, dwdx, dpdx, denCell, denBoundary, uCell, vCell, wCell, massFlux, p, ar, su,
                           rface, small, icinl, icout, icsym, icwal, solveTurbEnergy, solveTurbDiss, solveVisc,
                           solveEnthalpy, teCell, teBoundary, edCell, edBoundary, viseffCell, viseffBoundary, tCell,
                           tBoundary):
    cnt = np.array([icinl[0], icout[0], icsym[0], icwal[0]], np.int32)
    rc = m.call("mass", dict(dudx=dudx, dvdx=dvdx, dwdx=dwdx, dpdx=dpdx, denCell=denCell, denBoundary=denBoundary,
                             uCell=uCell, vCell=vCell, wCell=wCell, massFlux=massFlux, p=p, ar=ar, su=su,
                             rface=rface, counters=cnt, teCell=teCell, teBoundary=teBoundary, edCell=edCell,
                             edBoundary=edBoundary, viseffCell=viseffCell, viseffBoundary=viseffBoundary,
                             tCell=tCell, tBoundary=tBoundary),
                dict(small=small, solveTurbEnergy=int(solveTurbEnergy), solveTurbDiss=int(solveTurbDiss),
                     solveVisc=int(solveVisc), solveEnthalpy=int(solveEnthalpy)))
    for i, c in enumerate((icinl, icout, icsym, icwal)):
        c[0] = cnt[i]
    return rc


def FluxMassDolfynBoundaryLoop1(m, massFlux, flowin):
    return m.call("mass_bl1", dict(massFlux=massFlux, flowin=flowin), {})


def FluxMassDolfynBoundaryLoop2(m, massFlux, flowRegion, flowout):
    return m.call("mass_bl2", dict(massFlux=massFlux, flowRegion=flowRegion, flowout=flowout), {})


def FluxMassDolfynBoundaryLoop3(m, flowin, ratearea):
    return m.call("mass_bl3", dict(ratearea=ratearea), dict(flowin=flowin))


def FluxMassDolfynBoundaryLoop4(m, massFlux, uBoundary, vBoundary, wBoundary, denBoundary, ratearea, flowout):
    return m.call("mass_bl4", dict(massFlux=massFlux, uBoundary=uBoundary, vBoundary=vBoundary, wBoundary=wBoundary,
                                   denBoundary=denBoundary, flowout=flowout), dict(ratearea=ratearea))


def FluxMassDolfynBoundaryLoop5(m, massFlux, su, uBoundary, vBoundary, wBoundary, fact, solveU, solveV, solveW,
                                flowFact, flowout2):
    return m.call("mass_bl5", dict(massFlux=massFlux, su=su, uBoundary=uBoundary, vBoundary=vBoundary,
                                   wBoundary=wBoundary, flowFact=flowFact, flowout2=flowout2),
                  dict(fact=fact, solveU=int(solveU), solveV=int(solveV), solveW=int(solveW)))


def FluxMassDolfynRegionLoop(m, flowFact, flowRegion, flowIn):
    return m.call("mass_region", dict(flowFact=flowFact, flowRegion=flowRegion), dict(flowIn=flowIn))


def FluxScalarDolfynFaceLoop(m, PhiCell, PhiBoundary, VisEff, Au, Su, Den, MassFlux, TE, CpBoundary, visEffBoundary,
                             RFace, dPhidx, SolveTurb, SolveEnthalpy, vislam, ivar, VarT, Sigma_T, Prandtl, VarTE,
                             Sigma_k, VarED, Sigma_e, Sigma_s, Schmidt, GammaBlend, Small, Large, TMCmu):
    return m.call("scalar", dict(PhiCell=PhiCell, PhiBoundary=PhiBoundary, VisEff=VisEff, Au=Au, Su=Su, Den=Den,
                                 MassFlux=MassFlux, TE=TE, CpBoundary=CpBoundary, visEffBoundary=visEffBoundary,
                                 RFace=RFace, dPhidx=dPhidx),
                  dict(SolveTurb=int(SolveTurb), SolveEnthalpy=int(SolveEnthalpy), vislam=vislam, ivar=int(ivar),
                       VarT=int(VarT), Sigma_T=Sigma_T, Prandtl=Prandtl, VarTE=int(VarTE), Sigma_k=Sigma_k,
                       VarED=int(VarED), Sigma_e=Sigma_e, Sigma_s=Sigma_s, Schmidt=Schmidt, GammaBlend=GammaBlend,
                       Small=Small, Large=Large, TMCmu=TMCmu))


def FluxUVWDolfynFaceLoop1(m, gammaBlend, small, uCell, vCell, wCell, uBoundary, vBoundary, wBoundary, visEffCell,
                           visEffBoundary, massFlux, dudx, dvdx, dwdx, rFace, su, sv, sw, au, av, aw):
    return m.call("uvw", dict(uCell=uCell, vCell=vCell, wCell=wCell, uBoundary=uBoundary, vBoundary=vBoundary,
                              wBoundary=wBoundary, visEffCell=visEffCell, visEffBoundary=visEffBoundary,
                              massFlux=massFlux, dudx=dudx, dvdx=dvdx, dwdx=dwdx, rFace=rFace, su=su, sv=sv, sw=sw,
                              au=au, av=av, aw=aw), dict(gammaBlend=gammaBlend, small=small))


def FluxUVWDolfynRegionLoop1(m):
    return m.call("uvw_region1", {}, {})


def FluxUVWDolfynBndsLoop1(m):
    return m.call("uvw_bnds1", {}, {})


def calculateViscosityDolfynCellLoop1(m, small, cmu, visURF, visLam, TE, ED, den, visEff):
    return m.call("visc_cell1", dict(TE=TE, ED=ED, den=den, visEff=visEff),
                  dict(small=small, cmu=cmu, visURF=visURF, visLam=visLam))


def calculateViscosityDolfynRegionLoop(kappa, m):
    return m.call("visc_region", {}, dict(kappa=kappa))


def calculateViscosityDolfynBoundaryLoop(m, tmcmu, large, small, kappa, visLam, TE, den, visEffCell, visEffBoundary):
    return m.call("visc_boundary", dict(TE=TE, den=den, visEffCell=visEffCell, visEffBoundary=visEffBoundary),
                  dict(tmcmu=tmcmu, large=large, small=small, kappa=kappa, visLam=visLam))


def calculateViscosityDolfynCellLoop2(m, visLam, visEffCell, visEffBoundary):
    return m.call("visc_cell2", dict(visEffCell=visEffCell, visEffBoundary=visEffBoundary), dict(visLam=visLam))

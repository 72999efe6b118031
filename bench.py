"""Benchmark of the finite-volume hot path on one MI355X.

A step is one sweep of CUP-CFD's four face-loop kernels — GradientPhiGaussDolfyn,
FluxMassDolfynFaceLoop, FluxScalarDolfynFaceLoop, FluxUVWDolfynFaceLoop1 —
over a StructGen 128^3 hex mesh (2,097,152 cells, 6,340,608 faces) whose cell
and face ids are relabelled by a seeded permutation, in FP64, with the
inputs already resident in HBM (resident C ABI, internal numbering).

value = cells / (step time) / 1e9 [Gcell-sweeps/s]. Per kernel the line also
reports Gcell-sweeps/s and the algorithmic-byte bandwidth (bytes each
reference statement must move, counted once) against the 8 TB/s HBM peak;
the dominant kernel's figure is the `roofline` object, timed with HIP events
on the stream the kernels are launched on.

cpu_baseline: the sequential C restatement of the same kernels
(oracle/liboracle_fvm.so, -O2, no FMA, one core) timed on this host over a
bounded sample of the same workload.

Further fields of the same line (each skippable):
  batched  K = 64 scalars through gradient + scalar flux on a relabelled 96^3
           mesh: one batched call of each against K single-scalar calls;
  adjoint  gradient + scalar flux forward against forward + adjoint w.r.t.
           (phi, massFlux) on a relabelled 64^3 mesh;
  tets     the tetrahedral stress mesh (100^3 hexes x 6 jittered tets,
           relabelled): the plain sweep, batched K = 16 and forward + adjoint.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
       [--no-cpu] [--no-batched] [--no-adjoint] [--no-tets]
       [--batched-only] [--tets-only] [--schedule S]
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("cup-cfd_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(C, F, Fi, B, s):
    """Bytes each kernel's reference statement reads and writes, each array
    element counted once (read-modify-write twice); v = 3s; indices 4 B
    (faceCell1 + faceBoundaryID per face, faceCell2 per interior face)."""
    v = 3 * s
    idx = 8 * F + 4 * Fi
    return {
        # phi, vol, phiB in; lambda, S; dPhidx and dPhidxo out (nGradient = 1:
        # dPhidxo and cell centres are structurally zero and not read)
        "gradient": idx + s * Fi + v * F + 2 * s * C + s * B + 2 * v * C,
        # Xf, S, xpac, xnac, lambda, area per interior face; massFlux, 2 rface out;
        # 4 gradients, centre, rho, u, v, w, p, ar, vol per cell
        "mass": idx + 4 * B + Fi * (s + 4 * v + s + s + 2 * s) + C * (4 * v + v + 6 * s + s),
        # lambda, rlencos, massFlux, S per interior face; 2 RFace out;
        # phi, VisEff, dPhidx, centre per cell; Su read-modify-write
        "scalar": idx + 4 * B + Fi * (3 * s + v + 2 * s) + C * (2 * s + 2 * v + 2 * s),
        # lambda, rlencos, massFlux, S per interior face; 2 rFace out; du, dv, dw,
        # centre, u, v, w, visEff per cell; su, sv, sw read-modify-write
        "uvw": idx + 4 * B + Fi * (3 * s + v + 2 * s) + C * (3 * v + v + 4 * s + 6 * s),
    }


def quiesce():
    """Collect Python garbage now, before a timed region. main() keeps the collector off otherwise: a
    generation-2 collection in this process takes 20-50 ms, and one that lands inside a timed region leaves
    the GPU idle behind the host, which the region's events then count as kernel time."""
    gc.collect()


def stream_copy_gbs(torch, nbytes=1 << 30, reps=10):
    quiesce()
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    return 2 * nbytes / t / 1e9


def stream_read_gbs(torch, nbytes=1 << 30, reps=10):
    """Read-only streaming rate: a sum over nbytes (torch's reduction, one read of every byte)."""
    quiesce()
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    a.sum()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        a.sum()
    e1.record()
    torch.cuda.synchronize()
    return nbytes / (e0.elapsed_time(e1) / 1e3 / reps) / 1e9


def oracle_sweeps(host, fields, budget_s, max_reps=5):
    """Median seconds per kernel of the oracle's gradient+mass+scalar+uvw sweep on ``host``, repeated until
    about budget_s seconds of CPU work or max_reps sweeps; returns (per-kernel seconds, sweeps run)."""
    import pyoracle as po
    from kernel_runs import run

    m = po.OracleMesh(host, 64)
    times = {}
    reps = 0
    t_start = time.perf_counter()
    while True:
        for k in ("gradient", "mass", "scalar", "uvw"):
            f = fields.copy()
            t0 = time.perf_counter()
            run(k, po, m, f, lambda: {})
            times.setdefault(k, []).append(time.perf_counter() - t0)
        reps += 1
        if time.perf_counter() - t_start > budget_s or reps >= max_reps:
            break
    return {k: float(np.median(v)) for k, v in times.items()}, reps


def cpu_baseline(host, fields, budget_s=20.0):
    """Time the oracle (one core) on the same mesh and inputs: one sweep of
    the four kernels, repeated until about budget_s seconds of CPU work."""
    per, reps = oracle_sweeps(host, fields, budget_s)
    t_sweep = sum(per.values())
    return {
        "value": host.nCells / t_sweep / 1e9,
        "unit": "Gcell-sweeps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{reps} sweep(s) of gradient+mass+scalar+uvw on the same {host.params['nx']}^3 relabelled "
                  f"mesh, median per kernel (includes per-call input copies outside the timed calls)",
        "per_kernel_s": per,
    }


def cpu_worker(spec):
    """One partition of the all-cores CPU baseline, run in a process of its own (no torch, no GPU): the
    oracle sweep on a relabelled StructGen mesh of nx*ny*nz cells; prints its cells and median sweep time."""
    import cupcfd_hip as ch
    from cupcfd_hip import workload as W

    nx, ny, nz, seed, budget = spec.split(",")
    host = ch.HostMesh(int(nx), int(ny), int(nz), seed=int(seed), permute_cells=True, permute_faces=True)
    fields = W.Fields(host, seed=777)
    print("ready", flush=True)
    sys.stdin.readline()  # every partition starts sweeping at the same time
    per, reps = oracle_sweeps(host, fields, float(budget), max_reps=1000)
    print(json.dumps({"cells": host.nCells, "sweep_s": sum(per.values()), "reps": reps}), flush=True)


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_all_cores(n, budget_s):
    """The oracle on every host core this process may use (at most 16), one partition per core as the
    reference runs one MPI rank per core: the n^3 grid is cut into P slabs of n*n*(n/P) cells, each swept
    by its own process at the same time. Aggregate rate = all cells / the slowest partition's sweep."""
    import subprocess

    P = max(1, min(len(os.sched_getaffinity(0)), 16))
    nz = max(1, round(n / P))
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               f"{n},{n},{nz},{1000 + i},{budget_s}"], stdin=subprocess.PIPE,
                              stdout=subprocess.PIPE, text=True)
             for i in range(P)]
    for pr in procs:
        if pr.stdout.readline().strip() != "ready":
            raise RuntimeError("CPU baseline worker failed to start")
    for pr in procs:
        pr.stdin.write("go\n")
        pr.stdin.flush()
    res = []
    for pr in procs:
        out, _ = pr.communicate(timeout=600)
        if pr.returncode != 0:
            raise RuntimeError(f"CPU baseline worker exited with {pr.returncode}")
        res.append(json.loads(out.strip().splitlines()[-1]))
    cells = sum(r["cells"] for r in res)
    slowest = max(r["sweep_s"] for r in res)
    return {
        "value": cells / slowest / 1e9,
        "unit": "Gcell-sweeps/s",
        "cores": P,
        "kind": "port",
        "sample": f"{P} concurrent processes, each sweeping gradient+mass+scalar+uvw on its own relabelled "
                  f"{n}x{n}x{nz} StructGen partition ({cells} cells in all), all started together and sweeping for "
                  f"about {budget_s:g} s, median per kernel",
        "slowest_partition_sweep_s": slowest,
    }


def oracle_outputs(host, f, kernels=("gradient", "mass", "scalar", "uvw")):
    """The oracle's outputs of each kernel on fresh copies of ``f`` (mesh numbering), FP64."""
    import pyoracle as po
    from kernel_runs import run

    m = po.OracleMesh(host, 64)
    return {k: run(k, po, m, f.copy(), lambda: {})[1] for k in kernels}


def parity_vs_oracle(dev, f, oracle):
    """Each kernel once through the drop-in C ABI on fresh copies of ``f`` against the oracle's outputs on the
    same inputs: per output, bit identity and the max-norm relative difference. Also the wall time of each
    drop-in call (host arrays in and out over PCIe, synchronous): the reference's calling convention."""
    from cupcfd_hip import fvm
    from kernel_runs import compare, run

    res, all_bits, wall = {}, True, {}
    run("gradient", fvm, dev, f.copy(), lambda: {})  # first call: the host staging's one-time set-up
    for k, o in oracle.items():
        fk = f.copy()
        t0 = time.perf_counter()
        _, g = run(k, fvm, dev, fk, lambda: {})
        wall[k] = (time.perf_counter() - t0) * 1e3
        res[k] = {}
        for name, want in o.items():
            a, b = np.ascontiguousarray(g[name]), np.ascontiguousarray(want)
            bits = a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))
            res[k][name] = {"bit_identical": bool(bits), "max_rel": compare(a, b)[1]}
            all_bits = all_bits and bits
    return {"all_bit_identical": bool(all_bits), "kernels": res, "dropin_ms": wall,
            "dropin_sweep_ms": sum(wall.values())}


def bench_config0(torch, iters=100):
    """The reference's CPU benchmark case: StructGen 32^3 over [0,1]^3 (one DEFAULT region), FP64,
    ``iters`` iterations of GradientPhiGaussDolfyn + FluxMassDolfynFaceLoop. Timed on the GPU resident
    (HIP events around all iterations), through the drop-in C ABI (host arrays, synchronous: the
    reference's calling convention) and on one host core with the oracle."""
    import cupcfd_hip as ch
    import pyoracle as po
    from cupcfd_hip import fvm
    from cupcfd_hip import workload as W
    from cupcfd_hip.fvm import resident as R

    host = ch.HostMesh(32)
    dev = ch.DeviceMesh(host, 64, reorder=True)
    f = W.Fields(host, seed=0)
    C, F = host.nCells, host.nFaces
    kernels, d = make_sweep(torch, host, dev, f, arrays=True)
    k = dict(kernels)
    stream = torch.cuda.current_stream()

    def step():
        k["gradient"]()
        k["mass"]()

    quiesce()
    step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(iters):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    t_gpu = e0.elapsed_time(e1) / 1e3
    # drop-in: host arrays, one synchronous call each
    cnt = np.zeros(4, np.int32)
    c = [cnt[i:i + 1] for i in range(4)]
    g, go = np.zeros((C, 3)), np.zeros((C, 3))
    mf, rface, su = f.massFlux.copy(), np.zeros(2 * F), f.su.copy()
    q = W.MASS

    def dropin(api, mesh):
        api.GradientPhiGaussDolfyn(mesh, 1, f.phi, f.phiB, g, go)
        api.FluxMassDolfynFaceLoop(mesh, f.dudx, f.dvdx, f.dwdx, f.dpdx, f.den, f.denB, f.u, f.v, f.w, mf, f.p, f.ar,
                                   su, rface, q["small"], c[0], c[1], c[2], c[3], 1, 1, 1, 1, f.te, f.teB, f.ed,
                                   f.edB, f.visEff, f.visEffB, f.t, f.tB)

    quiesce()
    dropin(fvm, dev)
    t0 = time.perf_counter()
    for _ in range(iters):
        dropin(fvm, dev)
    t_drop = time.perf_counter() - t0
    om = po.OracleMesh(host, 64)
    dropin(po, om)
    t0 = time.perf_counter()
    for _ in range(iters):
        dropin(po, om)
    t_cpu = time.perf_counter() - t0
    rate = lambda t: C * iters / t / 1e9  # noqa: E731
    return {"workload": f"StructGen 32^3 over [0,1]^3, one DEFAULT region, FP64, {iters} x (gradient + mass flux)",
            "cells": C, "iterations": iters,
            "gpu_resident_ms": t_gpu * 1e3, "gpu_resident_gcell_iters_per_s": rate(t_gpu),
            "gpu_dropin_ms": t_drop * 1e3, "gpu_dropin_gcell_iters_per_s": rate(t_drop),
            "cpu_oracle_1core_ms": t_cpu * 1e3, "cpu_oracle_1core_gcell_iters_per_s": rate(t_cpu),
            "note": "32^3 fits the 256 MB Infinity Cache: a latency point, not an HBM point"}


def bench_norms(torch, dev, fields, steps):
    """Residual norms (max |x|, sqrt(sum x^2), FP64) of the sweep's source terms, timed over ``steps``
    calls; algorithmic bytes = one read of each field."""
    from cupcfd_hip.fvm import resident as R

    stream = torch.cuda.current_stream()
    out = torch.zeros((len(fields), 2), dtype=torch.float64, device="cuda")
    R.residual_norms(dev, fields, out, stream)
    torch.cuda.synchronize()
    quiesce()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        R.residual_norms(dev, fields, out, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / steps
    nbytes = sum(x.numel() * x.element_size() for x in fields)
    return {"fields": ["scalar Su", "UVW su", "UVW sv", "UVW sw"], "us": t * 1e6, "algorithmic_bytes": nbytes,
            "hbm_frac": nbytes / t / 1e9 / HBM_PEAK_GBS, "values": out.cpu().numpy().tolist()}


def batched_fields(host, K, seed):
    """K scalar fields phi_k = sum_{m=1..8} a_m sin(k_m . x + theta_m) + 0.05 U(-1,1) at cell centres (and
    boundary face centres), with seeded wave vectors, phases and amplitudes; returned scalar-innermost."""
    rng = np.random.default_rng(seed)
    cc = host.a["cellCenter"]
    fc = host.a["faceCenter"][host.a["bndFace"]]
    phi = np.empty((host.nCells, K))
    phiB = np.empty((host.nBoundaries, K))
    for k in range(K):
        kv = rng.uniform(-4 * np.pi, 4 * np.pi, (8, 3))
        th = rng.uniform(0, 2 * np.pi, 8)
        am = rng.uniform(0.1, 1.0, 8)
        phi[:, k] = np.sin(cc @ kv.T + th) @ am + 0.05 * rng.uniform(-1, 1, host.nCells)
        phiB[:, k] = np.sin(fc @ kv.T + th) @ am + 0.05 * rng.uniform(-1, 1, host.nBoundaries)
    return phi, phiB


def make_sweep(torch, host, dev, f, arrays=False):
    """The four face-loop kernels of one sweep on resident copies of ``f`` (internal numbering, in the
    precision of ``dev``), as (name, callable) pairs launching on the current stream; with ``arrays`` also
    the dict of device arrays they read and write."""
    from cupcfd_hip import workload as W
    from cupcfd_hip.fvm import resident as R

    C, F = host.nCells, host.nFaces
    co, fo = dev.cell_order, dev.face_order
    cuda = torch.device("cuda:0")
    dt = torch.float64 if dev.precision == 64 else torch.float32

    def cells(a):
        return torch.from_numpy(np.ascontiguousarray(a[co])).to(cuda, dt)

    def faces(a, k=1):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(F, k)[fo].ravel())).to(cuda, dt)

    def bnd(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(cuda, dt)

    d = dict(phi=cells(f.phi), phiB=bnd(f.phiB), u=cells(f.u), v=cells(f.v), w=cells(f.w), den=cells(f.den),
             p=cells(f.p), ar=cells(f.ar), visEff=cells(f.visEff), visEffB=bnd(f.visEffB),
             dudx=cells(f.dudx), dvdx=cells(f.dvdx), dwdx=cells(f.dwdx), dpdx=cells(f.dpdx), dphidx=cells(f.dphidx),
             mf=faces(f.massFlux), te=cells(f.te), ed=cells(f.ed), t=cells(f.t), teB=bnd(f.teB), edB=bnd(f.edB),
             tB=bnd(f.tB), denB=bnd(f.denB), cpB=bnd(f.cpB), uB=bnd(f.uB), vB=bnd(f.vB), wB=bnd(f.wB),
             su=cells(f.su), sv=cells(f.sv), sw=cells(f.sw), au=cells(f.au), av=cells(f.av), aw=cells(f.aw))
    d["g"] = torch.empty(C * 3, dtype=dt, device=cuda)
    d["go"] = torch.empty_like(d["g"])
    d["rface"] = torch.zeros(2 * F, dtype=dt, device=cuda)
    d["rface2"] = torch.zeros(2 * F, dtype=dt, device=cuda)
    d["rface3"] = torch.zeros(2 * F, dtype=dt, device=cuda)
    d["mf_out"] = d["mf"].clone()
    d["Su"] = d["su"].clone()
    d["Au"] = d["au"].clone()
    d["phiB2"] = d["phiB"].clone()
    stream = torch.cuda.current_stream()
    mass_flags = (W.MASS["solveTurbEnergy"], W.MASS["solveTurbDiss"], W.MASS["solveVisc"], W.MASS["solveEnthalpy"])

    def k_gradient():
        R.gradient(dev, 1, d["phi"], d["phiB"], d["g"], d["go"], stream)

    def k_mass():
        R.mass(dev, d["dudx"], d["dvdx"], d["dwdx"], d["dpdx"], d["den"], d["denB"], d["u"], d["v"], d["w"],
               d["mf_out"], d["p"], d["ar"], d["su"], d["rface"], flags=mass_flags, teC=d["te"], teB=d["teB"],
               edC=d["ed"], edB=d["edB"], visC=d["visEff"], visB=d["visEffB"], tC=d["t"], tB=d["tB"], stream=stream)

    def k_scalar():
        R.scalar(dev, d["phi"], d["phiB2"], d["visEff"], d["Au"], d["Su"], d["den"], d["mf"], d["te"], d["cpB"],
                 d["visEffB"], d["rface2"], d["dphidx"], W.SCALAR, stream)

    def k_uvw():
        R.uvw(dev, W.UVW["gammaBlend"], W.UVW["small"], d["u"], d["v"], d["w"], d["uB"], d["vB"], d["wB"],
              d["visEff"], d["visEffB"], d["mf"], d["dudx"], d["dvdx"], d["dwdx"], d["rface3"], d["su"], d["sv"],
              d["sw"], d["au"], d["av"], d["aw"], stream)

    kernels = [("gradient", k_gradient), ("mass", k_mass), ("scalar", k_scalar), ("uvw", k_uvw)]
    return (kernels, d) if arrays else kernels


# outputs of one sweep compared between precisions: (kernel, output, array name in make_sweep's dict)
SWEEP_OUTPUTS = [("gradient", "dPhidx", "g"), ("mass", "massFlux", "mf_out"), ("mass", "RFace", "rface"),
                 ("scalar", "Su", "Su"), ("scalar", "RFace", "rface2"), ("scalar", "PhiB", "phiB2"),
                 ("uvw", "su", "su"), ("uvw", "sv", "sv"), ("uvw", "sw", "sw"), ("uvw", "rFace", "rface3")]


# kernel_runs output name of each SWEEP_OUTPUTS entry (mesh numbering)
ORACLE_NAME = {("mass", "RFace"): "rface"}


def oracle_internal(dev, host, kern, name, a):
    """An oracle output (mesh numbering) of the sweep in the device mesh's internal order, flattened."""
    co, fo, F = dev.cell_order, dev.face_order, host.nFaces
    a = np.asarray(a)
    if kern == "mass" and name == "massFlux":
        return a[fo].ravel()
    if name in ("RFace", "rFace"):
        return a.reshape(F, 2)[fo].ravel()
    if name == "PhiB":
        return a.ravel()
    return a[co].ravel()


def bench_f32(torch, host, f, steps, warmup, schedule, oracle=None):
    """The same sweep in FP32 (T = float throughout: the reference's float instantiation, FP32 mesh and
    fields): per-kernel time and algorithmic-byte fraction of the HBM peak at s = 4, and the max-norm relative
    error of each output after one sweep against the FP64 oracle on the same inputs (``oracle``, outputs of
    oracle_outputs), or, without it, against the FP64 sweep on the GPU."""
    import cupcfd_hip as ch

    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    Fi = int((host.a["faceBnd"] < 0).sum())
    d32 = ch.DeviceMesh(host, 32, reorder=True)
    d64 = ch.DeviceMesh(host, 64, reorder=True)
    if schedule >= 0:
        d32.set_schedule(schedule)
        d64.set_schedule(schedule)
        sched = [schedule] * 3
    else:
        sched = d32.autotune()
        d64.autotune()
    k32, a32 = make_sweep(torch, host, d32, f, arrays=True)
    k64, a64 = make_sweep(torch, host, d64, f, arrays=True)
    for (_, k), (_, k2) in zip(k32, k64):
        k()
        k2()
    torch.cuda.synchronize()
    err = {}
    for kern, name, key in SWEEP_OUTPUTS:
        if oracle is not None:
            o = oracle[kern][ORACLE_NAME.get((kern, name), name)]
            ref = torch.from_numpy(np.ascontiguousarray(oracle_internal(d32, host, kern, name, o))).cuda()
        else:
            ref = a64[key]
        scale = ref.abs().max().item()
        err[f"{kern}.{name}"] = ((a32[key].double().ravel() - ref.ravel()).abs().max().item() / scale) \
            if scale > 0 else 0.0
    del k64, a64, d64
    per, step_s, _ = time_sweep(torch, k32, steps, warmup)
    ab = algorithmic_bytes(C, F, Fi, B, 4)
    kern = {}
    for name, t in per.items():
        gbs = ab[name] / t / 1e9
        kern[name] = {"ms": t * 1e3, "gcell_sweeps_per_s": C / t / 1e9, "algorithmic_bytes": ab[name],
                      "algorithmic_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
    return {"workload": f"the same sweep in FP32 (float mesh, fields and arithmetic), StructGen "
                        f"{host.params['nx']}^3 relabelled", "gcell_sweeps_per_s": C / step_s / 1e9,
            "ms_per_step": step_s * 1e3, "gather_schedules": dict(zip(("gradient", "scalar", "uvw"), sched)),
            "kernels": kern, "rel_err_vs_f64": err,
            "rel_err_reference": "FP64 oracle on the same inputs" if oracle is not None else "FP64 GPU sweep"}


def time_sweep(torch, kernels, steps, warmup):
    """Mean seconds per kernel and per sweep over ``steps`` timed sweeps (HIP events on the launch stream)."""
    quiesce()
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        for _, k in kernels:
            k()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(kernels) + 1)] for _ in range(steps)]
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    for s in range(steps):
        ev[s][0].record(stream)
        for j, (_, k) in enumerate(kernels):
            k()
            ev[s][j + 1].record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    per = {name: float(np.mean([ev[s][j].elapsed_time(ev[s][j + 1]) for s in range(steps)])) / 1e3
           for j, (name, _) in enumerate(kernels)}
    step_s = float(np.mean([ev[s][0].elapsed_time(ev[s][-1]) for s in range(steps)])) / 1e3
    return per, step_s, t_wall


def kernel_report(C, F, Fi, B, per):
    """Per-kernel time, Gcell-sweeps/s and algorithmic-byte bandwidth against the HBM peak."""
    ab = algorithmic_bytes(C, F, Fi, B, 8)
    out = {}
    for name, t in per.items():
        gbs = ab[name] / t / 1e9
        out[name] = {"ms": t * 1e3, "gcell_sweeps_per_s": C / t / 1e9, "algorithmic_bytes": ab[name],
                     "algorithmic_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
    return out


def loop_bytes(host, s):
    """Algorithmic bytes of the boundary, region and viscosity loops of one BenchmarkKernels repetition: the
    ids and values each touches per boundary of the region type(s) it acts on (4-byte ids), per region and
    per cell; read-modify-write counted twice, v = 3s."""
    v = 3 * s
    C, B, R = host.nCells, host.nBoundaries, host.nRegions
    bt = host.a["regType"][host.a["bndRegion"]]
    n_in, n_out, n_wall = (int((bt == t).sum()) for t in (1, 2, 4))
    return {
        "mass_bl1": n_in * (8 + s),                                     # INLET list, face id; massFlux
        "mass_bl2": n_out * (12 + s) + 2 * s * R,                       # + flowRegion read-modify-write
        "mass_bl3": n_out * (8 + s),                                    # area
        "mass_bl4": n_out * (12 + s + v + s + s + 3 * s) + s * R,       # area, S, denB; massFlux, uB/vB/wB out
        "mass_bl5": n_out * (12 + 2 * s + 6 * s + 2 * s + s),           # massFlux, uB/vB/wB, su RMW; flowFact
        "mass_region": R * (4 + 3 * s),
        "uvw_region1": R * v,
        "uvw_bnds1": n_wall * (8 + v) + 2 * v * R,
        "visc_cell1": 5 * s * C,                                        # TE, ED, den; visEff RMW
        "visc_region": R * (4 + 4 * s),
        "visc_boundary": B * (12 + 4 * s),                              # ids; cell TE, den, visEff; visEffB out
        "visc_cell2": 2 * s * (C + B),                                  # visEff, visEffB RMW
    }


def bench_all_kernels(torch, n, steps, warmup, schedule):
    """All 16 kernels of BenchmarkKernels::runBenchmark (BenchmarkKernels.cpp:55-115), resident, each timed on
    its own (HIP events around `steps` back-to-back calls) on a typed n^3 relabelled StructGen mesh, FP64.
    Typed (INLET x-low, OUTLET x-high, WALL y, SYMP z) so the boundary and region loops have work. The O(B)
    and O(R) loops move a few hundred KB and are latency-bound; their fractions are not bandwidth figures."""
    import cupcfd_hip as ch
    from cupcfd_hip import workload as W
    from cupcfd_hip.fvm import resident as R

    host = ch.HostMesh(n, seed=31, typed=True, permute_cells=True, permute_faces=True)
    dev = ch.DeviceMesh(host, 64, reorder=True)
    if schedule >= 0:
        dev.set_schedule(schedule)
    else:
        dev.autotune()
    f = W.Fields(host, seed=321)
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    Fi = int((host.a["faceBnd"] < 0).sum())
    sweep, d = make_sweep(torch, host, dev, f, arrays=True)
    cuda = torch.device("cuda:0")

    def one(x):
        return torch.full((1,), x, dtype=torch.float64, device=cuda)

    plain = {k: torch.from_numpy(np.ascontiguousarray(getattr(f, k))).to(cuda)
             for k in ("flowRegion", "flowFact")}
    fl = [one(0.0) for _ in range(4)]
    q, vc = W.MASS_LOOPS, W.VISC
    ve2 = d["visEff"].clone()
    vb2 = d["visEffB"].clone()
    calls = dict(sweep)
    calls.update({
        "mass_bl1": lambda: R.mass_bl1(dev, d["mf_out"], fl[0]),
        "mass_bl2": lambda: R.mass_bl2(dev, d["mf_out"], plain["flowRegion"], fl[1]),
        "mass_bl3": lambda: R.mass_bl3(dev, q["flowin"], fl[2]),
        "mass_bl4": lambda: R.mass_bl4(dev, d["mf_out"], d["uB"], d["vB"], d["wB"], d["denB"], q["ratearea"], fl[1]),
        "mass_bl5": lambda: R.mass_bl5(dev, d["mf_out"], d["su"], d["uB"], d["vB"], d["wB"], q["fact"], 1, 1, 1,
                                       plain["flowFact"], fl[3]),
        "mass_region": lambda: R.mass_region(dev, plain["flowFact"], plain["flowRegion"], 0.0),
        "uvw_region1": lambda: R.uvw_region1(dev),
        "uvw_bnds1": lambda: R.uvw_bnds1(dev),
        "visc_cell1": lambda: R.visc_cell1(dev, vc["cell1"]["small"], vc["cell1"]["cmu"], vc["cell1"]["visURF"],
                                           vc["cell1"]["visLam"], d["te"], d["ed"], d["den"], ve2),
        "visc_region": lambda: R.visc_region(dev, vc["region"]["kappa"]),
        "visc_boundary": lambda: R.visc_boundary(dev, vc["boundary"]["tmcmu"], vc["boundary"]["large"],
                                                 vc["boundary"]["small"], vc["boundary"]["kappa"],
                                                 vc["boundary"]["visLam"], d["te"], d["den"], ve2, vb2),
        "visc_cell2": lambda: R.visc_cell2(dev, vc["cell2"]["visLam"], ve2, vb2),
    })
    order = ["gradient", "mass", "mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5", "mass_region", "scalar",
             "uvw", "uvw_region1", "uvw_bnds1", "visc_cell1", "visc_region", "visc_boundary", "visc_cell2"]
    nbytes = dict(algorithmic_bytes(C, F, Fi, B, 8))
    nbytes.update(loop_bytes(host, 8))
    stream = torch.cuda.current_stream()
    quiesce()
    out = {}

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    for name in order:
        t = timed(calls[name])
        gbs = nbytes[name] / t / 1e9
        out[name] = {"us": t * 1e6, "algorithmic_bytes": nbytes[name], "algorithmic_GBps": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS, "bound": "hbm" if nbytes[name] > (32 << 20) else "latency"}
    # the loops with ordered sums again, their sums taken as a fixed-shape parallel tree (not the reference's
    # bits; the tests bound the difference)
    tree = {}
    dev.set_ordered_sums(False)
    try:
        for name in ("mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5", "uvw_bnds1"):
            tree[name] = {"us": timed(calls[name]) * 1e6, "sequential_us": out[name]["us"]}
    finally:
        dev.set_ordered_sums(True)
    return {"workload": f"the 16 BenchmarkKernels kernels, resident, one at a time, StructGen {n}^3 relabelled "
                        f"with INLET/OUTLET/WALL/SYMP regions, FP64", "cells": C, "boundaries": B,
            "repetition_us": sum(v["us"] for v in out.values()), "kernels": out,
            "tree_sums": {"repetition_us": sum(v["us"] for v in out.values())
                          + sum(v["us"] - v["sequential_us"] for v in tree.values()), "kernels": tree}}


def bench_batched(torch, host, dev, K, steps, warmup, label, oracle_sample=None):
    """K scalars through gradient + scalar flux on ``host`` (FP64): one batched call of each against K
    single-scalar calls on K separate arrays. The scalars share the mass flux of one FluxMassDolfynFaceLoop
    call. Returns per-scalar cell throughputs and their ratio. The FP32-storage errors are measured against
    the FP64 oracle on the scalars of ``oracle_sample`` (default 0, K/2, K-1)."""
    from cupcfd_hip import workload as W
    from cupcfd_hip.fvm import resident as R

    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    f = W.Fields(host, seed=99)
    co, fo = dev.cell_order, dev.face_order
    cuda = torch.device("cuda:0")
    dt = torch.float64

    def cells(a):
        return torch.from_numpy(np.ascontiguousarray(a[co])).to(cuda)

    # shared mass flux from one mass face loop
    mf = torch.from_numpy(np.ascontiguousarray(f.massFlux[fo])).to(cuda)
    su0 = cells(f.su)
    rf = torch.zeros(2 * F, dtype=dt, device=cuda)
    R.mass(dev, cells(f.dudx), cells(f.dvdx), cells(f.dwdx), cells(f.dpdx), cells(f.den),
           torch.from_numpy(f.denB).to(cuda), cells(f.u), cells(f.v), cells(f.w), mf, cells(f.p), cells(f.ar), su0, rf)
    phi, phiB = batched_fields(host, K, seed=2024)
    phi_b = cells(phi)                                   # (C, K)
    phiB_b = torch.from_numpy(phiB).to(cuda)             # (B, K)
    g_b = torch.zeros((C, 3, K), dtype=dt, device=cuda)
    go_b = torch.zeros_like(g_b)
    su_b = torch.zeros((C, K), dtype=dt, device=cuda)
    au_b = torch.ones((C, K), dtype=dt, device=cuda)
    pB_b = phiB_b.clone()
    sh = dict(visEff=cells(f.visEff), den=cells(f.den), te=cells(f.te),
              cpB=torch.from_numpy(f.cpB).to(cuda), visEffB=torch.from_numpy(f.visEffB).to(cuda))
    rface = torch.zeros(2 * F, dtype=dt, device=cuda)
    # K separate single-scalar arrays (no shared cache footprint between calls)
    one = [dict(phi=phi_b[:, k].contiguous(), phiB=phiB_b[:, k].contiguous(),
                g=torch.zeros((C, 3), dtype=dt, device=cuda), go=torch.zeros((C, 3), dtype=dt, device=cuda),
                su=torch.zeros(C, dtype=dt, device=cuda), au=torch.ones(C, dtype=dt, device=cuda),
                pB=phiB_b[:, k].contiguous()) for k in range(K)]
    stream = torch.cuda.current_stream()

    def run_batched():
        R.gradient_batched(dev, 1, K, phi_b, phiB_b, g_b, go_b, stream)
        R.scalar_batched(dev, K, phi_b, pB_b, sh["visEff"], au_b, su_b, sh["den"], mf, sh["te"], sh["cpB"],
                         sh["visEffB"], rface, g_b, W.SCALAR, stream)

    def run_plain():
        for o in one:
            R.gradient(dev, 1, o["phi"], o["phiB"], o["g"], o["go"], stream)
            R.scalar(dev, o["phi"], o["pB"], sh["visEff"], o["au"], o["su"], sh["den"], mf, sh["te"], sh["cpB"],
                     sh["visEffB"], rface, o["g"], W.SCALAR, stream)

    def timed(fn):
        quiesce()
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    tb = timed(run_batched)
    tp = timed(run_plain)
    # the batched results are those of the single-scalar calls, bit for bit (both ran the same number of
    # accumulating calls)
    same = all(torch.equal(g_b[:, :, k], one[k]["g"]) and torch.equal(su_b[:, k], one[k]["su"])
               for k in (0, K // 2, K - 1))
    tbg = timed(lambda: R.gradient_batched(dev, 1, K, phi_b, phiB_b, g_b, go_b, stream))
    # FP32 storage, FP64 arithmetic: time, and max-norm error against the FP64 gradient of the FP64 inputs
    phi32, phiB32 = phi_b.float(), phiB_b.float()
    g32 = torch.zeros((C, 3, K), dtype=torch.float32, device=cuda)
    go32 = torch.zeros_like(g32)
    tbg32 = timed(lambda: R.gradient_batched_f32s(dev, 1, K, phi32, phiB32, g32, go32, stream))
    R.gradient_batched(dev, 1, K, phi_b, phiB_b, g_b, go_b, stream)
    torch.cuda.synchronize()
    err32 = ((g32.double() - g_b).abs().max() / g_b.abs().max()).item()
    # batched adjoint w.r.t. (phi_k, massFlux), seeded N(0,1) cotangents
    rng2 = np.random.default_rng(99)
    sbar = torch.from_numpy(rng2.normal(size=(C, K))).to(cuda)
    rbar = torch.from_numpy(rng2.normal(size=2 * F)).to(cuda)
    pbar = torch.empty((C, K), dtype=dt, device=cuda)
    mbar = torch.empty(F, dtype=dt, device=cuda)
    tadj = timed(lambda: R.scalar_adjoint_batched(dev, K, phi_b, sh["visEff"], mf, sbar, rbar, W.SCALAR, pbar, mbar,
                                                  stream))
    # the same with FP32 storage of phi, SuBar and PhiBar: time, and max-norm error of PhiBar and MassFluxBar
    # against the FP64 adjoint of the FP64 inputs
    sbar32 = sbar.float()
    pbar32 = torch.empty((C, K), dtype=torch.float32, device=cuda)
    mbar32 = torch.empty(F, dtype=dt, device=cuda)
    tadj32 = timed(lambda: R.scalar_adjoint_batched_f32s(dev, K, phi32, sh["visEff"], mf, sbar32, rbar, W.SCALAR,
                                                         pbar32, mbar32, stream))
    torch.cuda.synchronize()
    err_adj32 = {"PhiBar": ((pbar32.double() - pbar).abs().max() / pbar.abs().max()).item(),
                 "MassFluxBar": ((mbar32 - mbar).abs().max() / mbar.abs().max()).item()}
    tbs = timed(lambda: R.scalar_batched(dev, K, phi_b, pB_b, sh["visEff"], au_b, su_b, sh["den"], mf, sh["te"],
                                         sh["cpB"], sh["visEffB"], rface, g_b, W.SCALAR, stream))
    # FP32-storage scalar flux (per-scalar arrays float, FP64 arithmetic) on the FP32-storage gradient: time, and
    # max-norm error of Su, Au and PhiB against the FP64 chain on the FP64 inputs, from the same starting arrays
    su32, au32, pB32 = su_b.float(), au_b.float(), phiB32.clone()
    tbs32 = timed(lambda: R.scalar_batched_f32s(dev, K, phi32, pB32, sh["visEff"], au32, su32, sh["den"], mf,
                                                sh["te"], sh["cpB"], sh["visEffB"], rface, g32, W.SCALAR, stream))
    R.gradient_batched(dev, 1, K, phi_b, phiB_b, g_b, go_b, stream)
    R.gradient_batched_f32s(dev, 1, K, phi32, phiB32, g32, go32, stream)
    s64 = dict(su=torch.zeros((C, K), dtype=dt, device=cuda), au=torch.ones((C, K), dtype=dt, device=cuda),
               pB=phiB_b.clone())
    s32 = {k: v.float() for k, v in s64.items()}
    R.scalar_batched(dev, K, phi_b, s64["pB"], sh["visEff"], s64["au"], s64["su"], sh["den"], mf, sh["te"],
                     sh["cpB"], sh["visEffB"], rface, g_b, W.SCALAR, stream)
    R.scalar_batched_f32s(dev, K, phi32, s32["pB"], sh["visEff"], s32["au"], s32["su"], sh["den"], mf, sh["te"],
                          sh["cpB"], sh["visEffB"], rface, g32, W.SCALAR, stream)
    torch.cuda.synchronize()
    err_s32 = {n: ((s32[n].double() - s64[n]).abs().max() / s64[n].abs().max()).item() for n in ("su", "au", "pB")}
    # the same FP32-storage errors against the FP64 oracle on a sample of scalars (mesh numbering)
    import pyoracle as po
    sample = sorted(set(oracle_sample if oracle_sample is not None else (0, K // 2, K - 1)))
    om = po.OracleMesh(host, 64)
    mf_mesh = np.empty(F)
    mf_mesh[fo] = mf.cpu().numpy()
    sb_h, rb_h = sbar.cpu().numpy(), rbar.cpu().numpy()
    rb_mesh = np.empty((F, 2))
    rb_mesh[fo] = rb_h.reshape(F, 2)
    g32h, su32h, au32h, pB32h = g32.cpu().numpy(), s32["su"].cpu().numpy(), s32["au"].cpu().numpy(), \
        s32["pB"].cpu().numpy()
    pbar32h = pbar32.cpu().numpy()
    oerr = {"gradient": 0.0, "Su": 0.0, "Au": 0.0, "PhiB": 0.0, "PhiBar": 0.0}

    def rel(a, b):
        scale = float(np.abs(b).max())
        return float(np.abs(np.asarray(a, np.float64) - b).max()) / scale if scale > 0 else 0.0

    p = W.SCALAR
    for k in sample:
        gk, gok = np.zeros((C, 3)), np.zeros((C, 3))
        po.GradientPhiGaussDolfyn(om, 1, phi[:, k].copy(), phiB[:, k].copy(), gk, gok)
        oerr["gradient"] = max(oerr["gradient"], rel(g32h[:, :, k], gk[co]))
        suk, auk, pbk = np.zeros(C), np.ones(C), phiB[:, k].copy()
        po.FluxScalarDolfynFaceLoop(om, phi[:, k].copy(), pbk, f.visEff.copy(), auk, suk, f.den.copy(), mf_mesh.copy(),
                                    f.te.copy(), f.cpB.copy(), f.visEffB.copy(), np.zeros(2 * F), gk, p["SolveTurb"],
                                    p["SolveEnthalpy"], p["vislam"], p["ivar"], p["VarT"], p["Sigma_T"], p["Prandtl"],
                                    p["VarTE"], p["Sigma_k"], p["VarED"], p["Sigma_e"], p["Sigma_s"], p["Schmidt"],
                                    p["GammaBlend"], p["Small"], p["Large"], p["TMCmu"])
        oerr["Su"] = max(oerr["Su"], rel(su32h[:, k], suk[co]))
        oerr["Au"] = max(oerr["Au"], rel(au32h[:, k], auk[co]))
        oerr["PhiB"] = max(oerr["PhiB"], rel(pB32h[:, k], pbk))
        sb_mesh = np.empty(C)
        sb_mesh[co] = sb_h[:, k]
        _, pbar_k, _ = po.scalar_adjoint(om, phi[:, k].copy(), f.visEff.copy(), mf_mesh.copy(), sb_mesh,
                                         rb_mesh.ravel().copy(), p)
        oerr["PhiBar"] = max(oerr["PhiBar"], rel(pbar32h[:, k], pbar_k[co]))
    return {
        "workload": f"gradient + scalar flux of K={K} scalars, {label}, FP64, shared mass flux",
        "cells": C, "K": K,
        "batched_ms": tb * 1e3, "plain_ms": tp * 1e3,
        "batched_gcell_scalar_per_s": C * K / tb / 1e9, "plain_gcell_scalar_per_s": C * K / tp / 1e9,
        "speedup": tp / tb, "bit_identical_sample": bool(same),
        "batched_gradient_ms": tbg * 1e3, "batched_scalar_ms": tbs * 1e3,
        "batched_gradient_f32_storage_ms": tbg32 * 1e3, "f32_storage_gradient_rel_err": err32,
        "batched_scalar_f32_storage_ms": tbs32 * 1e3,
        "f32_storage_scalar_rel_err": {"Su": err_s32["su"], "Au": err_s32["au"], "PhiB": err_s32["pB"]},
        "batched_adjoint_ms": tadj * 1e3, "batched_forward_adjoint_ratio": (tb + tadj) / tb,
        "batched_adjoint_f32_storage_ms": tadj32 * 1e3, "f32_storage_adjoint_rel_err": err_adj32,
        "f32_storage_rel_err_vs_oracle": {"scalars": sample, **oerr,
                                          "note": "FP32-storage chain (gradient, scalar flux, adjoint) against the "
                                                  "FP64 oracle on the same inputs, max over the sampled scalars; the "
                                                  "_rel_err fields above are against the FP64 batched GPU results, "
                                                  "which equal the oracle's bit for bit (tests)"},
    }


def bench_adjoint(torch, host, dev, steps, warmup, label, nG=1, full=False, params=None):
    """Forward (gradient with nG sweeps + scalar flux) against forward + adjoint w.r.t. (phi, massFlux) on
    ``host`` in FP64, seeded N(0,1) cotangents on Su and RFace. ``full`` times cupcfd_dev_scalar_adjoint_full_f64
    (any mesh, nG sweeps), else cupcfd_dev_scalar_adjoint_f64 (nG = 1, all-DEFAULT mesh)."""
    from cupcfd_hip import workload as W
    from cupcfd_hip.fvm import resident as R

    C, F = host.nCells, host.nFaces
    f = W.Fields(host, seed=5)
    co, fo = dev.cell_order, dev.face_order
    cuda = torch.device("cuda:0")
    dt = torch.float64

    def cells(a):
        return torch.from_numpy(np.ascontiguousarray(a[co])).to(cuda)

    phi, visEff, den, te = cells(f.phi), cells(f.visEff), cells(f.den), cells(f.te)
    mf = torch.from_numpy(np.ascontiguousarray(f.massFlux[fo])).to(cuda)
    phiB = torch.from_numpy(f.phiB).to(cuda)
    cpB, visEffB = torch.from_numpy(f.cpB).to(cuda), torch.from_numpy(f.visEffB).to(cuda)
    g = torch.empty(3 * C, dtype=dt, device=cuda)
    go = torch.empty_like(g)
    su = torch.zeros(C, dtype=dt, device=cuda)
    au = torch.zeros(C, dtype=dt, device=cuda)
    rface = torch.zeros(2 * F, dtype=dt, device=cuda)
    pB = phiB.clone()
    rng = np.random.default_rng(77)
    sb = torch.from_numpy(rng.normal(size=C)).to(cuda)
    rb = torch.from_numpy(rng.normal(size=2 * F)).to(cuda)
    pbar = torch.empty(C, dtype=dt, device=cuda)
    mbar = torch.empty(F, dtype=dt, device=cuda)
    stream = torch.cuda.current_stream()
    P = params or W.SCALAR

    def fwd():
        R.gradient(dev, nG, phi, phiB, g, go, stream)
        R.scalar(dev, phi, pB, visEff, au, su, den, mf, te, cpB, visEffB, rface, g, P, stream)

    def adj():
        if full:
            R.scalar_adjoint_full(dev, nG, phi, visEff, mf, g, te, den, cpB, visEffB, sb, rb, P, pbar, mbar, stream)
        else:
            R.scalar_adjoint(dev, phi, visEff, mf, sb, rb, P, pbar, mbar, stream)

    def timed(fn):
        quiesce()
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    tf = timed(fwd)
    tfa = timed(lambda: (fwd(), adj()))
    return {
        "workload": f"gradient (nGradient={nG}) + scalar flux, forward and reverse w.r.t. (phi, massFlux), {label}, "
                    f"FP64" + (", typed regions, enthalpy wall" if full else ", one DEFAULT region"),
        "cells": C, "forward_ms": tf * 1e3, "forward_adjoint_ms": tfa * 1e3, "ratio": tfa / tf,
    }


def partition_sizes(nparts, seed):
    """Grid sizes of ``nparts`` StructGen partitions with cell counts N log-uniform in [8^3, 64^3]:
    n_x = n_y = round(N^(1/3)), n_z = round(N / n_x^2), each at least 8."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nparts):
        N = float(np.exp(rng.uniform(np.log(8**3), np.log(64**3))))
        nx = max(8, int(round(N ** (1.0 / 3.0))))
        nz = max(8, int(round(N / (nx * nx))))
        out.append((nx, nx, nz))
    return out


def bench_partitions(torch, nparts, K, steps, warmup):
    """The batched gradient (K scalars, nGradient = 1, FP64) over many uneven relabelled StructGen
    partitions: one launch per partition, those launches captured in a CUDA/HIP graph, one persistent
    launch over the whole mesh set, and that launch in a graph."""
    import cupcfd_hip as ch
    from cupcfd_hip.fvm import resident as R

    sizes = partition_sizes(nparts, seed=48)
    hosts = [ch.HostMesh(*sz, seed=100 + i, permute_cells=True, permute_faces=True) for i, sz in enumerate(sizes)]
    devs = [ch.DeviceMesh(h, 64, reorder=True) for h in hosts]
    mset = ch.DeviceMeshSet(devs)
    cuda = torch.device("cuda:0")
    rng = np.random.default_rng(7)
    phi = [torch.from_numpy(rng.uniform(-1, 1, (h.nCells, K))).to(cuda) for h in hosts]
    phiB = [torch.from_numpy(rng.uniform(-1, 1, (h.nBoundaries, K))).to(cuda) for h in hosts]
    g = [torch.zeros((h.nCells, 3, K), dtype=torch.float64, device=cuda) for h in hosts]
    go = [torch.zeros_like(x) for x in g]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device=cuda)  # noqa: E731
    pp, pb, pg, pgo = ptr(phi), ptr(phiB), ptr(g), ptr(go)
    C = sum(h.nCells for h in hosts)

    def per_mesh():
        for i, d in enumerate(devs):
            R.gradient_batched(d, 1, K, phi[i], phiB[i], g[i], go[i])

    def persistent():
        R.gradient_batched_set(mset, K, pp, pb, pg, pgo)

    def timed(fn):
        quiesce()
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps, (time.perf_counter() - t0) / steps

    def graphed(fn):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        gr = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gr):
            fn()
        return gr.replay

    res = {}
    for name, fn in (("per_partition_launches", per_mesh), ("persistent_launch", persistent)):
        res[name + "_ms"], res[name + "_host_ms"] = (x * 1e3 for x in timed(fn))
        res[name + "_graph_ms"], res[name + "_graph_host_ms"] = (x * 1e3 for x in timed(graphed(fn)))
    ref = [x.clone() for x in g]
    persistent()
    torch.cuda.synchronize()
    same = all(torch.equal(a, b) for a, b in zip(ref, g))
    best = min(res["per_partition_launches_ms"], res["persistent_launch_ms"], res["persistent_launch_graph_ms"],
               res["per_partition_launches_graph_ms"])
    res.update({"workload": f"batched gradient, K={K} scalars, {nparts} relabelled StructGen partitions with cell "
                            f"counts log-uniform in [8^3, 64^3], FP64",
                "partitions": nparts, "cells": C, "K": K, "persistent_matches_per_partition": bool(same),
                "best_gcell_scalar_per_s": C * K / (best / 1e3) / 1e9})
    del mset
    return res


def structgen(n, seed):
    import cupcfd_hip as ch
    host = ch.HostMesh(n, seed=seed, permute_cells=True, permute_faces=True)
    return host, ch.DeviceMesh(host, 64, reorder=True), f"StructGen {n}^3 relabelled"


def bench_tets(torch, n, K, steps, warmup):
    """The tetrahedral stress mesh (n^3 hexes x 6 jittered tets, ids relabelled, one DEFAULT region), FP64:
    the plain sweep, the batched K-scalar comparison and forward + adjoint."""
    import cupcfd_hip as ch
    from cupcfd_hip import workload as W

    params = dict(seed=606, tets=True, jitter=0.2, permute_cells=True, permute_faces=True)
    t0 = time.time()
    host = ch.HostMesh(n, **params)
    t_gen = time.time() - t0
    dev = ch.DeviceMesh(host, 64, reorder=True)
    t_mesh = time.time() - t0
    # the same mesh generated and repacked on the GPU
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, dgen = ch.generate_mesh(n, host_copy=False, **params)
    t_dev = time.perf_counter() - t1
    mesh_build = {"host_generate_s": t_gen, "host_repack_s": t_mesh - t_gen, "device_generate_s": t_dev,
                  "device_generate_identical": dgen.digest() == dev.digest(), "speedup_setup": t_mesh / t_dev}
    dgen = None
    sched = dev.autotune()
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    f = W.Fields(host, seed=808)
    kernels = make_sweep(torch, host, dev, f)
    per, step_s, _ = time_sweep(torch, kernels, steps, warmup)
    del kernels
    torch.cuda.empty_cache()
    label = f"{n}^3 hexes x 6 jittered tets relabelled"
    out = {"workload": f"plain FP64 gradient+mass+scalar+UVW sweep, {label} ({C} cells, {F} faces), one DEFAULT "
                       f"region", "cells": C, "faces": F, "boundaries": B, "mesh_setup_s": t_mesh,
           "mesh_build": mesh_build, "gcell_sweeps_per_s": C / step_s / 1e9, "ms_per_step": step_s * 1e3,
           "gather_schedules": dict(zip(("gradient", "scalar", "uvw"), sched)),
           "kernels": kernel_report(C, F, F - B, B, per)}
    out["batched"] = bench_batched(torch, host, dev, K, steps=max(2, min(steps, 5)), warmup=1, label=label,
                                   oracle_sample=(0, K - 1))
    torch.cuda.empty_cache()
    out["adjoint"] = bench_adjoint(torch, host, dev, steps=max(3, min(steps, 10)), warmup=1, label=label)
    torch.cuda.empty_cache()
    return out


def bench_mesh_build(torch, host, dev_host, t_host_gen, t_host_repack, label, reps=3):
    """Mesh setup on the host against the GPU, for the same mesh:
      - host generator (cupcfd_structgen_create) + host repack (cupcfd_hip_mesh_create_f64);
      - GPU repack of the host mesh's view (cupcfd_hip_mesh_create_device_f64);
      - GPU generator + GPU repack (cupcfd_hip_mesh_generate_f64), with and without a host copy.
    Wall time of each call (uploads, downloads and host-side mirrors included, best of `reps`), and
    whether each device mesh is identical to the host-built one (cupcfd_hip_mesh_digest)."""
    import cupcfd_hip as ch

    ref = dev_host.digest()
    params = host.params
    dims = dict(nx=params["nx"], ny=params["ny"], nz=params["nz"])
    rest = {k: v for k, v in params.items() if k not in dims}

    def best(make):
        times, obj = [], None
        for _ in range(reps):
            obj = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            obj = make()
            times.append(time.perf_counter() - t0)
        return min(times), obj

    t_rep, dev = best(lambda: ch.DeviceMesh(host, 64, reorder=True, on_device=True))
    same_rep = dev.digest() == ref
    dev = None
    t_gen, (_, dev) = best(lambda: ch.generate_mesh(host_copy=False, **dims, **rest))
    same_gen = dev.digest() == ref
    dev = None
    t_gen_h, (hcopy, dev) = best(lambda: ch.generate_mesh(host_copy=True, **dims, **rest))
    same_host = all(hcopy.a[k].tobytes() == host.a[k].tobytes() for k in host.a)
    dev = hcopy = None
    return {"mesh": label, "cells": host.nCells, "faces": host.nFaces,
            "host_generate_s": t_host_gen, "host_repack_s": t_host_repack,
            "device_repack_s": t_rep, "device_repack_identical": same_rep,
            "device_generate_s": t_gen, "device_generate_identical": same_gen,
            "device_generate_with_host_copy_s": t_gen_h, "host_copy_identical": same_host,
            "speedup_repack": t_host_repack / t_rep,
            "speedup_setup": (t_host_gen + t_host_repack) / t_gen}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched K-scalar comparison")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched K-scalar comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint comparison")
    ap.add_argument("--no-tets", action="store_true", help="skip the tetrahedral stress mesh")
    ap.add_argument("--no-partitions", action="store_true", help="skip the many-partition batched launch")
    ap.add_argument("--no-mesh-build", action="store_true", help="skip the host vs GPU mesh repack timing")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 sweep")
    ap.add_argument("--no-all-kernels", action="store_true", help="skip the 16-kernel typed-mesh timing")
    ap.add_argument("--partitions-only", action="store_true", help="run only the many-partition batched launch")
    ap.add_argument("--tets-only", action="store_true", help="run only the tetrahedral stress mesh")
    ap.add_argument("--tets-n", type=int, default=100)
    ap.add_argument("--tets-k", type=int, default=16)
    ap.add_argument("--adjoint-n", type=int, default=128)
    ap.add_argument("--batch-n", type=int, default=96)
    ap.add_argument("--batch-k", type=int, default=64)
    ap.add_argument("--schedule", type=int, default=-1,
                    help="gather schedule: -1 autotuned per kernel (default), 0 automatic, 1 thread per cell, "
                         "4/6/8 lanes per cell, 64 cell tiles")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding design is documented, not built)", file=sys.stderr)
        sys.exit(2)

    if args.cpu_worker:
        cpu_worker(args.cpu_worker)
        return

    import torch

    # no automatic collections from here on; quiesce() collects before each timed region
    gc.disable()
    import cupcfd_hip as ch
    from cupcfd_hip import workload as W
    from cupcfd_hip.fvm import resident as R

    if args.batched_only:
        hb, db, lb = structgen(args.batch_n, 4242)
        print(json.dumps({"batched": bench_batched(torch, hb, db, args.batch_k, steps=args.steps,
                                                   warmup=args.warmup, label=lb)}))
        return
    if args.partitions_only:
        print(json.dumps({"partitions": bench_partitions(torch, 48, 64, args.steps, args.warmup)}))
        return
    if args.tets_only:
        print(json.dumps({"tets": bench_tets(torch, args.tets_n, args.tets_k, args.steps, args.warmup)}))
        return
    n = args.n
    t0 = time.time()
    host = ch.HostMesh(n, seed=12345, permute_cells=True, permute_faces=True)
    t_gen = time.time() - t0
    dev = ch.DeviceMesh(host, 64, reorder=True)
    t_mesh = time.time() - t0
    t_host_repack = t_mesh - t_gen
    if args.schedule >= 0:
        dev.set_schedule(args.schedule)
        schedules = [args.schedule] * 3
    else:
        schedules = dev.autotune()
    C, F, B = host.nCells, host.nFaces, host.nBoundaries
    Fi = int((host.a["faceBnd"] < 0).sum())
    f = W.Fields(host, seed=777)

    kernels, sweep_arrays = make_sweep(torch, host, dev, f, arrays=True)
    per, step_s, t_wall = time_sweep(torch, kernels, args.steps, args.warmup)
    K = args.steps
    kern = kernel_report(C, F, Fi, B, per)
    dom = max(per, key=per.get)
    copy_gbs = stream_copy_gbs(torch)
    read_gbs = stream_read_gbs(torch)

    traffic = None
    # HBM bytes per launch of the dominant kernel, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload (scaled by the
    # copy-kernel calibration recorded beside them)
    tfile = os.path.join(ROOT, "profiles", f"fvm_kernels_structgen{n}_f64.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch", {}).get(dom)

    out = {
        "metric": "Gcell-sweeps/s & % HBM peak per FVM kernel; batched & fwd+adjoint vs plain",
        "value": C / step_s / 1e9,
        "unit": "Gcell-sweeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded fields on a seeded relabelled StructGen mesh)",
        "config": {"workload": f"plain FP64 gradient+mass+scalar+UVW sweep, StructGen {n}^3 hex, cell and face ids "
                               f"permuted (seed 12345), one DEFAULT region", "cells": C, "faces": F,
                   "boundaries": B, "parallelism": "single GPU",
                   "gather_schedules": dict(zip(("gradient", "scalar", "uvw"), schedules))},
        "kernels": kern,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": kern[dom]["algorithmic_GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kern[dom]["hbm_frac"], "traffic": traffic},
        "stream_copy_GBps": copy_gbs,
        "stream_read_GBps": read_gbs,
        "wall_s_timed_region": t_wall,
        "mesh_setup_s": t_mesh,
    }
    out["residual_norms"] = bench_norms(torch, dev, [sweep_arrays[k] for k in ("Su", "su", "sv", "sw")],
                                        steps=max(5, min(args.steps, 50)))
    del kernels, sweep_arrays
    oracle = None
    if not args.no_cpu:
        # the oracle's outputs on the same mesh and inputs, against one drop-in sweep on the GPU
        oracle = oracle_outputs(host, f)
        out["parity_vs_oracle"] = parity_vs_oracle(dev, f, oracle)
        out["parity_vs_oracle"]["workload"] = "each kernel once on fresh copies of the sweep's inputs, same mesh"
    if not args.no_all_kernels:
        out["all_kernels"] = bench_all_kernels(torch, n, steps=max(5, min(args.steps, 20)), warmup=2,
                                               schedule=args.schedule)
    if not args.no_f32:
        out["f32"] = bench_f32(torch, host, f, steps=max(5, min(args.steps, 20)), warmup=2, schedule=args.schedule,
                               oracle=oracle)
    del oracle
    if not args.no_mesh_build:
        out["mesh_build"] = bench_mesh_build(torch, host, dev, t_gen, t_host_repack, f"StructGen {n}^3 relabelled")
    if not args.no_batched:
        hb, db, lb = structgen(args.batch_n, 4242)
        out["batched"] = bench_batched(torch, hb, db, args.batch_k, steps=max(2, min(args.steps, 10)), warmup=1,
                                       label=lb)
        del hb, db
    if not args.no_adjoint:
        ha, da, la = structgen(args.adjoint_n, 12345)
        out["adjoint"] = bench_adjoint(torch, ha, da, steps=max(5, min(args.steps, 20)), warmup=2, label=la)
        del ha, da
        # any mesh, several sweeps: typed relabelled mesh, enthalpy wall, nGradient = 2
        import cupcfd_hip as ch
        from cupcfd_hip import workload as W
        ht = ch.HostMesh(args.adjoint_n, seed=12345, permute_cells=True, permute_faces=True, typed=True,
                         face_extras=True)
        dt_ = ch.DeviceMesh(ht, 64, reorder=True)
        pw = dict(W.SCALAR, ivar=W.SCALAR["VarT"], SolveEnthalpy=1)
        out["adjoint"]["typed_full"] = bench_adjoint(torch, ht, dt_, steps=max(5, min(args.steps, 20)), warmup=2,
                                                     label=f"StructGen {args.adjoint_n}^3 relabelled", nG=2,
                                                     full=True, params=pw)
        del ht, dt_
    if not args.no_partitions:
        out["partitions"] = bench_partitions(torch, 48, 64, steps=max(3, min(args.steps, 10)), warmup=2)
    if not args.no_tets:
        torch.cuda.empty_cache()
        out["tets"] = bench_tets(torch, args.tets_n, args.tets_k, steps=max(3, min(args.steps, 10)), warmup=1)
    if not args.no_cpu:
        out["config0"] = bench_config0(torch)
        out["cpu_baseline"] = cpu_baseline(host, f, args.cpu_budget)
        out["cpu_baseline"]["all_cores"] = cpu_baseline_all_cores(n, args.cpu_budget / 2)
        out["cpu_baseline"]["host_cpu"] = host_cpu_model()
        out["cpu_baseline"]["nproc"] = os.cpu_count()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""The boundary loops' sums as a fixed-shape parallel tree (cupcfd_hip_mesh_set_ordered_sums(mesh, 0)), GPU.

On a typed relabelled 64^3 mesh (4096 boundaries per side, so the sequential
sums run over several LDS chunks): every output of mass boundary loops 1-5
and the UVW boundary loop is within 1e-12 (FP64) / 1e-4 (FP32) of the oracle's
sequential result (max-norm relative), two calls give the same bits, and
switching back to sequential order gives the oracle's bits again.
"""
import zlib

import numpy as np
import pytest

import cupcfd_hip as ch
import pyoracle as po
from cupcfd_hip import fvm
from cupcfd_hip import workload as W
from kernel_runs import compare, run, seed_state

KERNELS = ["mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5", "uvw_bnds1"]
_mesh = {}


def mesh(precision):
    if precision not in _mesh:
        host = ch.HostMesh(64, typed=True, permute_cells=True, permute_faces=True, seed=17)
        _mesh[precision] = (host, ch.DeviceMesh(host, precision=precision, reorder=True), po.OracleMesh(host, precision))
    return _mesh[precision]


def one(kernel, precision, tree):
    host, dev, om = mesh(precision)
    dt = np.float64 if precision == 64 else np.float32
    fields = W.Fields(host, seed=zlib.crc32(f"tree/{kernel}".encode()) % 1000)
    st0 = {k: v.astype(dt) for k, v in seed_state(host).items()}
    host.set_state(st0, precision)
    dev.set_ordered_sums(not tree)
    try:
        rc, out = run(kernel, fvm, dev, fields.cast(dt), lambda: host.state(precision))
    finally:
        dev.set_ordered_sums(True)
    assert rc == ch.E_SUCCESS
    om.set_state(st0)
    rc_o, want = run(kernel, po, om, fields.cast(dt), om.state)
    assert rc_o == ch.E_SUCCESS
    return out, want


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("kernel", KERNELS)
def test_tree_sums_within_tolerance_and_deterministic(kernel, precision):
    print(f"[tree sums] {kernel} fp{precision}", flush=True)
    out1, want = one(kernel, precision, True)
    out2, _ = one(kernel, precision, True)
    tol = 1e-12 if precision == 64 else 1e-4
    for k in want:
        _, rel = compare(out1[k], want[k])
        assert rel <= tol, f"{kernel}/{precision}: {k} max rel {rel:.3e}"
        a, b = np.ascontiguousarray(out1[k]), np.ascontiguousarray(out2[k])
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"{kernel}/{precision}: {k} not repeatable"
    seq, want = one(kernel, precision, False)
    for k in want:
        assert compare(seq[k], want[k])[0], f"{kernel}/{precision}: {k} sequential order lost"


@pytest.mark.gpu
def test_tree_mode_is_in_effect():
    """The tree's order differs from the sequential one: over the 4096 INLET boundaries of the 64^3 mesh the
    FP64 inflow sum changes in its last bits (the mode switch reaches the kernel)."""
    tree, want = one("mass_bl1", 64, True)
    seq, _ = one("mass_bl1", 64, False)
    assert compare(seq["flowin"], want["flowin"])[0]
    assert not compare(tree["flowin"], want["flowin"])[0]

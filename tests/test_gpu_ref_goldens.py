"""The HIP kernels against CUP-CFD's own outputs (GPU, through the drop-in C ABI).

Each case runs one kernel of libcupcfd_hip on a mesh of
tests/golden/ref_kernels.npz with the fixture's inputs and mesh state, and
compares every output and every piece of mesh state with what the reference's
own fvm kernels, compiled from its sources, produced on the same mesh arrays.

Bars (max-norm relative, ||gpu - reference||inf / ||reference||inf):
  FP64: kernels without transcendental calls bit-identical; the others
        (scalar WALL-enthalpy, viscosity region and boundary loops: log, pow,
        sqrt of the wall functions) within 1e-12.
  FP32: the same non-transcendental kernels bit-identical, every case within 2e-6.
"""
import numpy as np
import pytest

import cupcfd_hip as ch
import ref_cases as rc
from cupcfd_hip import fvm
from kernel_runs import compare, run
from test_ref_goldens import GOLD, bits, cases, fixture_case

BITEXACT = {"gradient", "gradient3", "mass", "mass_bl1", "mass_bl2", "mass_bl3", "mass_bl4", "mass_bl5",
            "mass_region", "scalar", "scalar_t", "uvw", "uvw_region1", "uvw_bnds1", "visc_cell1", "visc_cell2"}
FP32_BITEXACT = BITEXACT

_dev = {}


def _device(name, prec):
    if (name, prec) not in _dev:
        mesh, f, st = fixture_case(name)
        _dev[(name, prec)] = (mesh, ch.DeviceMesh(mesh, prec), f, st)
    return _dev[(name, prec)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,prec,kern", cases())
def test_hip_matches_reference(name, prec, kern):
    mesh, dev, f, st0 = _device(name, prec)
    dt = np.float64 if prec == 64 else np.float32
    mesh.set_state({k: v.astype(dt) for k, v in st0.items()}, prec)
    code, out = run(kern, fvm, dev, f.cast(dt), lambda: mesh.state(prec))
    key = f"{name}/{prec}/{kern}"
    assert code == int(GOLD[key + "/rc"][0]) == ch.E_SUCCESS
    expect = {k[len(key) + 5:]: GOLD[k] for k in GOLD.files if k.startswith(key + "/out/")}
    assert expect.keys() == out.keys()
    for k, v in expect.items():
        same, rel = compare(out[k], v)
        if (prec == 64 and kern in BITEXACT) or (prec == 32 and kern in FP32_BITEXACT):
            assert np.array_equal(bits(out[k]), bits(v)), f"{key}: {k} not bit-identical (max rel {rel:.3e})"
        tol = 1e-12 if prec == 64 else 2e-6
        assert rel <= tol, f"{key}: {k} max rel {rel:.3e} > {tol}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(rc.STRUCTGEN))
def test_device_generator_matches_reference_structgen(name):
    """The GPU generator's host copy equals the reference's StructGen mesh bit for bit."""
    nx, ny, nz, bounds = rc.STRUCTGEN[name]
    host, dev = ch.generate_mesh(nx, ny, nz, bounds=bounds)
    for k in rc.SG_KEYS:
        assert np.array_equal(bits(host.a[k]), bits(GOLD[f"{name}/{k}"])), f"{name}: {k}"

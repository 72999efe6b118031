"""Argument checks of the resident (device-tensor) bindings, on the CPU.

cupcfd_hip.fvm.resident checks every tensor against the mesh before the C
call: dtype of the mesh's precision, contiguous, long enough, on the GPU. A
violation raises CupCfdError and nothing reaches the library (the stand-in
mesh below has no device handle, so any call that got through would fail
differently)."""
import pytest

import cupcfd_hip as ch
from cupcfd_hip.fvm import resident as R

torch = pytest.importorskip("torch")


class StandInMesh:
    """Counts and precision of a 4x4x4 hex mesh; no device mesh behind it."""
    nCells, nFaces, nBoundaries, nRegions = 64, 240, 96, 1
    precision = 64
    sfx = "f64"
    handle = None


def _grad(phi=None, phiB=None, g=None, go=None):
    m = StandInMesh
    d = torch.float64
    phi = torch.zeros(m.nCells, dtype=d) if phi is None else phi
    phiB = torch.zeros(m.nBoundaries, dtype=d) if phiB is None else phiB
    g = torch.zeros(3 * m.nCells, dtype=d) if g is None else g
    go = torch.zeros(3 * m.nCells, dtype=d) if go is None else go
    R.gradient(m, 1, phi, phiB, g, go)


def _code(fn, *a, **k):
    with pytest.raises(ch.CupCfdError) as e:
        fn(*a, **k)
    return e.value.code, str(e.value)


def test_wrong_dtype_is_refused():
    code, msg = _code(_grad, phi=torch.zeros(64, dtype=torch.float32))
    assert code == 90 and "phi" in msg and "dtype" in msg


def test_non_contiguous_view_is_refused():
    x = torch.zeros((64, 2), dtype=torch.float64)
    code, msg = _code(_grad, phi=x[:, 1])
    assert code == 90 and "contiguous" in msg


def test_short_tensor_is_refused():
    code, msg = _code(_grad, g=torch.zeros(3 * 64 - 1, dtype=torch.float64))
    assert code == ch.E_ARRAY_SIZE_MISMATCH and "dPhidx" in msg


def test_batched_sizes_scale_with_k():
    m = StandInMesh
    d = torch.float64
    K = 5
    ok = dict(phi=torch.zeros(m.nCells * K, dtype=d), phiB=torch.zeros(m.nBoundaries * K, dtype=d),
              g=torch.zeros(3 * m.nCells * K, dtype=d), go=torch.zeros(3 * m.nCells * K, dtype=d))
    short = dict(ok, phiB=torch.zeros(m.nBoundaries * K - 1, dtype=d))
    assert _code(R.gradient_batched, m, 1, K, **short)[0] == ch.E_ARRAY_SIZE_MISMATCH
    # every check passes; the host tensors are then refused as not on the GPU
    code, msg = _code(R.gradient_batched, m, 1, K, **ok)
    assert code == 90 and "GPU" in msg


def test_f32_storage_arrays_must_be_float32():
    m = StandInMesh
    K = 2
    f64 = lambda n: torch.zeros(n, dtype=torch.float64)  # noqa: E731
    code, msg = _code(R.gradient_batched_f32s, m, 1, K, f64(m.nCells * K), f64(m.nBoundaries * K),
                      f64(3 * m.nCells * K), f64(3 * m.nCells * K))
    assert code == 90 and "float32" in msg

"""The BenchmarkKernels-style C++ driver (cup-cfd_amd/cupcfd_bench) runs the
16 kernels through the cupcfd::fvm-shaped shim on a reference-schema JSON."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cup-cfd_amd", "cupcfd_bench")
CFG = os.path.join(ROOT, "cup-cfd_amd", "driver", "cupcfd.json")

KERNELS = ["GradientPhiGaussDolfynBenchmark", "FluxMassDolfynFaceLoopBenchmark",
           "FluxMassDolfynBoundaryLoop1Benchmark", "FluxMassDolfynBoundaryLoop2Benchmark",
           "FluxMassDolfynBoundaryLoop3Benchmark", "FluxMassDolfynBoundaryLoop4Benchmark",
           "FluxMassDolfynBoundaryLoop5Benchmark", "FluxMassDolfynRegionLoopBenchmark",
           "FluxScalarDolfynFaceLoopBenchmark", "FluxUVWDolfynFaceLoop1Benchmark", "FluxUVWDolfynRegionLoop1Benchmark",
           "FluxUVWDolfynBndsLoop1Benchmark", "calculateViscosityDolfynCellLoop1Benchmark",
           "calculateViscosityDolfynRegionLoopBenchmark", "calculateViscosityDolfynBoundaryLoopBenchmark",
           "calculateViscosityDolfynCellLoop2Benchmark"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_driver_runs_reference_schema(tmp_path, precision):
    cfg = json.load(open(CFG))
    cfg["DataTypes"]["FloatDataType"] = precision
    p = tmp_path / "cupcfd.json"
    p.write_text(json.dumps(cfg))
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout)
    names = [b["BenchmarkName"] for b in rep["benchmarks"]]
    assert names == ["KernelTest", "KernelTestTypedPermuted"]  # BenchmarkExchange is skipped
    assert "BenchmarkExchange" in r.stderr
    for b in rep["benchmarks"]:
        assert b["failures"] == 0
        assert b["LocalCells"] == 11 * 12 * 14
        assert sorted(b["kernels"]) == sorted(KERNELS)
        for k in KERNELS:
            assert b["kernels"][k]["calls"] == b["Repetitions"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_driver_graph_mode(tmp_path, precision):
    """"Graph": the 16 kernels on resident inputs, captured once into a hipGraph; one replay equals one eager
    repetition from the same device state, byte for byte, on a DEFAULT and on a typed relabelled mesh."""
    cfg = json.load(open(CFG))
    cfg["DataTypes"]["FloatDataType"] = precision
    for b in cfg["Benchmarks"]:
        if "BenchmarkKernels" in b:
            b["BenchmarkKernels"]["Graph"] = 1
            b["BenchmarkKernels"]["Repetitions"] = 3
    p = tmp_path / "cupcfd.json"
    p.write_text(json.dumps(cfg))
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout)
    assert len(rep["benchmarks"]) == 2
    for b in rep["benchmarks"]:
        g = b["graph"]
        assert g["ok"] and g["replay_matches_eager"], b
        assert g["graph_ms_per_repetition"] > 0 and g["eager_ms_per_repetition"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_driver_mesh_build_host_and_device(tmp_path, precision):
    """"MeshBuild": the device mesh is repacked on the GPU by default and on the host with "host"; both builds
    hash to the same mesh digest, the kernels run without failures on either, and the repack time is reported
    apart from the kernels'."""
    reps = {}
    for build in ("host", "device"):
        cfg = json.load(open(CFG))
        cfg["DataTypes"]["FloatDataType"] = precision
        for b in cfg["Benchmarks"]:
            if "BenchmarkKernels" in b:
                b["BenchmarkKernels"]["Repetitions"] = 2
                if build == "host":
                    b["BenchmarkKernels"]["MeshBuild"] = "host"
        p = tmp_path / f"cupcfd_{build}.json"
        p.write_text(json.dumps(cfg))
        r = subprocess.run([BIN, str(p)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        reps[build] = json.loads(r.stdout)["benchmarks"]
    for h, d in zip(reps["host"], reps["device"]):
        assert h["BenchmarkName"] == d["BenchmarkName"]
        assert (h["mesh_build"], d["mesh_build"]) == ("host", "device")
        assert h["mesh_digest"] == d["mesh_digest"], h["BenchmarkName"]
        assert h["failures"] == 0 and d["failures"] == 0
        assert h["mesh_setup_s"] > 0 and d["mesh_setup_s"] > 0
    # the plain and the typed relabelled meshes differ
    assert reps["device"][0]["mesh_digest"] != reps["device"][1]["mesh_digest"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_driver_reads_hdf5_mesh_file(tmp_path, precision):
    # "MeshSourceFile" {"FileFormat": "HDF5", "FilePath"} in place of the StructGen grid
    cfg = json.load(open(CFG))
    cfg["DataTypes"]["FloatDataType"] = precision
    cfg["Mesh"]["MeshSource"] = {"MeshSourceFile": {"FileFormat": "HDF5",
                                                    "FilePath": os.path.join(ROOT, "tests", "golden",
                                                                             "MeshHDF5.hdf5")}}
    cfg["Benchmarks"] = [b for b in cfg["Benchmarks"] if "BenchmarkKernels" in b][:1]
    p = tmp_path / "cupcfd.json"
    p.write_text(json.dumps(cfg))
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    (b,) = json.loads(r.stdout)["benchmarks"]
    assert (b["LocalCells"], b["LocalFaces"], b["LocalBounds"]) == (4, 20, 17)
    assert b["failures"] == 0
    assert sorted(b["kernels"]) == sorted(KERNELS)
    # a path that does not exist: E_FILE_MISSING (10) reported, non-zero exit
    cfg["Mesh"]["MeshSource"]["MeshSourceFile"]["FilePath"] = str(tmp_path / "absent.hdf5")
    p.write_text(json.dumps(cfg))
    r = subprocess.run([BIN, str(p)], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "returned 10" in r.stderr

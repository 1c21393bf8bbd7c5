"""The C++ API (include/panostitcher/CudaInterface.hpp): its test program,
tests/cpp/facade_test.cpp, built with the library.  Host mode runs without a
GPU (rig files, memory estimate, types, exceptions); GPU mode runs the
reference's per-camera flow (maps, cudaTransform, cudaReprojectTo16S,
CudaTilingMultibandBlendFast) and CudaPanoramaRender2 on one rig and requires
the same canvas byte for byte, plus the error conventions."""
import os
import subprocess

import pytest

import panostitcher_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_test")
RIGS = os.path.join(ROOT, "tests", "golden", "rigs")


def _binary():
    if not os.path.exists(BIN):
        P.build()
    return BIN


def test_facade_host(tmp_path):
    r = subprocess.run([_binary(), "host", RIGS], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade host checks passed" in r.stdout


@pytest.mark.gpu
def test_facade_gpu(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    r = subprocess.run([_binary(), "gpu", RIGS], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "blend vs render: 0 differing bytes; tile vs blend: 0" in r.stdout

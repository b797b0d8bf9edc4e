"""Host-side checks of traversal arithmetic (hipcc builds a CPU program; no GPU)."""
import os
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "bin" / "hipcc"


@pytest.mark.skipif(not HIPCC.exists(), reason="hipcc not installed")
def test_triangle_prefilter_never_rejects_a_hit():
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "prefilter_fuzz"
        build = subprocess.run([str(HIPCC), "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                                "-o", str(exe), str(ROOT / "tests" / "native" / "prefilter_fuzz.cpp")],
                               capture_output=True, text=True)
        assert build.returncode == 0, build.stderr[-2000:]
        run = subprocess.run([str(exe), "2000000"], capture_output=True, text=True, timeout=300)
        print(run.stdout)
        assert run.returncode == 0, run.stdout
        assert "false_rejections 0" in run.stdout


@pytest.mark.skipif(not HIPCC.exists(), reason="hipcc not installed")
def test_slab_test_never_rejects_a_box_the_ray_meets():
    """Axis-parallel rays (exactly zero direction components) included: an
    infinite reciprocal in the fused slab form gives NaN / wrongly signed
    infinities and used to reject boxes such rays pass through."""
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "slab_fuzz"
        build = subprocess.run([str(HIPCC), "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                                "-o", str(exe), str(ROOT / "tests" / "native" / "slab_fuzz.cpp")],
                               capture_output=True, text=True)
        assert build.returncode == 0, build.stderr[-2000:]
        run = subprocess.run([str(exe), "2000000"], capture_output=True, text=True, timeout=300)
        print(run.stdout)
        assert run.returncode == 0, run.stdout
        assert "false_rejections 0" in run.stdout


@pytest.mark.skipif(not HIPCC.exists(), reason="hipcc not installed")
def test_packed_nodes_decode_and_stay_conservative():
    """48-byte nodes: child ids carried in four planes' low bytes decode
    exactly, the byte-carrying planes only move outward, and the slab test on
    a packed node enters every box the exact planes' test enters."""
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "node_fuzz"
        build = subprocess.run([str(HIPCC), "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                                "-o", str(exe), str(ROOT / "tests" / "native" / "node_fuzz.cpp")],
                               capture_output=True, text=True)
        assert build.returncode == 0, build.stderr[-2000:]
        run = subprocess.run([str(exe), "2000000"], capture_output=True, text=True, timeout=300)
        print(run.stdout)
        assert run.returncode == 0, run.stdout
        assert "node_fuzz ok" in run.stdout

"""Host-side check of the kernels' index arithmetic (tests/host/index_check.hip):
pull sources in bounds, gather a bijection, fast division exact, XCD remap a
permutation.  Compiled with hipcc and run on the CPU; no GPU involved."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_index_arithmetic_on_host():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "index_check")
        subprocess.run([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "d3q19-single-phase_amd", "csrc"),
                        os.path.join(ROOT, "tests", "host", "index_check.hip"), "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
        assert "ok remap" in r.stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_particle_boxes_stay_inside_the_lattice():
    """tests/host/box_check.hip: sphere_box for centres inside, outside, huge,
    infinite and NaN enumerates only lattice nodes and covers every node
    within R of a centre inside the channel."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "box_check")
        subprocess.run([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "d3q19-single-phase_amd", "csrc"),
                        "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host", "box_check.hip"),
                        "-o", exe], check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
        assert "ok boxes" in r.stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_blocked_two_step_sweep_schedule():
    """tests/host/block_check.hip: the launch list of the two-step blocked
    sweep, emulated with the kernels' addressing on tagged values, gathers
    exactly what the single-step pull gathers and writes every output slot
    once, over whole slabs and in row windows; and lbm_create's default sweep
    (block_choice) is the measured optimum on the workloads it was tuned on,
    with the environment overrides."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "block_check")
        subprocess.run([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "d3q19-single-phase_amd", "csrc"),
                        os.path.join(ROOT, "tests", "host", "block_check.hip"), "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
        assert "ok sweep choice" in r.stdout
        assert "ok blocked sweep" in r.stdout
        assert "ok row windows" in r.stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_fp16c_storage_format():
    """tests/host/fp16c_check.hip: every FP16C word decodes to the format's
    value and encodes back; words are ordered; floats encode to the nearest
    word (brute force, ties to even), saturating above the largest."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "fp16c_check")
        subprocess.run([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "d3q19-single-phase_amd", "csrc"),
                        os.path.join(ROOT, "tests", "host", "fp16c_check.hip"), "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
        assert "ok fp16c" in r.stdout

"""Host-side AddressSanitizer and UndefinedBehaviorSanitizer over the C ABI's host code.

Builds every library source with the sanitizers on the host side only (`-Xarch_host`;
the device code is compiled as usual and never launched) together with
tests/host/abi_validation_driver.cpp, then runs the driver on the CPU: each entry point
refuses NULL or inconsistent descriptors and sizes through its status code before any
device work, gx_workspace_size answers valid queries (up to the largest its caps allow)
without overflow, and ASan / UBSan report nothing (halt_on_error, leak detection on).
"""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")
CSRC = os.path.join(ROOT, "giwaxsim_amd", "csrc")
SOURCES = ["gx_workspace.hip", "gx_frames.hip", "gx_detector.hip", "gx_adjoint.hip", "gx_compare.hip",
           "gx_slab_build.hip"]
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not available")
def test_abi_validation_under_asan_ubsan(tmp_path):
    common = ["hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + SAN
    objs, procs = [], []
    for src in SOURCES:
        obj = str(tmp_path / (src + ".o"))
        objs.append(obj)
        procs.append(subprocess.Popen(common + ["-c", "-o", obj, os.path.join(CSRC, src)],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    drv = str(tmp_path / "driver.o")
    procs.append(subprocess.Popen(common + ["-c", "-o", drv, os.path.join(ROOT, "tests", "host",
                                                                      "abi_validation_driver.cpp")],
                                  stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, out.decode(errors="replace")[-3000:]
    exe = str(tmp_path / "driver")
    link = subprocess.run(["hipcc", "--offload-arch=gfx950"] + SAN + ["-o", exe, drv] + objs,
                          capture_output=True, text=True, timeout=300)
    assert link.returncode == 0, link.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    assert run.returncode == 0, (run.stdout + run.stderr)[-3000:]
    assert "checks passed:" in run.stdout

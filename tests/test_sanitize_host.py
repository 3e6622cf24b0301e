"""The library's host code and the island tasks' logic under AddressSanitizer
and UndefinedBehaviorSanitizer, on the CPU.

tests/sanitize/host_check.cpp drives the plan builder (both conventions, both
schedules, team sizing with estimates, the general-matrix plan, an invalid
island), the CPU emulation of the LSCM / ARAP / SLIM / CG / plan-PCG tasks for
three team shapes, the sequential flip repair and the chart regrouping.  The
sources are compiled for the host only (the same ones libuvart_emu.so is
built from) with both sanitizers; any report ends the run with a non-zero
status.  A fixed input set, about 15 s to build and 30 s to run.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "ultimate_uv_artist_amd" / "csrc"
SOURCES = [CSRC / "kernels.hip", CSRC / "setup.cpp", CSRC / "pcg.hip", CSRC / "validate.hip",
           ROOT / "tests" / "sanitize" / "host_check.cpp"]
# host-only objects, so each -fsanitize= goes to the host compilation
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    return None


@pytest.mark.timeout(600)
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    flags = ["--cuda-host-only", "-O0", "-g", "-std=c++17", "-DUVA_EMULATOR", "-pthread", *SAN,
             "-fno-omit-frame-pointer", "-Wno-unused-command-line-argument", "-I", str(ROOT / "include")]

    def compile_one(src):
        obj = tmp_path / (src.stem + ".o")
        r = subprocess.run([hipcc, *flags, "-c", str(src), "-o", str(obj)], capture_output=True, text=True)
        assert r.returncode == 0, f"{src.name}:\n{r.stderr[-4000:]}"
        return str(obj)

    with ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    exe = tmp_path / "host_check"
    r = subprocess.run([hipcc, "--cuda-host-only", "-pthread", *SAN, *objs, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               UVART_HOST_THREADS="4")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=480)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "host_check ok" in r.stdout, (r.stdout[-2000:] + r.stderr[-6000:])
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-6000:]

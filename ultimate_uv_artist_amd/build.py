"""Build libuvart.so (gfx950 HIP kernels + C++ host code) in-tree with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libuvart.so"
SOURCES = ["kernels.hip", "capi.cpp", "setup.cpp", "validate.hip", "pcg.hip"]
ARCH = os.environ.get("UVA_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


FLAGS = ["-O3", "-std=c++17", "-fPIC",
         # packed FP32 instructions stay off (hipcc forms them from float code)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-pthread"]


def build(verbose=False):
    """Compile every source for gfx950 (one hipcc per source, in parallel) and
    link libuvart.so next to this file."""
    tmp = LIB.with_suffix(".so.tmp")
    with tempfile.TemporaryDirectory(prefix="uvart_build_") as td:
        def compile_one(src):
            obj = str(Path(td) / (src + ".o"))
            cmd = [hipcc(), f"--offload-arch={ARCH}", *FLAGS, "-I", str(HERE.parent / "include"), "-c",
                   str(CSRC / src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            return obj, subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            results = list(ex.map(compile_one, SOURCES))
        for (obj, r), src in zip(results, SOURCES):
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed compiling {src}")
        cmd = [hipcc(), f"--offload-arch={ARCH}", *FLAGS, "-shared", *[o for o, _ in results], "-o", str(tmp)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libuvart.so")
    os.replace(tmp, LIB)
    return LIB


EMU_LIB = HERE / "libuvart_emu.so"


def build_emulator(verbose=False):
    """libuvart_emu.so: the island tasks compiled for the host with one thread
    per lane (test-only CPU emulation of the kernels; no GPU code inside)."""
    tmp = EMU_LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared", "-DUVA_EMULATOR", "-pthread",
           "-Wno-unused-command-line-argument", "-I", str(HERE.parent / "include"),
           str(CSRC / "kernels.hip"), str(CSRC / "setup.cpp"), str(CSRC / "validate.hip"), str(CSRC / "pcg.hip"),
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libuvart_emu.so")
    os.replace(tmp, EMU_LIB)
    return EMU_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_emulator(verbose=True))

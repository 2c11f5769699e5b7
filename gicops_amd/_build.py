"""Build recipes for the native libraries (host C++, HIP for gfx950, oracle).

Every library is built in-tree so that it travels with the repository:

  gicops_amd/lib/libgicops_host.so  index construction + synthetic data (g++/OpenMP)
  gicops_amd/lib/libgicops_hip.so   HIP kernels + the C ABI of include/gicops_amd.h
  oracle/liboracle.so               CPU restatement used only by tests and the
                                    CPU baseline (built from oracle/Makefile)

Floating-point contraction is off everywhere: the reference is built for
baseline x86-64 (no FMA) and parity of float bins and hyperscores depends on
unfused multiply/add.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "gicops_amd"
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
INCLUDE = ROOT / "include"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"

HOST_SOURCES = [CSRC / "host" / "slm_build.cpp", CSRC / "host" / "ms2.cpp"]
HIP_SOURCES = [
    CSRC / "hip" / "search.hip",
    CSRC / "hip" / "batched.hip",
    CSRC / "hip" / "capi.hip",
    CSRC / "hip" / "mlp.hip",
    CSRC / "hip" / "mlp_capi.hip",
    CSRC / "hip" / "learned.hip",
    CSRC / "hip" / "index_build.hip",
    CSRC / "hip" / "preprocess.hip",
    CSRC / "hip" / "partials.hip",
    CSRC / "hip" / "bandwidth.hip",
    CSRC / "hip" / "noalloc.hip",
]
HIP_HEADERS = sorted((CSRC / "hip").glob("*.h")) + sorted(INCLUDE.glob("*.h"))


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_host() -> Path:
    LIBDIR.mkdir(parents=True, exist_ok=True)
    out = LIBDIR / "libgicops_host.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-std=c++17", "-O3", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
          "-Wall", "-Wextra", "-Wno-unused-parameter",
          f"-I{INCLUDE}", f"-I{CSRC / 'host'}", *HOST_SOURCES, "-o", tmp])
    os.replace(tmp, out)
    return out


HIP_FLAGS = [f"--offload-arch={OFFLOAD_ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-parameter", "-Wno-unused-command-line-argument"]


def build_hip() -> Path:
    """Each source compiled on its own (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    LIBDIR.mkdir(parents=True, exist_ok=True)
    objdir = ROOT / "_build" / "hip"
    objdir.mkdir(parents=True, exist_ok=True)
    out = LIBDIR / "libgicops_hip.so"
    tmp = out.with_suffix(".so.tmp")
    objs = [objdir / (src.stem + ".o") for src in HIP_SOURCES]
    jobs = max(1, min(len(HIP_SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC / 'hip'}", "-c", so[0],
                                     "-o", so[1]]), zip(HIP_SOURCES, objs)))
    _run([HIPCC, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_adapter() -> Path:
    """C++ host side with the reference's types above the C ABI, and the
    end-to-end driver the GPU tests run."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    out = LIBDIR / "libgicops_adapter.so"
    tmp = out.with_suffix(".so.tmp")
    adir = CSRC / "adapter"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-Wextra",
          f"-I{INCLUDE}", f"-I{adir}", adir / "dslim_amd.cpp", "-o", tmp,
          f"-L{LIBDIR}", "-lgicops_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    exe = LIBDIR / "adapter_e2e"
    _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{INCLUDE}", f"-I{adir}", ROOT / "tests" / "cpp" / "adapter_e2e.cpp",
          "-o", exe, f"-L{LIBDIR}", "-lgicops_adapter", "-lgicops_host", "-lgicops_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle() -> Path:
    _run(["make", "-C", str(ORACLE), "-B"])
    return ORACLE / "liboracle.so"


def build_all() -> None:
    build_host()
    build_oracle()
    build_hip()
    build_adapter()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"host": build_host, "hip": build_hip, "oracle": build_oracle, "adapter": build_adapter,
         "all": build_all}[w]()

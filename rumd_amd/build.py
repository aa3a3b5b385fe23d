"""Builds the native parts of rumd_amd in-tree.

- ``rumd_amd/librumdx.so``: the gfx950 HIP kernels behind the C ABI declared in
  ``include/rumdx.h`` (hipcc, --offload-arch=gfx950).
- ``rumd_amd/_rumd*.so``: the C++ host classes (Sample, PairPotential,
  NeighborList, IntegratorNVT, ...) and their pybind11 Python surface, linked
  against librumdx.so.

Run ``python -m rumd_amd.build`` (or ``__graft_entry__.build()``).  Every call
recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
OBJ = PKG / "_build"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 VALU forms are kept out of the code objects; see the module docs
# of csrc/*.hip for why every kernel is plain FP32/FP64 VALU code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
HOST_FLAGS = [
    "-O2",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{INCLUDE}",
    f"-I{HOST}",
    "-I/opt/rocm/include",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    out = proc.stdout.strip()
    # hipcc's host pass does not know the packed-fp32 feature flag; that note is harmless.
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def _compile_hip(src: Path) -> Path:
    obj = OBJ / (src.stem + ".hip.o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_kernels(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    lib = PKG / "librumdx.so"
    tmp = PKG / "librumdx.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, lib)
    return lib


def _pybind_includes() -> list[str]:
    import pybind11

    return [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]


def _compile_host(src: Path) -> Path:
    obj = OBJ / (src.stem + ".host.o")
    _run(["g++", *HOST_FLAGS, *_pybind_includes(), "-fvisibility=hidden", "-c", str(src), "-o", str(obj)])
    return obj


def build_host(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(HOST.glob("*.cpp"))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_host, srcs))
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    lib = PKG / f"_rumd{ext}"
    tmp = PKG / f"_rumd{ext}.tmp"
    _run(["g++", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), f"-L{PKG}", "-lrumdx",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, lib)
    return lib


def build(jobs: int = 8) -> None:
    build_kernels(jobs)
    if any(HOST.glob("*.cpp")):
        build_host(jobs)


if __name__ == "__main__":
    build()
    print("built", PKG / "librumdx.so")

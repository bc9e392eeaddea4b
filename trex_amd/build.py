"""Build the native parts of trex_amd in-tree.

* ``kernels/<name>.hsaco`` -- one gfx950 code object per ``csrc/<name>.hip``
  kernel file (hipcc, device code only); the runtime loads every code object
  in that folder with ``hipModuleLoad``;
* ``libtrexhip.so`` -- the C ABI (``include/trex_hip.h``), host C++ built
  against the HIP runtime.

Run ``python -m trex_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

KERNEL_DIR = PKG / "kernels"
LIB = PKG / "libtrexhip.so"
OBJ_DIR = PKG / "_build"   # intermediate objects of the library (not shipped)

# Packed FP32 instructions are kept out of every kernel.
HIP_DEVICE_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-munsafe-fp-atomics"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-Wno-unused-result",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build_kernels(jobs: int | None = None) -> list[Path]:
    """Compiles every kernel file, up to ``jobs`` hipcc processes at once
    (default: the CPU count, at most 8)."""
    from concurrent.futures import ThreadPoolExecutor
    KERNEL_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))

    def one(src: Path) -> Path:
        out = KERNEL_DIR / (src.stem + ".hsaco")
        tmp = out.with_suffix(".hsaco.tmp")
        _run([ROCM / "bin" / "hipcc", "--genco", "--no-gpu-bundle-output", *HIP_DEVICE_FLAGS,
              "-I", CSRC, "-o", tmp, src])
        os.replace(tmp, out)
        return out

    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        built = list(pool.map(one, sources))
    for stale in KERNEL_DIR.glob("*.hsaco"):
        if stale not in built:
            stale.unlink()
    return built


def build_static_objects() -> list[Path]:
    """HIP sources of the library itself (host code plus static device
    arrays, ``csrc/static/*.hip``), compiled by hipcc into objects that are
    linked into ``libtrexhip.so``."""
    OBJ_DIR.mkdir(exist_ok=True)
    objs = []
    for src in sorted((CSRC / "static").glob("*.hip")):
        obj = OBJ_DIR / (src.stem + ".o")
        _run([ROCM / "bin" / "hipcc", "-x", "hip", "-c", "-fPIC", *HIP_DEVICE_FLAGS, "-Wall", "-Wextra",
              "-Wno-unused-parameter", f"-I{ROOT / 'include'}", f"-I{CSRC}", "-o", obj, src])
        objs.append(obj)
    return objs


def build_runtime() -> Path:
    sources = sorted(CSRC.glob("*.cpp"))
    objs = build_static_objects()
    tmp = LIB.with_suffix(".so.tmp")
    _run(["g++", *HOST_FLAGS, "-shared", "-o", tmp, *sources, *objs,
          f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}", "-lpthread", "-ldl"])
    os.replace(tmp, LIB)
    return LIB


def build_all() -> None:
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=2) as pool:
        lib = pool.submit(build_runtime)
        build_kernels()
        lib.result()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

"""Build the MI355X (gfx950) kernel library in-tree.

Compiles every ``csrc/*.hip`` with hipcc for gfx950 into objects under
``_build/`` and links ``libceed_amd/lib/libceed-mi355.so``, plus ``libceed-mi355-kernels.so`` (the
kernel ABI alone, for linking the backend into a libCEED build).  The host-side C
layer (``host/*.c``, the libCEED-API core and the /gpu/hip/mi355 backend) is
linked into the same library when present.

Usage: ``python -m libceed_amd.build [-j N]``
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
BUILD = ROOT / "_build"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libceed-mi355.so"
KERNELS_LIB = LIB_DIR / "libceed-mi355-kernels.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-std=c++17",
    # no packed FP32 instructions: some of their forms have given wrong results on MI355X
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{PKG / 'csrc'}",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", f"-I{ROOT / 'include'}", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def _needs(obj: Path, src: Path, deps: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(p.stat().st_mtime > t for p in [src, *deps])


def build(jobs: int | None = None, force: bool = True, verbose: bool = False) -> Path:
    """Compile all sources and link the shared library; returns its path."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    headers = sorted((PKG / "csrc").glob("*.h*")) + sorted((ROOT / "include").rglob("*.h"))
    hip_srcs = sorted((PKG / "csrc").glob("*.hip"))
    c_srcs = sorted((PKG / "host").glob("*.c")) if (PKG / "host").exists() else []
    jobs = jobs or min(16, os.cpu_count() or 4)

    tasks = []
    objs = []
    for src in hip_srcs:
        obj = BUILD / (src.stem + ".o")
        objs.append(obj)
        if force or _needs(obj, src, headers):
            tasks.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in c_srcs:
        obj = BUILD / ("host_" + src.stem + ".o")
        objs.append(obj)
        if force or _needs(obj, src, headers):
            tasks.append(["gcc", *C_FLAGS, "-c", str(src), "-o", str(obj)])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in tasks]:
            if verbose:
                print(" ".join(cmd[:1] + cmd[-3:]), flush=True)
            fut.result()
    link = ["-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *link])
    os.replace(tmp, LIB)
    # kernels only (no libCEED API core): what a libCEED build links the backend against
    kobjs = [o for o in objs if not o.name.startswith("host_")]
    tmp = KERNELS_LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, kobjs), *link])
    os.replace(tmp, KERNELS_LIB)
    build_c_programs()
    return LIB


def build_c_programs() -> list[Path]:
    """C programs written against the libCEED C API (tests/c/*.c), linked to the library."""
    src_dir = ROOT / "tests" / "c"
    out = []
    if not src_dir.exists():
        return out
    bin_dir = src_dir / "bin"
    bin_dir.mkdir(exist_ok=True)
    for src in sorted(src_dir.glob("*.c")):
        exe = bin_dir / src.stem
        _run(["gcc", "-O2", "-std=c11", "-Wall", f"-I{ROOT / 'include'}", f"-I{src_dir}", str(src.resolve()), f"-L{LIB_DIR}", "-lceed-mi355",
              "-Wl,-rpath,$ORIGIN/../../../libceed_amd/lib", "-lm", "-o", str(exe)])
        out.append(exe)
    return out


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    path = build(jobs=j, force="--incremental" not in sys.argv, verbose=True)
    print(path)

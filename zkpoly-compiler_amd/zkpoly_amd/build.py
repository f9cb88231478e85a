"""In-tree build of the HIP libraries for gfx950 (hipcc, no CMake).

Libraries land in <package>/lib/, where Libs (runtime.py) and the reference's
loader convention ($ZKPOLY_COMPILER_PROJECT_ROOT/lib/<name>) look for them.
Every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

from .runtime import PACKAGE_ROOT

CSRC = os.path.join(PACKAGE_ROOT, "csrc")
LIB = os.path.join(PACKAGE_ROOT, "lib")
ARCH = os.environ.get("ZKP_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

# MSM window sizes shipped prebuilt (libmsm_bn254_<w>_<w>_false.so); other
# configurations are built on demand by msm.ensure_msm_lib().
PREBUILT_MSM_WINDOWS = (12, 13, 14, 15, 16, 17, 18)


def hipcc() -> str:
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise FileNotFoundError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")


def build_shared(sources: list[str], out_name: str, defines: tuple[str, ...] = (), extra: tuple[str, ...] = ()) -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, out_name)
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", *[f"-D{d}" for d in defines], *extra, *sources, "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_ntt() -> str:
    return build_shared([os.path.join(CSRC, "ntt", "ntt.hip")], "libntt.so")


def build_intpeak() -> str:
    os.makedirs(os.path.join(PACKAGE_ROOT, "build"), exist_ok=True)
    out = os.path.join(PACKAGE_ROOT, "build", "intpeak")
    _run([hipcc(), *HIP_FLAGS, os.path.join(CSRC, "bench", "intpeak.hip"), "-o", out])
    return out


def targets() -> list:
    return [build_ntt, build_intpeak]


def build_all(jobs: int | None = None) -> list[str]:
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(t) for t in targets()]
        return [f.result() for f in futs]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)

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

# MsmConfig (window, target_windows) pairs shipped prebuilt as
# libmsm_bn254_<w>_<t>_false.so: windows 12..18 without precomputation
# (target = ceil(254 / w)) and with one precomputed table (target = half).
# Other configurations are built on demand by msm.ensure_msm_lib().
def _prebuilt_msm_configs():
    out = []
    for w in range(12, 19):
        windows = -(-254 // w)
        out.append((w, windows))
        out.append((w, -(-windows // 2)))
    return tuple(sorted(set(out)))


PREBUILT_MSM_CONFIGS = _prebuilt_msm_configs()


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


def build_poly() -> str:
    return build_shared([os.path.join(CSRC, "poly", "poly.hip"), os.path.join(CSRC, "poly", "permute.hip")],
                        "libpoly.so")


def build_msm_core() -> str:
    return build_shared([os.path.join(CSRC, "msm", "msm.hip")], "libzkp_msm.so")


def msm_lib_name(window: int, target: int, debug: bool = False) -> str:
    """Library name the reference derives from an MsmConfig (core/src/msm.rs:100-111)."""
    return f"libmsm_bn254_{window}_{target}_{'true' if debug else 'false'}.so"


def build_msm_abi(window: int, target: int, debug: bool = False) -> str:
    """Host-only wrapper binding msm / msm_precompute to one configuration."""
    os.makedirs(LIB, exist_ok=True)
    name = msm_lib_name(window, target, debug)
    out = os.path.join(LIB, name)
    tmp = out + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [
        shutil.which("g++") or "g++",
        "-O2",
        "-fPIC",
        "-shared",
        "-std=c++17",
        "-D__HIP_PLATFORM_AMD__",
        f"-I{rocm}/include",
        f"-DZKP_MSM_WINDOW={window}",
        f"-DZKP_MSM_TARGET={target}",
        f"-DZKP_MSM_DEBUG={1 if debug else 0}",
        os.path.join(CSRC, "msm", "msm_abi.cpp"),
        f"-L{LIB}",
        "-lzkp_msm",
        "-Wl,-rpath,$ORIGIN",
        "-o",
        tmp,
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_fused_codegen() -> str:
    """The C++ ArithGraph -> HIP generator (host-only library)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libzkp_fused_codegen.so")
    tmp = out + ".tmp"
    _run([shutil.which("g++") or "g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall",
          os.path.join(CSRC, "fused", "codegen.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_fused_demo() -> str:
    """Prebuilt fused-kernel library (the workload's a*b + gamma*c - z and test graphs)."""
    from . import fused
    from .runtime import Libs

    return fused.build_library(Libs(), fused.demo_kernels(), fused.DEMO_DIR)


def build_fused_large() -> str:
    """Prebuilt compiler-scale fused kernels (seeded random 800-vertex graphs) for the tests."""
    from . import fused
    from .runtime import Libs

    return fused.build_library(Libs(), fused.large_kernels(), fused.LARGE_DIR)


def build_fused_pipeline() -> str:
    """Host-only driver of the pipelined fused kernel (HIP runtime API only)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libzkp_fused_pipeline.so")
    tmp = out + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run([shutil.which("g++") or "g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
          f"-I{rocm}/include", os.path.join(CSRC, "fused", "pipeline.cpp"), f"-L{rocm}/lib", "-lamdhip64", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_serve() -> str:
    """Host-only serving engine; links libntt.so and libzkp_msm.so."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libzkp_serve.so")
    tmp = out + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run([shutil.which("g++") or "g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
          f"-I{rocm}/include", os.path.join(CSRC, "serve", "serve.cpp"), f"-L{LIB}", "-lntt", "-lzkp_msm",
          f"-L{rocm}/lib", "-lamdhip64", "-lpthread", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_pinned_pool() -> str:
    """Host-only buddy allocator over pinned host slabs (HIP runtime API only)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libzkp_pinned_pool.so")
    tmp = out + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run([shutil.which("g++") or "g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
          f"-I{rocm}/include", os.path.join(CSRC, "runtime", "pinned_pool.cpp"), f"-L{rocm}/lib", "-lamdhip64",
          "-o", tmp])
    os.replace(tmp, out)
    return out


def targets() -> list:
    return [build_ntt, build_poly, build_msm_core, build_fused_codegen, build_fused_pipeline, build_pinned_pool]


def build_all(jobs: int | None = None) -> list[str]:
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(t) for t in targets()]
        out = [f.result() for f in futs]
        # per-configuration MSM wrappers link against libzkp_msm.so; the fused
        # demo library needs the generator
        futs = [ex.submit(build_msm_abi, w, t, False) for (w, t) in PREBUILT_MSM_CONFIGS]
        futs.append(ex.submit(build_fused_demo))
        futs.append(ex.submit(build_fused_large))
        futs.append(ex.submit(build_serve))
        out += [f.result() for f in futs]
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)

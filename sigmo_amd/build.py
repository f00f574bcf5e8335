"""Build recipe for the native libraries (no cmake/ninja needed).

* ``sigmo_amd/libsigmo_hip.so`` -- the product: gfx950 HIP kernels + C ABI + host
  generator/.dat I/O.  HIP sources are compiled by ``hipcc --offload-arch=gfx950`` with
  packed-FP32 instructions disabled; host-only C++ is compiled by ``g++``.
* ``oracle/liboracle.so`` -- the CPU restatement used by the tests and the CPU baseline.
* ``sigmo_amd/sigmo_cli`` -- a C++ driver with the reference CLI, linked to the product.
* ``sigmo_amd/_build_san/`` -- the host sources and the oracle under ASan + UBSan, and the
  serving scheduler under TSan in a stand-alone harness (run by ``tests/test_sanitizers.py``).

Run ``python -m sigmo_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "sigmo_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsigmo_hip.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"
CLI = PKG / "sigmo_cli"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists() else "hipcc")
CXX = os.environ.get("CXX", "g++")
CLANGXX = "/opt/rocm/llvm/bin/clang++" if Path("/opt/rocm/llvm/bin/clang++").exists() else "clang++"
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-march=x86-64-v2"]

HIP_SOURCES = ["kernels.hip", "capi.hip", "batched.hip"]
HOST_SOURCES = ["host_graphs.cpp", "generator.cpp", "serve.cpp", "stream.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print("+", " ".join(str(c) for c in cmd), flush=True)
    proc = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(str(c) for c in cmd)}")
    # the host pass of hipcc warns about the GPU-only target feature; keep other output
    noise = "not a recognized feature for this target"
    out = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if noise not in l and l.strip())
    if verbose and out:
        print(out)


def build_product(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        _run([CXX, *HOST_FLAGS, "-c", CSRC / src, "-o", obj], verbose)
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", CSRC / src, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([CXX, "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wextra",
          ROOT / "oracle" / "oracle.cpp", "-o", tmp], verbose)
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_cli(verbose: bool = False) -> Path | None:
    src = PKG / "cli" / "sigmo_cli.cpp"
    if not src.exists():
        return None
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I", ROOT / "include", "-I", PKG / "include", src, "-o", CLI,
          "-L", PKG, "-lsigmo_hip", f"-Wl,-rpath,$ORIGIN"], verbose)
    return CLI


# ---- sanitizer builds (host code only; tests/test_sanitizers.py runs them) -------------------
SAN = PKG / "_build_san"
SAN_LIB = SAN / "libsigmo_hip.so"
SAN_ORACLE = SAN / "liboracle.so"
TSAN_HARNESS = SAN / "serve_schedule_race"
ASAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def build_asan(verbose: bool = False) -> None:
    """The product library with its host sources (.dat I/O, generator, serving scheduler, stream
    loop) and the oracle under AddressSanitizer + UndefinedBehaviorSanitizer (g++).  The HIP
    objects of the product build are linked in unchanged: their device code and host glue are not
    instrumented.  The CPU test suite runs against these with the sanitizer runtime preloaded."""
    SAN.mkdir(exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = SAN / (src + ".o")
        _run([CXX, "-fPIC", "-std=c++17", "-Wall", *ASAN_FLAGS, "-c", CSRC / src, "-o", obj], verbose)
        objs.append(obj)
    objs += [BUILD / (src + ".o") for src in HIP_SOURCES]
    _run([CXX, "-shared", *ASAN_FLAGS, *objs, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", SAN_LIB], verbose)
    _run([CXX, "-fopenmp", "-fPIC", "-shared", "-std=c++17", *ASAN_FLAGS, ROOT / "oracle" / "oracle.cpp", "-o", SAN_ORACLE], verbose)


def build_tsan(verbose: bool = False) -> None:
    """The serving scheduler (serve.cpp, host_graphs.cpp) under ThreadSanitizer in a stand-alone
    harness (tests/native/serve_schedule_race.cpp) that drives the scheduled and the online form
    from several threads with a stand-in executor."""
    # clang's runtime: it intercepts pthread_cond_clockwait (std::condition_variable::wait_until),
    # which GCC 11's libtsan does not, so GCC's TSan misses the unlock inside a timed wait and
    # reports the mutex as locked twice
    SAN.mkdir(exist_ok=True)
    _run([CLANGXX, "-std=c++17", "-fsanitize=thread", "-g", "-O1", "-pthread", ROOT / "tests" / "native" / "serve_schedule_race.cpp",
          CSRC / "serve.cpp", CSRC / "host_graphs.cpp", "-L", PKG, "-lsigmo_hip", f"-Wl,-rpath,{PKG}", "-o", TSAN_HARNESS], verbose)


def build_all(verbose: bool = False) -> None:
    build_product(verbose)
    build_oracle(verbose)
    build_cli(verbose)
    build_asan(verbose)
    build_tsan(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
    print("built", LIB, ORACLE_LIB)

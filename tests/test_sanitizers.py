"""CPU runs under the sanitizers (built by sigmo_amd/build.py: build_asan, build_tsan).

* The CPU test suite (everything not marked gpu or sanitize) against the product library whose
  host sources -- .dat parsing, the generator, the serving scheduler, the stream loop -- and the
  oracle are built with -fsanitize=address,undefined (UBSan errors fatal), the ASan runtime and
  libstdc++ preloaded into the Python process.
* The serving scheduler under ThreadSanitizer: tests/native/serve_schedule_race.cpp drives the
  scheduled and the online form from several worker and client threads with a stand-in executor.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

from sigmo_amd import build

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.sanitize


def _gcc_runtime(name: str) -> str:
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()


@pytest.mark.skipif(not build.SAN_LIB.exists() or not build.SAN_ORACLE.exists(), reason="sanitizer build missing (build_asan)")
def test_cpu_suite_under_address_and_undefined_sanitizers(tmp_path):
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{_gcc_runtime('libasan.so')} {_gcc_runtime('libstdc++.so.6')}",
        "ASAN_OPTIONS": f"detect_leaks=0:log_path={tmp_path / 'asan'}",
        "UBSAN_OPTIONS": f"print_stacktrace=1:halt_on_error=1:log_path={tmp_path / 'ubsan'}",
        "SIGMO_HIP_LIB": str(build.SAN_LIB),
        "SIGMO_ORACLE_LIB": str(build.SAN_ORACLE),
    })
    p = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests"), "-m", "not gpu and not sanitize", "-q", "-x",
                        "-p", "no:cacheprovider"], capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    reports = "".join(f.read_text() for f in tmp_path.iterdir())
    assert p.returncode == 0 and not reports, p.stdout[-3000:] + p.stderr[-3000:] + reports[-6000:]


@pytest.mark.skipif(not build.TSAN_HARNESS.exists(), reason="sanitizer build missing (build_tsan)")
def test_serving_scheduler_under_thread_sanitizer():
    p = subprocess.run([str(build.TSAN_HARNESS)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0 and "ThreadSanitizer" not in p.stderr, p.stdout + p.stderr[-6000:]
    assert "serve_schedule_race: ok" in p.stdout

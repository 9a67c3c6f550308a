"""The library's host C++ (bricking, partitioning, file formats, synthetic
inputs, camera projection, mesh split, BVH build) and the CPU oracle, built
with AddressSanitizer and UndefinedBehaviorSanitizer by g++/gcc (no HIP code
is compiled) and driven through their entry points and error paths by
tests/sanitize/host_check.cpp and oracle_check.c.  Any sanitizer report
fails the run (-fno-sanitize-recover, exit code of the report)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "aloha_amd" / "csrc"
HERE = ROOT / "tests" / "sanitize"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def _env():
    env = dict(os.environ)
    # the sanitizer runtime need not come first in the load order (an
    # unrelated preloaded library may); leak checking stays on
    env["ASAN_OPTIONS"] = "verify_asan_link_order=0:detect_leaks=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    return env


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(map(str, cmd))}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs g++ and gcc")
def test_host_code_and_oracle_under_asan_ubsan(tmp_path):
    host = tmp_path / "host_check"
    srcs = sorted(str(p) for p in CSRC.glob("host_*.cpp"))
    _run(["g++", "-std=c++17", *SAN, f"-I{ROOT / 'include'}", f"-I{CSRC}",
          str(HERE / "host_check.cpp"), *srcs, "-o", str(host), "-lpthread"])
    orc = tmp_path / "oracle_check"
    _run(["gcc", "-std=c11", *SAN, "-fopenmp", "-ffp-contract=off", str(HERE / "oracle_check.c"),
          "-o", str(orc), "-lm"])
    out = _run([str(host), str(tmp_path)], env=_env(), timeout=600)
    assert "host ok" in out.stdout
    out = _run([str(orc), str(tmp_path)], env=_env(), timeout=600)
    assert "oracle ok" in out.stdout

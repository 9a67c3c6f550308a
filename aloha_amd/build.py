"""Build the aloha_amd shared library (HIP for gfx950 + host C++) in-tree.

    python -m aloha_amd.build          # -> aloha_amd/libaloha_amd.so

Every HIP source is compiled with contraction off (the kernels reproduce the
CPU restatement's rounding) and without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libaloha_amd.so"
TOOLS = PKG / "tools"
BIN = PKG / "bin"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ALOHA_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-result",
]
HIP_ONLY = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fgpu-rdc" if os.environ.get("ALOHA_RDC") else "",
]


def _sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, "-x", "hip", *COMMON, *[f for f in HIP_ONLY if f], "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, "-x", "c++", *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(srcs)} sources")
    build_tools(verbose)
    return LIB


def build_tools(verbose: bool = True) -> None:
    """Command-line tools over the C ABI (aloha_amd/bin/), linked to the
    in-tree library with an $ORIGIN rpath."""
    BIN.mkdir(exist_ok=True)
    srcs = sorted(TOOLS.glob("*.cpp"))

    def one(src: Path):
        exe = BIN / src.stem
        cmd = [HIPCC, "-x", "hip", *COMMON, f"--offload-arch={ARCH}", str(src), "-o", str(exe),
               f"-L{PKG}", "-laloha_amd", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"tool build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return exe

    with ThreadPoolExecutor(max_workers=max(1, min(8, len(srcs)))) as ex:
        exes = list(ex.map(one, srcs))
    if verbose:
        print(f"built {len(exes)} tools in {BIN.relative_to(ROOT)}")


def build_oracle(verbose: bool = True) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    out = ROOT / "oracle" / "libaloha_oracle.so"
    if verbose:
        print(f"built {out.relative_to(ROOT)}")
    return out


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()

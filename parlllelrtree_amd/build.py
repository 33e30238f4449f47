"""Build recipe of libprt.so (HIP for gfx950) and of the CPU oracle library.

``python -m parlllelrtree_amd.build`` (or ``build()``) compiles every source
under ``csrc/`` with hipcc into ``_build/`` and links ``libprt.so`` next to
this file, then the command-line drivers under ``drivers/`` (host C++ over
the C-ABI only) into ``bin/``, linked against that library; the oracle (test infrastructure, never linked into libprt) is
built by ``oracle/Makefile`` into ``oracle/_build/liboracle.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libprt.so"
DRIVERS = PKG / "drivers"
BIN = PKG / "bin"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", f"{ROCM}/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}"]
# HIP translation units: gfx950 code objects; packed-FP32 instruction
# selection stays off.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n"
                           f"{res.stdout}\n{res.stderr}")


def _compile(src: Path) -> Path:
    obj = OUT / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_lib(jobs: int = 8) -> Path:
    OUT.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_drivers(jobs: int = 4) -> list[Path]:
    """prt_bu_rtree, prt_td_rtree, prt_overlay: host programs that include
    only prt.h and the HIP runtime, found next to libprt.so by rpath."""
    BIN.mkdir(exist_ok=True)

    def one(src: Path) -> Path:
        exe = BIN / f"prt_{src.stem}"
        tmp = exe.with_suffix(".tmp")
        _run([HIPCC, "-O2", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
              f"-I{ROOT / 'include'}", f"-I{ROCM}/include", f"-I{DRIVERS}", str(src),
              "-o", str(tmp), f"-L{PKG}", "-lprt", "-Wl,-rpath,$ORIGIN/..",
              f"-L{ROCM}/lib", "-lamdhip64"])
        os.replace(tmp, exe)
        return exe

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(one, sorted(DRIVERS.glob("*.cpp"))))


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "liboracle.so"


def build() -> None:
    build_lib()
    build_drivers()
    build_oracle()


if __name__ == "__main__":
    build()
    print(f"built {LIB}", file=sys.stderr)

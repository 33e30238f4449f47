"""CPU: the command-line drivers build, link against libprt.so through the
C-ABI only (every prt_* symbol they import is declared in include/prt.h),
and reject bad invocations before touching a device."""
import re
import subprocess
from pathlib import Path

import pytest

from parlllelrtree_amd import build

ROOT = Path(__file__).resolve().parents[1]
NAMES = ("prt_bu_rtree", "prt_td_rtree", "prt_overlay")


@pytest.fixture(scope="module")
def bins():
    if not build.LIB.exists():
        pytest.skip("libprt.so is not built")
    if not all((build.BIN / n).exists() for n in NAMES):
        build.build_drivers()
    return {n: build.BIN / n for n in NAMES}


def declared():
    text = (ROOT / "include" / "prt.h").read_text()
    return set(re.findall(r"\b(prt_[a-z0-9_]+)\s*\(", text))


def test_drivers_import_only_declared_symbols(bins):
    decl = declared()
    for name, exe in bins.items():
        out = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True,
                             text=True, check=True).stdout
        used = {ln.split()[-1].split("@")[0] for ln in out.splitlines() if " prt_" in ln}
        assert used, name
        assert used <= decl, (name, used - decl)


@pytest.mark.parametrize("name", NAMES)
def test_usage_and_missing_file(bins, tmp_path, name):
    r = subprocess.run([str(bins[name])], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    args = [tmp_path / "missing.txt"] * (2 if name == "prt_overlay" else 1)
    r = subprocess.run([str(bins[name]), *map(str, args)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "failed" in r.stderr

"""The nissa-side shims (nissa_amd/shims/) compile against include/nissa_amd.h and
against the names and signatures of the nissa interface they use (declared in
tests/shim_mock/ for this check only), and every C-ABI function they call is
exported by the library.  CPU only: no device calls."""
import os
import re
import shutil
import subprocess

import pytest

import nissa_amd as nq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIMS = os.path.join(ROOT, "nissa_amd", "shims")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_shims_compile(tmp_path):
    out = tmp_path / "nissa_shims.o"
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-c",
           "-I", os.path.join(ROOT, "tests", "shim_mock"), "-I", os.path.join(ROOT, "include"),
           os.path.join(SHIMS, "nissa_shims.cpp"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.stat().st_size > 0


def test_shims_call_only_exported_functions():
    src = "".join(open(os.path.join(SHIMS, f)).read() for f in ("nq_bridge.hpp", "nissa_shims.cpp"))
    called = set(re.findall(r"\b(nq_[a-z0-9_]+)\s*\(", src))
    called = {n for n in called if n not in ("nq_bridge", "nq_shims")}
    assert len(called) >= 20, sorted(called)
    lib = nq.lib()
    assert not [n for n in called if not hasattr(lib, n)]

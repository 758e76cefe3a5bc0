"""The C++ facade header compiles for a plain host C++17 caller."""
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_facade_header_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-I", str(ROOT / "include"),
                        str(ROOT / "tests" / "cpp" / "facade_demo.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

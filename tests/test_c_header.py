"""The public headers are plain C: a C99 translation unit including both
compiles with -Wall -Wextra -Werror -pedantic, and so do the C consumers of
the ABI (tests/c_abi/*.c)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fsyntax-only",
         "-I", os.path.join(ROOT, "include")]


@pytest.fixture
def gcc():
    g = shutil.which("gcc")
    if not g:
        pytest.skip("no C compiler")
    return g


def test_headers_compile_as_strict_c99(gcc, tmp_path):
    src = tmp_path / "h.c"
    src.write_text('#include "phylo_hip.h"\n#include "phylo_synth.h"\n'
                   "int main(void) { return PHY_OK; }\n")
    r = subprocess.run([gcc, *FLAGS, str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("name", ["abi_demo.c", "aa_demo.c", "sankoff_demo.c"])
def test_c_consumers_compile(gcc, name):
    r = subprocess.run([gcc, *FLAGS, os.path.join(ROOT, "tests", "c_abi", name)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

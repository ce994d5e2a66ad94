"""The OCaml binding (ocaml/): the stub file compiles as strict C99 against
test doubles of OCaml's headers (tests/mock_caml/caml: OCaml's names,
signatures and value layout), every `external` of the two NodeData modules
names a stub of the same arity, the stubs take at most five arguments (no
bytecode/native pair needed), and the stub object references nothing but
the C ABI, OCaml's runtime and libc.  Running the stubs is the GPU test
tests/test_gpu_ocaml_stubs.py."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STUBS = os.path.join(ROOT, "ocaml", "phylo_hip_stubs.c")
MLS = [os.path.join(ROOT, "ocaml", f) for f in ("nonAdditive_c.ml", "likelihood_c.ml")]
INC = ["-I", os.path.join(ROOT, "tests", "mock_caml"), "-I", os.path.join(ROOT, "include")]


@pytest.fixture
def gcc():
    g = shutil.which("gcc")
    if not g:
        pytest.skip("no C compiler")
    return g


def _externals():
    out = {}
    for path in MLS:
        src = open(path).read()
        for m in re.finditer(r'external\s+(\w+)\s*:\s*(.*?)\s*=\s*"(\w+)"', src, re.S):
            typ = m.group(2)
            out.setdefault(m.group(3), set()).add(len(re.findall(r"->", typ)))
    return out


def _stubs():
    src = open(STUBS).read()
    out = {}
    for m in re.finditer(r"^value\s+(\w+)\(([^)]*)\)\s*\{", src, re.M):
        args = [a for a in m.group(2).split(",") if a.strip()]
        out[m.group(1)] = len(args)
    return out


def test_stubs_compile_as_strict_c99(gcc):
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-fsyntax-only", *INC, STUBS], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_every_external_has_a_stub_of_its_arity():
    ext, stubs = _externals(), _stubs()
    assert len(ext) >= 20
    for name, arities in ext.items():
        assert name in stubs, f"external {name} has no stub"
        assert arities == {stubs[name]}, (name, arities, stubs[name])
    assert max(stubs.values()) <= 5
    # the MlModel externals of the reference bind to these stubs unchanged
    # (lib/mlModel.ml:73-90)
    for name, n in (("likelihood_CAML_diagonalize_gtr", 3), ("likelihood_CAML_diagonalize_sym", 2),
                    ("likelihood_CAML_compose_gtr", 4), ("likelihood_CAML_compose_sym", 3)):
        assert stubs.get(name) == n, name


def test_stub_object_references_only_the_abi_runtime_and_libc(gcc, tmp_path):
    obj = tmp_path / "stubs.o"
    r = subprocess.run([gcc, "-std=c99", "-O2", "-c", *INC, STUBS, "-o", str(obj)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True).stdout
    undef = [line.split()[-1] for line in nm.splitlines() if line.strip()]
    header = open(os.path.join(ROOT, "include", "phylo_hip.h")).read()
    libc = {"malloc", "calloc", "free", "memcpy", "memset", "floor", "__stack_chk_fail"}
    for sym in undef:
        if sym.startswith("phy_"):
            assert re.search(r"\b" + sym + r"\(", header), f"{sym} is not in phylo_hip.h"
        else:
            assert sym.startswith("caml_") or sym in libc, sym


def test_driver_builds_against_the_mock_runtime():
    # the GPU test runs it; here: it links (the stubs and the runtime agree)
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "mock_caml")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.access(os.path.join(ROOT, "tests", "mock_caml", "stub_driver"), os.X_OK)

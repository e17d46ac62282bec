"""The bindings INTEGRATION.md shows a maintainer (host only): its C snippets
compile against include/wtc.h inside a small prelude that declares what the
reference's `main` has in scope, its raw ctypes stub of the tc_GPU seam
matches the C structs' layout, and the names its Python example uses exist
in the binding."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

import wtc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")
INCLUDE = os.path.join(ROOT, "include")


def _blocks(lang):
    text = open(DOC).read()
    return re.findall(r"```%s\n(.*?)```" % lang, text, flags=re.S)


# What the snippets take from their surroundings: the reference's loop
# variables and timer (tc.cu's main) for the first, a caller's graphs and
# buffers for the second.
PRELUDE = {
    0: ("typedef uint32_t UINT_t;\n"
        "static double get_seconds(void) { return 0.0; }\n"
        "void snippet(const void *graph, UINT_t loop_cnt, uint32_t spread, uint32_t adjacency_matrix_len) {\n"),
    1: ("void snippet(const wtc_big_graph *oriented, uint32_t n, uint64_t m, uint64_t W, uint64_t w0, uint64_t w1,\n"
        "             uint64_t seed, const uint32_t *pairs, uint64_t npairs, const wtc_big_graph *const *graphs,\n"
        "             size_t k, uint64_t *counts, uint64_t tag, wtc_big_graph *g) {\n"),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("i", [0, 1])
def test_c_snippets_compile_against_the_header(tmp_path, i):
    blocks = _blocks("c")
    assert len(blocks) == 2, "INTEGRATION.md's C snippets changed: update PRELUDE"
    src = ("#include <stdio.h>\n#include <stdlib.h>\n#include <stdint.h>\n#include \"wtc.h\"\n"
           + PRELUDE[i] + blocks[i].replace('#include "wtc.h"', "") + "\n}\n")
    f = tmp_path / f"snippet{i}.c"
    f.write_text(src)
    r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Werror=implicit-function-declaration",
                        "-Werror=incompatible-pointer-types", "-Werror=int-conversion", "-I", INCLUDE, str(f)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_raw_ctypes_stub_matches_the_c_layout():
    """The stub's Graph32 and GpuTime have the sizes and field offsets of
    wtc_graph and wtc_gpu_time (the module's own structs mirror the header)."""
    stub = [b for b in _blocks("python") if "C.CDLL" in b]
    assert len(stub) == 1
    ns = {}
    code = stub[0].replace('lib = C.CDLL("wedge-parallel-triangle-counting_amd/libwtc.so")', "lib = None")
    code = "\n".join(line for line in code.splitlines() if not line.startswith("lib."))
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    assert C.sizeof(ns["Graph32"]) == C.sizeof(wtc._Graph32)
    for name, _ in ns["Graph32"]._fields_:
        assert getattr(ns["Graph32"], name).offset == getattr(wtc._Graph32, name).offset
    assert C.sizeof(ns["GpuTime"]) == C.sizeof(wtc.GpuTime)


def test_python_example_names_exist():
    example = [b for b in _blocks("python") if "import wtc" in b]
    assert len(example) == 1
    for name in re.findall(r"wtc\.([A-Za-z_][A-Za-z0-9_]*)", example[0]):
        assert hasattr(wtc, name), name
    for meth in ("upload",):
        assert hasattr(wtc.Context, meth)
    for meth in ("count", "count_arrow", "count_baseline"):
        assert hasattr(wtc.DeviceGraph, meth), meth

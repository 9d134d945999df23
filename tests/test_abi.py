"""The C ABI of include/mcs.h without a GPU: the library loads, exports every
declared function, the Python binding declares each one, and the ctypes
structures have the header's sizes and field offsets (measured by a small C
program compiled with gcc against the header)."""
import ctypes
import os
import re
import subprocess

import pytest

from mcstreamer_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcs.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char \*)\s*(mcs_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_functions():
    names = declared_functions()
    assert "mcs_advance" in names and "mcs_swarm_batch_step" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    path = _lib.LIB_PATH
    if not os.path.exists(path):
        pytest.skip("libmcs.so not built")
    so = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(so, n)]
    assert not missing, missing


def test_binding_declares_every_function():
    names = set(declared_functions()) - {"mcs_last_error"}
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


STRUCTS = {
    "mcs_field": _lib.Field,
    "mcs_step_stats": _lib.StepStats,
    "mcs_moments": _lib.Moments,
    "mcs_resample_stats": _lib.ResampleStats,
    "mcs_mg_config": _lib.MGConfigC,
    "mcs_swarm_params": _lib.SwarmParams,
    "mcs_swarm_tally": _lib.SwarmTally,
}


def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mcs.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            c, f, v = ln.split()
            got[(c, f)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)

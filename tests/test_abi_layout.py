"""The ctypes mirrors of the C ABI's structs have the layout the header
gives them: every field's offset and every struct's size, as a C compiler
lays them out from include/nido_amd.h (CPU only)."""
import os
import shutil
import subprocess

import pytest

import nido_amd as na

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

STRUCTS = {
    "nido_phase_stats": na.PhaseStats,
    "nido_run_stats": na.RunStats,
    "nido_tuning": na.Tuning,
    "nido_serve_config": na.ServeConfig,
}


def _c_layout(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "nido_amd.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    return got


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_ctypes_structs_match_the_header(tmp_path):
    c = _c_layout(tmp_path)
    for cname, py in STRUCTS.items():
        assert c[(cname, "size")] == na.C.sizeof(py), cname
        for f, _ in py._fields_:
            assert c[(cname, f)] == getattr(py, f).offset, f"{cname}.{f}"

"""The ctypes mirrors of the C ABI's structs (parcel_hip.Param, Stats,
BatchStats) against include/parcel_hip.h: a small C program compiled with the
host compiler prints sizeof and offsetof of every field the binding declares,
and each must equal the ctypes layout.  A field added to the header but not to
the binding (or in a different place) would otherwise be read at the wrong
offset without an error."""
import os
import shutil
import subprocess

import pytest

import parcel_hip as ph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = (("parcel_hip_param", ph.Param), ("parcel_hip_stats", ph.Stats),
           ("parcel_hip_batch_stats", ph.BatchStats))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs the host C compiler")
def test_ctypes_structs_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "parcel_hip.h"', 'int main(void) {']
    for cname, py in STRUCTS:
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for cname, py in STRUCTS:
        assert got[(cname, "sizeof")] == ctypes_sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def ctypes_sizeof(t):
    import ctypes
    return ctypes.sizeof(t)

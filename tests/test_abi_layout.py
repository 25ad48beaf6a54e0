"""The ctypes mirrors of the C ABI's structures (sdft_amd/native.py) against
include/sdft_hip.h as a C compiler lays it out: every field's offset and each
structure's size, so a field added to one side only fails here on the CPU
instead of corrupting a call on the GPU."""
import ctypes as ct
import os
import shutil
import subprocess

import pytest

from sdft_amd import native

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [("sdft_hip_operator_desc", native.OperatorDesc), ("sdft_hip_davidson_params", native.DavidsonParams),
         ("sdft_hip_davidson_result", native.DavidsonResult), ("sdft_hip_snapshot_info", native.SnapshotInfo)]


def _c_layout(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "sdft_hip.h"', "int main(void) {"]
    for cname, py in PAIRS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-Wall", f"-I{os.path.join(REPO, 'include')}", str(src), "-o", str(exe)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-4000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout
    table = {}
    for line in out.splitlines():
        a, b, c = line.split()
        table[(a, b)] = int(c)
    return table


def test_ctypes_structures_match_c_layout(tmp_path):
    table = _c_layout(tmp_path)
    for cname, py in PAIRS:
        assert ct.sizeof(py) == table[(cname, "size")], cname
        for field, _ in py._fields_:
            assert getattr(py, field).offset == table[(cname, field)], (cname, field)

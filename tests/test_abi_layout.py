"""The Python mirrors of the C-ABI structs against the headers' own layout:
sizes and field offsets printed by a program compiled from include/ with gcc,
compared with the ctypes structures (a field missing from a header fails the
compile)."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = [("greenstein2002_amd._lib", "GrwConfig", "grw_config"),
           ("greenstein2002_amd._lib", "GrwStats", "grw_stats"),
           ("greenstein2002_amd.cell", "GrwProtocol", "GrwProtocol"),
           ("greenstein2002_amd.cell", "GrwRkStats", "GrwRkStats"),
           ("greenstein2002_amd.cell", "GrwTune", "GrwTune"),
           ("greenstein2002_amd.batch", "GrwBatchConfig", "grw_batch_config"),
           ("greenstein2002_amd.batch", "GrwCellSpec", "grw_cell_spec")]


def _mirrors():
    import importlib

    return [(getattr(importlib.import_module(mod), cls), cname) for mod, cls, cname in MIRRORS]


def _c_layout(tmp_path, mirrors) -> dict:
    """sizeof and offsetof of every mirrored field, from the headers."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "grw_fru.h"', '#include "grw_cell.h"',
             '#include "grw_batch.h"', "int main(void) {"]
    for T, cname in mirrors:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in T._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    lay = {}
    for line in out.splitlines():
        t, m, v = line.split()
        lay[(t, m)] = int(v)
    return lay


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_mirrors_match_the_headers(tmp_path):
    mirrors = _mirrors()
    lay = _c_layout(tmp_path, mirrors)
    for T, cname in mirrors:
        assert C.sizeof(T) == lay[(cname, "size")], cname
        for f in T._fields_:
            assert getattr(T, f[0]).offset == lay[(cname, f[0])], (cname, f[0])

"""The ctypes mirror of include/lqt.h agrees with the header as a C compiler
reads it: every struct has the same size and field offsets (compiled with
gcc here, no GPU needed), and every declared function takes as many
arguments as its ctypes binding passes."""
import ctypes as C
import os
import re
import subprocess

import pytest

import lqt1_amd as A
from lqt1_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lqt.h")

STRUCTS = {
    "lqt_params": L.Params,
    "lqt_pacing": L.Pacing,
    "lqt_geometry": L.Geometry,
    "lqt_stim_protocol": L.StimProtocol,
    "lqt_strand_desc": L.StrandDesc,
    "lqt_strand_result": L.StrandResult,
    "lqt_vw_desc": L.VWDesc,
    "lqt_vw_result": L.VWResult,
    "lqt_sweep_desc": L.SweepDesc,
    "lqt_sweep_stats": L.SweepStats,
}


def _header_text():
    txt = open(HEADER).read()
    return re.sub(r"/\*.*?\*/", "", txt, flags=re.S)


def _c_layout(tmp_path):
    """sizeof and offsetof of every field, as gcc lays out the header."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "lqt.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    lay = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        lay[(s, f)] = int(v)
    return lay


def test_struct_layouts_match_the_header(tmp_path):
    lay = _c_layout(tmp_path)
    for cname, py in STRUCTS.items():
        assert C.sizeof(py) == lay[(cname, "size")], cname
        for fname, _ in py._fields_:
            assert getattr(py, fname).offset == lay[(cname, fname)], (cname, fname)
    # every member the header declares is mirrored (no field left out)
    txt = _header_text()
    for cname, py in STRUCTS.items():
        body = re.search(r"typedef struct \{([^}]*)\}\s*" + cname + r"\s*;", txt, flags=re.S).group(1)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            for part in decl.split(","):
                m = re.search(r"(\w+)\s*(\[\s*\d+\s*\])?\s*$", part.strip())
                names.append(m.group(1))
        assert names == [f for f, _ in py._fields_], (cname, names)


def _prototypes():
    txt = _header_text()
    out = {}
    for m in re.finditer(r"\b(lqt_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", txt, flags=re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_bindings_pass_every_declared_argument():
    protos = _prototypes()
    assert len(protos) > 40
    lib = A.lib()
    for name, nargs in protos.items():
        f = getattr(lib, name)
        assert f.argtypes is not None, f"{name} has no ctypes signature"
        assert len(f.argtypes) == nargs, (name, len(f.argtypes), nargs)

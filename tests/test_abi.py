"""CPU checks of the boundary: the C-ABI library loads and exports every
entry point include/ia.h declares; the host mirror validates arguments
before anything reaches the device.  (No compute without a GPU.)"""
import ctypes as C
import os
import re
import subprocess

import pytest

from __graft_entry__ import ROOT, load_package


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ia.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ia_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_path():
    names = declared_symbols()
    for must in ["ia_create", "ia_destroy", "ia_mom_step", "ia_advect", "ia_project", "ia_cfl",
                 "ia_set_field", "ia_get_field", "ia_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    ia = load_package()
    L = ia.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_library_is_gfx950_only():
    ia = load_package()
    data = open(ia.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def _c_layout(tmp_path, structs):
    """sizeof and offsetof of every field, as a C compiler lays out include/ia.h."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ia.h"', "int main(void) {"]
    for cname, fields in structs.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for ln in out.splitlines():
        cname, f, v = ln.split()
        res[(cname, f)] = int(v)
    return res


def test_struct_layouts_match_header(tmp_path):
    """Every field offset and every struct size of the ctypes mirror equals
    what a C compiler gives for include/ia.h."""
    ia = load_package()
    mirror = {"ia_params": ia.IAParams, "ia_solve_stats": ia.IASolveStats, "ia_step_stats": ia.IAStepStats,
              "ia_learned_info": ia.IALearnedInfo}
    structs = {c: [f[0] for f in py._fields_] for c, py in mirror.items()}
    lay = _c_layout(tmp_path, structs)
    for cname, py in mirror.items():
        assert lay[(cname, "sizeof")] == C.sizeof(py), cname
        for f in structs[cname]:
            assert lay[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_host_rejects_unsupported_plugins():
    ia = load_package()
    with pytest.raises(ValueError):
        ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), η=-1.0)
    with pytest.raises(ValueError):
        ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), exitBC=True, g=(0.5, 0.0))
    with pytest.raises(KeyError):
        ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), λ="quick")

"""CPU checks of the boundary: the C-ABI library loads and exports every
entry point include/ia.h declares; the host mirror validates arguments
before anything reaches the device.  (No compute without a GPU.)"""
import ctypes as C
import os
import re

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


def test_params_struct_matches_header():
    ia = load_package()
    # D, N[3], dtype, perdir (6 ints), then 6 doubles (u_bc, g) ...
    assert C.sizeof(ia.IAParams) >= 6 * 4 + 6 * 8
    assert ia.IAParams.u_bc.offset % 8 == 0


def test_host_rejects_unsupported_plugins():
    ia = load_package()
    with pytest.raises(ValueError):
        ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), η=-1.0)
    with pytest.raises(ValueError):
        ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), exitBC=True)
    with pytest.raises(KeyError):
        ia.TwoPhaseSimulation((8, 8), (0.0, 0.0), λ="quick")

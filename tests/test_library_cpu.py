"""CPU-side checks of the product library: it loads without a GPU and exports
every entry point include/gfsph.h declares; host parameter mapping."""
import ctypes
import os
import re

import numpy as np

import gfsph
from gfsph import api, cases

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "gfsph.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gfsph_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(api.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_every_symbol():
    assert os.path.exists(gfsph.LIB), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(gfsph.LIB)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_params_struct_layout_matches_header():
    # gfsph_params: 12 int32, 3 doubles, 111 doubles, then the pore-pressure
    # column block: 4 int32 (nuaux, bc types, pwp0_type) and 3 doubles
    assert ctypes.sizeof(api.CParams) == 12 * 4 + 3 * 8 + 111 * 8 + 4 * 4 + 3 * 8 + 2 * 4 + 2 * 8
    assert api.CParams.dt.offset == 48 and api.CParams.cst.offset == 72
    assert api.CParams.nuaux.offset == 960 and api.CParams.pwp_bc_val.offset == 976
    assert api.CParams.xirel.offset == 992
    assert api.CParams.ic_k0.offset == 1000 and api.CParams.k0_active.offset == 1008
    assert api.CDem.Z.offset == 56


def test_material_constants_are_float32_rounded():
    p = cases.Params(const=cases.frictional(28.0, cv=0.01))
    c = p.cst()
    assert c[9] == float(np.float32(np.tan(np.radians(28.0))))
    assert c[13] == float(np.float32(np.pi ** 2 * 0.01 / 4))
    assert p.nfric == 5

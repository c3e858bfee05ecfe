"""The C-ABI library loads and exports every function declared in
include/openqcd_hip.h (no device calls: this runs without a GPU)."""
import ctypes as C
import os

import pyoqh


def test_library_exports_header_symbols():
    assert os.path.exists(pyoqh.LIBPATH), "run __graft_entry__.build() first"
    lib = C.CDLL(pyoqh.LIBPATH)
    names = pyoqh.header_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_names_reference_entry_points():
    names = set(pyoqh.header_functions())
    for n in ("Dw", "Dw_dble", "Dwhat", "Dwhat_dble", "Dw_blk", "sap_gcr", "fgcr", "sap",
              "blk_mres", "sw_term", "spinor_prod_dble", "acc_qflt", "add_qflt"):
        assert n in names

"""The C-ABI library loads and exports every entry point include/afx.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "afx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(afx_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "afx_cocg13_single" in names and "afx_solve_device" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    from actfemtide_amd import afx
    lib = ctypes.CDLL(afx.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from actfemtide_amd import afx
    assert sorted(afx.SIGNATURES) == declared()


def test_default_opts_without_gpu():
    from actfemtide_amd import afx
    o = afx.default_opts()
    assert o.precond == afx.PC["SSOR"] and o.tol == 1e-8 and o.nset == 1 and o.precision == afx.FP64


def declared_nn():
    src = open(os.path.join(ROOT, "include", "afx_nn.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(afx_nn_\w+)\s*\(", src, re.M)))


def test_nn_header_exports_and_binding():
    from actfemtide_amd import afx, nn
    names = declared_nn()
    assert "afx_nn_gemm_nt" in names and len(names) >= 15
    lib = ctypes.CDLL(afx.LIB_PATH)
    assert not [n for n in names if not hasattr(lib, n)]
    assert sorted(nn.SIGNATURES) == names

"""The C-ABI library loads and exports every symbol include/cdj.h declares (no GPU needed; no compute calls)."""
import ctypes
import os
import re

from cutanddisplacejulia_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cdj.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cdj_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_binding_declares_signatures_without_gpu():
    L = _lib.load()
    from cutanddisplacejulia_amd import learned
    learned._bind()
    for name in _lib.EXPORTS:
        assert getattr(L, name).argtypes is not None or name == "cdj_last_error"
    assert L.cdj_last_kernel_ms(None) == -1.0


def test_report_and_opts_layout():
    # cdj_opts: int32, int32, double, int32, double -> 32 bytes with natural alignment
    assert ctypes.sizeof(_lib.Opts) == 32
    # cdj_report: int32, int32, double x3, int64 -> 40 bytes
    assert ctypes.sizeof(_lib.Report) == 40

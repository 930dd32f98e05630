"""The C-ABI library loads without a GPU and exports every function that
include/codens.h declares; the ctypes records match the C layouts."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "codens.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(codens_\w+)\s*\(", txt)))


def test_library_exports_header():
    path = os.path.join(ROOT, "codens_amd", "libcodens_gpu.so")
    assert os.path.exists(path), "build the library first"
    lib = ctypes.CDLL(path)
    names = declared()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n


def test_python_bindings_cover_header():
    from codens_amd import _lib
    names = set(declared()) - {"codens_last_error"}
    assert names <= set(_lib.SIGNATURES), names - set(_lib.SIGNATURES)


def test_record_layouts():
    from codens_amd._lib import CodensBlock, CodensPhys, CodensScheme
    assert ctypes.sizeof(CodensBlock) == 64
    assert ctypes.sizeof(CodensPhys) == 72
    assert ctypes.sizeof(CodensScheme) == 16

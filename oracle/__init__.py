"""CPU checkers for the residual-and-update path.  TEST INFRASTRUCTURE ONLY.

Two libraries live here:

* ``liboracle.so``: the C restatement of the reference algorithm (oracle/*.c),
  each function citing the reference file:line it follows.
* ``_ref/libcodens_ref.so``: the reference's own Fortran routines compiled
  from the reference sources by oracle/Makefile, wrapped by ref_harness.f90.
  It validates the restatement.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package.  The product (codens_amd) never does.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libcodens_ref.so")
_DROPIN = os.path.join(HERE, "_ref", "libcodens_dropin.so")

_cache = {}


def _load(path):
    if path not in _cache:
        _cache[path] = ctypes.CDLL(path) if os.path.exists(path) else None
    return _cache[path]


def oracle_lib():
    lib = _load(_ORACLE)
    if lib is None:
        raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
    return lib


def ref_lib():
    """The flang build of the reference routines, or None where it was not built."""
    return _load(_REF)


def dropin_lib():
    """The reference harness relinked against the GPU drop-in modules of
    codens_amd/fortran (same entry points as ref_lib), or None.  The GPU
    library is loaded first through codens_amd._lib, which settles the HIP
    runtime of the process."""
    if not os.path.exists(_DROPIN):
        return None
    import codens_amd._lib  # noqa: F401
    return _load(_DROPIN)


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)

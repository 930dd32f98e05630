"""Loader of the gfx950 library and ctypes signatures of include/codens.h.

There is no fallback: if ``libcodens_gpu.so`` is missing or fails to load,
importing this module raises, so nothing silently runs elsewhere.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcodens_gpu.so")

c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class CodensBlock(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("npc", "id1", "id2", "jd1", "jd2", "kd1", "kd2", "ii1", "ii2", "jj1", "jj2", "kk1",
                 "kk2", "nnn", "pad0", "pad1")]


class CodensPhys(ctypes.Structure):
    _fields_ = [("gam", c_dbl), ("gam1", c_dbl), ("pinfl", c_dbl), ("pr", c_dbl), ("prt", c_dbl),
                ("cp", c_dbl), ("reynz", c_dbl), ("tnz", c_dbl), ("iflu", c_int), ("pad", c_int)]


class CodensScheme(ctypes.Structure):
    _fields_ = [("muscl", c_int), ("ilim", c_int), ("xk", c_dbl)]


P = ctypes.POINTER

# name -> argtypes (restype int32 for all but codens_last_error)
SIGNATURES = {
    "codens_zpres": [P(CodensBlock), P(CodensPhys), c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "codens_zvismo": [P(CodensBlock), P(CodensPhys), c_vp, c_vp, c_vp, c_vp],
    "codens_chronos": [P(CodensBlock), P(CodensPhys), c_dbl, c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                       c_vp, c_vp],
    "codens_chrono_global": [P(CodensBlock), c_int, c_dbl, c_vp, P(c_dbl), c_vp],
    "codens_sch_roe_euler": [P(CodensBlock), P(CodensPhys), P(CodensScheme), c_int, c_vp, c_vp, c_i64, c_vp,
                             c_vp, c_vp, c_vp],
    "codens_sch_roe": [P(CodensBlock), P(CodensPhys), P(CodensScheme), c_int, c_vp, c_vp, c_i64, c_vp, c_vp,
                       c_vp, c_vp, c_vp],
    "codens_zvisqc": [P(CodensBlock), P(CodensPhys), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_dissip_jameson": [P(CodensBlock), c_dbl, c_dbl, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_roe2o": [P(CodensBlock), c_vp, c_i64, c_vp, c_vp, c_vp],
    "codens_met_roe": [P(CodensBlock), c_vp, c_i64, c_vp, c_vp, c_vp],
    "codens_teq_gradv": [P(CodensBlock), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp],
    "codens_met_prod": [P(CodensBlock), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
    "codens_met_gradtr": [P(CodensBlock), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp],
    "codens_met_smsa": [P(CodensBlock), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                        c_vp],
    "codens_met_smko": [P(CodensBlock), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_flud": [P(CodensBlock), c_int, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "codens_sch_turb": [P(CodensBlock), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_sch_expli": [P(CodensBlock), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
    "codens_mesh_metrics": [P(CodensBlock), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_rk_stage": [P(CodensBlock), c_dbl, c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp],
    "codens_implimf_3d": [P(CodensBlock), P(CodensPhys), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_int,
                          c_vp, c_vp, c_dbl, c_int, c_int, c_vp, c_vp, c_vp],
    "codens_adj_residual_jvp": [P(CodensBlock), P(CodensPhys), P(CodensScheme), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp,
                                c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_adj_residual_vjp": [P(CodensBlock), P(CodensPhys), P(CodensScheme), c_vp, c_i64, c_vp, c_vp, c_vp, c_i64,
                                c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_adj_metric_vjp": [P(CodensBlock), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_sweep_tile": [c_int, c_vp],
    "codens_residual_path": [c_int, c_vp],
    "codens_restart_write": [ctypes.c_char_p, c_vp, c_int, c_int, c_int, c_vp, c_i64, c_vp],
    "codens_restart_read": [ctypes.c_char_p, c_vp, c_int, c_int, c_int, c_vp, c_i64, c_vp],
    "codens_mesh_write": [ctypes.c_char_p, c_vp, c_int, c_i64, c_vp, c_vp, c_vp],
    "codens_mesh_read": [ctypes.c_char_p, c_vp, c_int, c_i64, c_vp, c_vp, c_vp],
    "codens_sweep_tile_of": [P(CodensBlock), c_int, c_int, c_vp],
    "codens_met_smke": [P(CodensBlock), c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_kemut": [P(CodensBlock), c_dbl, c_int, c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_smklsas": [P(CodensBlock), c_vp, c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_vp, c_vp, c_vp],
    "codens_workspace_set": [c_int],
    "codens_sch_roe_batch": [c_int, c_vp, P(CodensPhys), P(CodensScheme), c_int, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp,
                             c_vp, c_vp],
    "codens_implimf_3d_batch": [c_int, c_vp, P(CodensPhys), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_int,
                                c_vp, c_vp, c_dbl, c_int, c_int, c_vp, c_vp],
    "codens_implimf_3d_f32": [P(CodensBlock), P(CodensPhys), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_int,
                          c_vp, c_vp, c_dbl, c_int, c_int, c_vp, c_vp, c_vp],
    "codens_implimf_3d_noalloc": [P(CodensBlock), P(CodensPhys), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                                  c_int, c_vp, c_vp, c_dbl, c_int, c_int, c_vp, c_vp, c_vp],
    "codens_bc_coincident": [c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "codens_bc_wall": [c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "codens_bc_farfield": [c_i64, c_vp, c_vp, c_vp, P(c_dbl), P(CodensPhys), c_vp, c_i64, c_vp, c_vp, c_vp,
                           c_vp],
    "codens_residu": [c_vp, c_int, c_int, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
    "codens_check_state": [c_vp, c_int, c_int, P(CodensPhys), c_vp, c_i64, c_vp, c_vp],
    "codens_bc_sym": [c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp],
    "codens_rfve": [c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "codens_utnrd": [c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, P(CodensPhys), P(c_dbl)],
    "codens_bc_tq_copy": [c_i64, c_vp, c_vp, c_vp, c_i64, c_vp],
    "codens_bc_tq_coincident": [c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp],
    "codens_sumsq_rate": [P(CodensBlock), c_int, c_vp, c_vp, c_vp, P(c_dbl), c_vp],
    "codens_timing_enable": [c_int],
    "codens_timing_read": [ctypes.c_char_p, P(c_dbl), P(c_i64), c_int],
    "codens_implimf_eu": [P(CodensBlock), P(CodensPhys), c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_int, c_vp, c_vp,
                          c_dbl, c_int, c_int, c_vp, c_vp, c_vp],
    "codens_impli2_eqt_3d": [P(CodensBlock), c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp],
    "codens_impli2_eqt_3d_batch": [c_int, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp],
    "codens_met_samut": [P(CodensBlock), P(c_dbl), c_vp, c_i64, c_vp, c_vp, c_vp],
    "codens_met_kocmut": [P(CodensBlock), P(c_dbl), c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_smkor": [P(CodensBlock), P(c_dbl), c_dbl, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_smch": [P(CodensBlock), c_dbl, c_dbl, c_dbl, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                        c_vp],
    "codens_met_cut": [P(CodensBlock), c_int, c_dbl, c_dbl, c_vp, c_i64, c_vp],
    "codens_met_smdes": [P(CodensBlock), P(c_dbl), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp,
                         c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_smsasas": [P(CodensBlock), P(c_dbl), c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_vp],
    "codens_met_smkl": [P(CodensBlock), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                        c_vp, c_vp],
    "codens_met_klmut": [P(CodensBlock), c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp],
    "codens_met_mtcorf1": [c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_i64, c_vp],
    "codens_met_bord": [c_int, c_i64, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp, c_i64, c_vp],
    "codens_met_parko": [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp, c_i64, c_vp],
    "codens_met_smmt": [P(CodensBlock), c_vp, c_int, c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                        c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "codens_met_smmtr": [P(CodensBlock), c_vp, c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                         c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "codens_met_fludmt": [P(CodensBlock), c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp],
    "codens_met_komut": [P(CodensBlock), c_dbl, c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "codens_met_rbse": [c_i64, c_vp, c_vp, c_vp, c_i64, c_vp],
    "codens_met_rbsc": [c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp],
    "codens_host_mirror": [c_vp, c_i64, c_i64, c_i64, c_i64, c_int, c_int, P(c_vp), c_vp],
    "codens_host_fetch": [c_vp, c_i64, c_i64, c_i64, c_int, c_vp],
    "codens_host_bundle": [P(c_vp), c_int, c_i64, c_i64, c_i64, c_int, P(c_vp), c_vp],
    "codens_host_bundle_fetch": [P(c_vp), c_int, c_i64, c_i64, c_i64, c_vp],
    "codens_host_residency": [c_int, P(c_int), c_vp],
    "codens_host_sync": [c_vp, c_vp],
    "codens_host_sync_all": [c_vp],
    "codens_host_modified": [c_vp],
    "codens_host_traffic": [P(c_i64), P(c_i64)],
}


def _one_hip_runtime():
    """torch ships its own libamdhip64.so.7 with the same soname as ROCm's;
    whichever is loaded first serves the whole process.  Importing torch
    before this library keeps a single runtime for torch and the kernels
    whatever the caller imported first (torch is optional for the library)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path: str = LIB_PATH):
    _one_hip_runtime()
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -m codens_amd.build` (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    lib.codens_last_error.restype = ctypes.c_char_p
    lib.codens_last_error.argtypes = []
    return lib


lib = load()

"""Python view of the C-ABI: the same entry points, with torch device tensors.

Each function mirrors one reference routine (same name without the
``codens_`` prefix, same argument meaning) and raises :class:`CodensError`
where the C call returns a non-zero status, the analogue of the reference's
``stop``/``abort`` paths.  Tensors must be float64 (int64 for face lists),
contiguous and on the current HIP device.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from ._lib import CodensBlock, CodensPhys, CodensScheme, lib


class CodensError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _chk(rc):
    if rc != 0:
        raise CodensError(rc, lib.codens_last_error().decode())


def _p(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if not t.is_cuda:
        raise ValueError("tensor must be on the GPU")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _s(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def block(b) -> CodensBlock:
    """codens_block of a :class:`codens_amd.layout.Block`."""
    return CodensBlock(*b.ext16())


def phys(gam=1.4, pinfl=0.0, pr=0.72, prt=1.0, reynz=1e5, tnz=288.15, iflu=1) -> CodensPhys:
    return CodensPhys(gam, gam - 1.0, pinfl, pr, prt, 1.0 / (gam - 1.0), reynz, tnz, iflu, 0)


def scheme(muscl=1, ilim=0, xk=1.0 / 3.0) -> CodensScheme:
    return CodensScheme(muscl, ilim, xk)


def ld(v):
    return v.shape[-1]


def zpres(b, ph, v, ps, temp, cson, stream=None):
    _chk(lib.codens_zpres(ctypes.byref(b), ctypes.byref(ph), _p(v), ld(v), _p(ps), _p(temp), _p(cson), _s(stream)))


def zvismo(b, ph, temp, mu, nbad=None, stream=None):
    _chk(lib.codens_zvismo(ctypes.byref(b), ctypes.byref(ph), _p(temp), _p(mu), _p(nbad), _s(stream)))


def chronos(b, ph, eta, ns, v, mu, mut, sn_blk, vol, cson, dt, stream=None):
    _chk(lib.codens_chronos(ctypes.byref(b), ctypes.byref(ph), eta, 1 if ns else 0, _p(v), ld(v), _p(mu),
                            _p(mut), _p(sn_blk), _p(vol), _p(cson), _p(dt), _s(stream)))


def chrono_global(blocks, dt1min, dt, stream=None):
    """Returns the smallest local step; raises CodensError(4) below dt1min."""
    arr = (CodensBlock * len(blocks))(*blocks)
    out = ctypes.c_double(0.0)
    _chk(lib.codens_chrono_global(arr, len(blocks), dt1min, _p(dt), ctypes.byref(out), _s(stream)))
    return out.value


def sch_roe_euler(b, ph, sc, ityprk, u, v, ff, sn_blk, ps, stream=None):
    _chk(lib.codens_sch_roe_euler(ctypes.byref(b), ctypes.byref(ph), ctypes.byref(sc), ityprk, _p(u), _p(v), ld(v),
                                  _p(ff), _p(sn_blk), _p(ps), _s(stream)))


def sch_roe(b, ph, sc, ityprk, u, v, ff, tq, sn_blk, ps, stream=None):
    _chk(lib.codens_sch_roe(ctypes.byref(b), ctypes.byref(ph), ctypes.byref(sc), ityprk, _p(u), _p(v), ld(v),
                            _p(ff), _p(tq), _p(sn_blk), _p(ps), _s(stream)))


def zvisqc(b, ph, sn_blk, vol, v, temp, mu, mut, tq, stream=None):
    _chk(lib.codens_zvisqc(ctypes.byref(b), ctypes.byref(ph), _p(sn_blk), _p(vol), _p(v), ld(v), _p(temp), _p(mu),
                           _p(mut), _p(tq), _s(stream)))


def dissip_jameson(b, ki2, ki4, v, d, sn_blk, ps, cson, stream=None):
    _chk(lib.codens_dissip_jameson(ctypes.byref(b), ki2, ki4, _p(v), ld(v), _p(d), _p(sn_blk), _p(ps), _p(cson),
                                   _s(stream)))


def met_roe2o(b, t, d, sn_blk, stream=None):
    _chk(lib.codens_met_roe2o(ctypes.byref(b), _p(t), ld(t), _p(d), _p(sn_blk), _s(stream)))


def met_roe(b, t, d, sn_blk, stream=None):
    _chk(lib.codens_met_roe(ctypes.byref(b), _p(t), ld(t), _p(d), _p(sn_blk), _s(stream)))


def _tc(tc):
    """16 turbulence constants as a host double array (cb1 ... alfae)."""
    return (ctypes.c_double * 16)(*[float(x) for x in tc])


def teq_gradv(b, sn_blk, vol, v, g9, stream=None):
    _chk(lib.codens_teq_gradv(ctypes.byref(b), _p(sn_blk), _p(vol), _p(v), ld(v), _p(g9), _s(stream)))


def met_prod(b, mut, v, g9, tprod, stream=None):
    _chk(lib.codens_met_prod(ctypes.byref(b), _p(mut), _p(v), ld(v), _p(g9), _p(tprod), _s(stream)))


def met_gradtr(b, sn_blk, vol, v, fd, stream=None):
    _chk(lib.codens_met_gradtr(ctypes.byref(b), _p(sn_blk), _p(vol), _p(v), ld(v), _p(fd), _s(stream)))


def met_smsa(b, tc, sn_blk, vol, v, mu, dist, g9, fd, cfke, ts6, ts7, stream=None):
    _chk(lib.codens_met_smsa(ctypes.byref(b), _tc(tc), _p(sn_blk), _p(vol), _p(v), ld(v), _p(mu), _p(dist), _p(g9),
                             _p(fd), _p(cfke), _p(ts6), _p(ts7), _s(stream)))


def met_smko(b, tc, v, tprod, cfke, ts6, ts7, cson, stream=None):
    _chk(lib.codens_met_smko(ctypes.byref(b), _tc(tc), _p(v), ld(v), _p(tprod), _p(cfke), _p(ts6), _p(ts7), _p(cson),
                             _s(stream)))


def met_flud(b, kfludis, tc, v, mu, mut, fd, stream=None):
    _chk(lib.codens_met_flud(ctypes.byref(b), kfludis, _tc(tc), _p(v), ld(v), _p(mu), _p(mut), _p(fd), _s(stream)))


def sch_turb(b, u, v, d, fd, ts6, ts7, sn_blk, vol, stream=None):
    _chk(lib.codens_sch_turb(ctypes.byref(b), _p(u), _p(v), ld(v), _p(d), _p(fd), _p(ts6), _p(ts7), _p(sn_blk),
                             _p(vol), _s(stream)))


def impli2_eqt_3d(b, u, dt, v, mu, mut, cfke, sn_blk, vol, cmt, stream=None):
    """impli2_eqt_3d.f90:4; u and v hold 7 planes (transport equations 6-7)."""
    assert u.shape[0] >= 7 and v.shape[0] >= 7 and ld(u) == ld(v)
    _chk(lib.codens_impli2_eqt_3d(ctypes.byref(b), _p(u), ld(v), _p(dt), _p(v), _p(mu), _p(mut), _p(cfke),
                                  _p(sn_blk), _p(vol), cmt, _s(stream)))


def impli2_eqt_3d_batch(blocks_arr, nblk, u, dt, v, mu, mut, cfke, sn_all, vol, cmt, stream=None):
    """impli2_eqt_3d.f90:4 for all blocks in one launch per pass."""
    assert u.shape[0] >= 7 and v.shape[0] >= 7 and ld(u) == ld(v)
    _chk(lib.codens_impli2_eqt_3d_batch(int(nblk), blocks_arr, _p(u), ld(v), _p(dt), _p(v), _p(mu), _p(mut), _p(cfke),
                                        _p(sn_all), _p(vol), cmt, _s(stream)))


def met_samut(b, tc, v, mu, mut, stream=None):
    _chk(lib.codens_met_samut(ctypes.byref(b), _tc(tc), _p(v), ld(v), _p(mu), _p(mut), _s(stream)))


def met_kocmut(b, tc, kwsst, sn_blk, vol, v, mu, mut, dist, g9, stream=None):
    _chk(lib.codens_met_kocmut(ctypes.byref(b), _tc(tc), kwsst, _p(sn_blk), _p(vol), _p(v), ld(v), _p(mu), _p(mut),
                               _p(dist), _p(g9), _s(stream)))


def met_smkor(b, tc, sigmd, v, fd, g9, tprod, cfke, ts6, ts7, stream=None):
    _chk(lib.codens_met_smkor(ctypes.byref(b), _tc(tc), sigmd, _p(v), ld(v), _p(fd), _p(g9), _p(tprod), _p(cfke),
                              _p(ts6), _p(ts7), _s(stream)))


def met_smch(b, cke1, cke2, utaumin, v, mu, dist, mnpar, utau, tprod, ts6, ts7, stream=None):
    assert mnpar.dtype == torch.int32
    _chk(lib.codens_met_smch(ctypes.byref(b), cke1, cke2, utaumin, _p(v), ld(v), _p(mu), _p(dist), _p(mnpar),
                             _p(utau), _p(tprod), _p(ts6), _p(ts7), _s(stream)))


def met_cut(b, which, epsk, epse, v, stream=None):
    """which = 1: met_cut.f90, 2: met_cutsa.f90."""
    _chk(lib.codens_met_cut(ctypes.byref(b), which, epsk, epse, _p(v), ld(v), _s(stream)))


def met_smdes(b, tc, x, y, z, delta, sn_blk, vol, v, mu, dist, g9, fd, cfke, ts6, ts7, stream=None):
    """met_smdes.f90:4; delta (nnn, written) gets met_dist's longest edges."""
    _chk(lib.codens_met_smdes(ctypes.byref(b), _tc(tc), _p(x), _p(y), _p(z), _p(delta), _p(sn_blk), _p(vol), _p(v),
                              ld(v), _p(mu), _p(dist), _p(g9), _p(fd), _p(cfke), _p(ts6), _p(ts7), _s(stream)))


def met_smsasas(b, tc, ischema, sn_blk, vol, v, mu, dist, g9, fd, cfke, ts6, ts7, stream=None):
    """met_smsasas.f90:4 (ischema = 1: met_laplaciens's plain form)."""
    _chk(lib.codens_met_smsasas(ctypes.byref(b), _tc(tc), ischema, _p(sn_blk), _p(vol), _p(v), ld(v), _p(mu),
                                _p(dist), _p(g9), _p(fd), _p(cfke), _p(ts6), _p(ts7), _s(stream)))


def _kl(kc):
    """k-l constants cklb1, ckle2, sigmal, xkappa as a host array."""
    kc = [float(x) for x in kc]
    assert len(kc) == 4
    return (ctypes.c_double * 4)(*kc)


def met_smkl(b, kc, sn_blk, vol, v, mu, mut, dist, fd, g9, tprod, cfke, ts6, ts7, stream=None):
    _chk(lib.codens_met_smkl(ctypes.byref(b), _kl(kc), _p(sn_blk), _p(vol), _p(v), ld(v), _p(mu), _p(mut), _p(dist),
                             _p(fd), _p(g9), _p(tprod), _p(cfke), _p(ts6), _p(ts7), _s(stream)))


def met_klmut(b, kc, v, mu, mut, dist, stream=None):
    _chk(lib.codens_met_klmut(ctypes.byref(b), _kl(kc), _p(v), ld(v), _p(mu), _p(mut), _p(dist), _s(stream)))


def _menter(mc):
    """The 10 Menter constants as a host array (betae, okappa, sigme1,
    sigma1, beta1, wsig1, sigme2, sigma2, beta2, wsig2)."""
    mc = [float(x) for x in mc]
    assert len(mc) == 10
    return (ctypes.c_double * 10)(*mc)


def _starts(face_start):
    fs = [int(x) for x in (face_start if face_start is not None else [0])]
    return (ctypes.c_int64 * len(fs))(*fs), len(fs) - 1


def met_mtcorf1(face_start, lines, imxclko, mnpar, frac, stream=None):
    """face_start: host int64 array (walls + 1); lines: device uint8 tensor
    holding codens_wall_line records; mnpar: device int32."""
    fs, nf = _starts(face_start)
    assert mnpar.dtype == torch.int32
    _chk(lib.codens_met_mtcorf1(nf, fs, _p(lines), imxclko, _p(mnpar), _p(frac), frac.numel(), _s(stream)))


def met_bord(kind, g, i, r, rokinf, roeinf, v, stream=None):
    """met_bord.f90:4 on one frontier list: kind 0 coincident, 1 wall
    (kparoi 0), 2 wall (kparoi 1), 3 far field, 4 symmetry."""
    _chk(lib.codens_met_bord(int(kind), g.numel(), _p(g), _p(i), _p(r), float(rokinf), float(roeinf), _p(v), ld(v),
                             _s(stream)))


def met_parko(face_start, lines, ghost, mnpar, mu, dist, beta1, v, stream=None):
    """met_parko.f90:4: face_start host int64 (walls + 1); lines device uint8
    (codens_wall_line records); ghost int64, mnpar int32 device tensors."""
    fs, nf = _starts(face_start)
    assert mnpar.dtype == torch.int32 and ghost.dtype == torch.int64
    _chk(lib.codens_met_parko(nf, fs, _p(lines), _p(ghost), _p(mnpar), _p(mu), _p(dist), float(beta1), _p(v), ld(v),
                              _s(stream)))


def met_smmt(b, mc, kfracom, imxclko, v, mu, mut, dist, fd, tprod, cfke, gkgo, frac, ts6, ts7, face_start=None,
             lines=None, mnpar=None, stream=None):
    fs, nf = _starts(face_start)
    _chk(lib.codens_met_smmt(ctypes.byref(b), _menter(mc), kfracom, imxclko, _p(v), ld(v), _p(mu), _p(mut),
                             _p(dist), _p(fd), _p(tprod), _p(cfke), _p(gkgo), _p(frac), _p(ts6), _p(ts7), nf, fs,
                             _p(lines), _p(mnpar), _s(stream)))


def met_smmtr(b, mc, imxclko, v, mu, mut, dist, fd, g9, tprod, cfke, frac, ts6, ts7, face_start=None, lines=None,
              mnpar=None, stream=None):
    fs, nf = _starts(face_start)
    _chk(lib.codens_met_smmtr(ctypes.byref(b), _menter(mc), imxclko, _p(v), ld(v), _p(mu), _p(mut), _p(dist), _p(fd),
                              _p(g9), _p(tprod), _p(cfke), _p(frac), _p(ts6), _p(ts7), nf, fs, _p(lines), _p(mnpar),
                              _s(stream)))


def met_fludmt(b, mc, kfracom, mu, mut, frac, fd, stream=None):
    _chk(lib.codens_met_fludmt(ctypes.byref(b), _menter(mc), kfracom, _p(mu), _p(mut), _p(frac), _p(fd), ld(fd),
                               _s(stream)))


def met_komut(b, betae, komsst, sn_blk, vol, v, mu, mut, dist, g9, stream=None):
    _chk(lib.codens_met_komut(ctypes.byref(b), betae, komsst, _p(sn_blk), _p(vol), _p(v), ld(v), _p(mu), _p(mut),
                              _p(dist), _p(g9), _s(stream)))


def met_rbse(g, i, fd, stream=None):
    _chk(lib.codens_met_rbse(g.numel(), _p(g), _p(i), _p(fd), ld(fd), _s(stream)))


def met_rbsc(g, i, r, fd, stream=None):
    _chk(lib.codens_met_rbsc(g.numel(), _p(g), _p(i), _p(r), _p(fd), ld(fd), _s(stream)))


def sch_expli(b, u, v, dt, vol, stream=None):
    _chk(lib.codens_sch_expli(ctypes.byref(b), _p(u), _p(v), ld(v), _p(dt), _p(vol), _s(stream)))


def mesh_metrics(b, x, y, z, sn_blk, vol, stream=None):
    """Ghost nodes (overwritten in x, y, z), sn_blk (9 nnn) and vol of one
    block from its node coordinates (extmhg.f90, norm.f90, svol.f90)."""
    _chk(lib.codens_mesh_metrics(ctypes.byref(b), _p(x), _p(y), _p(z), _p(sn_blk), _p(vol), _s(stream)))


def rk_stage(b, alpha, v0, u, v, dt, vol, nvar=5, stream=None):
    """v = v0 - alpha dt/vol u on the real cells (one stage of the 4-stage
    explicit scheme; v0, u and v share one leading dimension)."""
    if ld(v0) != ld(v) or ld(u) != ld(v):
        raise ValueError("rk_stage: v0, u and v must share one leading dimension")
    _chk(lib.codens_rk_stage(ctypes.byref(b), float(alpha), _p(v0), _p(u), _p(v), ld(v), int(nvar), _p(dt), _p(vol),
                             _s(stream)))


def implimf_3d(b, ph, u, dt, v, d, ff, mu, mut, lmx, sn_blk, vol, ps, cson, dtpas=1.0, kfmg=0, ityprk=0,
               stream=None):
    _chk(lib.codens_implimf_3d(ctypes.byref(b), ctypes.byref(ph), _p(u), _p(dt), _p(v), ld(v), _p(d), _p(ff),
                               _p(mu), _p(mut), lmx, _p(sn_blk), _p(vol), dtpas, kfmg, ityprk, _p(ps), _p(cson),
                               _s(stream)))


def implimf_3d_f32(b, ph, u, dt, v, d, ff, mu, mut, lmx, sn_blk, vol, ps, cson, dtpas=1.0, kfmg=0, ityprk=0,
                   stream=None):
    """FP32 defect-correction form of implimf_3d (same arguments and results;
    the sub-iterations in FP32, d and v += d in FP64)."""
    _chk(lib.codens_implimf_3d_f32(ctypes.byref(b), ctypes.byref(ph), _p(u), _p(dt), _p(v), ld(v), _p(d), _p(ff),
                                   _p(mu), _p(mut), lmx, _p(sn_blk), _p(vol), dtpas, kfmg, ityprk, _p(ps), _p(cson),
                                   _s(stream)))


def implimf_3d_noalloc(b, ph, u, dt, v, d, ff, mu, mut, lmx, sn_blk, vol, ps, cson, dtpas=1.0, kfmg=0, ityprk=0,
                       stream=None):
    """implimf_3d with the library's static scratch: allocates nothing (for
    callers under graph capture); blocks of up to 2,400,000 storage entries."""
    _chk(lib.codens_implimf_3d_noalloc(ctypes.byref(b), ctypes.byref(ph), _p(u), _p(dt), _p(v), ld(v), _p(d), _p(ff),
                                       _p(mu), _p(mut), lmx, _p(sn_blk), _p(vol), dtpas, kfmg, ityprk, _p(ps),
                                       _p(cson), _s(stream)))


def implimf_eu(b, ph, u, dt, v, d, ff, lmx, sn_blk, vol, ps, cson, dt1min=1.0, kfmg=0, ityprk=0, stream=None):
    """implimf_eu.f90:6; v and d hold nvar = v.shape[0] planes, all updated."""
    assert d.shape[0] >= v.shape[0]
    _chk(lib.codens_implimf_eu(ctypes.byref(b), ctypes.byref(ph), _p(u), _p(dt), _p(v), ld(v), v.shape[0], _p(d),
                               _p(ff), lmx, _p(sn_blk), _p(vol), dt1min, kfmg, ityprk, _p(ps), _p(cson), _s(stream)))


def bc_coincident(g, i, r, v, ps, temp, cson, stream=None):
    _chk(lib.codens_bc_coincident(g.numel(), _p(g), _p(i), _p(r), _p(v), ld(v), _p(ps), _p(temp), _p(cson),
                                  _s(stream)))


def bc_wall(g, i, v, ps, temp, cson, stream=None):
    _chk(lib.codens_bc_wall(g.numel(), _p(g), _p(i), _p(v), ld(v), _p(ps), _p(temp), _p(cson), _s(stream)))


def bc_farfield(g, i, nrm, w0, ph, v, ps, temp, cson, stream=None):
    w = (ctypes.c_double * 5)(*[float(x) for x in w0])
    _chk(lib.codens_bc_farfield(g.numel(), _p(g), _p(i), _p(nrm), w, ctypes.byref(ph), _p(v), ld(v), _p(ps),
                                _p(temp), _p(cson), _s(stream)))


def rfve(g, i, v, ps, temp, cson, stream=None):
    """rfve.f90: ghost = adjacent real cell (v, p, T, c)."""
    _chk(lib.codens_rfve(g.numel(), _p(g), _p(i), _p(v), ld(v), _p(ps), _p(temp), _p(cson), _s(stream)))


def bc_sym(g, i, nrm, v, stream=None):
    """clsym.f90: symmetry-plane ghosts (writes v only)."""
    _chk(lib.codens_bc_sym(g.numel(), _p(g), _p(i), _p(nrm), _p(v), ld(v), _s(stream)))


def utnrd(mach, alpha, beta, roa1, aa1, pa1, ph):
    """utnrd.f90: the far-field conservative state (5 floats) of Mach number,
    incidence angles and stagnation state."""
    w = (ctypes.c_double * 5)()
    _chk(lib.codens_utnrd(mach, alpha, beta, roa1, aa1, pa1, ctypes.byref(ph), w))
    return list(w)


def bc_tq_copy(g, i, tq, stream=None):
    _chk(lib.codens_bc_tq_copy(g.numel(), _p(g), _p(i), _p(tq), ld(tq), _s(stream)))


def bc_tq_coincident(g, i, r, tq, stream=None):
    _chk(lib.codens_bc_tq_coincident(g.numel(), _p(g), _p(i), _p(r), _p(tq), ld(tq), _s(stream)))


def sumsq_rate(blocks, a, bv, dt, stream=None):
    arr = (CodensBlock * len(blocks))(*blocks)
    out = ctypes.c_double(0.0)
    _chk(lib.codens_sumsq_rate(arr, len(blocks), _p(a), _p(bv), _p(dt), ctypes.byref(out), _s(stream)))
    return out.value


def residu(blocks_arr, nblk, neq, v, u, dt, out, stream=None):
    """residu.f90 / rscpsv.f90: steadiness norms of (v - u) / dt into the
    device array out ((nblk + 1) x neq x 7 doubles: L1, L2, Linf, block, i, j,
    k per block and equation, then the aggregate)."""
    _chk(lib.codens_residu(blocks_arr, nblk, neq, _p(v), _p(u), ld(v), _p(dt), _p(out), _s(stream)))


def check_state(blocks_arr, nblk, neq, ph, v, nbad, stream=None):
    """Adds to the device int nbad the real cells with rho <= 0, p <= 0 or a
    non-finite value in the first neq variables."""
    _chk(lib.codens_check_state(blocks_arr, nblk, neq, ctypes.byref(ph), _p(v), ld(v), _p(nbad), _s(stream)))


def timing_enable(on=True):
    """Bracket every launch of the main kernels with HIP events."""
    _chk(lib.codens_timing_enable(1 if on else 0))


def timing_read(name, reset=True):
    """(total ms, launch count) of one kernel since the last reset."""
    ms = ctypes.c_double(0.0)
    cnt = ctypes.c_int64(0)
    _chk(lib.codens_timing_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt), 1 if reset else 0))
    return ms.value, cnt.value


# ---- differentiation of the block residual (laminar NS, unlimited MUSCL)
def _bc(bcs):
    """(nwall, wg, wi, nfar, fg, fi, nrm) from the (wall ghost, wall inner,
    far ghost, far inner, far normals) device tensors."""
    wg, wi, fg, fi, nrm = bcs
    return (wg.numel(), _p(wg) if wg.numel() else None, _p(wi) if wi.numel() else None, fg.numel(),
            _p(fg) if fg.numel() else None, _p(fi) if fi.numel() else None, _p(nrm) if fg.numel() else None)


def adj_residual_jvp(b, ph, sc, v, vdot, sn_blk, sndot, vol, voldot, winf, bcs, u, udot, stream=None):
    nw, wg, wi, nf, fg, fi, nrm = _bc(bcs)
    _chk(lib.codens_adj_residual_jvp(ctypes.byref(b), ctypes.byref(ph), ctypes.byref(sc), _p(v), _p(vdot), ld(v),
                                     _p(sn_blk), _p(sndot), _p(vol), _p(voldot), _p(winf), nw, wg, wi, nf, fg, fi,
                                     nrm, _p(u), _p(udot), _s(stream)))


def adj_residual_vjp(b, ph, sc, v, sn_blk, vol, winf, bcs, ubar, vbar, snbar, volbar, stream=None):
    nw, wg, wi, nf, fg, fi, nrm = _bc(bcs)
    _chk(lib.codens_adj_residual_vjp(ctypes.byref(b), ctypes.byref(ph), ctypes.byref(sc), _p(v), ld(v),
                                     _p(sn_blk), _p(vol), _p(winf), nw, wg, wi, nf, fg, fi, nrm, _p(ubar), _p(vbar),
                                     _p(snbar), _p(volbar), _s(stream)))


def adj_metric_vjp(b, x, y, z, sg, snbar, volbar, xbar, ybar, zbar, stream=None):
    sgc = (ctypes.c_double * 3)(*[float(t) for t in sg])
    _chk(lib.codens_adj_metric_vjp(ctypes.byref(b), _p(x), _p(y), _p(z), sgc, _p(snbar), _p(volbar), _p(xbar),
                                   _p(ybar), _p(zbar), _s(stream)))


def sweep_tile(mode):
    """Tile shape of the Jacobi sub-iteration: -1 timed choice per block shape
    (default), 0 = 64 x 4, 1 = 32 x 8, 2 = 16 x 16.  Returns the previous mode."""
    prev = ctypes.c_int32(0)
    _chk(lib.codens_sweep_tile(int(mode), ctypes.byref(prev)))
    return prev.value


def _h(a):
    """Host (numpy float64, C-contiguous) array for the file-format entry points."""
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return ctypes.c_void_p(a.ctypes.data)


def restart_write(path, blocks_arr, nblk, v, mut=None, nturb=0):
    """writda.f90:4 records of every block (host arrays: v (nvar, ldv), mut (ldv) or None)."""
    _chk(lib.codens_restart_write(os.fsencode(path), blocks_arr, int(nblk), int(mut is not None), int(nturb), _h(v),
                                  v.shape[-1], _h(mut)))


def restart_read(path, blocks_arr, nblk, v, mut=None, nturb=0):
    """readda.f90:4: the records into the real cells of the host arrays v (and mut)."""
    _chk(lib.codens_restart_read(os.fsencode(path), blocks_arr, int(nblk), int(mut is not None), int(nturb), _h(v),
                                 v.shape[-1], _h(mut)))


def mesh_write(path, blocks_arr, nblk, x, y, z):
    """writdg.f90:4 records (nodes of every block) from host arrays."""
    _chk(lib.codens_mesh_write(os.fsencode(path), blocks_arr, int(nblk), x.shape[-1], _h(x), _h(y), _h(z)))


def mesh_read(path, blocks_arr, nblk, x, y, z):
    """readdg.f90:4 records into the nodes of host arrays."""
    _chk(lib.codens_mesh_read(os.fsencode(path), blocks_arr, int(nblk), x.shape[-1], _h(x), _h(y), _h(z)))


def residual_path(fused):
    """Kernel arrangement of sch_roe / sch_roe_euler: 1 fused (default), 0
    face passes + gather, 2 fused with the all-blocks launch persistent over
    a tile queue.  Returns the previous arrangement."""
    prev = ctypes.c_int32(0)
    _chk(lib.codens_residual_path(int(fused), ctypes.byref(prev)))
    return prev.value


# ---- host-array mirrors (the drop-in boundary) -----------------------------
def _host_ptr(a):
    """Address of a C-contiguous numpy array's first element."""
    if not (a.flags["C_CONTIGUOUS"] and a.flags["WRITEABLE"]):
        raise ValueError("host arrays must be C-contiguous and writeable numpy arrays")
    return ctypes.c_void_p(a.ctypes.data)


def host_mirror(a, offset, width, pitch=0, rows=1, to_device=1):
    """Device mirror of host array a (window in bytes); returns its address."""
    d = ctypes.c_void_p()
    _chk(lib.codens_host_mirror(_host_ptr(a), a.nbytes, offset, width, pitch or width, rows, to_device,
                                ctypes.byref(d), None))
    return d.value


def host_fetch(a, offset, width, pitch=0, rows=1):
    _chk(lib.codens_host_fetch(_host_ptr(a), offset, width, pitch or width, rows, None))


def host_residency(on):
    """1: the mirrors keep device-newer windows until host_sync; 0: every
    window moves both ways.  Returns the previous setting."""
    prev = ctypes.c_int32(0)
    _chk(lib.codens_host_residency(int(on), ctypes.byref(prev), None))
    return prev.value


def host_sync(a):
    _chk(lib.codens_host_sync(_host_ptr(a), None))


def host_sync_all():
    _chk(lib.codens_host_sync_all(None))


def host_modified(a):
    _chk(lib.codens_host_modified(_host_ptr(a)))


def host_traffic():
    """(bytes host -> device, bytes device -> host) moved by the mirrors so far."""
    h, d = ctypes.c_int64(0), ctypes.c_int64(0)
    _chk(lib.codens_host_traffic(ctypes.byref(h), ctypes.byref(d)))
    return h.value, d.value


def sweep_tile_of(b, eu=False, fp32=False):
    t = ctypes.c_int32(0)
    _chk(lib.codens_sweep_tile_of(ctypes.byref(b), int(eu), int(fp32), ctypes.byref(t)))
    return t.value


def met_smke(b, kc, variant, lparoi, ischema, sn_blk, vol, v, mu, mut, tprod, g9, cson, bark, bare, cfke, ts6, ts7,
             stream=None):
    """met_smke.f90:4 (k-epsilon sources; variant 0 smkes, 1 smker, 2 smkec, 3 smkesas)."""
    k = (ctypes.c_double * 3)(*[float(x) for x in kc])
    _chk(lib.codens_met_smke(ctypes.byref(b), k, int(variant), int(lparoi), int(ischema), _p(sn_blk), _p(vol), _p(v),
                             ld(v), _p(mu), _p(mut), _p(tprod), _p(g9), _p(cson), _p(bark), _p(bare), _p(cfke), _p(ts6),
                             _p(ts7), _s(stream)))


def met_kemut(b, cmu, form, lparoi, sn_blk, vol, v, mu, mut, dist, g9, stream=None):
    """met_kemut.f90:4 (form 0 Jones-Launder, 1 with SST limiter, 2 Launder-Sharma)."""
    _chk(lib.codens_met_kemut(ctypes.byref(b), float(cmu), int(form), int(lparoi), _p(sn_blk), _p(vol), _p(v), ld(v),
                              _p(mu), _p(mut), _p(dist), _p(g9), _s(stream)))


def met_smklsas(b, kc, ischema, sn_blk, vol, v, mu, mut, dist, fd, g9, tprod, cfke, ts6, ts7, stream=None):
    """met_smklsas.f90:4 (Smith k-l with the scale-adaptive term)."""
    k = (ctypes.c_double * 4)(*[float(x) for x in kc])
    _chk(lib.codens_met_smklsas(ctypes.byref(b), k, int(ischema), _p(sn_blk), _p(vol), _p(v), ld(v), _p(mu), _p(mut),
                                _p(dist), _p(fd), _p(g9), _p(tprod), _p(cfke), _p(ts6), _p(ts7), _s(stream)))


def workspace_set(k):
    """Select workspace set k (0..15) for the following calls."""
    _chk(lib.codens_workspace_set(int(k)))


def block_array(blocks):
    """ctypes array of codens_block records (for the batched entry points)."""
    arr = (CodensBlock * len(blocks))()
    for k, b in enumerate(blocks):
        arr[k] = b
    return arr


def implimf_3d_batch(blocks_arr, nblk, ph, u, dt, v, d, ff, mu, mut, lmx, sn_all, vol, cson, dtpas=1.0, kfmg=0,
                     ityprk=0, stream=None):
    """implimf_3d.f90:4 for all blocks in one launch per pass."""
    _chk(lib.codens_implimf_3d_batch(int(nblk), blocks_arr, ctypes.byref(ph), _p(u), _p(dt), _p(v), ld(v), _p(d),
                                     _p(ff), _p(mu), _p(mut), int(lmx), _p(sn_all), _p(vol), dtpas, kfmg, ityprk,
                                     _p(cson), _s(stream)))


def sch_roe_batch(blocks_arr, nblk, ph, sc, ityprk, u, v, ff, tq, sn_all, ps, stream=None):
    """sch_roe.f90:4 for all blocks in one launch per pass."""
    _chk(lib.codens_sch_roe_batch(int(nblk), blocks_arr, ctypes.byref(ph), ctypes.byref(sc), int(ityprk), _p(u), _p(v),
                                  ld(v), _p(ff), _p(tq), _p(sn_all), _p(ps), _s(stream)))

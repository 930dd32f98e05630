"""The turbulence boundary routines at the head of met_num.f90 (:145-162):
met_bord (with met_rbvc, met_rbve) on every frontier kind and met_parko
(the omega wall condition of the k-omega models), the oracle against the
reference's own routines (flang build), bit for bit."""
import ctypes

import numpy as np
import pytest

import cpu_ref as C
import oracle
from codens_amd import layout as L
from menter_case import walls
from oracle import ptr
from randcase import random_mesh, random_mesh_xyz

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(8, 6, 5), (6, 6, 4)]
FRONTIERS = [
    (0, "i1", "rc"), (0, "i2", "rc"),
    (1, "j1", "rc"), (1, "j2", "rc"),
    (0, "j1", "para"), (0, "j2", "nrd"), (0, "k1", "sym"), (0, "k2", "nrd"),
    (1, "i1", "para"), (1, "i2", "nrd"), (1, "k1", "nrd"), (1, "k2", "sym"),
]
KIND = {"rc": 0, "nrd": 3, "sym": 4}


def orc_met_bord(O, fr, kparoi, rk, re, v):
    """met_bord's order: the coincident frontiers (met_rbvc), then every
    frontier in listing order."""
    n = v.shape[1]
    L_ = C.L_
    lib = O.lib
    for f in fr:
        if f["kind"] == "rc":
            lib.orc_met_bord(0, L_(len(f["g"])), ptr(f["g"]), ptr(f["i"]), ptr(f["r"]), C.D(0), C.D(0), ptr(v), L_(n))
    for f in fr:
        if f["kind"] == "rc":
            continue
        kind = (1 + kparoi) if f["kind"] == "para" else KIND[f["kind"]]
        lib.orc_met_bord(kind, L_(len(f["g"])), ptr(f["g"]), ptr(f["i"]), ptr(f["i"]), C.D(rk), C.D(re), ptr(v), L_(n))


@pytest.mark.parametrize("kparoi", [0, 1])
def test_met_bord(kparoi):
    lay, sn, vol, xyz = random_mesh_xyz(DIMS, seed=17)
    fr = []
    for l, face, kind in FRONTIERS:
        g, i, fi = L.face_cells(lay.blocks[l], face)
        fr.append(dict(kind=kind, g=np.ascontiguousarray(g.astype(np.int64)), i=np.ascontiguousarray(i.astype(np.int64))))
    partner = [-1] * len(fr)
    for a in range(0, 4, 2):
        partner[a], partner[a + 1] = a + 1, a
        fr[a]["r"], fr[a + 1]["r"] = fr[a + 1]["i"], fr[a]["i"]
    R = C.Ref(lay, C.orc_phys())
    RF = C.RefFrontiers(R, [f["g"] for f in fr], [f["i"] for f in fr], partner)
    O = C.Orc(lay, C.orc_phys())
    rng = np.random.default_rng(19 + kparoi)
    v = np.ascontiguousarray(rng.standard_normal((7, lay.ip11)))
    vr, vo = v.copy(), v.copy()
    rk, re = 1.25e-4, 3.5
    RF.met_bord([f["kind"] for f in fr], kparoi, rk, re, vr)
    orc_met_bord(O, fr, kparoi, rk, re, vo)
    assert np.array_equal(vr, vo)
    assert not np.array_equal(vr[5:], v[5:]) and np.array_equal(vr[:5], v[:5])
    # every kind wrote its rule somewhere: walls 0 in v6, far fields rokinf
    wall = np.concatenate([f["g"] for f in fr if f["kind"] == "para"])
    far = np.concatenate([f["g"] for f in fr if f["kind"] == "nrd"])
    assert np.all(vr[5, wall] == 0.0) and np.all(vr[5, far] == rk) and np.all(vr[6, far] == re)


def test_met_parko():
    """omega wall condition on the j1 walls of two blocks (12 and 14 cells
    from the wall, so the 10-cell reach of the condition stays inside), with
    the wall-point numbers of every cell (mnpar) pointing at its own wall's
    column, except for a band of cells that belongs to no wall."""
    dims = [(7, 12, 4), (5, 14, 3)]
    lay, sn, vol = random_mesh(dims, seed=23)
    n = lay.ip11
    rng = np.random.default_rng(29)
    specs = [(0, "j1"), (1, "j1")]
    desc, ncin, lines, starts, owner = walls(lay, specs)
    ghost = np.ascontiguousarray((lines["nmin"] - lines["dm3"]).astype(np.int64))
    ncbd = np.ascontiguousarray((ghost + 1).astype(np.int32))
    mnpar = np.zeros(n, dtype=np.int32)
    for w, (l, face) in enumerate(specs):
        for p in range(starts[w], starts[w + 1]):
            ln = lines[p]
            for s in range(ln["m3max"] - ln["m3min"] + 1):
                mnpar[ln["nmin"] + s * ln["dm3"]] = ln["mpn0"] + (p - starts[w]) + 1
    # cells 6..7 above the wall of block 0 belong to no wall
    for p in range(starts[0], starts[1]):
        for s in (6, 7):
            mnpar[lines[p]["nmin"] + s * lines[p]["dm3"]] = 0
    v = np.ascontiguousarray(rng.standard_normal((7, n)))
    v[6] = np.abs(v[6]) * 50
    mu = 1e-3 * (1 + rng.random(n)); mut = rng.random(n)
    dist = 1e-3 + rng.random(n)
    R = C.Ref(lay, C.orc_phys())
    R.lib.ref_set_walls(len(specs), ptr(desc))
    vr, vo = v.copy(), v.copy()
    R.lib.ref_met_parko(ptr(ncin), ptr(ncbd), ptr(vr), ptr(mut), ptr(mu), ptr(dist), ptr(mnpar))
    O = C.Orc(lay, C.orc_phys())
    rc = O.lib.orc_met_parko(C.L_(len(lines)), lines.ctypes.data_as(ctypes.c_void_p), ptr(ghost), ptr(mnpar),
                             ptr(mu), ptr(dist), C.D(0.075), ptr(vo), C.L_(n))
    assert rc == 0
    assert np.array_equal(vr, vo)
    assert not np.array_equal(vr[6], v[6]) and np.array_equal(vr[:6], v[:6])

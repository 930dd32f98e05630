"""On-disk formats: the restart records of writda.f90 /
readda.f90 and the mesh records of writdg.f90 / readdg.f90, written and read
by the library's host code (codens_restart_*, codens_mesh_*), against the
reference's own routines (flang build, oracle/_ref): the files are identical
byte for byte, and each side reads the other's file back exactly.  Host
code only; no GPU."""
import os

import numpy as np
import pytest

import oracle
import cpu_ref as C
from randcase import random_mesh

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(7, 5, 4), (4, 6, 3), (9, 3, 5)]


@pytest.fixture(scope="module")
def lay():
    return random_mesh(DIMS, seed=41)[0]


def _blocks(lay):
    from codens_amd import abi
    return abi.block_array([abi.block(b) for b in lay.blocks]), len(lay.blocks)


def _real(lay):
    m = np.zeros(lay.ip11, dtype=bool)
    for b in lay.blocks:
        b.view(m)[b.kk1 - b.kd1:b.kk2 - b.kd1, b.jj1 - b.jd1:b.jj2 - b.jd1, b.ii1 - b.id1:b.ii2 - b.id1] = True
    return m


def _nodes(lay):
    m = np.zeros(lay.ip11, dtype=bool)
    for b in lay.blocks:
        b.view(m)[b.kk1 - b.kd1:b.kk2 - b.kd1 + 1, b.jj1 - b.jd1:b.jj2 - b.jd1 + 1,
                  b.ii1 - b.id1:b.ii2 - b.id1 + 1] = True
    return m


@pytest.mark.parametrize("equat,ns,nturb", [("eu3d   ", 0, 0), ("ns3d   ", 1, 0), ("ns3d ke", 1, 2)])
def test_restart_matches_writda_readda(lay, tmp_path, equat, ns, nturb):
    from codens_amd import abi
    R = C.Ref(lay, C.orc_phys())
    rng = np.random.default_rng(3)
    n = lay.ip11
    v = np.ascontiguousarray(rng.standard_normal((7, n)))
    mut = rng.standard_normal(n)
    arr, nb = _blocks(lay)
    ours, theirs = str(tmp_path / "ours.da"), str(tmp_path / "ref.da")
    abi.restart_write(ours, arr, nb, v, mut if ns else None, nturb)
    R.writda_all(theirs, equat, v, mut)
    a, b = open(ours, "rb").read(), open(theirs, "rb").read()
    ncell = sum(x.ncells for x in lay.blocks)
    assert len(a) == (5 + ns + nturb) * (8 * ncell + 8 * nb)
    assert a == b
    real = _real(lay)
    nv = 5 + nturb
    # the reference's readda on our file, our reader on the reference's file
    for reader in ("ref", "ours"):
        v2 = np.full((7, n), 7.0); m2 = np.full(n, 7.0)
        if reader == "ref":
            R.readda_all(ours, equat, v2, m2)
        else:
            abi.restart_read(theirs, arr, nb, v2, m2 if ns else None, nturb)
        assert np.array_equal(v2[:nv, real], v[:nv, real]), reader
        assert np.all(v2[:nv, ~real] == 7.0) and np.all(v2[nv:] == 7.0), reader
        if ns:
            assert np.array_equal(m2[real], mut[real]) and np.all(m2[~real] == 7.0), reader
        else:
            assert np.all(m2 == 7.0)


def test_mesh_matches_writdg_readdg(lay, tmp_path):
    from codens_amd import abi
    R = C.Ref(lay, C.orc_phys())
    rng = np.random.default_rng(5)
    n = lay.ip11
    x, y, z = (rng.standard_normal(n) for _ in range(3))
    arr, nb = _blocks(lay)
    ours, theirs = str(tmp_path / "ours.dg"), str(tmp_path / "ref.dg")
    abi.mesh_write(ours, arr, nb, x, y, z)
    R.writdg_all(theirs, x, y, z)
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    nodes = _nodes(lay)
    for reader in ("ref", "ours"):
        c = [np.full(n, 9.0) for _ in range(3)]
        if reader == "ref":
            R.readdg_all(ours, *c)
        else:
            abi.mesh_read(theirs, arr, nb, *c)
        for got, want in zip(c, (x, y, z)):
            assert np.array_equal(got[nodes], want[nodes]) and np.all(got[~nodes] == 9.0), reader


def test_restart_errors(lay, tmp_path):
    """A truncated file and a block list that does not match the records are
    reported (CODENS_EIO) with the record and block named, not read past."""
    from codens_amd import abi
    from codens_amd.abi import CodensError
    rng = np.random.default_rng(7)
    n = lay.ip11
    v = np.ascontiguousarray(rng.standard_normal((5, n)))
    mut = rng.standard_normal(n)
    arr, nb = _blocks(lay)
    path = str(tmp_path / "r.da")
    abi.restart_write(path, arr, nb, v, mut)
    data = open(path, "rb").read()
    short = str(tmp_path / "short.da")
    open(short, "wb").write(data[:-100])
    with pytest.raises(CodensError) as e:
        abi.restart_read(short, arr, nb, v.copy(), mut.copy())
    assert e.value.code == 5 and "block 3" in str(e.value)
    # records of the first two blocks only, read as three
    arr2, _ = _blocks(lay)
    with pytest.raises(CodensError) as e:
        abi.restart_read(path, arr2, nb, v.copy(), None)  # no mut record expected: lengths then disagree
    assert e.value.code == 5 and "block 2" in str(e.value)
    with pytest.raises(CodensError) as e:
        abi.restart_read(str(tmp_path / "missing.da"), arr, nb, v.copy(), mut.copy())
    assert e.value.code == 5

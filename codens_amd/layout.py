"""Block storage layout, meshes, metrics and boundary-face lists.

Storage follows the reference exactly, so arrays produced here can be handed
to the reference routines or to the device kernels without reshuffling:

* A block with ``ni x nj x nk`` nodes has real cells ``ii1..ii2-1`` (``ii1=1``,
  ``ii2=ni``) and ``nfi`` ghost layers on each side; nodes and cells share one
  storage of ``(ni+2nfi)(nj+2nfi)(nk+2nfi)`` entries, i fastest
  (crdms.f90:126-153).
* All blocks are concatenated; ``npc = npn`` is the offset of a block's first
  entry, so a flat index is ``npc + (i-id1) + (j-jd1)*nid + (k-kd1)*nid*njd``
  (0-based here; the reference adds 1, sch_roe.f90:1487-1491).
* Conservative variables are SoA ``v[var, n]`` with leading dimension ``ip11``
  (solve.f90:982); face metrics of block l are ``sn[9*npc(l) + (c*3+d)*nnn + m]``
  for face family ``d`` (i, j, k) and component ``c`` (snorm.f90:65-71,
  atsch_num.f90:623).

The mesh metrics repeat norm.f90 and svol.f90 operation for operation, and the
ghost nodes are the zero-thickness extrapolation ``exs1=1, exs2=0`` of
extmhg.f90 (solve.f90:384-385), so ghost cells have zero volume.
"""
from __future__ import annotations

import dataclasses
from typing import List, Sequence, Tuple

import numpy as np

NFI = 1


@dataclasses.dataclass
class Block:
    """Extents of one block (node counts ni, nj, nk) and its storage offset."""

    ni: int
    nj: int
    nk: int
    npc: int = 0

    @property
    def ii1(self):
        return 1

    @property
    def jj1(self):
        return 1

    @property
    def kk1(self):
        return 1

    @property
    def ii2(self):
        return self.ni

    @property
    def jj2(self):
        return self.nj

    @property
    def kk2(self):
        return self.nk

    @property
    def id1(self):
        return self.ii1 - NFI

    @property
    def jd1(self):
        return self.jj1 - NFI

    @property
    def kd1(self):
        return self.kk1 - NFI

    @property
    def id2(self):
        return self.ii2 + NFI

    @property
    def jd2(self):
        return self.jj2 + NFI

    @property
    def kd2(self):
        return self.kk2 + NFI

    @property
    def nid(self):
        return self.id2 - self.id1 + 1

    @property
    def njd(self):
        return self.jd2 - self.jd1 + 1

    @property
    def nkd(self):
        return self.kd2 - self.kd1 + 1

    @property
    def nnn(self):
        return self.nid * self.njd * self.nkd

    @property
    def ncells(self):
        return (self.ni - 1) * (self.nj - 1) * (self.nk - 1)

    def idx(self, i, j, k):
        """0-based flat storage index of cell/node (i, j, k)."""
        return (self.npc + (np.asarray(i) - self.id1) + (np.asarray(j) - self.jd1) * self.nid
                + (np.asarray(k) - self.kd1) * self.nid * self.njd)

    def view(self, arr):
        """[k, j, i] view of this block's slice of a flat storage array."""
        return arr[self.npc:self.npc + self.nnn].reshape(self.nkd, self.njd, self.nid)

    def ext16(self):
        """Extents in the order of the C-ABI ``codens_block`` record."""
        return [self.npc, self.id1, self.id2, self.jd1, self.jd2, self.kd1, self.kd2,
                self.ii1, self.ii2, self.jj1, self.jj2, self.kk1, self.kk2, self.nnn, 0, 0]

    def interior(self):
        """Flat indices of the real cells, k outer, i fastest."""
        k, j, i = np.meshgrid(np.arange(self.kk1, self.kk2), np.arange(self.jj1, self.jj2),
                              np.arange(self.ii1, self.ii2), indexing="ij")
        return self.idx(i, j, k).ravel()


class Layout:
    """Concatenated blocks (crdms.f90:148-153: npc/npn are running totals)."""

    def __init__(self, dims: Sequence[Tuple[int, int, int]]):
        """``dims`` are CELL counts per block (ni-1, nj-1, nk-1)."""
        self.blocks: List[Block] = []
        off = 0
        for (ci, cj, ck) in dims:
            b = Block(ci + 1, cj + 1, ck + 1, off)
            off += b.nnn
            self.blocks.append(b)
        self.ip11 = off
        self.ip00 = max(b.nnn for b in self.blocks)

    @property
    def ncells(self):
        return sum(b.ncells for b in self.blocks)

    def ext_array(self):
        return np.array([b.ext16() for b in self.blocks], dtype=np.int32)

    def interior(self):
        return np.concatenate([b.interior() for b in self.blocks])


# ---------------------------------------------------------------------------
# ghost nodes and metrics
# ---------------------------------------------------------------------------

def extrapolate_ghost_nodes(b: Block, x, y, z):
    """Zero-thickness ghost nodes: copy of the nearest real node.

    extmhg.f90 with ex1=1, ex2=0 (solve.f90:384-385): the k layers first, then
    the i layers over all k, then the j layers over all i and k.
    """
    for arr in (x, y, z):
        a = b.view(arr)
        i1, i2, j1, j2 = b.ii1 - b.id1, b.ii2 - b.id1, b.jj1 - b.jd1, b.jj2 - b.jd1
        k1, k2 = b.kk1 - b.kd1, b.kk2 - b.kd1
        a[k1 - 1, j1:j2 + 1, i1:i2 + 1] = a[k1, j1:j2 + 1, i1:i2 + 1]
        a[k2 + 1, j1:j2 + 1, i1:i2 + 1] = a[k2, j1:j2 + 1, i1:i2 + 1]
        a[:, j1:j2 + 1, i1 - 1] = a[:, j1:j2 + 1, i1]
        a[:, j1:j2 + 1, i2 + 1] = a[:, j1:j2 + 1, i2]
        a[:, j1 - 1, :] = a[:, j1, :]
        a[:, j2 + 1, :] = a[:, j2, :]


def _cross(d1, d2):
    dx1, dy1, dz1 = d1
    dx2, dy2, dz2 = d2
    return (dy1 * dz2 - dz1 * dy2, dz1 * dx2 - dx1 * dz2, dx1 * dy2 - dy1 * dx2)


def block_metrics(b: Block, x, y, z, sn, vol):
    """Face area vectors and cell volumes of one block (norm.f90, svol.f90).

    ``sn`` is the global metric array (9*ip11) and ``vol`` the global volume
    array (ip11); both are written for block ``b`` only.
    """
    X, Y, Z = b.view(x), b.view(y), b.view(z)
    nkd, njd, nid = X.shape
    snb = sn[9 * b.npc: 9 * b.npc + 9 * b.nnn].reshape(3, 3, nkd, njd, nid)  # [comp, dir, k, j, i]
    snb[...] = 0.0
    P = (X, Y, Z)

    def node(di, dj, dk, sl):
        ks, js, is_ = sl
        return tuple(a[ks.start + dk:ks.stop + dk, js.start + dj:js.stop + dj,
                       is_.start + di:is_.stop + di] for a in P)

    def sub(a, bb):
        return tuple(p - q for p, q in zip(a, bb))

    # orientation test at (i1+1, j1+1, k1+1)  (norm.f90:99-107)
    oi, oj, ok = b.ii1 + 1 - b.id1, b.jj1 + 1 - b.jd1, b.kk1 + 1 - b.kd1
    o = slice(ok, ok + 1), slice(oj, oj + 1), slice(oi, oi + 1)

    def orient(d1, d2, incs):
        t = _cross(d1, d2)
        w = [None, None, None]
        for c in range(3):
            acc = None
            for (a_, b_) in incs:
                term = node(*a_, o)[c] - node(*b_, o)[c]
                acc = term if acc is None else acc + term
            w[c] = acc
        ps = t[0] * w[0] + t[1] * w[1] + t[2] * w[2]
        return 0.5 if ps.item() >= 0 else -0.5

    # k faces: k = kmin..kmax, j = jmin..jmax-1, i = imin..imax-1
    s_k = slice(0, nkd), slice(0, njd - 1), slice(0, nid - 1)
    d1 = sub(node(1, 1, 0, o), node(0, 0, 0, o)); d2 = sub(node(0, 1, 0, o), node(1, 0, 0, o))
    sg = orient(d1, d2, [((0, 0, 1), (0, 0, 0)), ((1, 0, 1), (1, 0, 0)), ((0, 1, 1), (0, 1, 0)),
                         ((1, 1, 1), (1, 1, 0))])
    d1 = sub(node(1, 1, 0, s_k), node(0, 0, 0, s_k)); d2 = sub(node(0, 1, 0, s_k), node(1, 0, 0, s_k))
    t = _cross(d1, d2)
    for c in range(3):
        snb[c, 2, :, :njd - 1, :nid - 1] = sg * t[c]
    # j faces: j = jmin..jmax, k = kmin..kmax-1, i = imin..imax-1
    s_j = slice(0, nkd - 1), slice(0, njd), slice(0, nid - 1)
    d1 = sub(node(1, 0, 1, o), node(0, 0, 0, o)); d2 = sub(node(0, 0, 1, o), node(1, 0, 0, o))
    sg = orient(d1, d2, [((0, 1, 0), (0, 0, 0)), ((1, 1, 0), (1, 0, 0)), ((0, 1, 1), (0, 0, 1)),
                         ((1, 1, 1), (1, 0, 1))])
    d1 = sub(node(1, 0, 1, s_j), node(0, 0, 0, s_j)); d2 = sub(node(0, 0, 1, s_j), node(1, 0, 0, s_j))
    t = _cross(d1, d2)
    for c in range(3):
        snb[c, 1, :nkd - 1, :, :nid - 1] = sg * t[c]
    # i faces: i = imin..imax, k = kmin..kmax-1, j = jmin..jmax-1
    s_i = slice(0, nkd - 1), slice(0, njd - 1), slice(0, nid)
    d1 = sub(node(0, 1, 1, o), node(0, 0, 0, o)); d2 = sub(node(0, 0, 1, o), node(0, 1, 0, o))
    sg = orient(d1, d2, [((1, 0, 0), (0, 0, 0)), ((1, 1, 0), (0, 1, 0)), ((1, 0, 1), (0, 0, 1)),
                         ((1, 1, 1), (0, 1, 1))])
    d1 = sub(node(0, 1, 1, s_i), node(0, 0, 0, s_i)); d2 = sub(node(0, 0, 1, s_i), node(0, 1, 0, s_i))
    t = _cross(d1, d2)
    for c in range(3):
        snb[c, 0, :nkd - 1, :njd - 1, :] = sg * t[c]

    # volumes (svol.f90:77-129): face "centroid * S" sums, then divergence
    def fsum(corners, sl, d):
        acc = 0.0
        for c in range(3):
            pts = [node(*cc, sl)[c] for cc in corners]
            s = ((pts[0] + pts[1]) + pts[2]) + pts[3]
            term = s * snb[c, d][sl]
            acc = term if c == 0 else acc + term
        return acc

    skv = np.zeros((nkd, njd, nid)); sjv = np.zeros((nkd, njd, nid)); siv = np.zeros((nkd, njd, nid))
    sl = slice(0, nkd), slice(0, njd - 1), slice(0, nid - 1)
    skv[sl] = fsum([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], sl, 2)
    sl = slice(0, nkd - 1), slice(0, njd), slice(0, nid - 1)
    sjv[sl] = fsum([(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)], sl, 1)
    sl = slice(0, nkd - 1), slice(0, njd - 1), slice(0, nid)
    siv[sl] = fsum([(0, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1)], sl, 0)
    V = b.view(vol)
    c = slice(0, nkd - 1), slice(0, njd - 1), slice(0, nid - 1)
    V[c] = 0.25 * (siv[:nkd - 1, :njd - 1, 1:] - siv[c] + sjv[:nkd - 1, 1:, :nid - 1] - sjv[c]
                   + skv[1:, :njd - 1, :nid - 1] - skv[c]) / 3


def metrics(layout: Layout, x, y, z):
    """Ghost-node extrapolation, then sn (9*ip11) and vol (ip11) for all blocks."""
    sn = np.zeros(9 * layout.ip11)
    vol = np.zeros(layout.ip11)
    for b in layout.blocks:
        extrapolate_ghost_nodes(b, x, y, z)
        block_metrics(b, x, y, z, sn, vol)
    return sn, vol


# ---------------------------------------------------------------------------
# boundary faces
# ---------------------------------------------------------------------------

FACES = ("i1", "i2", "j1", "j2", "k1", "k2")


def face_cells(b: Block, face: str):
    """(ghost, interior, sn face index) of a block face, in the reference's order.

    initis.f90:71-150 lists i faces k-outer/j-inner, j faces k-outer/i-inner,
    k faces j-outer/i-inner; initbs.f90:57-68 puts the adjacent real cell one
    step inward; initns.f90 takes the normal of the face between them.
    """
    d = face[0]
    lo = face[1] == "1"
    I = np.arange(b.ii1, b.ii2)
    J = np.arange(b.jj1, b.jj2)
    K = np.arange(b.kk1, b.kk2)
    if d == "i":
        k, j = np.meshgrid(K, J, indexing="ij")
        g = b.ii1 - 1 if lo else b.ii2
        gi = b.idx(g, j, k).ravel()
        inc = 1
    elif d == "j":
        k, i = np.meshgrid(K, I, indexing="ij")
        g = b.jj1 - 1 if lo else b.jj2
        gi = b.idx(i, g, k).ravel()
        inc = b.nid
    else:
        j, i = np.meshgrid(J, I, indexing="ij")
        g = b.kk1 - 1 if lo else b.kk2
        gi = b.idx(i, j, g).ravel()
        inc = b.nid * b.njd
    inner = gi + inc if lo else gi - inc
    faceidx = gi + inc if lo else gi        # storage index carrying the face's sn
    return gi.astype(np.int64), inner.astype(np.int64), faceidx.astype(np.int64)


def inward_normals(b: Block, sn, face: str, faceidx):
    """Unit normals pointing into the block (initns.f90:56-163)."""
    d = "ijk".index(face[0])
    sgn = 1.0 if face[1] == "1" else -1.0
    m = faceidx - b.npc
    base = 9 * b.npc
    nx = sgn * sn[base + (0 * 3 + d) * b.nnn + m]
    ny = sgn * sn[base + (1 * 3 + d) * b.nnn + m]
    nz = sgn * sn[base + (2 * 3 + d) * b.nnn + m]
    nn = np.sqrt(nx * nx + ny * ny + nz * nz)
    nn = nn + (np.where(-np.abs(nn) < 0, -1.0, 1.0) + 1.0) * 1e-10
    return nx / nn, ny / nn, nz / nn

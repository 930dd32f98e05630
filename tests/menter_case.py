"""Menter k-omega test inputs: the model constants of '2MT' and wall faces
given both ways, as the reference's module `boundary` tables (for the
reference build's met_mtcorf1, which finds its lines with atindnor) and as
the line descriptors of the C-ABI (codens_wall_line).  Only walls on the
first face of a direction ('i1', 'j1', 'k1') are made: for the second face
the reference's line addressing runs outward from the wall
(met_mtcorf1.f90:98 with atindnor.f90:46-51), past the block's storage."""
import numpy as np

# betae, okappa, sigme1, sigma1, beta1, wsig1, sigme2, sigma2, beta2, wsig2
# (atlecdon.f90:255-278 for '2MT': sigme2 = 0.85)
MENTER = np.array([0.09, 0.41, 0.5, 0.5, 0.075, 0.0, 0.85, 0.856, 0.0828, 0.856])

LINE = np.dtype([("nmin", "<i8"), ("m3min", "<i4"), ("m3max", "<i4"), ("isens3", "<i4"), ("dm3", "<i4"),
                 ("mpn0", "<i4"), ("mpn1", "<i4")])
assert LINE.itemsize == 32

FACES = {"i1": 1, "j1": 3, "k1": 5}


def walls(lay, specs):
    """specs: list of (block index, face 'i1'|'j1'|'k1').  Returns the
    boundary tables desc (12 x nw, Fortran order), ncin (int32, 1-based cell
    per wall point), the lines of every wall (LINE array, walls in list
    order, points in the reference's order) and face_start (int64, nw+1),
    plus the block (0-based) of each wall."""
    desc = np.zeros((len(specs), 12), dtype=np.int32)
    ncin, lines, starts, owner = [], [], [0], []
    npts = 0
    for w, (l, face) in enumerate(specs):
        b = lay.blocks[l]
        nid = b.id2 - b.id1 + 1
        nijd = nid * (b.jd2 - b.jd1 + 1)
        lo = {"i": (b.ii1, b.ii2 - 1), "j": (b.jj1, b.jj2 - 1), "k": (b.kk1, b.kk2 - 1)}
        ax = face[0]
        rng = dict(lo)
        rng[ax] = (lo[ax][0], lo[ax][0])  # the wall's own layer
        (imin, imax), (jmin, jmax), (kmin, kmax) = rng["i"], rng["j"], rng["k"]
        pts = [(i, j, k) for k in range(kmin, kmax + 1) for j in range(jmin, jmax + 1) for i in range(imin, imax + 1)]
        desc[w] = [l + 1, FACES[face], imin, imax, jmin, jmax, kmin, kmax, len(pts), npts, npts, 0]
        m3min, m3max, dm3 = {"i": (imin, b.ii2, 1), "j": (jmin, b.jj2, nid), "k": (kmin, b.kk2, nijd)}[ax]
        for (i, j, k) in pts:
            n = b.npc + (i - b.id1) + (j - b.jd1) * nid + (k - b.kd1) * nijd  # 0-based
            ncin.append(n + 1)
            lines.append((n, m3min, m3max, 1, dm3, npts, npts + len(pts)))
        npts += len(pts)
        starts.append(len(lines))
        owner.append(l)
    return (np.asfortranarray(desc.T), np.array(ncin, dtype=np.int32), np.array(lines, dtype=LINE),
            np.array(starts, dtype=np.int64), owner)


def walls_of_block(lines, starts, owner, l):
    """The lines and face_start of the walls of block l (in list order)."""
    sel = [w for w, o in enumerate(owner) if o == l]
    parts = [lines[starts[w]:starts[w + 1]] for w in sel]
    st = np.cumsum([0] + [len(p) for p in parts]).astype(np.int64)
    ln = np.concatenate(parts) if parts else np.zeros(0, dtype=LINE)
    return ln, st

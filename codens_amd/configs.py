"""Synthetic cases of the residual-and-update path: mesh, boundary faces,
gas constants, switches and initial fields, all seeded and deterministic.

* ``periodic_euler``  -- 64x32x4 cells, uniform dx=dy=dz=1/64, coincident
  (periodic) faces in i, j and k, Euler, M=0.8, explicit forward Euler with a
  global step (kdtl=0) at CFL 0.8.  4 seeded sinusoidal modes of amplitude 1%
  on rho and u.
* ``wall_ns``  -- 256x128x64 cells by default, Cartesian with seeded
  sinusoidal node displacement of 5% of the local spacing and tanh stretching
  toward an adiabatic no-slip wall at j=1 (clpara); far field (clnrd) on the
  five other faces.  Laminar Navier-Stokes, Re=1e5 per unit length, M=0.5,
  Roe-MUSCL with minmod (ilim=1), implicit point-Jacobi (kmf=1, lmax=10),
  local time step (kdtl=1) with eta=10.
* ``wall_rans``  -- the wall block with a turbulence model (SA or Wilcox
  k-omega) on the transport equations 6-7, wall distance = y of the cell
  centres.

Seeds: numpy PCG64(20261201 + case index).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import layout as L

GAM = 1.4


@dataclasses.dataclass
class FaceSet:
    """Boundary faces of one kind, frontier after frontier in the reference's
    listing order (initis.f90): flat indices of the ghost cell, the adjacent
    real cell and, for coincident faces, the real cell across the interface.
    sizes[f] faces belong to frontier f (of block blocks[f]); coincident
    frontiers come in pairs 2m, 2m+1 that are each other's partner
    (ndcc, rbvc.f90)."""
    ghost: np.ndarray
    inner: np.ndarray
    remote: Optional[np.ndarray] = None
    normals: Optional[np.ndarray] = None  # (3, nf) inward unit normals
    sizes: Optional[List[int]] = None
    blocks: Optional[List[int]] = None

    def __post_init__(self):
        if self.sizes is None:
            self.sizes = [len(self.ghost)] if len(self.ghost) else []
        if self.blocks is None:
            self.blocks = [0] * len(self.sizes)

    def frontier(self, f):
        """Slice of frontier f in the face arrays."""
        o = int(sum(self.sizes[:f]))
        return slice(o, o + self.sizes[f])


@dataclasses.dataclass
class TurbCfg:
    """Turbulence model of the transport equations 6-7 (met_num.f90): model
    '1SA' (Spalart-Allmaras: met_smsa, met_fludcsa, met_cutsa, met_samut) or
    '2KO' (Wilcox k-omega: met_smko, met_fludko, met_cut, met_kocmut)."""
    model: str
    tc: list            # 16 constants cb1 cb2 sigma kappa cw1 cw2 cw3 cv1 ct3 ct4 betas beta sigmk sigmw alfak alfae
    dist: np.ndarray    # wall distance of every cell (ip11)
    klroe: int = 2      # dissipation: 1 met_roe, 2 met_roe2o (met_num.f90:623-641)
    kwsst: int = 0      # k-omega eddy viscosity: 1 with the SST limiter
    epsk: float = 1e-12
    epse: float = 1e-12
    rokinf: float = 0.0  # far-field ghost values of planes 6, 7 (met_bord.f90:204-214)
    roeinf: float = 0.0
    # met_parko's walls (k-omega): LINE records of the wall points, their
    # ghosts, the walls' point ranges and every cell's wall point (1-based,
    # 0 for none); None where the model has no omega wall condition
    wall_lines: Optional[np.ndarray] = None
    wall_ghost: Optional[np.ndarray] = None
    wall_start: Optional[np.ndarray] = None
    mnpar: Optional[np.ndarray] = None

    @property
    def kparoi(self):   # atlecdon.f90:191 (1SA), :360 (2KO)
        return {"1SA": 0, "2KO": 1}[self.model]

    @property
    def cmt(self):      # impli2_eqt_3d.f90:82-95
        return {"1SA": 1.0, "2KO": 2.0}[self.model]

    @property
    def kfludis(self):  # met_num.f90:528-556
        return {"1SA": 2, "2KO": 5}[self.model]


@dataclasses.dataclass
class Case:
    name: str
    layout: L.Layout
    sn: np.ndarray
    vol: np.ndarray
    v0: np.ndarray  # (5, ip11) initial conservative state, ghosts included
    ns: bool
    phys: Dict[str, float]
    scheme: Dict[str, float]
    eta: float
    kdtl: int
    dt1min: float
    kmf: int
    lmax: int
    winf: np.ndarray  # far-field conservative state (FarField.state of farspec)
    coincident: FaceSet
    wall: FaceSet
    farfield: FaceSet
    turb: Optional[TurbCfg] = None
    farspec: Optional[FarField] = None
    sym: Optional[FaceSet] = None  # symmetry planes (clsym.f90)

    def __post_init__(self):
        if self.sym is None:
            self.sym = _empty_faces()


def conservative(rho, u, v, w, p, gam=GAM, pinfl=0.0):
    return np.stack([rho, rho * u, rho * v, rho * w, (p - pinfl) / (gam - 1.0) + pinfl + 0.5 * rho * (u * u + v * v + w * w)])


def _empty_faces():
    z = np.zeros(0, dtype=np.int64)
    return FaceSet(z, z, z, np.zeros((3, 0)), [], [])


def _cat(parts: List[FaceSet]) -> FaceSet:
    if not parts:
        return _empty_faces()
    g = np.concatenate([p.ghost for p in parts])
    i = np.concatenate([p.inner for p in parts])
    r = np.concatenate([p.remote for p in parts]) if parts[0].remote is not None else None
    nrm = np.concatenate([p.normals for p in parts], axis=1) if parts[0].normals is not None else None
    sizes = [n for p in parts for n in p.sizes]
    blocks = [l for p in parts for l in p.blocks]
    return FaceSet(g, i, r, nrm, sizes, blocks)


@dataclasses.dataclass
class FarField:
    """A non-reflecting far field as the reference's deck gives it (boundary
    condition 'nrd', rbord.f90:300-323): Mach number, incidence angles alpha
    and beta in degrees (bceqt(:,1:3)) and the stagnation state roa1, aa1,
    pa1 (c_nzst.f90:56-59).  Only a non-rotating frame (omg = 0)."""
    mach: float
    alpha: float = 0.0
    beta: float = 0.0
    roa1: float = 1.0
    aa1: float = 1.0
    pa1: float = 1.0 / GAM

    def state(self, gam=GAM, pinfl=0.0):
        """The conservative far-field state of utnrd.f90:59-105 with omg = 0:
        static density, sound speed and pressure from the isentropic
        relations, velocity (cos a cos b, -sin b, sin a cos b) x M a.  With
        the same libm calls and operation order as utnrd, so the state is the
        reference's to the last bit."""
        gam1 = gam - 1.0
        gam2, gam4 = 0.5 * gam1, 1.0 / gam1
        degrad = math.atan2(1.0, 0.0) / 90.0
        f = 1.0 + gam2 * (self.mach * self.mach)
        ro = self.roa1 / math.pow(f, gam4)
        a = self.aa1 / math.pow(f, 0.5)
        p = self.pa1 / math.pow(f, gam / gam1)
        q = self.mach * a
        al, be = self.alpha * degrad, self.beta * degrad
        return np.array([ro, ro * q * math.cos(al) * math.cos(be), -ro * q * math.sin(be),
                         ro * q * math.sin(al) * math.cos(be), p / gam1 + pinfl + 0.5 * ro * (q * q)])


def static_farfield(mach, alpha=0.0, beta=0.0, gam=GAM):
    """The far field whose static state is rho = 1, c = 1 (p = 1/gam) at
    Mach `mach`: the stagnation state of the isentropic relations."""
    f = 1.0 + 0.5 * (gam - 1.0) * mach * mach
    return FarField(mach, alpha, beta, roa1=f ** (1.0 / (gam - 1.0)), aa1=math.sqrt(f),
                    pa1=f ** (gam / (gam - 1.0)) / gam)


def cell_centres(b: L.Block, x, y, z):
    X, Y, Z = b.view(x), b.view(y), b.view(z)
    def c(a):
        return 0.125 * (a[:-1, :-1, :-1] + a[:-1, :-1, 1:] + a[:-1, 1:, :-1] + a[:-1, 1:, 1:]
                        + a[1:, :-1, :-1] + a[1:, :-1, 1:] + a[1:, 1:, :-1] + a[1:, 1:, 1:])
    return c(X), c(Y), c(Z)  # shape (nkd-1, njd-1, nid-1); entry [k,j,i] = cell (i,j,k) in storage offsets


def periodic_euler(ci=64, cj=32, ck=4, seed=20261201, mach=0.8, cfl=0.8, amp=0.01, distort=0.0,
                   dt_factor=0.1) -> Case:
    lay = L.Layout([(ci, cj, ck)])
    b = lay.blocks[0]
    x = np.zeros(lay.ip11); y = np.zeros(lay.ip11); z = np.zeros(lay.ip11)
    h = 1.0 / 64.0
    K, J, I = np.meshgrid(np.arange(b.nkd), np.arange(b.njd), np.arange(b.nid), indexing="ij")
    # optional interior distortion that vanishes on the block faces, so
    # opposite faces stay congruent (a fraction `distort` of the spacing)
    bump = (np.sin(2 * np.pi * (I - b.ii1) / ci) * np.sin(2 * np.pi * (J - b.jj1) / cj)
            * np.sin(2 * np.pi * (K - b.kk1) / ck))
    b.view(x)[:] = (I - b.ii1) * h + distort * h * bump
    b.view(y)[:] = (J - b.jj1) * h + distort * h * bump
    b.view(z)[:] = (K - b.kk1) * h + distort * h * bump
    sn, vol = L.metrics(lay, x, y, z)
    Lx, Ly, Lz = ci * h, cj * h, ck * h
    rng = np.random.Generator(np.random.PCG64(seed))
    kv = rng.integers(1, 4, size=(4, 3))
    ph = rng.uniform(0, 2 * np.pi, size=4)
    xc, yc, zc = cell_centres(b, x, y, z)
    s = np.zeros_like(xc)
    for m in range(4):
        s += np.sin(2 * np.pi * (kv[m, 0] * xc / Lx + kv[m, 1] * yc / Ly + kv[m, 2] * zc / Lz) + ph[m])
    rho = np.ones(lay.ip11); u = np.full(lay.ip11, mach); vv = np.zeros(lay.ip11); w = np.zeros(lay.ip11)
    p = np.full(lay.ip11, 1.0 / GAM)
    Rv = b.view(rho); Uv = b.view(u)
    Rv[:-1, :-1, :-1] = 1.0 + amp * s
    Uv[:-1, :-1, :-1] = mach * (1.0 + amp * s)
    v0 = conservative(rho, u, vv, w, p)
    # periodic faces: each face's ghost averages its own adjacent cell with the
    # adjacent cell of the opposite face (rbvc.f90), same listing order
    parts = []
    for d in "ijk":
        g1, i1, _ = L.face_cells(b, d + "1")
        g2, i2, _ = L.face_cells(b, d + "2")
        parts.append(FaceSet(g1, i1, i2, None, [len(g1)], [0]))
        parts.append(FaceSet(g2, i2, i1, None, [len(g2)], [0]))
    coin = _cat(parts)
    coin.normals = np.zeros((3, len(coin.ghost)))
    # global step (kdtl=0): every stored dt becomes dt1min (chrono.f90:133-152).
    # sch_expli.f90:59-63 adds dt/vol * u with u the net OUTGOING flux, so the
    # explicit path runs backward in time and amplifies; at the full CFL-0.8
    # step the reference stops at step 22 (chrono's dt1min check).  dt1min is
    # dt_factor x the smallest CFL-0.8 local step of the initial field
    # (chronos.f90:150-170 on the uniform mesh: dism = h/sqrt(3)); 0.1 keeps
    # 50 steps bounded.
    inner = b.interior()
    r = v0[0, inner]
    q = np.sqrt((v0[1, inner] ** 2 + v0[2, inner] ** 2 + v0[3, inner] ** 2) / r ** 2)
    pr_ = (GAM - 1.0) * (v0[4, inner] - 0.5 * (v0[1, inner] ** 2 + v0[2, inner] ** 2 + v0[3, inner] ** 2) / r)
    c = np.sqrt(GAM * pr_ / r)
    dt1min = dt_factor * float(np.min(cfl * (h / np.sqrt(3.0)) / (q + c)))
    return Case("periodic_euler", lay, sn, vol, v0, False,
                dict(gam=GAM, pinfl=0.0, pr=0.72, prt=1.0, reynz=1e5, tnz=288.15, iflu=1),
                dict(muscl=1, ilim=0, xk=1.0 / 3.0), cfl, 0, dt1min, 0, 0,
                conservative(np.array(1.0), np.array(mach), np.array(0.0), np.array(0.0), np.array(1 / GAM)),
                coin, _empty_faces(), _empty_faces())


def wall_ns(ci=256, cj=128, ck=64, seed=20261202, mach=0.5, reynz=1e5, eta=10.0, lmax=10, ilim=1,
            beta=2.5, disp=0.05, span="farfield") -> Case:
    """span: the k1 and k2 faces are far field ("farfield") or symmetry
    planes ("sym", clsym.f90)."""
    lay = L.Layout([(ci, cj, ck)])
    b = lay.blocks[0]
    Lx, Ly, Lz = 2.0, 1.0, 0.5
    rng = np.random.Generator(np.random.PCG64(seed))
    xi = np.arange(ci + 1) / ci * Lx
    eta_ = np.arange(cj + 1) / cj
    yj = Ly * (1.0 - np.tanh(beta * (1.0 - eta_)) / np.tanh(beta))
    zk = np.arange(ck + 1) / ck * Lz
    Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
    hx, hz = Lx / ci, Lz / ck
    hy = np.gradient(yj)[None, :, None] * np.ones_like(Y)
    kv = rng.integers(1, 4, size=(3, 3))
    ph = rng.uniform(0, 2 * np.pi, size=3)
    arg = lambda m: 2 * np.pi * (kv[m, 0] * X / Lx + kv[m, 1] * Y / Ly + kv[m, 2] * Z / Lz) + ph[m]
    Xd = X + disp * hx * np.sin(arg(0))
    Yd = Y + disp * hy * np.sin(arg(1)) * (Y > 0)  # keep the wall plane y=0
    Zd = Z + disp * hz * np.sin(arg(2))
    x = np.zeros(lay.ip11); y = np.zeros(lay.ip11); z = np.zeros(lay.ip11)
    sl = (slice(b.kk1 - b.kd1, b.kk2 - b.kd1 + 1), slice(b.jj1 - b.jd1, b.jj2 - b.jd1 + 1),
          slice(b.ii1 - b.id1, b.ii2 - b.id1 + 1))
    b.view(x)[sl] = Xd; b.view(y)[sl] = Yd; b.view(z)[sl] = Zd
    sn, vol = L.metrics(lay, x, y, z)
    n = lay.ip11
    far = static_farfield(mach)
    winf = far.state()
    v0 = np.repeat(winf[:, None], n, axis=1)
    global _last_wall_distance
    yc = 0.125 * sum(Yd[dk:dk + ck, dj:dj + cj, di:di + ci] for dk in (0, 1) for dj in (0, 1) for di in (0, 1))
    dist = np.ones(n)
    b.view(dist)[b.kk1 - b.kd1:b.kk2 - b.kd1, b.jj1 - b.jd1:b.jj2 - b.jd1, b.ii1 - b.id1:b.ii2 - b.id1] = yc
    _last_wall_distance = dist
    if span not in ("farfield", "sym"):
        raise ValueError(f"span must be 'farfield' or 'sym', not {span!r}")
    wall_parts, far_parts, sym_parts = [], [], []
    for f in L.FACES:
        g, i, fi = L.face_cells(b, f)
        if f == "j1":
            wall_parts.append(FaceSet(g, i))
        else:
            nx, ny, nz = L.inward_normals(b, sn, f, fi)
            (sym_parts if span == "sym" and f[0] == "k" else far_parts).append(
                FaceSet(g, i, None, np.stack([nx, ny, nz])))
    return Case("wall_ns", lay, sn, vol, v0, True,
                dict(gam=GAM, pinfl=0.0, pr=0.72, prt=1.0, reynz=reynz, tnz=288.15, iflu=1),
                dict(muscl=1, ilim=ilim, xk=1.0 / 3.0), eta, 1, 0.0, 1, lmax, winf,
                _empty_faces(), _cat(wall_parts), _cat(far_parts), farspec=far, sym=_cat(sym_parts))


_last_wall_distance = None


def turb_constants(model="1SA"):
    """Spalart-Allmaras (atlecdon.f90:183-204) and Wilcox k-omega
    (atlecdon.f90:353-366) constants in the order of TurbCfg.tc."""
    cb1, sigma, cb2, kappa = 0.1355, 2.0 / 3.0, 0.622, 0.41
    cw1 = cb1 / kappa ** 2 + (1 + cb2) / sigma
    return [cb1, cb2, sigma, kappa, cw1, 0.3, 2.0, 7.1, 1.1, 2.0, 0.09, 0.075, 0.5, 0.5, 1.0, 1.3]


def wall_rans(ci=48, cj=32, ck=16, model="1SA", klroe=2, kwsst=0, nut_ratio=3.0, **kw) -> Case:
    """The curvilinear wall block of wall_ns with a turbulence model: nu~ =
    nut_ratio x the laminar kinematic viscosity (SA), or k = 1e-6 U^2 and
    omega with mut/mu = nut_ratio (k-omega), uniform; wall distance = y of
    the cell centre."""
    c = wall_ns(ci, cj, ck, **kw)
    return with_turbulence(c, _last_wall_distance.copy(), model, klroe, kwsst, nut_ratio)


def with_turbulence(c: Case, dist, model="1SA", klroe=2, kwsst=0, nut_ratio=3.0) -> Case:
    """Case c with a turbulence model added: planes 6-7 of the state (SA:
    nu~ = nut_ratio x the laminar kinematic viscosity; k-omega: k = 1e-6
    |U|^2 and omega with mut/mu = nut_ratio), uniform, and the wall distance
    `dist` (storage array)."""
    n = c.layout.ip11
    mach = float(np.sqrt(np.sum(c.winf[1:4] ** 2)) / c.winf[0])
    mu_inf = 1.0 / c.phys["reynz"]  # Sutherland at T = 1 (zvismo.f90:90-104)
    v0 = np.zeros((7, n)); v0[:5] = c.v0
    if model == "1SA":
        v0[5] = nut_ratio * mu_inf
    else:
        k = 1e-6 * mach * mach
        v0[5] = k; v0[6] = k / (nut_ratio * mu_inf)
    c.v0 = v0
    c.name = f"{c.name}_{model}"
    c.turb = TurbCfg(model, turb_constants(model), dist, klroe=klroe, kwsst=kwsst, rokinf=float(v0[5, 0]),
                     roeinf=float(v0[6, 0]))
    if c.turb.kparoi == 1:
        c.turb.wall_lines, c.turb.wall_ghost, c.turb.wall_start, c.turb.mnpar = j1_wall_lines(c)
    return c


# record of one wall point for met_parko / met_mtcorf1 (codens_wall_line)
WALL_LINE = np.dtype([("nmin", "<i8"), ("m3min", "<i4"), ("m3max", "<i4"), ("isens3", "<i4"), ("dm3", "<i4"),
                      ("mpn0", "<i4"), ("mpn1", "<i4")])


def j1_wall_lines(c: Case):
    """The wall points of a single block's j1 wall (c.wall, in its listing
    order), as atindnor.f90 describes them: first cell nmin, the j index from
    jj1 to jj2 upward (dm3 = nid), wall points numbered 1..npts; every real
    cell's wall point is the one of its (i, k) column (mnpar, 1-based)."""
    b = c.layout.blocks[0]
    assert len(c.layout.blocks) == 1
    npts = len(c.wall.ghost)
    ln = np.zeros(npts, dtype=WALL_LINE)
    ln["nmin"] = c.wall.inner
    ln["m3min"], ln["m3max"], ln["isens3"], ln["dm3"] = b.jj1, b.jj2, 1, b.nid
    ln["mpn0"], ln["mpn1"] = 0, npts
    mnpar = np.zeros(c.layout.ip11, dtype=np.int32)
    ni = b.ii2 - b.ii1
    for j in range(b.jj1, b.jj2):
        k, i = np.meshgrid(np.arange(b.kk1, b.kk2), np.arange(b.ii1, b.ii2), indexing="ij")
        mnpar[b.idx(i, j, k).ravel()] = ((k - b.kk1) * ni + (i - b.ii1)).ravel() + 1
    assert np.array_equal(mnpar[c.wall.inner], np.arange(1, npts + 1))
    return ln, np.asarray(c.wall.ghost, dtype=np.int64), np.array([0, npts], dtype=np.int64), mnpar


def inclined_rans(ci=32, cj=16, ck=8, model="2KO", klroe=2, kwsst=0, nut_ratio=3.0, eta=10.0, **kw) -> Case:
    """The inclined-free-stream wall block of adjoint_wall (no bump, no state
    perturbation: the case that converges laminar) with a turbulence model;
    wall distance = y of the cell centre."""
    c = adjoint_wall(ci, cj, ck, eta=eta, amp=0.0, bump=0.0, **kw)
    b = c.layout.blocks[0]
    x, y, z = c.nodes
    _, yc, _ = cell_centres(b, x, y, z)
    dist = np.ones(c.layout.ip11)
    r = (slice(b.kk1 - b.kd1, b.kk2 - b.kd1), slice(b.jj1 - b.jd1, b.jj2 - b.jd1), slice(b.ii1 - b.id1, b.ii2 - b.id1))
    b.view(dist)[r] = yc[r]
    return with_turbulence(c, dist, model, klroe, kwsst, nut_ratio)


def _face_grid(b: L.Block, face: str):
    """Cell counts of a block face in face_cells' order (outer, inner)."""
    ci, cj, ck = b.ni - 1, b.nj - 1, b.nk - 1
    return {"i": (ck, cj), "j": (ck, ci), "k": (cj, ci)}[face[0]]


def _opposite(face: str) -> str:
    return face[0] + ("1" if face[1] == "2" else "2")


def multiblock_dims(nblocks=48, lo=16, hi=192, total=(38_000_000, 42_000_000), seed=20261203):
    """Seeded block dimensions, each axis uniform in lo..hi cells, drawn again
    (all blocks together) until the total cell count falls in `total`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(1_000_000):
        dims = rng.integers(lo, hi + 1, size=(nblocks, 3))
        tot = int(np.prod(dims, axis=1).sum())
        if total is None or total[0] <= tot <= total[1]:
            return [tuple(int(x) for x in d) for d in dims], rng
    raise RuntimeError("no block dimensions in the requested total")


def multiblock_sa(nblocks=48, lo=16, hi=192, total=(38_000_000, 42_000_000), seed=20261203, mach=0.7,
                  reynz=5e6, nut_ratio=3.0, h=1.0 / 128.0, eta=10.0, lmax=10, ilim=1) -> Case:
    """Uneven multi-block RANS case: `nblocks` Cartesian blocks of spacing h
    joined 1-to-1 through coincident face patches (rbvc.f90) along a seeded
    adjacency tree, Spalart-Allmaras, implicit (kmf=1) with local steps.

    Tree: block b (1..n-1) attaches to a free face of a block drawn among
    0..b-1 through its opposite face, so every block sits on one integer
    lattice of spacing h and the two faces' cells coincide.  The patch is the
    low corner of both faces, min(extent) in each tangential direction; the
    rest of every face, and every face left free, is far field (clnrd), except
    the j1 face of block 0, which is an adiabatic no-slip wall (clpara) and
    the plane the wall distance is measured from.  nu~ = nut_ratio x the
    laminar kinematic viscosity everywhere."""
    dims, rng = multiblock_dims(nblocks, lo, hi, total, seed)
    lay = L.Layout(dims)
    B = lay.blocks
    free = [set(L.FACES) for _ in B]
    free[0].discard("j1")
    origin = np.zeros((nblocks, 3), dtype=np.int64)
    links = []
    for b in range(1, nblocks):
        cands = [(p, f) for p in range(b) for f in L.FACES if f in free[p]]
        p, f = cands[int(rng.integers(len(cands)))]
        d = "ijk".index(f[0])
        g = _opposite(f)
        origin[b] = origin[p]
        origin[b, d] = origin[p, d] + dims[p][d] if f[1] == "2" else origin[p, d] - dims[b][d]
        free[p].discard(f)
        free[b].discard(g)
        links.append((p, f, b, g))
    n = lay.ip11
    x = np.zeros(n); y = np.zeros(n); z = np.zeros(n)
    for l, b in enumerate(B):
        ci, cj, ck = dims[l]
        zk = (origin[l, 2] + np.arange(ck + 1)) * h
        yj = (origin[l, 1] + np.arange(cj + 1)) * h
        xi = (origin[l, 0] + np.arange(ci + 1)) * h
        Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
        sl = (slice(b.kk1 - b.kd1, b.kk2 - b.kd1 + 1), slice(b.jj1 - b.jd1, b.jj2 - b.jd1 + 1),
              slice(b.ii1 - b.id1, b.ii2 - b.id1 + 1))
        b.view(x)[sl] = X; b.view(y)[sl] = Y; b.view(z)[sl] = Z
    sn, vol = L.metrics(lay, x, y, z)
    # faces: coincident patches first, the rest far field / wall
    used = {(l, f): np.zeros(_face_grid(B[l], f), dtype=bool) for l in range(nblocks) for f in L.FACES}
    cells = {}
    def fc(l, f):
        if (l, f) not in cells:
            g, i, fi = L.face_cells(B[l], f)
            s = _face_grid(B[l], f)
            cells[(l, f)] = (g.reshape(s), i.reshape(s), fi.reshape(s))
        return cells[(l, f)]
    coin = []
    for p, f, b, g in links:
        gp, ip_, _ = fc(p, f)
        gb, ib, _ = fc(b, g)
        a0, a1 = min(gp.shape[0], gb.shape[0]), min(gp.shape[1], gb.shape[1])
        P = (slice(0, a0), slice(0, a1))
        coin.append(FaceSet(gp[P].ravel(), ip_[P].ravel(), ib[P].ravel(), None, [a0 * a1], [p]))
        coin.append(FaceSet(gb[P].ravel(), ib[P].ravel(), ip_[P].ravel(), None, [a0 * a1], [b]))
        used[(p, f)][P] = True
        used[(b, g)][P] = True
    wall_parts, far_parts = [], []
    for l, b in enumerate(B):
        for f in L.FACES:
            g, i, fi = fc(l, f)
            m = ~used[(l, f)]
            if not m.any():
                continue
            g, i, fi = g[m], i[m], fi[m]
            if l == 0 and f == "j1":
                wall_parts.append(FaceSet(g, i, None, None, [len(g)], [l]))
            else:
                nx, ny, nz = L.inward_normals(b, sn, f, fi)
                far_parts.append(FaceSet(g, i, None, np.stack([nx, ny, nz]), [len(g)], [l]))
    coinc = _cat(coin)
    coinc.normals = np.zeros((3, len(coinc.ghost)))
    mu_inf = 1.0 / reynz  # Sutherland at T = 1 (zvismo.f90:90-104)
    far = static_farfield(mach)
    w = far.state()  # far-field state of the mean flow (clnrd)
    v0 = np.zeros((7, n))
    v0[:5] = w[:, None]
    v0[5] = nut_ratio * mu_inf
    y0 = origin[0, 1] * h
    dist = np.ones(n)
    for l, b in enumerate(B):
        ci, cj, ck = dims[l]
        yc = (origin[l, 1] + np.arange(cj) + 0.5) * h
        b.view(dist)[b.kk1 - b.kd1:b.kk2 - b.kd1, b.jj1 - b.jd1:b.jj2 - b.jd1, b.ii1 - b.id1:b.ii2 - b.id1] = \
            np.abs(yc - y0)[None, :, None]
    c = Case(f"multiblock_sa_{nblocks}", lay, sn, vol, v0, True,
             dict(gam=GAM, pinfl=0.0, pr=0.72, prt=1.0, reynz=reynz, tnz=288.15, iflu=1),
             dict(muscl=1, ilim=ilim, xk=1.0 / 3.0), eta, 1, 0.0, 1, lmax, w,
             coinc, _cat(wall_parts), _cat(far_parts), farspec=far)
    c.turb = TurbCfg("1SA", turb_constants("1SA"), dist, klroe=2, rokinf=float(v0[5, 0]), roeinf=float(v0[6, 0]))
    c.links = links
    c.origin = origin
    return c


def adjoint_wall(ci=128, cj=64, ck=32, seed=20261205, mach=0.5, reynz=1e5, alpha=5.0, beta=-10.0, amp=0.01,
                 bump=0.01, tanh_beta=2.5, disp=0.05, eta=1.0, nwall_nodes=4096) -> Case:
    """Viscous wall block for the discrete adjoint: the curvilinear mesh of
    wall_ns (seeded 5% node displacement, tanh stretching toward the no-slip
    adiabatic wall at j=1, far field elsewhere) with a smooth wall bump
    y += bump sin(2 pi x / Lx) cos(2 pi z / Lz) (1 - eta_j)^2, so that the
    wall faces have an x component; laminar Navier-Stokes, Roe + unlimited
    MUSCL (ilim = 0); free stream at Mach `mach` with the incidence angles
    alpha and beta of utnrd.f90 (direction (cos a cos b, -sin b, sin a cos b);
    the defaults point 10 degrees away from the wall and 5 across the span),
    so that no far-field face has a zero normal velocity (the switch of
    clnrd.f90); initial state = free stream x (1 + amp x 4
    seeded sinusoidal modes) on every variable.  Local steps at CFL `eta`.

    Extra attributes: nodes (x, y, z storage arrays, ghost nodes
    extrapolated), wall_nodes (flat indices of `nwall_nodes` seeded wall
    nodes of the (ci+1) x (ck+1) on j=1)."""
    lay = L.Layout([(ci, cj, ck)])
    b = lay.blocks[0]
    Lx, Ly, Lz = 2.0, 1.0, 0.5
    rng = np.random.Generator(np.random.PCG64(seed))
    xi = np.arange(ci + 1) / ci * Lx
    eta_ = np.arange(cj + 1) / cj
    yj = Ly * (1.0 - np.tanh(tanh_beta * (1.0 - eta_)) / np.tanh(tanh_beta))
    zk = np.arange(ck + 1) / ck * Lz
    Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
    E = np.meshgrid(zk, eta_, xi, indexing="ij")[1]
    hx, hz = Lx / ci, Lz / ck
    hy = np.gradient(yj)[None, :, None] * np.ones_like(Y)
    kv = rng.integers(1, 4, size=(3, 3))
    ph = rng.uniform(0, 2 * np.pi, size=3)
    arg = lambda m: 2 * np.pi * (kv[m, 0] * X / Lx + kv[m, 1] * Y / Ly + kv[m, 2] * Z / Lz) + ph[m]
    Xd = X + disp * hx * np.sin(arg(0))
    Yd = Y + disp * hy * np.sin(arg(1)) * (Y > 0)
    Zd = Z + disp * hz * np.sin(arg(2))
    Yd = Yd + bump * np.sin(2 * np.pi * X / Lx) * np.cos(2 * np.pi * Z / Lz) * (1.0 - E) ** 2
    x = np.zeros(lay.ip11); y = np.zeros(lay.ip11); z = np.zeros(lay.ip11)
    sl = (slice(b.kk1 - b.kd1, b.kk2 - b.kd1 + 1), slice(b.jj1 - b.jd1, b.jj2 - b.jd1 + 1),
          slice(b.ii1 - b.id1, b.ii2 - b.id1 + 1))
    b.view(x)[sl] = Xd; b.view(y)[sl] = Yd; b.view(z)[sl] = Zd
    sn, vol = L.metrics(lay, x, y, z)  # also extrapolates the ghost nodes
    n = lay.ip11
    far = static_farfield(mach, alpha, beta)
    winf = far.state()
    xc, yc, zc = cell_centres(b, x, y, z)
    kvs = rng.integers(1, 4, size=(4, 3))
    phs = rng.uniform(0, 2 * np.pi, size=(4, 5))
    v0 = np.repeat(winf[:, None], n, axis=1).astype(np.float64)
    for e in range(5):
        s = np.zeros_like(xc)
        for m in range(4):
            s += np.sin(2 * np.pi * (kvs[m, 0] * xc / Lx + kvs[m, 1] * yc / Ly + kvs[m, 2] * zc / Lz) + phs[m, e])
        vb = b.view(v0[e])
        vb[:-1, :-1, :-1] = winf[e] * (1.0 + amp * s)
    wall_parts, far_parts = [], []
    for f in L.FACES:
        g, i, fi = L.face_cells(b, f)
        if f == "j1":
            wall_parts.append(FaceSet(g, i))
        else:
            nx, ny, nz = L.inward_normals(b, sn, f, fi)
            far_parts.append(FaceSet(g, i, None, np.stack([nx, ny, nz])))
    c = Case("adjoint_wall", lay, sn, vol, v0, True,
             dict(gam=GAM, pinfl=0.0, pr=0.72, prt=1.0, reynz=reynz, tnz=288.15, iflu=1),
             dict(muscl=1, ilim=0, xk=1.0 / 3.0), eta, 1, 0.0, 1, 10, winf,
             _empty_faces(), _cat(wall_parts), _cat(far_parts), farspec=far)
    c.nodes = (x, y, z)
    ii, kk = np.meshgrid(np.arange(b.ii1, b.ii2 + 1), np.arange(b.kk1, b.kk2 + 1), indexing="ij")
    allw = b.idx(ii.ravel(), b.jj1, kk.ravel()).astype(np.int64)
    pick = np.sort(rng.choice(len(allw), size=min(nwall_nodes, len(allw)), replace=False))
    c.wall_nodes = allw[pick]
    return c

"""Seeded FASOLT3 problem generator: brick meshes, boundary tables, texture, material.

This stands in for the reference preprocessor FS3PRE at the sizes the GPU
path runs (the preprocessor's node merge is O(n^2) and its static limits stop
at 8500 elements).  The arrays it produces follow the reference's module
layout (src/fasolt/fas-com-mod.f90:9-80) and .fsd record contents
(src/fasolt/Faslt_io_nowin.f90:29-128):

* elements are generated x-fastest, then y, then z (src/prepro/fs3pre.f90:68-141)
  with the brick node order of src/prepro/fs3pre.f90:193-196;
* boundary codes follow the SETBOUNDS face coding (src/prepro/fs3p-bkc.f90:32-171):
  ``ib = 100*table`` for a kinematic table, ``1`` for a free surface node;
* kinematic tables follow SETKINC (src/prepro/fs3p-bkc.f90:182-305): entries 1-6
  the constraint pattern, 7/11/15 the prescribed displacement increments, which
  are ``eps_inc * current length`` so that the imposed strain is logarithmic,
  and ``tiem = sqrt(sum eps_inc^2)`` (unit reference slip rate);
* orientations are Bunge angles in radians, ``phi1 = 2 pi u1, Phi = acos(2 u2 - 1),
  phi2 = 2 pi u3`` (src/prepro/fs3p-gps.f90:105-110), one per element.

Every real the generator emits is rounded to float32 so that the REAL(4)
reference and the FP64 path read identical inputs.
"""
from __future__ import annotations

import dataclasses
import itertools
import math
from typing import Dict, List, Optional

import numpy as np

MAXSS = 60   # fas-com-mod.f90:55
MAXSV = 63   # fas-com-mod.f90:14
NTAB = 99    # x_spec(1:15, 1:99, inc)

# 8-node brick: 1 point (reference SHAPE_12) or 2x2x2 points (new element type)
LT_BRICK_1GP = 12
LT_BRICK_8GP = 13
LT_BRICK_20 = 22     # 20-node brick, 2x2x2 points (SHAPE_22, Shapef.f90:210)
LT_TET10 = 21        # 10-node tetrahedron, 4 points (SHAPE_21, Shapef.f90:104)
LT_WEDGE15 = 23      # 15-node pentahedron, 6 points (SHAPE_23, Shapef.f90:360)

# local coordinates of the 20 nodes of SHAPE_22 (Shapef.f90:191-200): bottom
# corners 1,3,5,7 and mid-sides 2,4,6,8 anticlockwise, 9-12 halfway up the
# vertical edges, top face 13-20 like the bottom
BRICK20 = [(-1, -1, -1), (0, -1, -1), (1, -1, -1), (1, 0, -1), (1, 1, -1), (0, 1, -1), (-1, 1, -1), (-1, 0, -1),
           (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0),
           (-1, -1, 1), (0, -1, 1), (1, -1, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1), (-1, 1, 1), (-1, 0, 1)]


def f32(a):
    """Round to float32 and return float64 (identical inputs for FP32 and FP64)."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@dataclasses.dataclass
class Material:
    nss: int
    e: np.ndarray            # (9,) compliances S11 S22 S33 S23 S31 S12 S44 S55 S66, MPa^-1
    ssang: np.ndarray        # (nss,)
    ab: np.ndarray           # (nss, 3) slip directions (crystal axes)
    an: np.ndarray           # (nss, 3) plane normals
    ampar: np.ndarray        # (nss, 10) m, theta0, thetaIV, alpha, tauS, n, ...
    smatrx: np.ndarray       # (nss, nss) smatrx[i, j] = Fortran smatrx(j, i)
    crss: np.ndarray         # (nss,) initial slip resistances sv(4:3+nss)
    twinlim: int = 0
    twinprob: float = 0.0


@dataclasses.dataclass
class Control:
    """fafcdf.dat contents (Fasolt3.f90:216-218) plus the SOR settings."""
    toler: float = 1.0e-3
    maxit: int = 100
    alpha: float = 20.0
    afactor: float = 1.0
    efix: float = 1.0e20
    smalp: float = 1.0e-1
    smalf: float = 10.0
    smald: float = 1.0e-5
    smale: float = 0.5
    smalr: float = 0.5
    sor_omega: float = 1.2       # Stif_sol.f90:173
    sor_maxit: int = 1000        # Stif_sol.f90:173
    sor_tol: float = 1.0e-3      # reference: toler
    sor_kind: int = 0            # 0: reference |dx|^2/|x|^2 < tol^2, 1: |D^-1(b-Ax)|/|D^-1 b| < tol, D = diag(A)
    stall_tol: float = 0.0       # 0: CPSTRS to maxit as the reference; > 0: stop at a stationary stress


@dataclasses.dataclass
class Problem:
    nlmnt: int
    nnod: int
    ngps: int
    nsv: int
    le: np.ndarray           # (nlmnt,) material code, negative = elastic only
    lt: np.ndarray           # (nlmnt,) element type
    nd: np.ndarray           # (nlmnt,) nodes per element
    np_: np.ndarray          # (nlmnt,) points per element
    ln: np.ndarray           # (nlmnt, 20) 1-based node numbers
    lg: np.ndarray           # (nlmnt, 8) 1-based point numbers
    ib: np.ndarray           # (nnod,) boundary codes
    x: np.ndarray            # (3 nnod,)
    dx: np.ndarray           # (3 nnod,)
    sv: np.ndarray           # (ngps, 63)
    tiem: np.ndarray         # (ninc1,)
    x_spec: np.ndarray       # (ninc1, 99, 15)
    f_spec: np.ndarray       # (ninc1, 99, 3)
    materials: List[Material]
    control: Control
    nxs: int = 0
    nfs: int = 0
    dims: tuple = (0, 0, 0)

    @property
    def ndof(self) -> int:
        return 3 * self.nnod

    @property
    def ninc1(self) -> int:
        return int(self.tiem.shape[0])

    def zero_state(self) -> Dict[str, np.ndarray]:
        """Initial point state of a fresh .fsd file (all increments zero)."""
        g = self.ngps
        return dict(
            eps=np.zeros(g), dep=np.zeros(g),
            sv=self.sv.copy(), dsv=np.zeros((g, MAXSV)),
            ep=np.zeros((g, 6)), de=np.zeros((g, 6)),
            st=np.zeros((g, 6)), ds=np.zeros((g, 6)),
            rt=np.zeros((g, 3)), dr=np.zeros((g, 3)),
            srate=np.zeros((g, MAXSS)), cumslip=np.zeros((g, MAXSS)),
            twinsys=np.zeros(g, np.int32), timestwinned=np.zeros(g, np.int32),
        )


# ---------------------------------------------------------------------------
# materials
# ---------------------------------------------------------------------------

def cubic_compliance(c11: float, c12: float, c44: float) -> np.ndarray:
    """Cubic stiffness (GPa) -> the reference's e(1:9) compliances in MPa^-1."""
    c11, c12, c44 = (1e3 * c11, 1e3 * c12, 1e3 * c44)
    den = (c11 - c12) * (c11 + 2.0 * c12)
    s11 = (c11 + c12) / den
    s12 = -c12 / den
    s44 = 1.0 / c44
    return f32([s11, s11, s11, s12, s12, s12, s44, s44, s44])


def hexagonal_compliance(c11, c12, c13, c33, c44) -> np.ndarray:
    """Transversely isotropic stiffness (GPa, c axis = crystal 3) -> e(1:9)."""
    c = np.zeros((6, 6))
    c[0, 0] = c[1, 1] = c11
    c[0, 1] = c[1, 0] = c12
    c[0, 2] = c[2, 0] = c[1, 2] = c[2, 1] = c13
    c[2, 2] = c33
    c[3, 3] = c[4, 4] = c44
    c[5, 5] = 0.5 * (c11 - c12)
    s = np.linalg.inv(1e3 * c)   # engineering-shear compliance, MPa^-1
    return f32([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[2, 0], s[0, 1], s[3, 3], s[4, 4], s[5, 5]])


def fcc_slip_systems() -> tuple:
    """The 12 {111}<110> systems in the order of sample_input_files/aluminium2.fmt:5-16."""
    n = np.array([[1, 1, 1]] * 3 + [[1, 1, -1]] * 3 + [[-1, 1, 1]] * 3 + [[1, -1, 1]] * 3, float)
    d = np.array([[0, 1, -1], [-1, 0, 1], [1, -1, 0],
                  [0, 1, 1], [-1, 0, -1], [1, -1, 0],
                  [0, 1, -1], [1, 0, 1], [-1, -1, 0],
                  [0, -1, -1], [-1, 0, 1], [1, 1, 0]], float)
    assert np.allclose((n * d).sum(1), 0.0)
    return f32(d / np.linalg.norm(d, axis=1)[:, None]), f32(n / np.linalg.norm(n, axis=1)[:, None])


def fcc_aluminium(tau0=16.0, taus=60.0, h0=200.0, latent=1.4, m=0.02) -> Material:
    """FCC aluminium: C11=108.2, C12=61.3, C44=28.5 GPa, Voce hardening."""
    ab, an = fcc_slip_systems()
    nss = 12
    ampar = np.zeros((nss, 10))
    ampar[:, 0] = m        # rate sensitivity
    ampar[:, 1] = h0       # theta0
    ampar[:, 2] = 0.0      # thetaIV
    ampar[:, 3] = 1.0      # alpha index
    ampar[:, 4] = taus     # saturation stress at unit rate
    ampar[:, 5] = 0.0      # saturation rate sensitivity
    sm = np.full((nss, nss), latent)
    np.fill_diagonal(sm, 1.0)
    return Material(nss=nss, e=cubic_compliance(108.2, 61.3, 28.5), ssang=np.zeros(nss),
                    ab=ab, an=an, ampar=f32(ampar), smatrx=f32(sm), crss=f32(np.full(nss, tau0)))


def _mb_direction(u, v, t, w, ca):
    """Miller-Bravais direction [uvtw] -> Cartesian (a1 along x, c along z)."""
    a1 = np.array([1.0, 0.0, 0.0])
    a2 = np.array([-0.5, math.sqrt(3) / 2, 0.0])
    a3 = np.array([-0.5, -math.sqrt(3) / 2, 0.0])
    c = np.array([0.0, 0.0, ca])
    return u * a1 + v * a2 + t * a3 + w * c


def _mb_plane(h, k, i, l, ca):
    """Miller-Bravais plane (hkil) -> Cartesian normal."""
    # reciprocal-lattice normal of the hexagonal cell with a = 1
    n = np.array([h, (2 * k + h) / math.sqrt(3), l / ca])
    return n


def bcc_iron(tau0=50.0, taus=150.0, h0=300.0, latent=1.4, m=0.02) -> Material:
    """BCC iron: the 12 {110}<111> and 12 {112}<111> systems (24, a count with no
    kernel of its own); C11=231.4, C12=134.7, C44=116.4 GPa, Voce hardening."""
    dirs, norms = [], []
    for d in ((1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1)):
        d = np.array(d, float)
        for fam in ((1, 1, 0), (1, 1, 2)):
            cand = set()
            for p in itertools.permutations(fam):
                for sg in itertools.product((1, -1), repeat=3):
                    nrm = tuple(a * b for a, b in zip(p, sg))
                    if np.dot(nrm, d) == 0 and tuple(-x for x in nrm) not in cand:
                        cand.add(nrm)
            for nrm in sorted(cand, reverse=True):
                dirs.append(d)
                norms.append(np.array(nrm, float))
    d = np.array(dirs)
    n = np.array(norms)
    d /= np.linalg.norm(d, axis=1)[:, None]
    n /= np.linalg.norm(n, axis=1)[:, None]
    nss = len(d)
    assert nss == 24 and np.allclose((d * n).sum(1), 0.0)
    ampar = np.zeros((nss, 10))
    ampar[:, 0] = m
    ampar[:, 1] = h0
    ampar[:, 3] = 1.0
    ampar[:, 4] = taus
    sm = np.full((nss, nss), latent)
    np.fill_diagonal(sm, 1.0)
    return Material(nss=nss, e=cubic_compliance(231.4, 134.7, 116.4), ssang=np.zeros(nss), ab=f32(d), an=f32(n),
                    ampar=f32(ampar), smatrx=f32(sm), crss=f32(np.full(nss, tau0)))


def hcp_magnesium(ca=1.624, pyramidal_a=False) -> Material:
    """Mg-like HCP: 3 basal, 3 prismatic, 6 pyramidal <c+a>, 6 {10-12} tensile twins.

    Elastic C11=59.7 C12=26.2 C13=21.7 C33=61.7 C44=16.4 GPa; CRSS 5/40/80/12 MPa.
    Twins carry ssang = 86.3 degrees (crys-cons.f90:165 treats |ssang| > 1.5 as twin).
    pyramidal_a adds the 6 {10-11}<11-20> pyramidal <a> systems (CRSS 60 MPa)
    after the prismatic ones: 24 systems, a count with no kernel of its own.
    """
    dirs, norms, kind = [], [], []
    basal_d = [(2, -1, -1, 0), (-1, 2, -1, 0), (-1, -1, 2, 0)]
    for d in basal_d:
        dirs.append(_mb_direction(*d, ca)); norms.append(_mb_plane(0, 0, 0, 1, ca)); kind.append(0)
    prism = [((0, 1, -1, 0), (2, -1, -1, 0)), ((1, 0, -1, 0), (-1, 2, -1, 0)), ((1, -1, 0, 0), (1, 1, -2, 0))]
    for p, d in prism:
        dirs.append(_mb_direction(*d, ca)); norms.append(_mb_plane(*p, ca)); kind.append(1)
    if pyramidal_a:
        for p in ((1, 0, -1, 1), (0, 1, -1, 1), (-1, 1, 0, 1), (-1, 0, 1, 1), (0, -1, 1, 1), (1, -1, 0, 1)):
            nv = _mb_plane(*p, ca)
            dv = [v for v in (_mb_direction(*b, ca) for b in basal_d) if abs(np.dot(v, nv)) < 1e-12]
            dirs.append(dv[0]); norms.append(nv); kind.append(4)
    pyr = [((1, 1, -2, 2), (-1, -1, 2, 3)), ((-1, 2, -1, 2), (1, -2, 1, 3)), ((-2, 1, 1, 2), (2, -1, -1, 3)),
           ((-1, -1, 2, 2), (1, 1, -2, 3)), ((1, -2, 1, 2), (-1, 2, -1, 3)), ((2, -1, -1, 2), (-2, 1, 1, 3))]
    for p, d in pyr:
        dirs.append(_mb_direction(*d, ca)); norms.append(_mb_plane(*p, ca)); kind.append(2)
    twin = [((1, 0, -1, 2), (-1, 0, 1, 1)), ((0, 1, -1, 2), (0, -1, 1, 1)), ((-1, 1, 0, 2), (1, -1, 0, 1)),
            ((-1, 0, 1, 2), (1, 0, -1, 1)), ((0, -1, 1, 2), (0, 1, -1, 1)), ((1, -1, 0, 2), (-1, 1, 0, 1))]
    for p, d in twin:
        dirs.append(_mb_direction(*d, ca)); norms.append(_mb_plane(*p, ca)); kind.append(3)
    d = np.array(dirs); n = np.array(norms)
    d /= np.linalg.norm(d, axis=1)[:, None]
    n /= np.linalg.norm(n, axis=1)[:, None]
    assert np.allclose((d * n).sum(1), 0.0, atol=1e-12)
    nss = len(kind)
    crss_by_kind = [5.0, 40.0, 80.0, 12.0, 60.0]
    ampar = np.zeros((nss, 10))
    ampar[:, 0] = 0.02
    ampar[:, 1] = 200.0
    ampar[:, 3] = 1.0
    ampar[:, 4] = [3.0 * crss_by_kind[k] + 20.0 for k in kind]
    ssang = np.array([86.3 if k == 3 else 0.0 for k in kind])
    sm = np.full((nss, nss), 1.4)
    np.fill_diagonal(sm, 1.0)
    return Material(nss=nss, e=hexagonal_compliance(59.7, 26.2, 21.7, 61.7, 16.4), ssang=f32(ssang),
                    ab=f32(d), an=f32(n), ampar=f32(ampar), smatrx=f32(sm),
                    crss=f32([crss_by_kind[k] for k in kind]), twinlim=1, twinprob=1.0)


# ---------------------------------------------------------------------------
# meshes and boundary tables
# ---------------------------------------------------------------------------

BRICK = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                  (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])


def random_euler(rng: np.random.Generator, n: int) -> np.ndarray:
    u = rng.random((n, 3))
    return np.stack([2 * math.pi * u[:, 0], np.arccos(2 * u[:, 1] - 1), 2 * math.pi * u[:, 2]], 1)


def basal_fibre_euler(rng: np.random.Generator, n: int, spread_deg: float = 15.0) -> np.ndarray:
    u = rng.random((n, 2))
    phi = np.abs(rng.normal(0.0, math.radians(spread_deg), n))
    return np.stack([2 * math.pi * u[:, 0], phi, 2 * math.pi * u[:, 1]], 1)


def _tet10_cells():
    """The 6 tetrahedra of a cell around its (0,0,0)-(2,2,2) diagonal, nodes in
    SHAPE_21 order (corners 1,3,5,10; 2,4,6 between 1-3, 3-5, 1-5; 7,8,9
    between 10 and 1, 3, 5), oriented so that SHAPE_21's weight is positive."""
    out = []
    for perm in itertools.permutations(range(3)):
        v = [np.zeros(3, int)]
        for ax in perm:
            nxt = v[-1].copy()
            nxt[ax] += 2
            v.append(nxt)
        c1, c3, c5, c10 = v[0], v[1], v[2], v[3]
        # SHAPE_21's local axes are (L10, L3, L1) with L5 = 1 - sum: weight =
        # -det/24 with det = (x10 - x5) . ((x3 - x5) x (x1 - x5))
        if np.dot(c10 - c5, np.cross(c3 - c5, c1 - c5)) > 0:
            c1, c3 = c3, c1
        mid = lambda a, b: (a + b) // 2
        out.append([tuple(x) for x in (c1, mid(c1, c3), c3, mid(c3, c5), c5, mid(c1, c5),
                                       mid(c1, c10), mid(c3, c10), mid(c5, c10), c10)])
    return out


def _wedge15_cells():
    """The 2 prisms of a cell on its xy diagonal, nodes in SHAPE_23 order (base
    corners 1,3,5 anticlockwise seen from above, top corners 10,12,14; 2,4,6 on
    the base edges, 7,8,9 on the vertical ones, 11,13,15 on the top edges)."""
    out = []
    for tri in (((0, 0), (2, 0), (2, 2)), ((0, 0), (2, 2), (0, 2))):
        b = [np.array([x, y, 0]) for x, y in tri]
        t = [p + np.array([0, 0, 2]) for p in b]
        mid = lambda a, c: (a + c) // 2
        out.append([tuple(x) for x in (b[0], mid(b[0], b[1]), b[1], mid(b[1], b[2]), b[2], mid(b[2], b[0]),
                                       mid(b[0], t[0]), mid(b[1], t[1]), mid(b[2], t[2]),
                                       t[0], mid(t[0], t[1]), t[1], mid(t[1], t[2]), t[2], mid(t[2], t[0]))])
    return out


def brick_problem(n, length=1.0, *, lt=LT_BRICK_8GP, material: Optional[Material] = None,
                  faces: Optional[Dict[str, object]] = None, eps_inc=(0.0, 0.0, 1.0e-3),
                  ninc=50, seed=0, texture="random", control: Optional[Control] = None,
                  elastic_elements: Optional[np.ndarray] = None) -> Problem:
    """Structured n^3 (or (nx,ny,nz)) brick of hexahedra: 8-node (lt 12, 13) or
    20-node (lt 22, corner and mid-edge nodes of a doubled grid).

    ``faces`` maps 'xmin','xmax','ymin','ymax','zmin','zmax' to 'fixed' (zero
    normal displacement), 'spec' (normal displacement increment
    eps_inc[axis] * current length) or 'free'.  Default: uniaxial straining
    along z with symmetry planes at the min faces.
    """
    nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
    lx, ly, lz = (length,) * 3 if np.isscalar(length) else tuple(length)
    material = material or fcc_aluminium()
    control = control or Control()
    faces = faces or dict(xmin="fixed", ymin="fixed", zmin="fixed", zmax="spec")
    rng = np.random.default_rng(seed)

    nel = nx * ny * nz
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # element number x fastest
    eidx = (ei + nx * (ej + ny * ek)).ravel()
    base = np.zeros((nel, 3), int)
    base[eidx] = np.stack([ei.ravel(), ej.ravel(), ek.ravel()], 1)
    ln = np.zeros((nel, 20), np.int32)
    if lt in (LT_TET10, LT_WEDGE15):
        # each cell of the doubled grid's 3x3x3 points split into 6 tetrahedra
        # around its (0,0,0)-(1,1,1) diagonal or 2 prisms on its xy diagonal;
        # every edge midpoint is a grid point, and the splits conform
        gx, gy, gz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
        cells = []
        for e in range(nel):
            b = 2 * base[e]
            for loc in (_tet10_cells() if lt == LT_TET10 else _wedge15_cells()):
                cells.append([(b[0] + p[0], b[1] + p[1], b[2] + p[2]) for p in loc])
        cells = np.array(cells)                                 # (nel * k, nn, 3) grid indices
        flat = cells[..., 0] + gx * (cells[..., 1] + gy * cells[..., 2])
        order = np.unique(flat)                                 # x fastest over the used points
        num = np.full(gx * gy * gz, -1, np.int64)
        num[order] = np.arange(len(order))
        nnod = len(order)
        coords = np.zeros((nnod, 3))
        coords[:, 0] = (order % gx) * lx / (2 * nx)
        coords[:, 1] = ((order // gx) % gy) * ly / (2 * ny)
        coords[:, 2] = (order // (gx * gy)) * lz / (2 * nz)
        nd_el = flat.shape[1]
        per = flat.shape[0] // nel
        nel = flat.shape[0]
        ln = np.zeros((nel, 20), np.int32)
        ln[:, :nd_el] = 1 + num[flat]
        base = np.repeat(base, per, axis=0)
    elif lt == LT_BRICK_20:
        # grid of half cells; nodes are the points with at most one odd index
        gx, gy, gz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
        I, J, K = np.meshgrid(np.arange(gx), np.arange(gy), np.arange(gz), indexing="ij")
        gidx = (I + gx * (J + gy * K)).ravel()
        keep = ((I % 2) + (J % 2) + (K % 2)).ravel() <= 1
        num = np.full(gx * gy * gz, -1, np.int64)
        order = np.sort(gidx[keep])                  # x fastest over the kept points
        num[order] = np.arange(len(order))
        nnod = len(order)
        coords = np.zeros((nnod, 3))
        oi, oj, ok = order % gx, (order // gx) % gy, order // (gx * gy)
        coords[:, 0] = oi * lx / (2 * nx)
        coords[:, 1] = oj * ly / (2 * ny)
        coords[:, 2] = ok * lz / (2 * nz)
        for a, (ci, cj, ck) in enumerate(BRICK20):
            gi = 2 * base[:, 0] + 1 + ci
            gj = 2 * base[:, 1] + 1 + cj
            gk = 2 * base[:, 2] + 1 + ck
            ln[:, a] = 1 + num[gi + gx * (gj + gy * gk)]
        nd_el = 20
    else:
        nnx, nny, nnz = nx + 1, ny + 1, nz + 1
        nnod = nnx * nny * nnz
        ii, jj, kk = np.meshgrid(np.arange(nnx), np.arange(nny), np.arange(nnz), indexing="ij")
        # node number = i + nnx*(j + nny*k), x fastest
        coords = np.zeros((nnod, 3))
        idx = (ii + nnx * (jj + nny * kk)).ravel()
        coords[idx, 0] = (ii.ravel() * lx / nx)
        coords[idx, 1] = (jj.ravel() * ly / ny)
        coords[idx, 2] = (kk.ravel() * lz / nz)
        for a, (di, dj, dk) in enumerate(BRICK):
            ln[:, a] = 1 + (base[:, 0] + di) + nnx * ((base[:, 1] + dj) + nny * (base[:, 2] + dk))
        nd_el = 8
    coords = f32(coords)
    npe = {LT_BRICK_1GP: 1, LT_TET10: 4, LT_WEDGE15: 6}.get(lt, 8)
    ngps = nel * npe
    lg = np.zeros((nel, 8), np.int32)
    lg[:, :npe] = 1 + np.arange(ngps, dtype=np.int32).reshape(nel, npe)

    # texture: one orientation per element, repeated on its points
    if texture == "random":
        eul = random_euler(rng, nel)
    elif texture == "basal":
        eul = basal_fibre_euler(rng, nel)
    else:
        eul = np.zeros((nel, 3))
    nss = material.nss
    nsv = 3 + nss
    sv = np.zeros((ngps, MAXSV))
    sv[:, 0:3] = np.repeat(eul, npe, axis=0)
    sv[:, 3:3 + nss] = material.crss[None, :]
    sv = f32(sv)

    le = np.ones(nel, np.int32)
    if elastic_elements is not None:
        le[np.asarray(elastic_elements)] = -1

    # boundary codes and kinematic tables
    eps_inc = np.asarray(eps_inc, float)
    ext = np.array([[0.0, lx], [0.0, ly], [0.0, lz]])
    tol = 1e-6 * max(lx, ly, lz)
    ib = np.zeros(nnod, np.int32)
    on_surface = np.zeros(nnod, bool)
    cons = {}    # node -> tuple of (axis, 'fixed'|'spec')
    for ax, name in enumerate("xyz"):
        lo = np.abs(coords[:, ax] - ext[ax, 0]) < tol
        hi = np.abs(coords[:, ax] - ext[ax, 1]) < tol
        on_surface |= lo | hi
        for mask, side in ((lo, "min"), (hi, "max")):
            kind = faces.get(name + side, "free")
            if kind == "free":
                continue
            for node in np.nonzero(mask)[0]:
                cons.setdefault(int(node), {})[ax] = kind
    combos = {}
    for node, c in sorted(cons.items()):
        key = tuple(sorted(c.items()))
        if key not in combos:
            combos[key] = len(combos) + 1
        ib[node] = 100 * combos[key]
    ib[(ib == 0) & on_surface] = 1
    assert len(combos) < 50, "absolute kinematic tables only"

    ninc1 = ninc
    nxs = max(combos.values()) if combos else 0
    tiem = np.zeros(ninc1)
    x_spec = np.zeros((ninc1, NTAB, 15))
    f_spec = np.zeros((ninc1, NTAB, 3))
    cur = np.array([lx, ly, lz], dtype=np.float32)
    pat = {0: 0, 1: 2, 2: 5}       # diagonal slots of the upper-triangular pattern
    disp = {0: 6, 1: 10, 2: 14}    # x_spec(7), (11), (15)
    e32 = eps_inc.astype(np.float32)
    for inc in range(ninc1):
        tiem[inc] = np.sqrt(np.float32(e32[0] ** 2 + e32[1] ** 2 + e32[2] ** 2))
        dcur = (e32 * cur).astype(np.float32)
        cur = (cur + dcur).astype(np.float32)
        for key, t in combos.items():
            for ax, kind in key:
                x_spec[inc, t - 1, pat[ax]] = 1.0
                x_spec[inc, t - 1, disp[ax]] = dcur[ax] if kind == "spec" else 0.0
    prob = Problem(nlmnt=nel, nnod=nnod, ngps=ngps, nsv=nsv,
                   le=le, lt=np.full(nel, lt, np.int32), nd=np.full(nel, nd_el, np.int32),
                   np_=np.full(nel, npe, np.int32), ln=ln, lg=lg, ib=ib,
                   x=coords.reshape(-1).copy(), dx=np.zeros(3 * nnod),
                   sv=sv, tiem=f32(tiem), x_spec=f32(x_spec), f_spec=f32(f_spec),
                   materials=[material], control=control, nxs=nxs, nfs=0, dims=(nx, ny, nz))
    return prob


# ---------------------------------------------------------------------------
# the benchmark configurations
# ---------------------------------------------------------------------------

def with_second_phase(prob: Problem, material: Material, fraction: float = 0.5, seed: int = 0) -> Problem:
    """The mesh with a random fraction of its elements given material code 2
    (a second material set, as FS3PRE's material fractions assign them); their
    points start from that material's slip resistances."""
    rng = np.random.default_rng(seed)
    le = prob.le.copy()
    pick = (rng.random(prob.nlmnt) < fraction) & (prob.np_ > 0)
    le[pick] = 2 * np.sign(le[pick])
    sv = prob.sv.copy()
    nss_max = max(prob.materials[0].nss, material.nss)
    for e in np.nonzero(pick)[0]:
        for k in range(prob.np_[e]):
            gpt = prob.lg[e, k] - 1
            sv[gpt, 3:3 + nss_max] = 0.0
            sv[gpt, 3:3 + material.nss] = material.crss
    return dataclasses.replace(prob, le=le, sv=f32(sv), nsv=3 + nss_max,
                               materials=[prob.materials[0], material])


def fcc_tension(n, ninc=50, seed=0, lt=LT_BRICK_8GP, **kw) -> Problem:
    """FCC aluminium, uniaxial tension along z, 1e-3 per increment, symmetry planes."""
    return brick_problem(n, lt=lt, material=fcc_aluminium(), ninc=ninc, seed=seed,
                         faces=dict(xmin="fixed", ymin="fixed", zmin="fixed", zmax="spec"),
                         eps_inc=(0.0, 0.0, 1.0e-3), **kw)


def fcc_plane_strain(n, ninc=100, seed=2, lt=LT_BRICK_8GP, **kw) -> Problem:
    """FCC aluminium, plane-strain compression: eps_zz = -1e-3 per increment, eps_yy = 0."""
    return brick_problem(n, lt=lt, material=fcc_aluminium(), ninc=ninc, seed=seed,
                         faces=dict(xmin="fixed", ymin="fixed", ymax="fixed", zmin="fixed", zmax="spec"),
                         eps_inc=(0.0, 0.0, -1.0e-3), **kw)


def hcp_compression(n, ninc=60, seed=3, lt=LT_BRICK_8GP, **kw) -> Problem:
    """Mg-like HCP with twins, basal fibre along z, compression along x."""
    return brick_problem(n, lt=lt, material=hcp_magnesium(), ninc=ninc, seed=seed, texture="basal",
                         faces=dict(xmin="fixed", xmax="spec", ymin="fixed", zmin="fixed"),
                         eps_inc=(-1.0e-3, 0.0, 0.0), **kw)

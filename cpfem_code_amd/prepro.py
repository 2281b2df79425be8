"""Pre-processing: the structured-brick meshes, boundary codes, load tables and
state of the reference's pre-processor FS3PRE (src/prepro), written as .fsd.

FS3PRE (fs3pre.f90:7-147) asks its questions on stdin; :func:`run` takes the
same answers as keyword arguments and :func:`main` reads them from stdin in
FS3PRE's order, with the same directory file (``dirname``: material,
orientation and output directories) and the same input files (``.fmt``
material files, ``.ori`` orientation files).  The steps and their sources:

* elements ``x`` fastest, then ``y``, then ``z``: 8-node bricks with one point
  (lt 12, ADDEL_8NQ fs3pre.f90:156-229) or 20-node bricks with 2x2x2 points
  (lt 22, ADDEL_20NQ fs3pre.f90:231-287); every element gets fresh nodes;
* SETSWTOL (fs3p-swp.f90:7-46): tolerance = 0.5 * sqrt(smallest squared
  distance between two nodes of one element);
* SWEEP (fs3p-swp.f90:54-109): a node within the tolerance of an earlier one
  is merged into the first such node; the remaining nodes keep their order.
  Node 1 is never a merge target (the search starts at node 2);
* SETBOUNDS (fs3p-bkc.f90:9-171): boundary code ``ib`` from the faces of the
  bounding box a node lies on, then the ``xfree`` options;
* SETSTATE with file orientations (fs3p-gps.f90:7-270, option 4): Bunge
  angles of point k from line k of the ``.ori`` file, all other state zero;
* SETMATC (fs3p-matl.f90:9-130): the ``.fmt`` file and the initial slip
  resistances of the points of each material;
* SETKINC (fs3p-bkc.f90:179-305): the 50 kinematic tables per increment;
* SETTIES (fs3p-bkc.f90:313-408): two-node tie elements between opposite
  faces;
* WTDATA (fs3p-io_nowin.f90:10-113): the .fsd record list.

Everything is computed in single precision with the reference's operation
order, so the file is byte-identical to FS3PRE's for the same answers.  What
the reference draws from the compiler's ``RANDOM_NUMBER`` (several
orientation sets or materials assigned at random, an elastic-only fraction)
is drawn here from a seeded numpy generator: those files are valid but not
byte-identical.  Orientation options 1, 2, 3 and 5 are not offered: the
reference's point loop (fs3p-gps.f90:248-260) takes every point's angles from
the file list, which those options never fill.  The 10-node tetrahedra of
element set 3 are not offered (the solver has no ``lt 21``).
"""
from __future__ import annotations

import math
import os
import sys
from decimal import Decimal
from fractions import Fraction
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import fsd
from .problem import BRICK20, MAXSS, MAXSV, Control, Material, Problem

F32 = np.float32

# local coordinates of ADDEL_8NQ and ADDEL_20NQ (fs3pre.f90:170-174, 244-248)
_BRICK8 = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
_BRICK20 = [tuple((c + 1) / 2 for c in p) for p in BRICK20]

# SETBOUNDS face code l_side (1 x min, 2 x max, 4 y min, 8 y max, 16 z min,
# 32 z max) -> ib, the 'full specification' defaults (fs3p-bkc.f90:44-103)
_IB_FULL = {1: 100, 4: 200, 16: 300, 5: 400, 17: 500, 20: 600, 21: 700,
            2: 800, 8: 900, 32: 1000, 10: 1100, 34: 1200, 40: 1300, 42: 1400,
            9: 1500, 33: 1600, 6: 1700, 36: 1800, 18: 1900, 24: 2000,
            26: 2100, 38: 2200, 41: 2300, 37: 2400, 25: 2500, 22: 2600}
# xfree = 2, free x faces (fs3p-bkc.f90:119-145)
_IB_FREE_X = {1: 1, 2: 1, 5: 200, 6: 200, 9: 900, 10: 900, 17: 300, 18: 300, 33: 1000, 34: 1000,
              22: 600, 25: 2000, 26: 2000, 37: 1800, 38: 1800, 41: 1300, 42: 1300}
# xfree = 3, only x faces (fs3p-bkc.f90:149-160)
_IB_X_ONLY = {**{k: 1 for k in (4, 8, 16, 20, 24, 32, 36, 40)},
              **{k: 100 for k in (1, 5, 9, 17, 25, 33, 37, 41)},
              **{k: 800 for k in (6, 10, 18, 22, 26, 34, 38, 42)}}


def real4(text: str) -> np.float32:
    """A REAL(4) read from text: the single-precision number nearest to the
    decimal value (ties to even), without a detour through double precision."""
    q = Fraction(Decimal(text.strip().replace("d", "e").replace("D", "e")))
    c = F32(float(q))
    best = c
    for cand in (np.nextafter(c, F32(-np.inf)), np.nextafter(c, F32(np.inf))):
        if not np.isfinite(cand):
            continue
        da, db = abs(Fraction(float(cand)) - q), abs(Fraction(float(best)) - q)
        if da < db or (da == db and (int(cand.view(np.int32)) & 1) == 0):
            best = cand
    return best


# ---------------------------------------------------------------------------
# list-directed input (the answers on stdin and the .fmt / .ori files)
# ---------------------------------------------------------------------------
class Records:
    """Fortran list-directed reading from a list of records (lines): a READ
    takes values from as many records as it needs and then moves to the next
    record; a READ without items consumes one record."""

    def __init__(self, lines: Sequence[str]):
        self.lines = list(lines)
        self.pos = 0

    def _tokens(self, line: str) -> List[str]:
        out, cur, quote = [], "", None
        for ch in line:
            if quote:
                if ch == quote:
                    quote = None
                else:
                    cur += ch
            elif ch in "'\"":
                quote = ch
            elif ch in " \t,":
                if cur:
                    out.append(cur)
                    cur = ""
            elif ch == "/":
                break
            else:
                cur += ch
        if cur:
            out.append(cur)
        return out

    def values(self, n: int) -> List[str]:
        got: List[str] = []
        while len(got) < n:
            if self.pos >= len(self.lines):
                raise EOFError("input ended before all values were read")
            got += self._tokens(self.lines[self.pos])
            self.pos += 1
        return got[:n]

    def ints(self, n: int) -> List[int]:
        return [int(v) for v in self.values(n)]

    def reals(self, n: int) -> List[np.float32]:
        return [real4(v) for v in self.values(n)]

    def word(self) -> str:
        return self.values(1)[0]

    def skip(self) -> None:
        self.pos += 1

    def line(self) -> str:
        """READ(*,'(A)'): the next record as it stands."""
        if self.pos >= len(self.lines):
            raise EOFError("input ended")
        self.pos += 1
        return self.lines[self.pos - 1]


def read_fmt(path: str) -> Dict:
    """A .fmt material file (fs3p-matl.f90:67-103): nmpar, nss, e(1:9),
    ssang(1:nss), nss lines of ab an, nss lines of ampar(1:nmpar), nss rows of
    the latent hardening matrix (row i is stored as smatrx(:, i))."""
    with open(path) as fh:
        r = Records(fh.read().splitlines())
    nmpar = r.ints(1)[0]
    nss = r.ints(1)[0]
    e = np.array(r.reals(9), F32)
    ssang = np.array(r.reals(nss), F32)
    abn = np.array([r.reals(6) for _ in range(nss)], F32)
    ampar = np.array([r.reals(nmpar) for _ in range(nss)], F32).reshape(nss, nmpar)
    smatrx = np.array([r.reals(nss) for _ in range(nss)], F32)
    return dict(nmpar=nmpar, nss=nss, e=e, ssang=ssang, ab=abn[:, :3], an=abn[:, 3:], ampar=ampar, smatrx=smatrx)


def read_ori(path: str, npts: int) -> np.ndarray:
    """The first npts lines of an .ori file: weight, phi1, Phi, phi2 (radians);
    returns the angles (fs3p-gps.f90:213-219)."""
    with open(path) as fh:
        r = Records(fh.read().splitlines())
    out = np.zeros((npts, 3), F32)
    for k in range(npts):
        v = r.reals(4)
        out[k] = v[1:]
    return out


# ---------------------------------------------------------------------------
# the steps
# ---------------------------------------------------------------------------
def _elements(nein, xl, arr_type):
    """Elements in FS3PRE's order with fresh nodes each (ADDEL_8NQ / ADDEL_20NQ)."""
    loc = _BRICK8 if arr_type == 1 else _BRICK20
    xyz = np.array(loc, F32)
    dxl = np.array([F32(xl[j]) / F32(nein[j]) for j in range(3)], F32)
    xs, conn = [], []
    for iz in range(1, nein[2] + 1):
        for iy in range(1, nein[1] + 1):
            for ix in range(1, nein[0] + 1):
                x0 = np.array([dxl[0] * F32(ix - 1), dxl[1] * F32(iy - 1), dxl[2] * F32(iz - 1)], F32)
                base = len(xs)
                for a in range(len(loc)):
                    xs.append([x0[j] + dxl[j] * xyz[a, j] for j in range(3)])
                conn.append(list(range(base, base + len(loc))))
    return np.array(xs, F32), conn


def sweep_tolerance(x: np.ndarray, conn) -> np.float32:
    """SETSWTOL: 0.5 * sqrt of the smallest squared intra-element distance."""
    lmin = None
    for el in conn:
        p = x[el]
        for j in range(len(el) - 1):
            for k in range(j + 1, len(el)):
                s = F32(0.0)
                for l in range(3):
                    d = F32(p[j, l] - p[k, l])
                    s = F32(s + d * d)
                if lmin is None or s < lmin:
                    lmin = s
    return F32(F32(0.5) * F32(np.sqrt(lmin)))


def sweep(x: np.ndarray, conn, tol: np.float32):
    """SWEEP: merge each node into the first earlier node (from node 2 on)
    within tol; returns the kept coordinates and the renumbered elements."""
    kept: List[int] = []          # indices into x of the kept nodes, in order
    new_of = np.zeros(len(x), np.int64)
    cell: Dict[tuple, List[int]] = {}
    h = float(tol)

    def key(p):
        return tuple(int(math.floor(float(v) / h)) for v in p)

    for i in range(len(x)):
        p = x[i]
        match = None
        if kept:
            kc = key(p)
            cands = []
            for dx_ in (-1, 0, 1):
                for dy_ in (-1, 0, 1):
                    for dz_ in (-1, 0, 1):
                        cands += cell.get((kc[0] + dx_, kc[1] + dy_, kc[2] + dz_), [])
            for pos in sorted(cands):
                if pos == 0:
                    continue          # the search starts at node 2
                q = x[kept[pos]]
                s = F32(0.0)
                for l in range(3):
                    d = F32(p[l] - q[l])
                    s = F32(s + d * d)
                if F32(np.sqrt(s)) < tol:
                    match = pos
                    break
        if match is None:
            new_of[i] = len(kept)
            cell.setdefault(key(p), []).append(len(kept))
            kept.append(i)
        else:
            new_of[i] = match
    return x[kept], [[int(new_of[a]) for a in el] for el in conn]


def boundary_codes(x: np.ndarray, tol: np.float32, xfree: int) -> np.ndarray:
    """SETBOUNDS."""
    lo, hi = x.min(0), x.max(0)
    ib = np.zeros(len(x), np.int32)
    for i, p in enumerate(x):
        side = 0
        for j in range(3):
            if abs(F32(p[j] - lo[j])) < tol:
                side += 2 ** (2 * j)
            if abs(F32(p[j] - hi[j])) < tol:
                side += 2 ** (2 * j + 1)
        if side == 0:
            continue
        code = _IB_FULL.get(side, 1)
        if xfree == 0:
            code = 5000
        elif xfree == 2:
            code = _IB_FREE_X.get(side, code)
        elif xfree == 3:
            code = _IB_X_ONLY.get(side, code)
        ib[i] = code
    return ib


def kinematic_tables(xl, xfree: int, eps_inc, ninc1: int):
    """SETKINC: tiem and the 50 tables x_spec(1:15, j, inc) per increment."""
    ex, ey, ez = (F32(v) for v in eps_inc)
    tiem = np.zeros(ninc1, F32)
    xs = np.zeros((ninc1, 99, 15), F32)
    cur = [F32(v) for v in xl]
    point = {1: (1, 0, 0, 0, 0, 0), 2: (0, 0, 1, 0, 0, 0), 3: (0, 0, 0, 0, 0, 1), 4: (1, 0, 1, 0, 0, 0),
             5: (1, 0, 0, 0, 0, 1), 6: (0, 0, 1, 0, 0, 1), 7: (1, 0, 1, 0, 0, 1)}
    rows = [point[1], point[2], point[3], point[4], point[5], point[6], point[7],
            point[1], point[2], point[3], point[4], point[5], point[6], point[7],
            point[4], point[5], point[4], point[6], point[5], point[6]] + [point[7]] * 6
    xmov = (8, 11, 12, 14, 17, 19, 21, 22, 26)
    ymov = (9, 11, 13, 14, 15, 20, 21, 23, 25)
    zmov = (10, 12, 13, 14, 16, 18, 22, 23, 24)
    for i in range(ninc1):
        tiem[i] = F32(np.sqrt(F32(F32(F32(ex * ex) + F32(ey * ey)) + F32(ez * ez))))
        for j, r in enumerate(rows):
            xs[i, j, :6] = r
        dxs, dys, dzs = F32(ex * cur[0]), F32(ey * cur[1]), F32(ez * cur[2])
        cur = [F32(cur[0] + dxs), F32(cur[1] + dys), F32(cur[2] + dzs)]
        for j in xmov:
            xs[i, j - 1, 6] = dxs
        for j in ymov:
            xs[i, j - 1, 10] = dys
        for j in zmov:
            xs[i, j - 1, 14] = dzs
        xs[i, 49, :6] = point[7]
        xs[i, 49, 6], xs[i, 49, 8], xs[i, 49, 10], xs[i, 49, 14] = ex, 0.0, ey, ez
    return tiem, xs


def ties(x: np.ndarray, ib: np.ndarray, tol: np.float32, tie) -> List[tuple]:
    """SETTIES: (lt, node1, node2) 1-based, in the reference's loop order."""
    lo, hi = x.min(0), x.max(0)
    out = []
    n = len(x)
    for i1 in range(n):
        for j1 in range(3):
            if not (tie[j1] and abs(F32(x[i1, j1] - lo[j1])) < tol):
                continue
            j2 = (j1 + 1) % 3
            j3 = (j2 + 1) % 3
            for i2 in range(n):
                if (abs(F32(x[i2, j1] - hi[j1])) < tol and abs(F32(x[i1, j2] - x[i2, j2])) < tol
                        and abs(F32(x[i1, j3] - x[i2, j3])) < tol):
                    c = int(ib[i1]) // 100
                    lt = (j1 + 4 if c in (1, 2, 3) else 1 if c in (6, 18, 20) else 2 if c in (5, 16, 19)
                          else 3 if c in (4, 15, 17) else None)
                    if lt is not None:
                        out.append((lt, i1 + 1, i2 + 1))
    return out


# ---------------------------------------------------------------------------
# the whole run
# ---------------------------------------------------------------------------
def run(nein: Sequence[int], xl: Sequence[float], arr_type: int, xfree: int, nss_max: int,
        orientations: np.ndarray, materials: Sequence[Dict], slip_resistances: Sequence[Sequence[float]],
        eps_inc: Sequence[float], ninc1: int, twinlim: int, twinprob: float, tie=(False, False, False),
        orientation_fractions: Optional[Sequence[float]] = None,
        material_fractions: Optional[Sequence[float]] = None,
        elastic_fraction: Optional[float] = None, seed: int = 0) -> Dict:
    """FS3PRE with file orientations.  orientations: (ngps, 3) Bunge angles of
    the points in .ori order; materials: read_fmt dicts; slip_resistances: the
    initial slip resistances of each material; eps_inc: (e11, e22, e33) (e11
    unused for xfree 2, e22 and e33 unused for xfree 3).  The *_fractions and
    elastic_fraction options use a seeded generator in place of the
    reference's RANDOM_NUMBER.  Returns the .fsd content as a dict."""
    if arr_type not in (1, 2):
        raise ValueError("element set 3 (10-node tetrahedra) is not offered")
    rng = np.random.default_rng(seed)
    x, conn = _elements(nein, xl, arr_type)
    tol = sweep_tolerance(x, conn)
    x, conn = sweep(x, conn, tol)
    nel = len(conn)
    npe = 1 if arr_type == 1 else 8
    ngps = nel * npe
    ib = boundary_codes(x, tol, xfree)
    nsv = nss_max + 3
    # SETSTATE, option 4: one orientation set per element ('grain'), then
    # sets assigned to elements by the fractions
    le = np.ones(nel, np.int32)
    nsets = len(orientation_fractions) if orientation_fractions else 1
    if nsets > 1:
        f = np.cumsum(np.array(orientation_fractions, F32) / F32(sum(orientation_fractions)))
        for e in range(nel):
            le[e] = 1 + int(np.sum(rng.random() > f))
    sv = np.zeros((ngps, MAXSV), F32)
    sv[:, :3] = np.asarray(orientations, F32)[:ngps]
    # SETMATC
    nmatl = len(materials)
    if int(le.max()) > nmatl:
        f = np.cumsum(np.array(material_fractions, F32) / F32(sum(material_fractions)))
        for e in range(nel):
            le[e] = 1 + int(np.sum(rng.random() > f))
    nmpar = materials[0]["nmpar"]
    for k, m in enumerate(materials):
        if m["nmpar"] != nmpar:
            raise ValueError("inconsistent plastic property list")
        ss = np.array([real4(str(v)) if isinstance(v, str) else F32(v) for v in slip_resistances[k]], F32)
        for e in range(nel):
            if le[e] == k + 1:
                for j in range(npe):
                    sv[e * npe + j, 3:3 + m["nss"]] = ss[:m["nss"]]
    if elastic_fraction is not None:
        for e in range(nel):
            if rng.random() <= elastic_fraction:
                le[e] = -le[e]
    tiem, x_spec = kinematic_tables(xl, xfree, eps_inc, ninc1)
    tl = ties(x, ib, tol, tie)
    lt = [12 if arr_type == 1 else 22] * nel + [t[0] for t in tl]
    nd = [8 if arr_type == 1 else 20] * nel + [2] * len(tl)
    np_ = [npe] * nel + [0] * len(tl)
    ln = [[a + 1 for a in el] for el in conn] + [[t[1], t[2]] for t in tl]
    lg = [[e * npe + j + 1 for j in range(npe)] for e in range(nel)] + [[] for _ in tl]
    le = list(le) + [0] * len(tl)
    return dict(ninc1=ninc1, nlmnt=len(ln), nnod=len(x), ngps=ngps, nsv=nsv, le=le, lt=lt, nd=nd, np=np_,
                ln=ln, lg=lg, ib=ib, x=x, sv=sv, nmpar=nmpar, materials=list(materials), twinlim=twinlim,
                twinprob=F32(twinprob), tiem=tiem, x_spec=x_spec, nxs=50, nfs=0)


def write(d: Dict, path: str) -> None:
    """WTDATA: the .fsd record list (fs3p-io_nowin.f90:23-107)."""
    i4, f4 = fsd._i, fsd._f
    with open(path, "wb") as fh:
        w = fsd._W(fh)
        w.rec(i4(d["ninc1"]))
        w.rec(i4(d["nlmnt"], d["nnod"], d["ngps"], d["nsv"]))
        for e in range(d["nlmnt"]):
            w.rec(i4(d["le"][e], d["lt"][e], d["nd"][e], d["np"][e]))
            w.rec(i4(*d["ln"][e]))
            w.rec(i4(*d["lg"][e]))
        z3 = np.zeros(3, F32)
        for i in range(d["nnod"]):
            w.rec(i4(d["ib"][i]))
            w.rec(f4(np.concatenate([d["x"][i], z3])))
            w.rec(f4(z3))
        nsv = d["nsv"]
        for g in range(d["ngps"]):
            w.rec(f4([0.0, 0.0]))
            w.rec(f4(d["sv"][g, :nsv]))
            w.rec(f4(np.zeros(nsv, F32)))
            for n in (6, 6, 6, 6, 3, 3):
                w.rec(f4(np.zeros(n, F32)))
        w.rec(i4(len(d["materials"]), d["nmpar"]))
        for m in d["materials"]:
            n = m["nss"]
            w.rec(i4(n))
            w.rec(i4(d["twinlim"]))
            w.rec(f4([d["twinprob"]]))
            w.rec(f4(m["e"]))
            w.rec(f4(m["ssang"][:n]))
            for i in range(n):
                w.rec(f4(np.concatenate([m["ab"][i], m["an"][i]])))
            for i in range(n):
                w.rec(f4(m["ampar"][i, :d["nmpar"]]))
            for i in range(n):
                w.rec(f4(m["smatrx"][i, :n]))
        w.rec(f4(z3))
        w.rec(i4(d["ninc1"], d["nxs"], d["nfs"]))
        for inc in range(d["ninc1"]):
            w.rec(f4([d["tiem"][inc]]))
            for j in range(d["nxs"]):
                w.rec(f4(d["x_spec"][inc, j]))


def to_problem(d: Dict, control: Optional[Control] = None) -> Problem:
    """The run as a Problem for the solver (double precision arrays)."""
    nel = d["nlmnt"]
    ln = np.zeros((nel, 20), np.int32)
    lg = np.zeros((nel, 8), np.int32)
    for e in range(nel):
        ln[e, :len(d["ln"][e])] = d["ln"][e]
        lg[e, :len(d["lg"][e])] = d["lg"][e]
    mats = []
    for m in d["materials"]:
        n = m["nss"]
        ampar = np.zeros((n, 10))
        ampar[:, :d["nmpar"]] = m["ampar"]
        ssang = np.zeros(MAXSS)
        ssang[:n] = m["ssang"]
        mats.append(Material(nss=n, e=np.asarray(m["e"], float), ssang=ssang, ab=np.asarray(m["ab"], float),
                             an=np.asarray(m["an"], float), ampar=ampar, smatrx=np.asarray(m["smatrx"], float),
                             crss=np.zeros(n), twinlim=int(d["twinlim"]), twinprob=float(d["twinprob"])))
    ninc1 = d["ninc1"]
    f_spec = np.zeros((ninc1, 99, 3))
    return Problem(nlmnt=nel, nnod=d["nnod"], ngps=d["ngps"], nsv=d["nsv"], le=np.array(d["le"], np.int32),
                   lt=np.array(d["lt"], np.int32), nd=np.array(d["nd"], np.int32),
                   np_=np.array(d["np"], np.int32), ln=ln, lg=lg, ib=np.asarray(d["ib"], np.int32),
                   x=np.asarray(d["x"], float).reshape(-1), dx=np.zeros(3 * d["nnod"]),
                   sv=np.asarray(d["sv"], float), tiem=np.asarray(d["tiem"], float),
                   x_spec=np.asarray(d["x_spec"], float), f_spec=f_spec, materials=mats,
                   control=control or Control(), nxs=d["nxs"], nfs=d["nfs"])


# ---------------------------------------------------------------------------
# the interactive form: FS3PRE's questions answered on stdin
# ---------------------------------------------------------------------------
def _dirs(workdir: str):
    path = os.path.join(workdir, "dirname")
    if not os.path.exists(path):
        return ".", ".", "."
    with open(path) as fh:
        v = Records(fh.read().splitlines()).values(3)
    return tuple(os.path.join(workdir, s) if not os.path.isabs(s) else s for s in v)


def answers(lines: Sequence[str], workdir: str = ".") -> Dict:
    """FS3PRE's answers in its order (fs3pre.f90:22-146 and the routines it
    calls), with the .ori and .fmt files they name read from the directories
    of ``dirname``: the keyword arguments of :func:`run`, plus ``name`` and
    ``outdir`` of the .fsd file."""
    r = Records(lines)
    nein = [r.ints(1)[0] for _ in range(3)]        # one READ statement per answer
    xl = [r.reals(1)[0] for _ in range(3)]
    arr_type = r.ints(1)[0]
    xfree = r.ints(1)[0]
    nss_max = r.ints(1)[0]
    iselect = r.ints(1)[0]
    if iselect != 4:
        raise ValueError(f"orientation option {iselect} is not offered (only file orientations, 4)")
    nsets = r.ints(1)[0]
    ofrac = [float(r.reals(1)[0]) for _ in range(nsets)]
    dmat, dori, dout = _dirs(workdir)
    ngps = nein[0] * nein[1] * nein[2] * (1 if arr_type == 1 else 8)
    ori = None
    for _ in range(nsets):
        r.skip()
        ori = read_ori(os.path.join(dori, r.word() + ".ori"), ngps)
    nmatl = r.ints(1)[0]
    mfrac = None
    if nsets > nmatl:
        mfrac = [float(r.reals(1)[0]) for _ in range(nmatl)]
    mats, ss = [], []
    for _ in range(nmatl):
        r.skip()
        m = read_fmt(os.path.join(dmat, r.word() + ".fmt"))
        mats.append(m)
        if r.word().lower() == "y":
            ss.append([r.reals(1)[0]] * m["nss"])
        else:
            ss.append(r.reals(m["nss"]))
    efrac = float(r.reals(1)[0]) if r.word().lower() == "y" else None
    if xfree in (0, 1):
        eps = r.reals(3)
    elif xfree == 2:
        eps = [F32(0.0)] + r.reals(2)
    else:
        eps = r.reals(1) + [F32(0.0), F32(0.0)]
    ninc1 = r.ints(1)[0]
    twinlim = r.ints(1)[0]
    twinprob = r.reals(1)[0]
    tie = tuple(r.line().strip()[:1] in ("y", "Y") for _ in range(3))
    name = r.word()
    return dict(nein=nein, xl=xl, arr_type=arr_type, xfree=xfree, nss_max=nss_max, orientations=ori,
                materials=mats, slip_resistances=ss, eps_inc=eps, ninc1=ninc1, twinlim=twinlim,
                twinprob=twinprob, tie=tie, orientation_fractions=ofrac if nsets > 1 else None,
                material_fractions=mfrac, elastic_fraction=efrac, name=name, outdir=dout)


def main(lines: Sequence[str], workdir: str = ".", seed: int = 0) -> str:
    """FS3PRE from its answers; returns the path of the .fsd written."""
    a = answers(lines, workdir)
    name, outdir = a.pop("name"), a.pop("outdir")
    out = os.path.join(outdir, name + ".fsd")
    write(run(**a, seed=seed), out)
    return out


if __name__ == "__main__":
    print(main(sys.stdin.read().splitlines()))

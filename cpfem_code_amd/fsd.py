"""FASOLT3 .fsd / .frs files: Fortran unformatted sequential, REAL(4) records.

The record list is the reference's (src/fasolt/Faslt_io_nowin.f90:29-128 read,
185-299 write; src/prepro/fs3p-io_nowin.f90:25-113 write): each WRITE is one
record framed by 4-byte little-endian length markers.  Also writes the small
text inputs the drivers read: fafcdf.dat (Fasolt3.f90:216-218), dirname
(Fasolt3.f90:204-206) and the optional cpfem.dat of the GPU driver.
"""
from __future__ import annotations

import os
import struct
from typing import Dict, List

import numpy as np

from .problem import MAXSS, MAXSV, Control, Material, Problem


class _W:
    def __init__(self, fh):
        self.fh = fh

    def rec(self, *parts):
        payload = b"".join(parts)
        n = struct.pack("<i", len(payload))
        self.fh.write(n + payload + n)


def _i(*v):
    return np.asarray(v, np.int32).tobytes()


def _f(a):
    return np.asarray(a, np.float32).ravel().tobytes()


def write_fsd(prob: Problem, path: str, state: Dict[str, np.ndarray] = None, ninc: int = None,
              first: int = None) -> None:
    """Write a problem (and optionally a state) as a .fsd/.frs file."""
    p = prob
    s = state if state is not None else p.zero_state()
    x = s.get("x", p.x)
    dx = s.get("dx", p.dx)
    fc = s.get("fc", np.zeros(p.ndof))
    with open(path, "wb") as fh:
        w = _W(fh)
        w.rec(_i(first if first is not None else (ninc if ninc is not None else p.ninc1)))
        w.rec(_i(p.nlmnt, p.nnod, p.ngps, p.nsv))
        for e in range(p.nlmnt):
            nd, npe = int(p.nd[e]), int(p.np_[e])
            w.rec(_i(p.le[e], p.lt[e], nd, npe))
            w.rec(_i(*p.ln[e, :nd]))
            w.rec(_i(*p.lg[e, :npe]))
        for i in range(p.nnod):
            w.rec(_i(p.ib[i]))
            w.rec(_f(np.concatenate([x[3 * i:3 * i + 3], dx[3 * i:3 * i + 3]])))
            w.rec(_f(fc[3 * i:3 * i + 3]))
        nsv = p.nsv
        for g in range(p.ngps):
            w.rec(_f([s["eps"][g], s["dep"][g]]))
            w.rec(_f(s["sv"][g, :nsv]))
            w.rec(_f(s["dsv"][g, :nsv]))
            for k in ("ep", "de", "st", "ds", "rt", "dr"):
                w.rec(_f(s[k][g]))
        nmpar = 6
        w.rec(_i(len(p.materials), nmpar))
        for m in p.materials:
            n = m.nss
            w.rec(_i(n))
            w.rec(_i(m.twinlim))
            w.rec(_f([m.twinprob]))
            w.rec(_f(m.e))
            w.rec(_f(m.ssang[:n]))
            for i in range(n):
                w.rec(_f(np.concatenate([m.ab[i], m.an[i]])))
            for i in range(n):
                w.rec(_f(m.ampar[i, :nmpar]))
            for i in range(n):
                w.rec(_f(m.smatrx[i, :n]))
        w.rec(_f([0.0, 0.0, 0.0]))
        w.rec(_i(p.ninc1, p.nxs, p.nfs))
        for inc in range(p.ninc1):
            w.rec(_f([p.tiem[inc]]))
            for j in range(p.nxs):
                w.rec(_f(p.x_spec[inc, j]))
            for j in range(p.nfs):
                w.rec(_f(p.f_spec[inc, j]))


class _R:
    def __init__(self, raw: bytes):
        self.raw = raw
        self.pos = 0

    def rec(self) -> bytes:
        (n,) = struct.unpack_from("<i", self.raw, self.pos)
        data = self.raw[self.pos + 4:self.pos + 4 + n]
        (n2,) = struct.unpack_from("<i", self.raw, self.pos + 4 + n)
        if n2 != n:
            raise ValueError("corrupt record markers")
        self.pos += 8 + n
        return data

    def ints(self):
        return np.frombuffer(self.rec(), np.int32)

    def reals(self):
        return np.frombuffer(self.rec(), np.float32).astype(np.float64)


def read_frs(path: str) -> Dict[str, np.ndarray]:
    """Read the mesh and state records of a .fsd/.frs file."""
    with open(path, "rb") as fh:
        r = _R(fh.read())
    out = {"inc": int(r.ints()[0])}
    nlmnt, nnod, ngps, nsv = (int(v) for v in r.ints())
    out.update(nlmnt=nlmnt, nnod=nnod, ngps=ngps, nsv=nsv)
    le, lt = np.zeros(nlmnt, np.int32), np.zeros(nlmnt, np.int32)
    nde, npe = np.zeros(nlmnt, np.int32), np.zeros(nlmnt, np.int32)
    ln, lg = np.zeros((nlmnt, 20), np.int32), np.zeros((nlmnt, 8), np.int32)
    for e in range(nlmnt):
        a = r.ints()
        le[e], lt[e], nde[e], npe[e] = a[0], a[1], a[2], a[3]
        ln[e, :a[2]] = r.ints()
        lg[e, :a[3]] = r.ints()
    ib = np.zeros(nnod, np.int32)
    x, dx, fc = np.zeros(3 * nnod), np.zeros(3 * nnod), np.zeros(3 * nnod)
    for i in range(nnod):
        ib[i] = r.ints()[0]
        v = r.reals()
        x[3 * i:3 * i + 3], dx[3 * i:3 * i + 3] = v[:3], v[3:]
        fc[3 * i:3 * i + 3] = r.reals()
    st = {k: np.zeros((ngps, w)) for k, w in (("ep", 6), ("de", 6), ("st", 6), ("ds", 6), ("rt", 3), ("dr", 3))}
    eps, dep = np.zeros(ngps), np.zeros(ngps)
    sv, dsv = np.zeros((ngps, MAXSV)), np.zeros((ngps, MAXSV))
    for g in range(ngps):
        eps[g], dep[g] = r.reals()
        sv[g, :nsv] = r.reals()
        dsv[g, :nsv] = r.reals()
        for k in ("ep", "de", "st", "ds", "rt", "dr"):
            st[k][g] = r.reals()
    out.update(le=le, lt=lt, ln=ln, lg=lg, ib=ib, x=x, dx=dx, fc=fc, eps=eps, dep=dep, sv=sv, dsv=dsv, **st)
    # element node and point counts, then the material records (Faslt_io_nowin.f90:77-107)
    out["nd"], out["np"] = nde, npe
    nmatl, nmpar = (int(v) for v in r.ints())
    mats = []
    for _ in range(nmatl):
        n = int(r.ints()[0])
        twinlim = int(r.ints()[0])
        twinprob = float(r.reals()[0])
        e = r.reals()
        ssang = r.reals()
        abn = np.array([r.reals() for _ in range(n)])
        ampar = np.array([r.reals() for _ in range(n)])
        smatrx = np.array([r.reals() for _ in range(n)])
        mats.append(dict(nss=n, twinlim=twinlim, twinprob=twinprob, e=e, ssang=ssang, ab=abn[:, :3], an=abn[:, 3:],
                         ampar=ampar, smatrx=smatrx))
    out["materials"] = mats
    return out


def write_control_files(dirpath: str, control: Control, solver_file: bool = True) -> None:
    """fafcdf.dat, dirname (all paths ./) and, optionally, cpfem.dat."""
    c = control
    with open(os.path.join(dirpath, "fafcdf.dat"), "w") as fh:
        fh.write(f"{c.toler:.6e} {c.maxit:d} {c.alpha:.6e} {c.afactor:.6e}\n")
        fh.write(f"{c.efix:.6e} {c.smalp:.6e} {c.smalf:.6e}\n")
        fh.write(f"{c.smald:.6e} {c.smale:.6e} {c.smalr:.6e}\n")
    with open(os.path.join(dirpath, "dirname"), "w") as fh:
        fh.write("'./' './' './' './' './'\n")
    if solver_file:
        with open(os.path.join(dirpath, "cpfem.dat"), "w") as fh:
            fh.write(f"{c.sor_omega:.6e} {c.sor_kind:d} {c.sor_tol:.6e} {c.sor_maxit:d} 0\n")

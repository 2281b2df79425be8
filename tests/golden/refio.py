"""Stream I/O for oracle/ref_harness.f90 (the compiled reference solver).

Writes the raw little-endian stream files the harness reads and parses what it
writes.  Used by the fixture scripts in this directory and by the tests that
run the reference where it is built (this container only).
"""
from __future__ import annotations

import io
import os
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
MAXSS, MAXSV = 60, 63


def available() -> bool:
    return os.path.exists(HARNESS)


def _i(buf, *v):
    buf.write(np.asarray(v, np.int32).tobytes())


def _f(buf, a):
    buf.write(np.asarray(a, np.float32).ravel(order="C").tobytes())


def write_material(buf, materials):
    _i(buf, len(materials), 10)
    for m in materials:
        n = m.nss
        _i(buf, n, m.twinlim)
        _f(buf, [m.twinprob])
        _f(buf, m.e)
        _f(buf, m.ssang[:n])
        _f(buf, m.ab[:n])
        _f(buf, m.an[:n])
        _f(buf, m.ampar[:n, :10])
        _f(buf, m.smatrx[:n, :n])


def write_control(buf, c):
    _f(buf, [c.toler]); _i(buf, c.maxit); _f(buf, [c.alpha, c.afactor])
    _f(buf, [c.efix, c.smalp, c.smalf])
    _f(buf, [c.smald, c.smale, c.smalr])


def write_mesh(buf, p, state):
    _i(buf, p.nlmnt, p.nnod, p.ngps, p.nsv)
    for e in range(p.nlmnt):
        _i(buf, p.le[e], p.lt[e], p.nd[e], p.np_[e])
        _i(buf, *p.ln[e])
        _i(buf, *p.lg[e])
    _i(buf, *p.ib)
    _f(buf, state["x"]); _f(buf, state["dx"])
    _f(buf, state["eps"]); _f(buf, state["dep"])
    _f(buf, state["sv"]); _f(buf, state["dsv"])
    for k in ("ep", "de", "st", "ds"):
        _f(buf, state[k])
    _f(buf, state["rt"]); _f(buf, state["dr"])
    _i(buf, *state["twinsys"]); _i(buf, *state["timestwinned"])
    _i(buf, p.ninc1, p.nxs, p.nfs)
    for inc in range(p.ninc1):
        _f(buf, [p.tiem[inc]])
        _f(buf, p.x_spec[inc])
        _f(buf, p.f_spec[inc])


class _Reader:
    def __init__(self, raw: bytes):
        self.raw = raw
        self.pos = 0

    def f(self, *shape):
        n = int(np.prod(shape))
        a = np.frombuffer(self.raw, np.float32, n, self.pos).astype(np.float64)
        self.pos += 4 * n
        return a.reshape(shape)

    def i(self, *shape):
        n = int(np.prod(shape))
        a = np.frombuffer(self.raw, np.int32, n, self.pos).copy()
        self.pos += 4 * n
        return a.reshape(shape)


def _run(mode, payload: bytes) -> bytes:
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(payload)
        subprocess.run([HARNESS, mode, fin, fout], check=True, cwd=td, capture_output=True)
        with open(fout, "rb") as fh:
            return fh.read()


def dstrss(materials, control, dt, iel, de, r, st, ds, sv, dsv, twinsys=None, timestwinned=None):
    n = len(iel)
    buf = io.BytesIO()
    write_material(buf, materials)
    write_control(buf, control)
    _i(buf, n); _f(buf, [dt])
    tw = np.zeros(n, np.int32) if twinsys is None else twinsys
    tt = np.zeros(n, np.int32) if timestwinned is None else timestwinned
    for g in range(n):
        _i(buf, iel[g])
        _f(buf, de[g]); _f(buf, r[g]); _f(buf, st[g]); _f(buf, ds[g])
        _f(buf, sv[g, :MAXSV]); _f(buf, dsv[g, :MAXSV])
        _i(buf, tw[g], tt[g])
    rd = _Reader(_run("dstrss", buf.getvalue()))
    out = dict(t=np.zeros((n, 6)), ds=np.zeros((n, 6)), chi=np.zeros((n, 6, 6)), deff=np.zeros(n),
               dsv=np.zeros((n, MAXSV)), srate=np.zeros((n, MAXSS)), twinsys=np.zeros(n, np.int32),
               err=np.zeros(n, np.int32))
    for g in range(n):
        out["t"][g] = rd.f(6)
        out["ds"][g] = rd.f(6)
        out["chi"][g] = rd.f(6, 6).T          # Fortran column-major -> [i, j]
        out["deff"][g] = rd.f(1)[0]
        out["dsv"][g] = rd.f(MAXSV)
        out["srate"][g] = rd.f(MAXSS)
        out["twinsys"][g], out["err"][g] = rd.i(2)
    out["sumcpiter"] = rd.f(1)[0]
    out["tflag"] = int(rd.i(1)[0])
    return out


def _read_state(rd, p):
    nd, g = p.ndof, p.ngps
    s = {}
    s["x"], s["dx"], s["fc"] = rd.f(nd), rd.f(nd), rd.f(nd)
    s["eps"], s["dep"] = rd.f(g), rd.f(g)
    s["sv"], s["dsv"] = rd.f(g, MAXSV), rd.f(g, MAXSV)
    for k in ("ep", "de", "st", "ds"):
        s[k] = rd.f(g, 6)
    s["rt"], s["dr"] = rd.f(g, 3), rd.f(g, 3)
    s["srate"], s["cumslip"] = rd.f(g, MAXSS), rd.f(g, MAXSS)
    s["twinsys"], s["timestwinned"] = rd.i(g), rd.i(g)
    s["tflag"] = int(rd.i(1)[0])
    return s


def nrstep(p, state, inc=1):
    buf = io.BytesIO()
    write_material(buf, p.materials)
    write_control(buf, p.control)
    write_mesh(buf, p, state)
    _i(buf, inc)
    rd = _Reader(_run("nrstep", buf.getvalue()))
    nd = p.ndof
    out = {}
    out["err"] = int(rd.i(1)[0])
    out["sumcpiter"] = rd.f(1)[0]
    out["fc"], out["rhs"] = rd.f(nd), rd.f(nd)
    out["is"] = rd.i(nd, 81)
    out["ss"] = rd.f(nd, 81)
    out["state"] = _read_state(rd, p)
    out["err_sor"] = int(rd.i(1)[0])
    out["ddx"] = rd.f(nd)
    return out


def increments(p, state, nsteps):
    buf = io.BytesIO()
    write_material(buf, p.materials)
    write_control(buf, p.control)
    write_mesh(buf, p, state)
    _i(buf, nsteps)
    rd = _Reader(_run("incs", buf.getvalue()))
    res = []
    for _ in range(nsteps):
        err = int(rd.i(1)[0])
        s = _read_state(rd, p)
        s["err"] = err
        res.append(s)
    return res


FASOLT3 = os.path.join(REPO, "oracle", "_ref", "fasolt3")
CSVMKR = os.path.join(REPO, "oracle", "_ref", "csvmkr")
FS3PRE = os.path.join(REPO, "oracle", "_ref", "fs3pre")


def fasolt3(workdir: str, name: str, save: int) -> None:
    """Run the reference driver on workdir/name.fsd (with fafcdf.dat and dirname
    there) with one intermediate save (Fasolt3.f90:28-32): it writes
    name-<save>.frs and, at the end, name-<last increment>.frs."""
    subprocess.run([FASOLT3], input=f"{name}\n1\n{save}\n\n", text=True, check=True, cwd=workdir,
                   capture_output=True, timeout=600)


def csvmkr(workdir: str, frs: str, element: bool, out: str) -> str:
    """Run CSVMKR on workdir/frs.frs; element means (mode 1) or point values (mode 2).
    Returns the text of the .csv it writes."""
    subprocess.run([CSVMKR], input=f"{frs}\n{1 if element else 2}\n{out}\n", text=True, check=True,
                   cwd=workdir, capture_output=True, timeout=120)
    with open(os.path.join(workdir, out + ".csv")) as fh:
        return fh.read()

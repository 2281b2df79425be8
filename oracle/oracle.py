"""ctypes front of the CPU oracle (oracle/cpfem_oracle.c) — test infrastructure only.

``Oracle(32)`` loads the float build (REAL(4) operation order of the
reference), ``Oracle(64)`` the double build that the HIP kernels are held to.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXSS, MAXSV = 60, 63


class OrcMaterial(C.Structure):
    _fields_ = [("nss", C.c_int32), ("twinlim", C.c_int32), ("twinprob", C.c_double),
                ("e", C.c_double * 9), ("ssang", C.c_double * MAXSS),
                ("ab", (C.c_double * 3) * MAXSS), ("an", (C.c_double * 3) * MAXSS),
                ("ampar", (C.c_double * 10) * MAXSS), ("smatrx", (C.c_double * MAXSS) * MAXSS)]


class OrcControl(C.Structure):
    _fields_ = [("toler", C.c_double), ("maxit", C.c_int32), ("sor_maxit", C.c_int32),
                ("alpha", C.c_double), ("afactor", C.c_double),
                ("efix", C.c_double), ("smalp", C.c_double), ("smalf", C.c_double),
                ("smald", C.c_double), ("smale", C.c_double), ("smalr", C.c_double),
                ("sor_omega", C.c_double), ("sor_tol", C.c_double),
                ("sor_kind", C.c_int32), ("keep_going", C.c_int32), ("stall_tol", C.c_double)]


class OrcCounters(C.Structure):
    _fields_ = [("flops", C.c_int64), ("transc", C.c_int64), ("outer", C.c_int32), ("inner", C.c_int32),
                ("max_nsa", C.c_int32), ("draws", C.c_int32),
                ("capped", C.c_int32), ("pad_", C.c_int32)]


_P = C.c_void_p


class OrcModel(C.Structure):
    _fields_ = [("nlmnt", C.c_int32), ("nnod", C.c_int32), ("ngps", C.c_int32), ("nsv", C.c_int32),
                ("nmatl", C.c_int32), ("inc", C.c_int32), ("ninc1", C.c_int32), ("nxs", C.c_int32),
                ("nfs", C.c_int32), ("maxofd", C.c_int32), ("tflag", C.c_int32), ("err", C.c_int32),
                ("mat_code", C.c_int32),
                ("le", _P), ("lt", _P), ("nd", _P), ("np", _P), ("ln", _P), ("lg", _P), ("ib", _P),
                ("x", _P), ("dx", _P), ("fc", _P), ("ddx", _P),
                ("eps", _P), ("dep", _P), ("st", _P), ("ep", _P), ("ds", _P), ("de", _P),
                ("rt", _P), ("dr", _P), ("sv", _P), ("dsv", _P), ("srate", _P), ("cumslip", _P),
                ("twinsys", _P), ("timestwinned", _P),
                ("tiem", _P), ("x_spec", _P), ("f_spec", _P), ("mat", _P), ("ctl", _P),
                ("is_", _P), ("ss", _P),
                ("sumcpiter", C.c_double), ("flops", C.c_int64), ("transc", C.c_int64),
                ("sor_sweeps", C.c_int32), ("sor_error", C.c_double), ("nit", C.c_int32),
                ("sdx", C.c_double), ("sfc", C.c_double), ("alp2", C.c_double), ("rng_calls", C.c_uint64),
                ("n987", C.c_int32), ("n987_capped", C.c_int32), ("nmaxit", C.c_int32), ("pad_", C.c_int32)]


def material_struct(m) -> OrcMaterial:
    s = OrcMaterial()
    s.nss = int(m.nss)
    s.twinlim = int(m.twinlim)
    s.twinprob = float(m.twinprob)
    for i in range(9):
        s.e[i] = float(m.e[i])
    for i in range(m.nss):
        s.ssang[i] = float(m.ssang[i])
        for k in range(3):
            s.ab[i][k] = float(m.ab[i, k])
            s.an[i][k] = float(m.an[i, k])
        for k in range(10):
            s.ampar[i][k] = float(m.ampar[i, k])
        for k in range(m.nss):
            s.smatrx[i][k] = float(m.smatrx[i, k])
    return s


def control_struct(c) -> OrcControl:
    s = OrcControl()
    for name in ("toler", "alpha", "afactor", "efix", "smalp", "smalf", "smald", "smale", "smalr",
                 "sor_omega", "sor_tol"):
        setattr(s, name, float(getattr(c, name)))
    s.maxit = int(c.maxit)
    s.sor_maxit = int(c.sor_maxit)
    s.sor_kind = int(c.sor_kind)
    s.keep_going = int(getattr(c, "keep_going", 0))
    s.stall_tol = float(getattr(c, "stall_tol", 0.0))
    return s


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


class Oracle:
    def __init__(self, bits: int = 64):
        path = os.path.join(HERE, "lib", f"liboracle{bits}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.bits = bits
        assert self.lib.orc_real_size() == bits // 8
        L = self.lib
        L.orc_spower.restype = C.c_double
        L.orc_spower.argtypes = [C.c_double, C.c_double]
        L.orc_set_batch_epoch.argtypes = [C.c_uint64]
        L.orc_twin_draw.restype = C.c_double
        L.orc_twin_draw.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.orc_fj2vom.restype = C.c_double
        L.orc_fj2vom.argtypes = [C.c_int, _P]
        L.orc_factor.argtypes = [C.c_int, _P]
        L.orc_solver.argtypes = [C.c_int, _P, _P]
        L.orc_ctmatrx.argtypes = [C.c_double] * 3 + [_P, _P]
        L.orc_elastc.argtypes = [C.POINTER(OrcMaterial)] + [C.c_double] * 3 + [_P]
        L.orc_tens_rot.argtypes = [_P, _P, _P]
        L.orc_invert_3.argtypes = [_P, _P]
        L.orc_shape.argtypes = [C.c_int, C.c_int, _P, _P, _P, _P]
        L.orc_dstrns.argtypes = [C.c_int, _P, _P, C.c_double, C.c_double, _P, _P, _P]
        L.orc_dstrss_batch.argtypes = [C.POINTER(OrcMaterial), C.c_int, C.POINTER(OrcControl), C.c_double,
                                       C.c_int, _P, _P, _P, _P, _P, _P, _P, C.c_int, _P, _P, _P, _P, _P, _P,
                                       _P, _P, _P, _P]
        for f in ("orc_setpta", "orc_elvals", "orc_suvals", "orc_nriter", "orc_doreorient"):
            getattr(L, f).argtypes = [C.POINTER(OrcModel)]
        L.orc_facsls.argtypes = [C.POINTER(OrcModel), C.c_int]
        L.orc_update_node.argtypes = [C.POINTER(OrcModel)]
        L.orc_update_gptv.argtypes = [C.POINTER(OrcModel)]

    # -- small routines --------------------------------------------------
    def spower(self, a, b):
        return self.lib.orc_spower(a, b)

    def fj2vom(self, lstress, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.orc_fj2vom(int(lstress), _ptr(x))

    def factor(self, w):
        w = np.asfortranarray(w, np.float64).copy(order="F")
        f = self.lib.orc_factor(w.shape[0], _ptr(w))
        return f, w

    def solver(self, w, x):
        w = np.asfortranarray(w, np.float64)
        x = np.array(x, np.float64)
        self.lib.orc_solver(w.shape[0], _ptr(w), _ptr(x))
        return x

    def ctmatrx(self, ph1, phi, ph2):
        a = np.zeros((3, 3), order="F")
        aug = np.zeros(3)
        self.lib.orc_ctmatrx(ph1, phi, ph2, _ptr(a), _ptr(aug))
        return a, aug

    def elastc(self, mat, ph1, phi, ph2):
        el = np.zeros((6, 6), order="F")
        ms = material_struct(mat)
        self.lib.orc_elastc(C.byref(ms), ph1, phi, ph2, _ptr(el))
        return el

    def tens_rot(self, r, a):
        r = np.ascontiguousarray(r, np.float64)
        a = np.ascontiguousarray(a, np.float64)
        b = np.zeros(6)
        self.lib.orc_tens_rot(_ptr(r), _ptr(a), _ptr(b))
        return b

    def invert_3(self, m):
        m = np.asfortranarray(m, np.float64).copy(order="F")
        det = np.zeros(1)
        self.lib.orc_invert_3(_ptr(m), _ptr(det))
        return m, det[0]

    def shape(self, lt, ig, xnode):
        """Shape functions of element type lt (12, 13: 8 nodes; 21: 10; 22: 20; 23: 15) at point ig (0-based)."""
        nn = {22: 20, 21: 10, 23: 15}.get(lt, 8)
        xnode = np.ascontiguousarray(xnode, np.float64).reshape(nn, 3)
        sf = np.zeros(nn)
        sfd = np.zeros((nn, 3))
        w = np.zeros(1)
        self.lib.orc_shape(lt, ig, _ptr(xnode), _ptr(sf), _ptr(sfd), _ptr(w))
        return sf, sfd, w[0]

    def dstrns(self, sfd, dxl, smale=0.5, smalr=0.5):
        sfd = np.ascontiguousarray(sfd, np.float64)
        dxl = np.ascontiguousarray(dxl, np.float64)
        n = sfd.shape[0]
        de = np.zeros(6)
        r = np.zeros(3)
        deff = np.zeros(1)
        err = self.lib.orc_dstrns(n, _ptr(sfd), _ptr(dxl), smale, smalr, _ptr(de), _ptr(r), _ptr(deff))
        return de, r, deff[0], err

    # -- per-point constitutive batch ---------------------------------------
    def dstrss_batch(self, materials, control, dt, iel, de, r, st, ds, sv, dsv, twinsys=None,
                     timestwinned=None, epoch=None) -> Dict[str, np.ndarray]:
        """epoch: stream epoch of this call's twin draws (None: the next of the oracle's own counter)."""
        n = len(iel)
        if epoch is not None:
            self.lib.orc_set_batch_epoch(int(epoch))
        mats = (OrcMaterial * len(materials))(*[material_struct(m) for m in materials])
        ctl = control_struct(control)
        iel = np.ascontiguousarray(iel, np.int32)
        de = np.ascontiguousarray(de, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        st = np.ascontiguousarray(st, np.float64)
        ds = np.array(ds, np.float64, order="C")
        sv = np.array(sv, np.float64, order="C")
        dsv = np.array(dsv, np.float64, order="C")
        ldsv = sv.shape[1]
        t = np.zeros((n, 6))
        chi = np.zeros((n, 36))
        deff = np.zeros(n)
        srate = np.zeros((n, MAXSS))
        tw = np.zeros(n, np.int32) if twinsys is None else np.array(twinsys, np.int32)
        tt = np.zeros(n, np.int32) if timestwinned is None else np.ascontiguousarray(timestwinned, np.int32)
        tflag = np.zeros(1, np.int32)
        sumcp = np.zeros(1)
        cnt = (OrcCounters * n)()
        err = np.zeros(n, np.int32)
        self.lib.orc_dstrss_batch(mats, len(materials), C.byref(ctl), float(dt), n, _ptr(iel), _ptr(de), _ptr(r),
                                  _ptr(st), _ptr(ds), _ptr(sv), _ptr(dsv), ldsv, _ptr(t), _ptr(chi), _ptr(deff),
                                  _ptr(srate), _ptr(tw), _ptr(tt), _ptr(tflag), _ptr(sumcp),
                                  C.cast(cnt, C.c_void_p), _ptr(err))
        flops = np.array([c.flops for c in cnt], np.int64)
        transc = np.array([c.transc for c in cnt], np.int64)
        outer = np.array([c.outer for c in cnt], np.int32)
        inner = np.array([c.inner for c in cnt], np.int32)
        max_nsa = np.array([c.max_nsa for c in cnt], np.int32)
        draws = np.array([c.draws for c in cnt], np.int32)
        return dict(max_nsa=max_nsa, draws=draws, t=t, ds=ds, chi=chi.reshape(n, 6, 6).transpose(0, 2, 1), deff=deff, dsv=dsv, srate=srate,
                    twinsys=tw, tflag=int(tflag[0]), sumcpiter=float(sumcp[0]), err=err, flops=flops,
                    transc=transc, outer=outer, inner=inner)


class OracleModel:
    """Module state of one FASOLT3 problem held in numpy arrays for the oracle."""

    def __init__(self, oracle: Oracle, prob, maxofd: int = 81):
        self.o = oracle
        self.prob = prob
        p = prob
        self.keep = []
        k = self.keep.append
        self.le = np.ascontiguousarray(p.le, np.int32); k(self.le)
        self.lt = np.ascontiguousarray(p.lt, np.int32); k(self.lt)
        self.nd = np.ascontiguousarray(p.nd, np.int32); k(self.nd)
        self.np_ = np.ascontiguousarray(p.np_, np.int32); k(self.np_)
        self.ln = np.ascontiguousarray(p.ln, np.int32); k(self.ln)
        self.lg = np.ascontiguousarray(p.lg, np.int32); k(self.lg)
        self.ib = np.ascontiguousarray(p.ib, np.int32); k(self.ib)
        self.x = np.array(p.x, np.float64)
        self.dx = np.array(p.dx, np.float64)
        self.fc = np.zeros(p.ndof)
        self.ddx = np.zeros(p.ndof)
        st = p.zero_state()
        for name, a in st.items():
            setattr(self, name, np.ascontiguousarray(a))
        self.tiem = np.ascontiguousarray(p.tiem, np.float64)
        self.x_spec = np.ascontiguousarray(p.x_spec, np.float64)
        self.f_spec = np.ascontiguousarray(p.f_spec, np.float64)
        self.mats = (OrcMaterial * len(p.materials))(*[material_struct(m) for m in p.materials])
        self.ctl = control_struct(p.control)
        self.maxofd = maxofd
        self.is_ = np.zeros((p.ndof, maxofd), np.int32)
        self.ss = np.zeros((p.ndof, maxofd))
        m = OrcModel()
        m.nlmnt, m.nnod, m.ngps, m.nsv = p.nlmnt, p.nnod, p.ngps, p.nsv
        m.nmatl, m.inc, m.ninc1, m.nxs, m.nfs = len(p.materials), 1, p.ninc1, p.nxs, p.nfs
        m.maxofd = maxofd
        m.mat_code = 1
        for name in ("le", "lt", "nd", "ln", "lg", "ib", "x", "dx", "fc", "ddx", "eps", "dep", "st", "ep", "ds",
                     "de", "rt", "dr", "sv", "dsv", "srate", "cumslip", "twinsys", "timestwinned", "tiem",
                     "x_spec", "f_spec", "is_", "ss"):
            setattr(m, name, _ptr(getattr(self, name)))
        m.np = _ptr(self.np_)
        m.mat = C.cast(self.mats, C.c_void_p)
        m.ctl = C.cast(C.pointer(self.ctl), C.c_void_p)
        self.m = m
        err = self.o.lib.orc_setpta(C.byref(m))
        if err:
            raise RuntimeError(f"SETPTA failed with code {err}")

    def set_inc(self, inc: int):
        self.m.inc = inc

    def nrstep(self):
        """fc = 0, ss = 0, ELVALS, SUVALS (one pass of NRITER up to the solve)."""
        self.m.sumcpiter = 0.0
        self.fc[:] = 0.0
        self.ss[:] = 0.0
        e1 = self.o.lib.orc_elvals(C.byref(self.m))
        e2 = self.o.lib.orc_suvals(C.byref(self.m))
        return max(e1, e2)

    def facsls(self):
        return self.o.lib.orc_facsls(C.byref(self.m), self.prob.ndof)

    def increment(self, inc: int):
        """UPDATE_NODE, NRITER, UPDATE_GPTV, DoReorient (Fasolt3.f90:38-46)."""
        self.m.inc = inc
        self.o.lib.orc_update_node(C.byref(self.m))
        err = self.o.lib.orc_nriter(C.byref(self.m))
        if err > 500 and not (self.ctl.keep_going and err in (987, 910, 904)):
            return err
        self.o.lib.orc_update_gptv(C.byref(self.m))
        if self.m.tflag != 0:
            self.o.lib.orc_doreorient(C.byref(self.m))
            self.m.tflag = 0
        return err

    def csr(self):
        """Assembled matrix as scipy CSR (1-based is lists -> 0-based columns)."""
        import scipy.sparse as sp
        n = self.prob.ndof
        rows, cols, vals = [], [], []
        for i in range(n):
            cnt = self.is_[i, 0]
            rows.append(i); cols.append(i); vals.append(self.ss[i, 0])
            for k in range(1, cnt + 1):
                rows.append(i); cols.append(self.is_[i, k] - 1); vals.append(self.ss[i, k])
        return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))

"""ctypes binding of the C ABI in include/cpfem.h (libcpfem.so).

Mirrors the reference's operator boundary for the hot path: the batch form of
DSTRSS (src/fasolt/Sig_eps.f90:84) and, on a device-resident problem, the
ELVALS / SUVALS / FACSLS / NRITER sequence (src/fasolt/Stif_sol.f90:10-88).
Error codes are the reference's FAFERR codes; a nonzero one raises
:class:`FafErr` unless the caller asks for the code.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict

import numpy as np

from .problem import MAXSS, MAXSV

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcpfem.so")

_P = C.c_void_p


class CpfemMaterial(C.Structure):
    _fields_ = [("nss", C.c_int32), ("twinlim", C.c_int32), ("twinprob", C.c_double),
                ("e", C.c_double * 9), ("ssang", C.c_double * MAXSS),
                ("ab", (C.c_double * 3) * MAXSS), ("an", (C.c_double * 3) * MAXSS),
                ("ampar", (C.c_double * 10) * MAXSS), ("smatrx", (C.c_double * MAXSS) * MAXSS)]


class CpfemControl(C.Structure):
    _fields_ = [("toler", C.c_double), ("maxit", C.c_int32), ("sor_maxit", C.c_int32),
                ("alpha", C.c_double), ("afactor", C.c_double),
                ("efix", C.c_double), ("smalp", C.c_double), ("smalf", C.c_double),
                ("smald", C.c_double), ("smale", C.c_double), ("smalr", C.c_double),
                ("sor_omega", C.c_double), ("sor_tol", C.c_double),
                ("sor_kind", C.c_int32), ("precision", C.c_int32), ("stall_tol", C.c_double)]


class FafErr(RuntimeError):
    """A FAFERR code returned by the library (reference Fasolt3.f90:299-323)."""

    def __init__(self, code: int, where: str):
        super().__init__(f"{where}: FAFNER error {code}")
        self.code = code


def material_struct(m) -> CpfemMaterial:
    s = CpfemMaterial()
    s.nss, s.twinlim, s.twinprob = int(m.nss), int(m.twinlim), float(m.twinprob)
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


def control_struct(c, precision: int = 0) -> CpfemControl:
    s = CpfemControl()
    for name in ("toler", "alpha", "afactor", "efix", "smalp", "smalf", "smald", "smale", "smalr",
                 "sor_omega", "sor_tol"):
        setattr(s, name, float(getattr(c, name)))
    s.maxit, s.sor_maxit, s.sor_kind = int(c.maxit), int(c.sor_maxit), int(c.sor_kind)
    s.precision = int(precision)
    s.stall_tol = float(getattr(c, "stall_tol", 0.0))
    return s


_lib = None


def lib() -> C.CDLL:
    """Load libcpfem.so; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build() (HIP library not built)")
        L = C.CDLL(LIB_PATH)
        L.cpfem_dstrss_batch.argtypes = [C.POINTER(CpfemMaterial), C.POINTER(CpfemControl), C.c_double, C.c_int32,
                                         _P, _P, _P, _P, _P, _P, _P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P,
                                         _P, _P]
        L.cpfem_device_info.argtypes = [C.c_char_p, C.c_int32, _P]
        L.cpfem_twin_stream.argtypes = [C.c_void_p, C.c_uint64]
        L.cpfem_dstrss_batch_noalloc.argtypes = L.cpfem_dstrss_batch.argtypes
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_info():
    buf = C.create_string_buffer(64)
    cus = np.zeros(1, np.int32)
    code = lib().cpfem_device_info(buf, 64, _ptr(cus))
    if code:
        raise RuntimeError(f"no GPU ({code})")
    return buf.value.decode(), int(cus[0])


def twin_stream(epoch: int) -> None:
    """Epoch of the twin draws of the next dstrss_batch call (each call advances it)."""
    _check(lib().cpfem_twin_stream(None, C.c_uint64(int(epoch))), "cpfem_twin_stream")


def dstrss_batch(material, control, dt, iel, de, r, st, ds, sv, dsv, twinsys=None, timestwinned=None,
                 raise_on_error=False, epoch=None, noalloc=False) -> Dict[str, np.ndarray]:
    """Batched DSTRSS on the GPU (host arrays in the reference's per-point layout).
    epoch: stream epoch of this call's twin draws (None: the next of the library's counter).
    noalloc: use cpfem_dstrss_batch_noalloc (no device allocation; at most 8192 points)."""
    if epoch is not None:
        twin_stream(epoch)
    n = len(iel)
    ms = material_struct(material)
    cs = control_struct(control)
    iel = np.ascontiguousarray(iel, np.int32)
    de = np.ascontiguousarray(de, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    st = np.ascontiguousarray(st, np.float64)
    ds = np.array(ds, np.float64, order="C")
    sv = np.ascontiguousarray(sv, np.float64)
    dsv = np.array(dsv, np.float64, order="C")
    t = np.zeros((n, 6))
    chi = np.zeros((n, 36))
    deff = np.zeros(n)
    srate = np.zeros((n, MAXSS))
    tw = np.zeros(n, np.int32) if twinsys is None else np.array(twinsys, np.int32)
    tt = np.zeros(n, np.int32) if timestwinned is None else np.ascontiguousarray(timestwinned, np.int32)
    tflag = np.zeros(1, np.int32)
    sumcp = np.zeros(1)
    flops = np.zeros(n, np.int64)
    err = np.zeros(n, np.int32)
    fn = lib().cpfem_dstrss_batch_noalloc if noalloc else lib().cpfem_dstrss_batch
    code = fn(C.byref(ms), C.byref(cs), float(dt), n, _ptr(iel), _ptr(de), _ptr(r), _ptr(st),
                                    _ptr(ds), _ptr(sv), _ptr(dsv), sv.shape[1], _ptr(t), _ptr(chi), _ptr(deff),
                                    _ptr(srate), _ptr(tw), _ptr(tt), _ptr(tflag), _ptr(sumcp), _ptr(flops), _ptr(err))
    if code < 0:
        raise RuntimeError(f"cpfem_dstrss_batch failed ({code})")
    if code and raise_on_error:
        raise FafErr(code, "cpfem_dstrss_batch")
    return dict(t=t, ds=ds, chi=chi.reshape(n, 6, 6).transpose(0, 2, 1), deff=deff, dsv=dsv, srate=srate,
                twinsys=tw, tflag=int(tflag[0]), sumcpiter=float(sumcp[0]), err=err, flops=flops, code=code)


# ---------------------------------------------------------------------------
# device-resident problem (NRITER call sequence)
# ---------------------------------------------------------------------------
class CpfemMesh(C.Structure):
    _fields_ = [("nlmnt", C.c_int32), ("nnod", C.c_int32), ("ngps", C.c_int32), ("nsv", C.c_int32),
                ("le", _P), ("lt", _P), ("nd", _P), ("np", _P), ("ln", _P), ("lg", _P), ("ib", _P),
                ("ninc1", C.c_int32), ("nxs", C.c_int32), ("nfs", C.c_int32),
                ("tiem", _P), ("x_spec", _P), ("f_spec", _P)]


class CpfemMaterialsR4(C.Structure):
    """The FAS_CPL material arrays (REAL(4), leading dimension maxss = 60) of cpfem_create_r4."""
    _fields_ = [("nmatl", C.c_int32), ("twinlim", C.c_int32), ("twinprob", C.c_float), ("nss", _P),
                ("e", _P), ("ssang", _P), ("ab", _P), ("an", _P), ("ampar", _P), ("smatrx", _P)]


def materials_r4(materials):
    """Module-layout REAL(4) arrays of the material sets (C order: [set][system][component])."""
    nm = 12
    a = dict(nss=np.zeros(nm, np.int32), e=np.zeros((nm, 9), np.float32), ssang=np.zeros((nm, MAXSS), np.float32),
             ab=np.zeros((nm, MAXSS, 3), np.float32), an=np.zeros((nm, MAXSS, 3), np.float32),
             ampar=np.zeros((nm, MAXSS, 10), np.float32), smatrx=np.zeros((nm, MAXSS, MAXSS), np.float32))
    for k, m in enumerate(materials):
        n = int(m.nss)
        a["nss"][k] = n
        a["e"][k] = m.e[:9]
        a["ssang"][k, :n] = m.ssang[:n]
        a["ab"][k, :n] = m.ab[:n]
        a["an"][k, :n] = m.an[:n]
        a["ampar"][k, :n] = m.ampar[:n, :10]
        a["smatrx"][k, :n, :n] = m.smatrx[:n, :n]
    return a


class CpfemStats(C.Structure):
    _fields_ = [("nit", C.c_int32), ("sor_sweeps", C.c_int32), ("sor_sweeps_total", C.c_int32), ("err", C.c_int32),
                ("sdx", C.c_double), ("sfc", C.c_double), ("sumcpiter", C.c_double), ("alp2", C.c_double),
                ("sor_error", C.c_double), ("t_const_ms", C.c_double), ("t_asm_ms", C.c_double),
                ("t_suvals_ms", C.c_double), ("t_sor_ms", C.c_double), ("t_update_ms", C.c_double),
                ("t_total_ms", C.c_double), ("flops_const", C.c_int64), ("transc_const", C.c_int64),
                ("gp_updates", C.c_int64), ("pts_continued", C.c_int64), ("pts_wide", C.c_int64),
                ("t_sweep_ms", C.c_double), ("sweep_launches", C.c_int64),
                ("sor_omega", C.c_double), ("t_tune_ms", C.c_double), ("sor_wpb", C.c_int32),
                ("tune_sweeps", C.c_int32), ("pts_987", C.c_int64), ("pts_987_last", C.c_int32),
                ("graph_launches", C.c_int32), ("t_const_pass_ms", C.c_double * 4),
                ("t_graph_ms", C.c_double)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["t_const_pass_ms"] = list(self.t_const_pass_ms)
        return d


_CTX = C.c_void_p


def _bind_mesh_api(L):
    L.cpfem_create.argtypes = [C.POINTER(CpfemMesh), C.POINTER(CpfemMaterial), C.c_int32, C.POINTER(CpfemControl),
                               C.POINTER(_CTX)]
    L.cpfem_destroy.argtypes = [_CTX]
    L.cpfem_set_state.argtypes = [_CTX] + [_P] * 6 + [C.c_int32] + [_P] * 8
    L.cpfem_get_state.argtypes = [_CTX] + [_P] * 7 + [C.c_int32] + [_P] * 10
    L.cpfem_create_r4.argtypes = [C.POINTER(CpfemMesh), C.POINTER(CpfemMaterialsR4), C.POINTER(CpfemControl),
                                  C.POINTER(_CTX)]
    L.cpfem_set_state_r4.argtypes = L.cpfem_set_state.argtypes
    L.cpfem_get_state_r4.argtypes = L.cpfem_get_state.argtypes
    L.cpfem_update_node.argtypes = [_CTX]
    L.cpfem_elvals.argtypes = [_CTX, C.c_int32, _P, _P]
    L.cpfem_suvals.argtypes = [_CTX, C.c_int32, _P]
    L.cpfem_facsls.argtypes = [_CTX, _P, _P]
    L.cpfem_nr_update.argtypes = [_CTX, _P, _P]
    L.cpfem_nriter.argtypes = [_CTX, C.c_int32, C.POINTER(CpfemStats)]
    L.cpfem_set_graph_mode.argtypes = [_CTX, C.c_int32]
    L.cpfem_update_gptv.argtypes = [_CTX]
    L.cpfem_reorient.argtypes = [_CTX, _P]
    L.cpfem_set_tflag.argtypes = [_CTX, C.c_int32]
    L.cpfem_cumslip.argtypes = [_CTX, C.c_int32]
    L.cpfem_volume_average.argtypes = [_CTX, _P]
    L.cpfem_increment.argtypes = [_CTX, C.c_int32, C.POINTER(CpfemStats)]
    L.cpfem_get_system.argtypes = [_CTX, _P, _P]
    L.cpfem_get_matrix.argtypes = [_CTX, _P, _P, _P, _P]
    L.cpfem_get_stats.argtypes = [_CTX, C.POINTER(CpfemStats)]
    L.cpfem_autotune.argtypes = [_CTX, C.c_int32, _P, _P]
    L.cpfem_autotune_points.argtypes = [_CTX, C.c_int32, _P, _P]
    L.cpfem_sweep_order.argtypes = [_CTX, _P, C.c_int32, C.c_int32]
    L.cpfem_calib_begin.argtypes = [_CTX, C.c_int32, C.c_int32, C.c_double, C.c_int32]
    L.cpfem_calib_record.argtypes = [_CTX, C.c_int32]
    L.cpfem_calib_history.argtypes = [_CTX, C.c_int32, _P]
    L.cpfem_calib_run.argtypes = [_CTX, _P, _P, C.c_int32, _P, _P, _P, C.POINTER(CpfemCalibStats)]
    L.cpfem_calib_exclude.argtypes = [_CTX, _P, _P]
    L.cpfem_calib_checkpoints.argtypes = [_CTX, C.c_int32]
    L.cpfem_calib_options.argtypes = [_CTX, C.c_int32]
    L.cpfem_calib_end.argtypes = [_CTX]


class CpfemCalibStats(C.Structure):
    _fields_ = [("points", C.c_int32), ("increments", C.c_int32), ("unconverged", C.c_int32),
                ("singular", C.c_int32), ("elastic", C.c_int32), ("equilibrium_solves", C.c_int32),
                ("t_forward_ms", C.c_double), ("t_reverse_ms", C.c_double), ("solve_sweeps", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# cpfem_calib_options flags
CALIB_EQUILIBRIUM = 1

# cpfem_autotune flags
TUNE_LAUNCH = 1
TUNE_OMEGA = 2
TUNE_ORDER = 4


# Codes after which the device Newton loop goes on (the caller decides, as the
# reference's FAFERR would): 987 slip Newton residual, 910 singular slip
# matrix, 904 SOR not converged.
LOCAL_CODES = (987, 910, 904)


def _check(code, where, ok_codes=()):
    if code < 0:
        raise RuntimeError(f"{where} failed ({code})")
    if code and code not in ok_codes and code > 500:
        raise FafErr(code, where)
    return code


class Solver:
    """A FASOLT3 problem resident on the GPU, driven through the C ABI.

    Methods follow the reference's NRITER call sequence (Stif_sol.f90:10-88):
    ``elvals``, ``suvals``, ``facsls``, ``nr_update``; ``nriter`` runs the whole
    damped Newton loop of one increment and ``increment`` adds UPDATE_NODE,
    UPDATE_GPTV and DoReorient (Fasolt3.f90:38-46).
    """

    def __init__(self, prob, precision: int = 0, control=None, real4: bool = False):
        """real4: create and fill the context through the REAL(4) entry points
        (cpfem_create_r4, cpfem_set_state_r4) from single-precision copies of
        the problem's arrays, as the reference's module arrays are."""
        L = lib()
        if not hasattr(L, "_mesh_bound"):
            _bind_mesh_api(L)
            L._mesh_bound = True
        self.L = L
        self.prob = prob
        p = prob
        keep = self._keep = {}

        def arr(name, a, dt):
            a = np.ascontiguousarray(a, dt)
            keep[name] = a
            return a.ctypes.data_as(C.c_void_p)

        m = CpfemMesh()
        m.nlmnt, m.nnod, m.ngps, m.nsv = p.nlmnt, p.nnod, p.ngps, p.nsv
        m.le, m.lt, m.nd = arr("le", p.le, np.int32), arr("lt", p.lt, np.int32), arr("nd", p.nd, np.int32)
        m.np = arr("np", p.np_, np.int32)
        m.ln, m.lg, m.ib = arr("ln", p.ln, np.int32), arr("lg", p.lg, np.int32), arr("ib", p.ib, np.int32)
        m.ninc1, m.nxs, m.nfs = p.ninc1, p.nxs, p.nfs
        m.tiem = arr("tiem", p.tiem, np.float64)
        m.x_spec = arr("x_spec", p.x_spec, np.float64)
        m.f_spec = arr("f_spec", p.f_spec, np.float64)
        self._mesh = m
        self._ctl = control_struct(control or p.control, precision)
        h = _CTX()
        self.real4 = real4
        if real4:
            m.tiem = arr("tiem4", p.tiem, np.float32)
            m.x_spec = arr("x_spec4", p.x_spec, np.float32)
            m.f_spec = arr("f_spec4", p.f_spec, np.float32)
            mt = materials_r4(p.materials)
            t = CpfemMaterialsR4()
            t.nmatl, t.twinlim, t.twinprob = len(p.materials), int(p.materials[-1].twinlim), \
                float(p.materials[-1].twinprob)
            for name in ("nss", "e", "ssang", "ab", "an", "ampar", "smatrx"):
                setattr(t, name, arr("m4_" + name, mt[name], mt[name].dtype))
            self._mats = t
            _check(L.cpfem_create_r4(C.byref(m), C.byref(t), C.byref(self._ctl), C.byref(h)), "cpfem_create_r4")
        else:
            self._mats = (CpfemMaterial * len(p.materials))(*[material_struct(mt) for mt in p.materials])
            _check(L.cpfem_create(C.byref(m), self._mats, len(p.materials), C.byref(self._ctl), C.byref(h)),
                   "cpfem_create")
        self.h = h
        st = p.zero_state()
        st["x"], st["dx"] = p.x, p.dx
        self.set_state(st)

    def close(self):
        if getattr(self, "h", None):
            self.L.cpfem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, s):
        ft = np.float32 if self.real4 else np.float64
        f = lambda k, dt=ft: None if k not in s else np.ascontiguousarray(s[k], dt)  # noqa: E731
        a = {k: f(k) for k in ("x", "dx", "eps", "dep", "sv", "dsv", "ep", "de", "st", "ds", "rt", "dr")}
        a["twinsys"] = f("twinsys", np.int32)
        a["timestwinned"] = f("timestwinned", np.int32)
        ldsv = a["sv"].shape[1] if a["sv"] is not None else MAXSV
        fn = self.L.cpfem_set_state_r4 if self.real4 else self.L.cpfem_set_state
        _check(fn(self.h, _ptr(a["x"]), _ptr(a["dx"]), _ptr(a["eps"]), _ptr(a["dep"]),
                                      _ptr(a["sv"]), _ptr(a["dsv"]), ldsv, _ptr(a["ep"]), _ptr(a["de"]),
                                      _ptr(a["st"]), _ptr(a["ds"]), _ptr(a["rt"]), _ptr(a["dr"]),
                                      _ptr(a["twinsys"]), _ptr(a["timestwinned"])), "cpfem_set_state")

    def get_state(self, real4=None):
        """The state in the reference's per-point layout; real4 (default: as
        the context was created) reads it through cpfem_get_state_r4 into
        single-precision arrays."""
        real4 = self.real4 if real4 is None else real4
        ft = np.float32 if real4 else np.float64
        p = self.prob
        g, d = p.ngps, p.ndof
        z = lambda *shape: np.zeros(shape, ft)  # noqa: E731
        s = dict(x=z(d), dx=z(d), fc=z(d), eps=z(g), dep=z(g), sv=z(g, MAXSV), dsv=z(g, MAXSV), ep=z(g, 6),
                 de=z(g, 6), st=z(g, 6), ds=z(g, 6), rt=z(g, 3), dr=z(g, 3), srate=z(g, MAXSS), cumslip=z(g, MAXSS),
                 twinsys=np.zeros(g, np.int32), timestwinned=np.zeros(g, np.int32))
        fn = self.L.cpfem_get_state_r4 if real4 else self.L.cpfem_get_state
        _check(fn(self.h, _ptr(s["x"]), _ptr(s["dx"]), _ptr(s["fc"]), _ptr(s["eps"]),
                                      _ptr(s["dep"]), _ptr(s["sv"]), _ptr(s["dsv"]), MAXSV, _ptr(s["ep"]),
                                      _ptr(s["de"]), _ptr(s["st"]), _ptr(s["ds"]), _ptr(s["rt"]), _ptr(s["dr"]),
                                      _ptr(s["srate"]), _ptr(s["cumslip"]), _ptr(s["twinsys"]),
                                      _ptr(s["timestwinned"])), "cpfem_get_state")
        return s

    def update_node(self):
        _check(self.L.cpfem_update_node(self.h), "cpfem_update_node")

    def elvals(self, inc, raise_on_error=False):
        sp = np.zeros(1)
        tf = np.zeros(1, np.int32)
        code = self.L.cpfem_elvals(self.h, inc, _ptr(sp), _ptr(tf))
        _check(code, "cpfem_elvals", ok_codes=() if raise_on_error else (code,))
        return code, float(sp[0]), int(tf[0])

    def suvals(self, inc):
        sfc = np.zeros(1)
        _check(self.L.cpfem_suvals(self.h, inc, _ptr(sfc)), "cpfem_suvals")
        return float(sfc[0])

    def facsls(self):
        sw = np.zeros(1, np.int32)
        err = np.zeros(1)
        code = _check(self.L.cpfem_facsls(self.h, _ptr(sw), _ptr(err)), "cpfem_facsls", ok_codes=(904,))
        return code, int(sw[0]), float(err[0])

    def autotune(self, launch=True, omega=True, order=False):
        """Tune the sweeps on the system of the last suvals: launch shape
        (waves per workgroup), relaxation factor and/or the order of the node
        colours within a sweep.  Returns (omega, waves)."""
        om = np.zeros(1)
        wpb = np.zeros(1, np.int32)
        flags = (TUNE_LAUNCH if launch else 0) | (TUNE_OMEGA if omega else 0) | (TUNE_ORDER if order else 0)
        _check(self.L.cpfem_autotune(self.h, flags, _ptr(om), _ptr(wpb)), "cpfem_autotune")
        return float(om[0]), int(wpb[0])

    def sweep_order(self, order=None):
        """The order of the node colours within a sweep; with order given, apply it."""
        buf = np.zeros(64, np.int32)
        if order is not None:
            buf = np.ascontiguousarray(order, np.int32)
            nc = self.L.cpfem_sweep_order(self.h, _ptr(buf), len(buf), 1)
        else:
            nc = self.L.cpfem_sweep_order(self.h, _ptr(buf), len(buf), 0)
        _check(min(nc, 0), "cpfem_sweep_order")
        return buf[:nc].tolist()

    def autotune_points(self, inc=0, first_iters=0, elems_per_block=0):
        """Constitutive first-pass budget and assembly elements per workgroup:
        timed on ELVALS of increment inc (> 0) or set (inc = 0).  Returns both."""
        fi = np.array([first_iters], np.int32)
        ep = np.array([elems_per_block], np.int32)
        _check(self.L.cpfem_autotune_points(self.h, int(inc), _ptr(fi), _ptr(ep)), "cpfem_autotune_points")
        return int(fi[0]), int(ep[0])

    def set_sweeps(self, omega=0.0, waves=0):
        """Apply a relaxation factor and/or waves per workgroup (0 keeps it)."""
        om = np.array([float(omega)])
        wpb = np.array([int(waves)], np.int32)
        _check(self.L.cpfem_autotune(self.h, 0, _ptr(om), _ptr(wpb)), "cpfem_autotune")
        return float(om[0]), int(wpb[0])

    def nr_update(self):
        sdx = np.zeros(1)
        alp2 = np.zeros(1)
        _check(self.L.cpfem_nr_update(self.h, _ptr(sdx), _ptr(alp2)), "cpfem_nr_update")
        return float(sdx[0]), float(alp2[0])

    def set_graph_mode(self, on=True):
        """Run each Newton iteration of nriter/increment as one graph launch
        (residual-rule solves only)."""
        _check(self.L.cpfem_set_graph_mode(self.h, 1 if on else 0), "cpfem_set_graph_mode")

    def nriter(self, inc):
        st = CpfemStats()
        code = _check(self.L.cpfem_nriter(self.h, inc, C.byref(st)), "cpfem_nriter", ok_codes=LOCAL_CODES)
        d = st.as_dict()
        d["code"] = code
        return d

    def cumslip(self, inc):
        _check(self.L.cpfem_cumslip(self.h, inc), "cpfem_cumslip")

    def update_gptv(self):
        _check(self.L.cpfem_update_gptv(self.h), "cpfem_update_gptv")

    def volume_average(self):
        """Volume, mean stress (6), mean strain ep (6) and mean eps over all points."""
        out = np.zeros(14)
        _check(self.L.cpfem_volume_average(self.h, _ptr(out)), "cpfem_volume_average")
        return dict(volume=float(out[0]), st=out[1:7].copy(), ep=out[7:13].copy(), eps=float(out[13]))

    def twin_stream(self, epoch):
        """Epoch of the twin draws of the next ELVALS (each ELVALS advances it)."""
        _check(self.L.cpfem_twin_stream(self.h, C.c_uint64(int(epoch))), "cpfem_twin_stream")

    def set_tflag(self, tflag):
        _check(self.L.cpfem_set_tflag(self.h, int(tflag)), "cpfem_set_tflag")

    def reorient(self):
        n = np.zeros(1, np.int32)
        _check(self.L.cpfem_reorient(self.h, _ptr(n)), "cpfem_reorient")
        return int(n[0])

    def increment(self, inc):
        st = CpfemStats()
        code = _check(self.L.cpfem_increment(self.h, inc, C.byref(st)), "cpfem_increment", ok_codes=LOCAL_CODES)
        d = st.as_dict()
        d["code"] = code
        return d

    # --- calibration: replay of a recorded load path and its adjoint -------
    def calib_begin(self, ninc, ckpt_every=10, toler=1e-10, maxit=60):
        """Start recording ninc increments from the current state."""
        _check(self.L.cpfem_calib_begin(self.h, int(ninc), int(ckpt_every), float(toler), int(maxit)),
               "cpfem_calib_begin")

    def calib_record(self, inc):
        """Keep the inputs of increment inc's last constitutive call (after
        nriter, before update_gptv)."""
        _check(self.L.cpfem_calib_record(self.h, int(inc)), "cpfem_calib_record")

    def calib_history(self, inc):
        """Recorded inputs of increment inc: de (6), dr (3), dsv(1:3), ds (6), weight."""
        rec = np.zeros((self.prob.ngps, 19))
        _check(self.L.cpfem_calib_history(self.h, int(inc), _ptr(rec)), "cpfem_calib_history")
        return dict(de=rec[:, 0:6], r=rec[:, 6:9], dg=rec[:, 9:12], ds=rec[:, 12:18], w=rec[:, 18])

    def calib_run(self, params, target, comp=2, grad=True):
        """Replay with params (tau0, taus, h0, q, m): curve, loss, gradient (or None), stats."""
        params = np.ascontiguousarray(params, np.float64)
        target = np.ascontiguousarray(target, np.float64)
        curve = np.zeros(len(target))
        loss = np.zeros(1)
        g = np.zeros(5) if grad else None
        st = CpfemCalibStats()
        code = self.L.cpfem_calib_run(self.h, _ptr(params), _ptr(target), int(comp), _ptr(curve), _ptr(loss),
                                      _ptr(g) if grad else None, C.byref(st))
        if code == -6:
            raise RuntimeError("cpfem_calib_run: singular Jacobian in a reverse step")
        _check(code, "cpfem_calib_run")
        return curve, float(loss[0]), g, st.as_dict()

    def calib_exclude(self, params):
        """Zero the weight of points whose replay at params does not converge;
        returns the number of points excluded so far."""
        params = np.ascontiguousarray(params, np.float64)
        n = np.zeros(1, np.int32)
        _check(self.L.cpfem_calib_exclude(self.h, _ptr(params), _ptr(n)), "cpfem_calib_exclude")
        return int(n[0])

    def calib_options(self, flags):
        """CALIB_EQUILIBRIUM: keep coordinates and Newton matrices for the
        equilibrium correction (before the first calib_record)."""
        _check(self.L.cpfem_calib_options(self.h, int(flags)), "cpfem_calib_options")

    def calib_checkpoints(self, ckpt_every):
        """Checkpoint spacing of the reverse sweep (<= 0: keep the whole trajectory)."""
        _check(self.L.cpfem_calib_checkpoints(self.h, int(ckpt_every)), "cpfem_calib_checkpoints")

    def calib_end(self):
        _check(self.L.cpfem_calib_end(self.h), "cpfem_calib_end")

    def stats(self):
        """Counters and timers accumulated since the last NRITER (or since creation)."""
        st = CpfemStats()
        _check(self.L.cpfem_get_stats(self.h, C.byref(st)), "cpfem_get_stats")
        return st.as_dict()

    def system(self):
        d = self.prob.ndof
        rhs, ddx = np.zeros(d), np.zeros(d)
        _check(self.L.cpfem_get_system(self.h, _ptr(rhs), _ptr(ddx)), "cpfem_get_system")
        return rhs, ddx

    def matrix(self):
        import scipy.sparse as sp
        nnz = np.zeros(1, np.int64)
        _check(self.L.cpfem_get_matrix(self.h, _ptr(nnz), None, None, None), "cpfem_get_matrix")
        n = int(nnz[0])
        rows, cols, vals = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        _check(self.L.cpfem_get_matrix(self.h, _ptr(nnz), _ptr(rows), _ptr(cols), _ptr(vals)), "cpfem_get_matrix")
        d = self.prob.ndof
        return sp.csr_matrix((vals, (rows, cols)), shape=(d, d))

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
                ("sor_kind", C.c_int32), ("precision", C.c_int32)]


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


def dstrss_batch(material, control, dt, iel, de, r, st, ds, sv, dsv, twinsys=None, timestwinned=None,
                 raise_on_error=False) -> Dict[str, np.ndarray]:
    """Batched DSTRSS on the GPU (host arrays in the reference's per-point layout)."""
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
    code = lib().cpfem_dstrss_batch(C.byref(ms), C.byref(cs), float(dt), n, _ptr(iel), _ptr(de), _ptr(r), _ptr(st),
                                    _ptr(ds), _ptr(sv), _ptr(dsv), sv.shape[1], _ptr(t), _ptr(chi), _ptr(deff),
                                    _ptr(srate), _ptr(tw), _ptr(tt), _ptr(tflag), _ptr(sumcp), _ptr(flops), _ptr(err))
    if code < 0:
        raise RuntimeError(f"cpfem_dstrss_batch failed ({code})")
    if code and raise_on_error:
        raise FafErr(code, "cpfem_dstrss_batch")
    return dict(t=t, ds=ds, chi=chi.reshape(n, 6, 6).transpose(0, 2, 1), deff=deff, dsv=dsv, srate=srate,
                twinsys=tw, tflag=int(tflag[0]), sumcpiter=float(sumcp[0]), err=err, flops=flops, code=code)

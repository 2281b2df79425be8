"""Finite-difference checks of the constitutive adjoint on the CPU.

The per-point reverse step (cpfem_code_amd/csrc/adjoint.hpp, host build in
tests/native/libadjoint_host.so) is held to central differences of the FP64
oracle's DSTRSS (oracle/cpfem_oracle.c, Sig_eps.f90:84-151) followed by the
commit of UPDATE_GPTV (Fasolt3.f90:254-284): for random cotangents (sb, phb,
s0b) the scalar
    J = sb . st' + phb . phi' + s0b . s0'
is differentiated with respect to the start stress, the angles, the slip
resistances and the parameters taus, h0, q and m.
"""
import copy
import ctypes as C
import os

import numpy as np

from cpfem_code_amd import capi
from cpfem_code_amd.problem import Control

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libadjoint_host.so")
TIGHT = Control(toler=1e-11, maxit=200)
PARAMS = ("taus", "h0", "q", "m")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def host_lib():
    if not os.path.exists(LIB):
        raise RuntimeError("tests/native/libadjoint_host.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    L.adj_point_reverse_host.restype = C.c_int
    L.adj_point_reverse_host.argtypes = [C.c_void_p, C.c_double, C.c_int] + [C.c_void_p] * 16
    return L


def with_params(mat, taus=None, h0=None, q=None, m=None):
    """Material with the calibration parameters replaced on every system."""
    mm = copy.deepcopy(mat)
    mm.ampar = np.array(mm.ampar, np.float64)
    mm.smatrx = np.array(mm.smatrx, np.float64)
    if taus is not None:
        mm.ampar[:, 4] = taus
    if h0 is not None:
        mm.ampar[:, 1] = h0
    if m is not None:
        mm.ampar[:, 0] = m
    if q is not None:
        d = np.diag(mm.smatrx).copy()
        mm.smatrx[:] = q
        np.fill_diagonal(mm.smatrx, d)
    return mm


def base_params(mat):
    return dict(taus=float(mat.ampar[0, 4]), h0=float(mat.ampar[0, 1]), q=float(mat.smatrx[0, 1]),
                m=float(mat.ampar[0, 0]))


def forward(oracle, mat, b, dt, ctl=TIGHT):
    """Per-point committed outputs (st', phi', s0') of one oracle DSTRSS call."""
    nss = mat.nss
    dsv = np.array(b["dsv"], np.float64)
    dsv[:, 3:] = 0.0
    o = oracle.dstrss_batch([mat], ctl, dt, b["iel"], b["de"], b["r"], b["st"], b["ds"], b["sv"], dsv)
    st1 = o["t"]
    ph1 = b["sv"][:, :3] + o["dsv"][:, :3]
    s01 = b["sv"][:, 3:3 + nss] + o["dsv"][:, 3:3 + nss]
    return st1, ph1, s01, o


def objective(oracle, mat, b, dt, cot, ctl=TIGHT):
    st1, ph1, s01, _ = forward(oracle, mat, b, dt, ctl)
    return (cot["sb"] * st1).sum(1) + (cot["phb"] * ph1).sum(1) + (cot["s0b"] * s01).sum(1)


def adjoint(L, mat, b, dt, cot, ctl=TIGHT, oracle=None):
    """Reverse step of every point: dict of per-point gradients."""
    st1, ph1, s01, o = forward(oracle, mat, b, dt, ctl)
    n, nss = len(b["iel"]), mat.nss
    ms = capi.material_struct(mat)
    out = dict(st=np.zeros((n, 6)), ph=np.zeros((n, 3)), s0=np.zeros((n, nss)), p=np.zeros((n, 5)),
               de=np.zeros((n, 6)), r=np.zeros((n, 3)), rc=np.zeros(n, int))
    for g in range(n):
        f = lambda a: np.ascontiguousarray(a, np.float64)
        st, php, s0 = f(b["st"][g]), f(b["sv"][g, :3]), f(b["sv"][g, 3:3 + nss])
        de, r, dg, dsg = f(b["de"][g]), f(b["r"][g]), f(b["dsv"][g, :3]), f(b["ds"][g])
        so, sr = f(st1[g]), f(o["srate"][g, :nss])
        sb = f(cot["sb"][g])
        phb = f(cot["phb"][g]).copy()
        s0b = f(cot["s0b"][g]).copy()
        stb = np.zeros(6)
        pb = np.zeros(5)
        deb, rb = np.zeros(6), np.zeros(3)
        rc = L.adj_point_reverse_host(C.byref(ms), dt, int(b["iel"][g]), _p(st), _p(php), _p(s0), _p(de), _p(r),
                                      _p(dg), _p(dsg), _p(so), _p(sr), _p(sb), _p(phb), _p(s0b), _p(stb), _p(pb),
                                      _p(deb), _p(rb))
        out["st"][g], out["ph"][g], out["s0"][g], out["p"][g], out["rc"][g] = stb, phb, s0b, pb, rc
        out["de"][g], out["r"][g] = deb, rb
    return out


def fd_gradients(oracle, mat, b, dt, cot, rel=1e-6, ctl=TIGHT):
    """Central differences of the per-point objective (step rel times the
    input's scale)."""
    n, nss = len(b["iel"]), mat.nss
    res = dict(st=np.zeros((n, 6)), ph=np.zeros((n, 3)), s0=np.zeros((n, nss)), p=np.zeros((n, 5)),
               de=np.zeros((n, 6)), r=np.zeros((n, 3)))

    def pert(field, cols, k, h):
        bb = {key: np.array(v, copy=True) for key, v in b.items()}
        bb[field][:, cols[k]] += h
        return bb

    for k in range(6):
        h = rel * np.abs(b["st"]).max()
        res["st"][:, k] = (objective(oracle, mat, pert("st", range(6), k, h), dt, cot, ctl)
                           - objective(oracle, mat, pert("st", range(6), k, -h), dt, cot, ctl)) / (2 * h)
    for k in range(6):
        h = rel * np.abs(b["de"]).max()
        res["de"][:, k] = (objective(oracle, mat, pert("de", range(6), k, h), dt, cot, ctl)
                           - objective(oracle, mat, pert("de", range(6), k, -h), dt, cot, ctl)) / (2 * h)
    for k in range(3):
        h = rel * np.abs(b["r"]).max()
        res["r"][:, k] = (objective(oracle, mat, pert("r", range(3), k, h), dt, cot, ctl)
                          - objective(oracle, mat, pert("r", range(3), k, -h), dt, cot, ctl)) / (2 * h)
    for k in range(3):
        h = rel
        res["ph"][:, k] = (objective(oracle, mat, pert("sv", range(3), k, h), dt, cot, ctl)
                           - objective(oracle, mat, pert("sv", range(3), k, -h), dt, cot, ctl)) / (2 * h)
    for k in range(nss):
        h = rel * np.abs(b["sv"][:, 3:3 + nss]).max()
        cols = list(range(3, 3 + nss))
        res["s0"][:, k] = (objective(oracle, mat, pert("sv", cols, k, h), dt, cot, ctl)
                           - objective(oracle, mat, pert("sv", cols, k, -h), dt, cot, ctl)) / (2 * h)
    bp = base_params(mat)
    for j, name in enumerate(PARAMS):
        h = rel * abs(bp[name])
        up = with_params(mat, **{name: bp[name] + h})
        dn = with_params(mat, **{name: bp[name] - h})
        res["p"][:, 1 + j] = (objective(oracle, up, b, dt, cot, ctl) - objective(oracle, dn, b, dt, cot, ctl)) / (2 * h)
    return res


def random_cotangents(rng, n, nss):
    return dict(sb=rng.normal(size=(n, 6)), phb=rng.normal(scale=100.0, size=(n, 3)),
                s0b=rng.normal(size=(n, nss)))


def point_materials(prob):
    """Material code of each point's element (1-based point numbers in lg)."""
    iel = np.zeros(prob.ngps, np.int32)
    for e in range(prob.nlmnt):
        for gp in prob.lg[e]:
            if gp > 0:
                iel[gp - 1] = prob.le[e]
    return iel


def replay_oracle(oracle, prob, params, hist, comp=2, ctl=TIGHT):
    """The replay of cpfem_calib_run restated with the FP64 oracle: every
    point's recorded increments through DSTRSS, then UPDATE_GPTV's commit."""
    from cpfem_code_amd.problem import MAXSV
    tau0, taus, h0, q, m = (float(v) for v in params)
    mat = with_params(prob.materials[0], taus=taus, h0=h0, q=q, m=m)
    n, nss = prob.ngps, mat.nss
    iel = point_materials(prob)
    st = np.zeros((n, 6))
    sv = np.array(prob.sv, np.float64)
    sv[:, 3:3 + nss] = tau0
    curve = np.zeros(len(hist))
    for k, h in enumerate(hist):
        dsv = np.zeros((n, MAXSV))
        dsv[:, :3] = h["dg"]
        o = oracle.dstrss_batch([mat], ctl, float(prob.tiem[k]), iel, h["de"], h["r"], st, h["ds"], sv, dsv)
        st = st + o["ds"]
        sv[:, :3 + nss] = sv[:, :3 + nss] + o["dsv"][:, :3 + nss]
        curve[k] = (h["w"] * st[:, comp]).sum() / h["w"].sum()
    return curve, st, sv


def record_oracle(oracle, prob, ninc):
    """The load path run by the oracle (UPDATE_NODE, NRITER, UPDATE_GPTV), with
    the inputs of each increment's last DSTRSS call kept as cpfem_calib_record
    keeps them (unit weights)."""
    from oracle import OracleModel
    om = OracleModel(oracle, prob)
    lib, m = om.o.lib, om.m
    hist = []
    for inc in range(1, ninc + 1):
        m.inc = inc
        lib.orc_update_node(C.byref(m))
        lib.orc_nriter(C.byref(m))
        hist.append(dict(de=om.de.copy(), r=om.dr.copy(), dg=om.dsv[:, :3].copy(), ds=om.ds.copy(),
                         w=np.ones(prob.ngps)))
        lib.orc_update_gptv(C.byref(m))
    return hist


def adjoint_replay_host(L, oracle, prob, params, hist, target, comp=2, ctl=TIGHT):
    """The reverse sweep of cpfem_calib_run on the CPU: the oracle's replay
    forward, then the host build of the per-point reverse step, increment by
    increment backwards.  Returns (loss, gradient, curve)."""
    from cpfem_code_amd.problem import MAXSV
    tau0, taus, h0, q, m = (float(v) for v in params)
    mat = with_params(prob.materials[0], taus=taus, h0=h0, q=q, m=m)
    ms = capi.material_struct(mat)
    n, nss = prob.ngps, mat.nss
    iel = point_materials(prob)
    st = np.zeros((n, 6))
    sv = np.array(prob.sv, np.float64)
    sv[:, 3:3 + nss] = tau0
    states, outs = [], []
    curve = np.zeros(len(hist))
    for k, h in enumerate(hist):
        states.append((st.copy(), sv.copy()))
        dsv = np.zeros((n, MAXSV))
        dsv[:, :3] = h["dg"]
        o = oracle.dstrss_batch([mat], ctl, float(prob.tiem[k]), iel, h["de"], h["r"], st, h["ds"], sv, dsv)
        outs.append((o["t"].copy(), o["srate"][:, :nss].copy()))
        st = st + o["ds"]
        sv[:, :3 + nss] = sv[:, :3 + nss] + o["dsv"][:, :3 + nss]
        curve[k] = (h["w"] * st[:, comp]).sum() / h["w"].sum()
    loss = float(((curve - target) ** 2).sum())
    stb, phb, s0b, pb = np.zeros((n, 6)), np.zeros((n, 3)), np.zeros((n, nss)), np.zeros(5)
    f = lambda a: np.ascontiguousarray(a, np.float64)
    for k in range(len(hist) - 1, -1, -1):
        h = hist[k]
        coef = 2.0 * (curve[k] - target[k]) / h["w"].sum()
        st0, sv0 = states[k]
        so_all, sr_all = outs[k]
        dt = float(prob.tiem[k])
        for g in range(n):
            sb = f(stb[g]).copy()
            sb[comp] += coef * h["w"][g]
            ph, s0v, stn, p = f(phb[g]).copy(), f(s0b[g]).copy(), np.zeros(6), np.zeros(5)
            deb, rb = np.zeros(6), np.zeros(3)
            L.adj_point_reverse_host(C.byref(ms), dt, int(iel[g]), _p(f(st0[g])), _p(f(sv0[g, :3])),
                                     _p(f(sv0[g, 3:3 + nss])), _p(f(h["de"][g])), _p(f(h["r"][g])),
                                     _p(f(h["dg"][g])), _p(f(h["ds"][g])), _p(f(so_all[g])), _p(f(sr_all[g])),
                                     _p(sb), _p(ph), _p(s0v), _p(stn), _p(p), _p(deb), _p(rb))
            stb[g], phb[g], s0b[g] = stn, ph, s0v
            pb += p
    grad = np.array([s0b.sum(), pb[1], pb[2], pb[3], pb[4]])
    return loss, grad, curve

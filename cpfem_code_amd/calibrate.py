"""Calibration harness: gradients of the macroscopic stress-strain curve with
respect to hardening and rate parameters.

The load path is run once on the GPU (UPDATE_NODE, NRITER, UPDATE_GPTV,
DoReorient; Fasolt3.f90:38-46) with the inputs of every point's last DSTRSS
call recorded (cpfem_calib_record).  The curve is then a function of the
parameters through the constitutive updates alone: each point replays its
recorded strain and rotation increments (cpfem_calib_run), and the loss
    L(p) = sum_n (curve_n(p) - target_n)^2,   p = (tau0, taus, h0, q, m)
is differentiated by the hand-written adjoint of the constitutive update
(reverse sweep with checkpoints).  With the recorded kinematics held fixed
the gradient is exact for the replay; with ``equilibrium=True`` the record
also keeps every increment's coordinates and Newton matrix, and the reverse
sweep adds the response of the converged displacement increments (one solve
per increment), which brings the gradient to that of the finite-element path
itself (within 1-3 % at 6^3, where the replay-only gradient is 30 % low).

``CalibLoss`` wraps the replay and its adjoint in a torch.autograd.Function,
so that torch optimisers can drive a fit; torch is used only here.
"""
from __future__ import annotations

import copy
import dataclasses
from typing import Optional

import numpy as np

from . import capi
from .problem import Problem

PARAM_NAMES = ("tau0", "taus", "h0", "q", "m")


def material_params(mat) -> np.ndarray:
    """(tau0, taus, h0, q, m) of a material with uniform parameters."""
    return np.array([float(mat.crss[0]), float(mat.ampar[0, 4]), float(mat.ampar[0, 1]),
                     float(mat.smatrx[0, 1]), float(mat.ampar[0, 0])])


@dataclasses.dataclass
class RecordedPath:
    solver: "capi.Solver"
    ninc: int
    fe_curve: np.ndarray          # volume-average stress component after each increment (finite-element run)
    comp: int
    codes: list


def record_path(prob: Problem, ninc: Optional[int] = None, comp: int = 2, ckpt_every: int = 10,
                toler: float = 1e-10, maxit: int = 60, precision: int = 0, control=None,
                equilibrium: bool = False) -> RecordedPath:
    """Run ninc increments of prob on the GPU, recording every point's last
    constitutive inputs; returns the solver holding the record.  toler and
    maxit are the replay's CPSTRS settings: converging points need well under
    60 outer iterations at 1e-10 (16^3 tension: the same loss and gradient,
    bit for bit, as with 200), and the ones that stop there are excluded."""
    ninc = ninc or prob.ninc1 - 1
    s = capi.Solver(prob, precision=precision, control=control)
    s.calib_begin(ninc, ckpt_every=ckpt_every, toler=toler, maxit=maxit)
    if equilibrium:
        s.calib_options(capi.CALIB_EQUILIBRIUM)
    curve, codes = np.zeros(ninc), []
    for inc in range(1, ninc + 1):
        s.update_node()
        r = s.nriter(inc)
        codes.append(r["code"])
        s.calib_record(inc)
        s.update_gptv()
        s.reorient()
        curve[inc - 1] = s.volume_average()["st"][comp]
    return RecordedPath(solver=s, ninc=ninc, fe_curve=curve, comp=comp, codes=codes)


def seeded_target(path: RecordedPath, p0: np.ndarray, seed: int = 4, spread: float = 0.1):
    """Target curve: the replay at parameters drawn around p0 (factor 1 +-
    spread, seeded).  Returns (target, p_target)."""
    rng = np.random.default_rng(seed)
    p_t = p0 * (1.0 + spread * rng.uniform(-1.0, 1.0, size=5))
    curve, _, _, _ = path.solver.calib_run(p_t, np.zeros(path.ninc), comp=path.comp, grad=False)
    return curve, p_t


def loss_and_grad(path: RecordedPath, params, target, grad: bool = True):
    curve, loss, g, st = path.solver.calib_run(params, target, comp=path.comp, grad=grad)
    return loss, g, curve, st


def central_differences(path: RecordedPath, params, target, rel: float = 1e-6, counts: Optional[list] = None):
    """Central differences of the replay loss (relative step rel per parameter).
    counts, if given, receives (unconverged at +h, unconverged at -h) per
    parameter: a change from the unperturbed count marks a point that crossed
    into or out of the non-converging regime, a jump in the loss."""
    p = np.asarray(params, np.float64)
    g = np.zeros(5)
    for k in range(5):
        h = rel * abs(p[k])
        up, dn = p.copy(), p.copy()
        up[k] += h
        dn[k] -= h
        _, lu, _, su = path.solver.calib_run(up, target, comp=path.comp, grad=False)
        _, ld, _, sd = path.solver.calib_run(dn, target, comp=path.comp, grad=False)
        g[k] = (lu - ld) / (2 * h)
        if counts is not None:
            counts.append((su["unconverged"], sd["unconverged"]))
    return g


def exclude_nonconverging(path: RecordedPath, params, rel: float = 0.0) -> int:
    """Give zero weight to the points whose replay stops at the iteration
    limit at params and, with rel > 0, at each params +- rel*|p_k| e_k (the
    stencil of central differences): points near the non-converging regime of
    CPSTRS cross it under small changes of the parameters, which puts jumps in
    the loss.  Exclusions accumulate; returns the number of points excluded."""
    p = np.asarray(params, np.float64)
    n = path.solver.calib_exclude(p)
    if rel > 0.0:
        for k in range(5):
            for sgn in (1.0, -1.0):
                q = p.copy()
                q[k] += sgn * rel * abs(p[k])
                n = path.solver.calib_exclude(q)
    return n


def problem_with_params(prob: Problem, params) -> Problem:
    """prob with material 1's five calibration parameters replaced (every
    system; tau0 as the initial slip resistance of every point)."""
    tau0, taus, h0, q, m = (float(v) for v in params)
    mat = copy.deepcopy(prob.materials[0])
    mat.ampar = np.array(mat.ampar, np.float64)
    mat.ampar[:, 0], mat.ampar[:, 1], mat.ampar[:, 4] = m, h0, taus
    sm = np.full((mat.nss, mat.nss), q)
    np.fill_diagonal(sm, np.diag(np.asarray(mat.smatrx, np.float64)))
    mat.smatrx = sm
    mat.crss = np.full(mat.nss, tau0)
    sv = np.array(prob.sv, np.float64)
    sv[:, 3:3 + mat.nss] = tau0
    return dataclasses.replace(prob, materials=[mat] + list(prob.materials[1:]), sv=sv)


def fe_curve(prob: Problem, params, ninc: int, comp: int = 2, mask=None) -> np.ndarray:
    """The finite-element path itself at params (not the replay): the stress
    component comp averaged over the points with mask (all if None) with their
    integration weights, after every increment."""
    s = capi.Solver(problem_with_params(prob, params))
    try:
        s.calib_begin(ninc, ckpt_every=0)
        c = np.zeros(ninc)
        for inc in range(1, ninc + 1):
            s.update_node()
            s.nriter(inc)
            s.calib_record(inc)
            s.update_gptv()
            s.reorient()
            w = s.calib_history(inc)["w"]
            if mask is not None:
                w = w * mask
            c[inc - 1] = (w * s.get_state()["st"][:, comp]).sum() / w.sum()
        return c
    finally:
        s.close()


def make_torch_loss():
    """torch.autograd.Function over the replay loss (imports torch lazily)."""
    import torch

    class CalibLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params, path, target):
            p = params.detach().cpu().double().numpy()
            loss, g, _, _ = loss_and_grad(path, p, target, grad=True)
            ctx.save_for_backward(torch.as_tensor(g, dtype=params.dtype))
            return params.new_tensor(loss)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g.to(grad_out.device), None, None

    return CalibLoss


def fit(path: RecordedPath, p_start, target, iterations: int = 20, lr: float = 1.0):
    """Fit the five parameters to target with torch's L-BFGS on the replay loss
    and its adjoint gradient (parameters scaled by p_start, so all start at 1).
    Returns (parameters, loss history per evaluation)."""
    import torch
    CalibLoss = make_torch_loss()
    scale = torch.tensor(np.asarray(p_start, np.float64))
    x = torch.ones(5, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([x], lr=lr, max_iter=iterations, tolerance_grad=1e-12, tolerance_change=1e-14,
                            history_size=10, line_search_fn="strong_wolfe")
    history = []

    def closure():
        opt.zero_grad()
        loss = CalibLoss.apply(x * scale, path, target)
        loss.backward()
        history.append(float(loss.detach()))
        return loss

    opt.step(closure)
    return (x.detach() * scale).numpy(), history

"""Parity in depth: the CPU-runnable configuration's whole load path.

FCC aluminium, 8^3 8-node bricks x 8 points (4,096 points), one orientation
per element from the seed, uniaxial tension along z in 50 increments of
1e-3, SOR omega 1.6 to the 1e-8 Jacobi-scaled residual, the reference's
Newton tolerance 1e-3 (Stif_sol.f90:75-76).  The GPU (FP64 path and the
mixed path) runs all 50 increments against the FP64 oracle.

The multicolour sweep order changes the SOR iterates, and the Newton loop
stops at sdx < 1e-3, so a converged increment carries a solver-dependent
error far above the SOR tolerance, and it accumulates along the path.  The
bar is therefore measured on the oracle itself: the same oracle run with the
SOR tolerance tightened to 1e-10, or with another relaxation factor at the
same tolerance (1.65, 1.7, 1.8 for the 1.6 of the configuration), is another
valid run of the same algorithm, and the largest distance of these runs from
the 1e-8 run at each increment is the spread that the solver alone causes on
this path.  The spread depends strongly on the perturbation because
increments 2-4 carry points held by SCALE_IT (FAFERR 987), whose stresses jump
under small changes of the strain increment: at increment 3 the oracle moves
by 2.4e-4 between tolerances 1e-8 and 1e-10, 1.5e-3 between omega 1.6 and
1.8 and 6.1e-3 between 1.6 and 1.65.  The GPU must stay within a small
multiple of the spread (with a floor at the solver level where the spread is
smaller than the tolerance itself), take the same Newton count in every
increment, and end with stresses and slip resistances within the oracle's
own spread.
"""
import numpy as np
import pytest

from cpfem_code_amd import capi
from cpfem_code_amd.problem import Control, fcc_tension
from oracle import Oracle, OracleModel

pytestmark = pytest.mark.gpu
NINC = 50


def _oracle_path(prob, tol, omega=None):
    import dataclasses
    ctl = dataclasses.replace(prob.control, sor_tol=tol)
    if omega is not None:
        ctl = dataclasses.replace(ctl, sor_omega=omega)
    p = dataclasses.replace(prob, control=ctl)
    om = OracleModel(Oracle(64), p)
    om.ctl.keep_going = 1
    us, nits = [], []
    for inc in range(1, NINC + 1):
        om.increment(inc)
        us.append(om.x - prob.x)
        nits.append(om.m.nit)
    return np.array(us), np.array(nits), om.st.copy(), om.sv.copy()


@pytest.fixture(scope="module")
def config0():
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=400000)
    prob = fcc_tension(8, ninc=NINC, seed=0, control=ctl)
    ua, na, sta, sva = _oracle_path(prob, 1e-8)
    variants = [_oracle_path(prob, 1e-10)] + [_oracle_path(prob, 1e-8, om) for om in (1.65, 1.7, 1.8)]
    return prob, (ua, na, sta, sva), variants


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision", [0, 2])
def test_config0_whole_path_within_oracle_spread(config0, precision):
    prob, (ua, na, sta, sva), variants = config0
    sv = capi.Solver(prob, precision=precision)
    ng = []
    worst = 0.0
    for inc in range(1, NINC + 1):
        st = sv.increment(inc)
        ng.append(st["nit"])
        u = sv.get_state()["x"] - prob.x
        d_gpu = _rel(u, ua[inc - 1])
        d_ref = max(_rel(v[0][inc - 1], ua[inc - 1]) for v in variants)
        ratio = d_gpu / max(d_ref, 1e-7)
        worst = max(worst, ratio)
        print(f"precision {precision} inc {inc}: nit gpu {st['nit']} oracle {na[inc - 1]} "
              f"(variants {[int(v[1][inc - 1]) for v in variants]}); displacement vs oracle {d_gpu:.2e}, "
              f"oracle spread {d_ref:.2e}", flush=True)
    g = sv.get_state()
    sv.close()
    es_g = np.abs(g["st"] - sta).max(1) / np.abs(sta).max()
    es_r = np.max([np.abs(v[2] - sta).max(1) for v in variants], axis=0) / np.abs(sta).max()
    ec_g = _rel(g["sv"][:, 3:prob.nsv], sva[:, 3:prob.nsv])
    ec_r = max(_rel(v[3][:, 3:prob.nsv], sva[:, 3:prob.nsv]) for v in variants)
    print(f"precision {precision}: worst displacement ratio {worst:.2f}; final stress max {es_g.max():.2e} "
          f"(oracle spread {es_r.max():.2e}) median {np.median(es_g):.2e} ({np.median(es_r):.2e}); slip "
          f"resistance {ec_g:.2e} ({ec_r:.2e}); Newton counts equal in {int(np.sum(np.array(ng) == na))} of {NINC}")
    assert np.array_equal(np.array(ng), na)
    assert worst < 3.0, worst
    assert es_g.max() < 3.0 * es_r.max() and np.median(es_g) < 3.0 * max(np.median(es_r), 1e-8)
    assert ec_g < 3.0 * ec_r

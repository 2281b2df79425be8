"""Graph mode of NRITER (cpfem_set_graph_mode) against the call-by-call path.

In graph mode each Newton iteration is one launch of a captured graph (ELVALS,
SUVALS, the solve's start, its first sweep batches with their residual checks
and the gated damped update), continued by replays of a graph of sweep
batches; every stopping decision of the residual-rule solve is taken on the
device.  The kernels and their order are those of the call-by-call path, so
the two must agree bit for bit: nodal fields, point state, Newton and sweep
counts, relaxation factors.  The call-by-call path itself is held to the
oracle by test_gpu_mesh.py.
"""
import dataclasses

import numpy as np
import pytest

from cpfem_code_amd import capi
from cpfem_code_amd.problem import Control, fcc_plane_strain, fcc_tension, hcp_compression

pytestmark = pytest.mark.gpu

FIELDS = ("x", "dx", "fc", "st", "sv", "dsv", "ep", "eps", "srate", "cumslip", "twinsys", "timestwinned")


def _run_pair(prob, precision, ninc, omega_auto=False):
    ctl = prob.control
    if omega_auto:
        ctl = dataclasses.replace(ctl, sor_omega=0.0)
    a = capi.Solver(prob, precision=precision, control=ctl)
    b = capi.Solver(prob, precision=precision, control=ctl)
    b.set_graph_mode(True)
    recs = []
    for inc in range(1, ninc + 1):
        sa = a.increment(inc)
        sb = b.increment(inc)
        print(f"inc {inc}: code {sa['code']}/{sb['code']} nit {sa['nit']}/{sb['nit']} "
              f"sweeps {sa['sor_sweeps_total']}/{sb['sor_sweeps_total']} omega {sa['sor_omega']:.4f}/"
              f"{sb['sor_omega']:.4f} graphs {sb['graph_launches']}", flush=True)
        recs.append((sa, sb))
    ga, gb = a.get_state(), b.get_state()
    a.close()
    b.close()
    return recs, ga, gb


def _assert_same(recs, ga, gb, graphs_expected=True):
    for sa, sb in recs:
        assert sb["code"] == sa["code"]
        assert sb["nit"] == sa["nit"]
        assert sb["sor_sweeps_total"] == sa["sor_sweeps_total"]
        assert sb["sor_omega"] == sa["sor_omega"]
        assert sb["sdx"] == sa["sdx"] and sb["alp2"] == sa["alp2"]
        assert sb["gp_updates"] == sa["gp_updates"] and sb["flops_const"] == sa["flops_const"]
        assert sb["pts_987"] == sa["pts_987"]
        assert sa["graph_launches"] == 0
    if graphs_expected:
        assert sum(sb["graph_launches"] for _, sb in recs) >= sum(sb["nit"] for _, sb in recs) - 1
    for k in FIELDS:
        if k in ga:
            assert np.array_equal(ga[k], gb[k]), k


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_graph_mode_matches_call_by_call(precision):
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=50000)
    prob = fcc_tension(4, ninc=3, seed=11, control=ctl)
    recs, ga, gb = _run_pair(prob, precision, 3)
    _assert_same(recs, ga, gb)


def test_graph_mode_plane_strain_automatic_omega():
    """Plane-strain compression with the relaxation factor tuned at the first
    solve (on the call-by-call path) and then following the solves' measured
    rates (the same sequence in both modes)."""
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=50000)
    prob = fcc_plane_strain(5, ninc=3, seed=2, control=ctl)
    recs, ga, gb = _run_pair(prob, 2, 3, omega_auto=True)
    _assert_same(recs, ga, gb)
    assert len({round(sb["sor_omega"], 12) for _, sb in recs}) >= 1


def test_graph_mode_hcp_twins_with_draws():
    """HCP with tensile twins at twinprob = 0.5: the twin draws' stream epoch
    advances on the device in graph mode and on the host otherwise."""
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=50000)
    prob = hcp_compression(3, ninc=2, seed=3, control=ctl)
    mats = [dataclasses.replace(m, twinprob=0.5) for m in prob.materials]
    prob = dataclasses.replace(prob, materials=mats)
    recs, ga, gb = _run_pair(prob, 0, 2)
    _assert_same(recs, ga, gb)
    assert ga["timestwinned"].sum() > 0


def test_graph_mode_needs_residual_rule():
    prob = fcc_tension(2, ninc=1, seed=0, control=Control(sor_kind=0))
    sv = capi.Solver(prob)
    with pytest.raises(Exception):
        sv.set_graph_mode(True)
    sv.close()

"""A system with NaN entries ends the solve at once with 904 and leaves the
nodal state as it was, through the C ABI.

The reference's SOR tests (Stif_sol.f90:200-212) compare with `>`, which a NaN
error passes silently: the solve would run to its sweep limit and NRITER would
apply a NaN correction to every node.  Here a non-finite stopping measure
stops the solve at the next check, FACSLS reports 904 (SOR not converged), and
NRITER ends without the update, in every stopping rule, precision and in graph
mode.  The NaN is planted from the host: one slip resistance of one point,
which makes that point's stress and tangent, and so the internal forces, the
matrix and the right-hand side, NaN.
"""
import numpy as np
import pytest

from cpfem_code_amd import capi
from cpfem_code_amd.problem import Control, fcc_tension
from gpbatch import mesh_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,precision,graph", [(0, 0, False), (1, 0, False), (1, 2, False), (1, 0, True)])
def test_nan_point_stress_ends_solve_with_904(kind, precision, graph):
    ctl = Control(sor_kind=kind, sor_tol=1e-8 if kind else 1e-6, sor_omega=1.6, sor_maxit=200000)
    prob = fcc_tension(4, ninc=3, seed=5, control=ctl)
    state = mesh_state(prob, seed=8)
    state["sv"][17, 3] = np.nan        # slip resistance of system 1
    sv = capi.Solver(prob, precision=precision)
    sv.set_state(state)
    if graph:
        sv.set_graph_mode(True)
    before = sv.get_state()
    st = sv.increment(1)
    g = sv.get_state()
    sv.close()
    print(f"kind {kind} precision {precision} graph {graph}: code {st['code']}, Newton iterations {st['nit']}, "
          f"sweeps {st['sor_sweeps_total']}, solve error {st['sor_error']}")
    assert st["code"] == 904
    assert st["nit"] == 1
    assert st["sor_sweeps_total"] <= 64          # stopped at the first check, not at the sweep limit
    assert not np.isfinite(st["sor_error"])
    assert np.isfinite(g["x"]).all()
    # the predictor of UPDATE_NODE is applied, the Newton correction is not
    assert np.array_equal(g["x"], before["x"] + before["dx"])

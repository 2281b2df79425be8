"""Sweep tuning (cpfem_autotune) through the C ABI.

The launch shape only packs row groups into workgroups, so the solve at 1, 2
and 4 waves per workgroup must give bitwise the same correction and sweep
count.  A tuned relaxation factor changes the iterates, not the solution: the
solve with it must match a direct sparse solve of the same system, in fewer
sweeps than the reference's compiled omega 1.2 (Stif_sol.f90:173), and a whole
increment with it must match one with a fixed factor at the solver tolerance.
"""
import dataclasses

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from cpfem_code_amd import capi
from cpfem_code_amd.problem import Control, fcc_plane_strain, fcc_tension
from gpbatch import mesh_state

pytestmark = pytest.mark.gpu


def _solver(prob, precision, seed=8):
    sv = capi.Solver(prob, precision=precision)
    sv.set_state(mesh_state(prob, seed=seed))
    sv.elvals(1)
    sv.suvals(1)
    return sv


@pytest.mark.parametrize("precision", [0, 1])
def test_launch_shape_keeps_iterates(precision):
    prob = fcc_plane_strain(8, seed=3, control=Control(sor_kind=1, sor_tol=1e-9, sor_omega=1.6, sor_maxit=20000))
    sv = _solver(prob, precision)
    print(f"autotune: launch shapes, precision {precision}", flush=True)
    out = {}
    for w in (4, 2, 1):
        assert sv.set_sweeps(waves=w) == (1.6, w)
        code, sweeps, err = sv.facsls()
        assert code == 0
        assert sv.stats()["sor_wpb"] == w
        out[w] = (sweeps, err, sv.system()[1].copy())
    for w in (2, 1):
        assert out[w][0] == out[4][0]
        assert out[w][1] == out[4][1]
        np.testing.assert_array_equal(out[w][2], out[4][2])
    om, w = sv.autotune(launch=True, omega=False)
    assert om == 1.6 and w in (1, 2, 4)
    code, sweeps, err = sv.facsls()
    assert sweeps == out[4][0]
    np.testing.assert_array_equal(sv.system()[1], out[4][2])
    sv.close()


def test_launch_shape_keeps_update_rule_solve():
    """The reference's relative-update rule, decided on the device from
    per-workgroup partial sums: the packing changes only their summation order."""
    prob = fcc_tension(6, ninc=3, seed=5, control=Control(sor_kind=0, sor_tol=1e-6, sor_omega=1.2, sor_maxit=5000))
    sv = _solver(prob, 0)
    res = []
    for w in (4, 1):
        sv.set_sweeps(waves=w)
        code, sweeps, err = sv.facsls()
        res.append((sweeps, sv.system()[1].copy()))
    assert abs(res[0][0] - res[1][0]) <= 1
    if res[0][0] == res[1][0]:
        np.testing.assert_array_equal(res[0][1], res[1][1])
    sv.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_tuned_omega_solves_system(precision):
    ctl = Control(sor_kind=1, sor_tol=1e-10, sor_omega=1.2, sor_maxit=50000)
    prob = fcc_tension(6, ninc=3, seed=5, control=ctl)
    sv = _solver(prob, precision)
    print(f"autotune: omega, precision {precision}", flush=True)
    code, sweeps_fixed, _ = sv.facsls()
    assert code == 0
    before = sv.system()[1].copy()
    om, w = sv.autotune()
    # the system and the last correction are left as they were
    np.testing.assert_array_equal(sv.system()[1], before)
    assert 1.0 <= om <= 1.98 and w in (1, 2, 4)
    st = sv.stats()
    assert st["sor_omega"] == om and st["tune_sweeps"] > 0
    code, sweeps, err = sv.facsls()
    assert code == 0 and err < 1e-10
    assert sv.stats()["sor_omega"] == om
    assert sweeps <= sweeps_fixed, (om, sweeps, sweeps_fixed)
    rhs, ddx = sv.system()
    A = sv.matrix().tocsc()
    direct = spla.spsolve(A, rhs)
    assert float(np.max(np.abs(ddx - direct)) / np.max(np.abs(direct))) < 1e-6
    sv.close()


def test_automatic_omega_increment_matches_fixed():
    """sor_omega <= 0 tunes at the first solve; the increments are the same as
    with a fixed factor to the solver tolerance.  Increments 1 and 2 of 4^3
    tension from the initial state converge without error codes (9 and 7
    Newton iterations), where the FP64 oracle itself moves by 7e-16 between
    omega 1.6 and 1.9; an increment with FAFERR 987 points is no test of this
    (from a random plastic state the oracle moves by 2e-9 between the same two
    factors: the held points amplify the solver's differences)."""
    base = fcc_tension(4, ninc=3, seed=7, control=Control(sor_kind=1, sor_tol=1e-11, sor_omega=1.6,
                                                          sor_maxit=50000))
    xs, stats = [], []
    for om in (1.6, 0.0):
        prob = dataclasses.replace(base, control=dataclasses.replace(base.control, sor_omega=om))
        sv = capi.Solver(prob, precision=0)
        st = [sv.increment(1), sv.increment(2)]
        stats.append(st)
        xs.append(sv.get_state()["x"].copy())
        sv.close()
    assert stats[0][0]["t_tune_ms"] == 0.0 and stats[0][0]["sor_omega"] == 1.6
    assert stats[1][0]["tune_sweeps"] > 0 and 1.0 <= stats[1][0]["sor_omega"] <= 1.98
    assert [s["nit"] for s in stats[1]] == [s["nit"] for s in stats[0]]
    assert all(s["code"] == 0 for s in stats[0] + stats[1])
    d = float(np.max(np.abs(xs[1] - xs[0])) / np.max(np.abs(xs[0])))
    assert d < 1e-8


def test_autotune_arguments():
    prob = fcc_tension(3, ninc=2, seed=1, control=Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.5))
    sv = capi.Solver(prob, precision=0)
    sv.set_state(mesh_state(prob, seed=1))
    with pytest.raises(RuntimeError):
        sv.autotune()                       # no system assembled yet
    with pytest.raises(RuntimeError):
        sv.set_sweeps(waves=3)
    with pytest.raises(RuntimeError):
        sv.set_sweeps(omega=2.0)
    assert sv.set_sweeps(omega=1.7, waves=2) == (1.7, 2)
    sv.close()
    bad = dataclasses.replace(prob, control=dataclasses.replace(prob.control, sor_omega=2.0))
    with pytest.raises(RuntimeError):
        capi.Solver(bad, precision=0)


def test_point_kernel_shapes_keep_results():
    """The constitutive first-pass budget and the assembly's elements per
    workgroup only pack the work: increments are bitwise the same at every
    setting, and tuning them on a state (cpfem_autotune_points) leaves that
    state as it was."""
    prob = fcc_plane_strain(4, ninc=3, seed=2, control=Control(sor_kind=1, sor_tol=1e-10, sor_omega=1.6,
                                                                sor_maxit=50000))
    runs = []
    for mode in ("default", (3, 1), (32, 2), "tuned"):
        sv = capi.Solver(prob, precision=0)
        if isinstance(mode, tuple):
            assert sv.autotune_points(0, *mode) == mode
        sv.increment(1)
        if mode == "tuned":
            fi, ep = sv.autotune_points(2)
            print(f"autotune points: first-pass budget {fi}, elements per workgroup {ep}", flush=True)
            assert fi in (4, 8, 16, 32) and ep in (1, 2, 4)
        sv.increment(2)
        runs.append(sv.get_state())
        sv.close()
    for r in runs[1:]:
        for k in ("x", "dx", "st", "sv", "ep", "srate", "cumslip"):
            assert np.array_equal(r[k], runs[0][k]), k


@pytest.mark.parametrize("precision", [0, 2])
def test_colour_order_changes_iterates_not_solution(precision):
    """The order of the node colours within a sweep (cpfem_sweep_order, tuned
    with CPFEM_TUNE_ORDER): every order solves the same system to the residual
    test (checked against a direct solve); order tuning leaves the system and
    the last correction as they were and keeps an order that needs no more
    sweeps than the colour-number order; non-permutations are refused."""
    ctl = Control(sor_kind=1, sor_tol=1e-10, sor_omega=1.85, sor_maxit=50000)
    prob = fcc_plane_strain(8, ninc=3, seed=2, control=ctl)
    sv = _solver(prob, precision)
    print(f"autotune: colour order, precision {precision}", flush=True)
    assert sv.sweep_order() == [0, 1, 5, 4, 6, 7, 3, 2]   # the default of an 8-colour mesh
    A = sv.matrix().tocsc()
    rhs = sv.system()[0].copy()
    direct = spla.spsolve(A, rhs)
    sweeps_of = {}
    for order in (list(range(8)), [0, 1, 5, 4, 6, 7, 3, 2], [7, 6, 5, 4, 3, 2, 1, 0]):
        assert sv.sweep_order(order) == order
        code, sweeps, err = sv.facsls()
        assert code == 0 and err < 1e-10
        ddx = sv.system()[1]
        assert float(np.max(np.abs(ddx - direct)) / np.max(np.abs(direct))) < 1e-6, order
        sweeps_of[tuple(order)] = sweeps
    print(f"sweeps per order: {sweeps_of}", flush=True)
    assert sweeps_of[(0, 1, 5, 4, 6, 7, 3, 2)] < sweeps_of[tuple(range(8))], sweeps_of
    sv.sweep_order(list(range(8)))
    sv.facsls()
    before = sv.system()[1].copy()
    sv.autotune(launch=False, omega=False, order=True)
    np.testing.assert_array_equal(sv.system()[1], before)
    chosen = sv.sweep_order()
    assert sorted(chosen) == list(range(8))
    code, sweeps, err = sv.facsls()
    assert code == 0 and err < 1e-10
    assert sweeps <= 1.05 * sweeps_of[tuple(range(8))], (chosen, sweeps, sweeps_of)
    with pytest.raises(RuntimeError):
        sv.sweep_order([0, 0, 1, 2, 3, 4, 5, 6])
    with pytest.raises(RuntimeError):
        sv.sweep_order([0, 1, 2])
    sv.close()

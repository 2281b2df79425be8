"""The persistent constitutive launches against one workgroup per 64 list
entries, bit for bit, through the C ABI.

A pass whose list needs more workgroups than the device holds at once runs as
a persistent launch: a grid of resident workgroups whose waves take the next
64 entries of the pass's point list from a device counter as soon as they
finish their last (the uneven work of CPSTRS, Cpstrs.f90:82-139, and of the
slip Newton, crys-cons.f90:24-104).  CPFEM_CONST_QUEUE=0, read when a context
or a batch call sets up its schedule, launches one workgroup per 64 entries
instead, drained by the hardware dispatcher.  A point's update does not
depend on which wave runs it, so both give the same results bit for bit.
The sizes are chosen above the resident capacity (about 1,024 workgroups of
64 points on 256 CUs), so that the persistent launches really run.
"""
import numpy as np
import pytest

from cpfem_code_amd import capi
from cpfem_code_amd.problem import Control, fcc_aluminium, fcc_plane_strain
from gpbatch import make_gp_batch

pytestmark = pytest.mark.gpu


def _with_queue(monkeypatch, on, fn):
    monkeypatch.setenv("CPFEM_CONST_QUEUE", "1" if on else "0")
    try:
        return fn()
    finally:
        monkeypatch.delenv("CPFEM_CONST_QUEUE")


def test_batch_persistent_matches_dispatch(monkeypatch):
    """200,000 FCC points near yield (5 % elastic): every output of
    cpfem_dstrss_batch equal bit for bit, and the same FLOP counts."""
    mat = fcc_aluminium()
    b = make_gp_batch(mat, 200_000, seed=11, elastic_frac=0.05, stress=40.0)

    def run():
        return capi.dstrss_batch(mat, Control(), 1e-3, b["iel"], b["de"], b["r"], b["st"], b["ds"], b["sv"],
                                 b["dsv"])

    q = _with_queue(monkeypatch, True, run)
    d = _with_queue(monkeypatch, False, run)
    assert int((q["deff"] > 0).sum()) > 100_000
    for k in ("t", "ds", "chi", "srate", "dsv", "deff", "flops", "err", "twinsys"):
        assert np.array_equal(q[k], d[k]), k


@pytest.mark.parametrize("graph", [False, True])
def test_mesh_increments_persistent_match_dispatch(monkeypatch, graph):
    """24^3 bricks x 8 points (110,592 points), plane-strain compression,
    increments 1-3 (the specimen yields in increment 2), call by call and with
    each Newton iteration as one graph launch (whose passes all size their
    launches by the point count): the states equal bit for bit."""
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.8, sor_maxit=100000)
    prob = fcc_plane_strain(24, ninc=3, seed=2, control=ctl)

    def run():
        sv = capi.Solver(prob)
        if graph:
            sv.set_graph_mode(True)
        st = [sv.increment(inc) for inc in (1, 2, 3)]
        out = sv.get_state()
        sv.close()
        return st, out

    st_q, q = _with_queue(monkeypatch, True, run)
    st_d, d = _with_queue(monkeypatch, False, run)
    assert [s["nit"] for s in st_q] == [s["nit"] for s in st_d]
    assert st_q[1]["pts_continued"] > 0
    for k in ("x", "st", "sv", "eps"):
        assert np.array_equal(q[k], d[k]), k

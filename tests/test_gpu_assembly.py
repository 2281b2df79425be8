"""Gather assembly against the colour-by-colour scatter (ELVALS, Elvals.f90:80-221).

On meshes of 8-node bricks the library stores every element's node-pair
blocks and node forces with plain stores, then sums each matrix block and
each node force from the list of its contributions in element-colour order.
With CPFEM_ASSEMBLY=scatter at creation it adds the blocks to the matrix
colour by colour through the scatter map instead.  Both add the same numbers
in the same order, so the matrix, the forces, the right-hand side and whole
increments must agree bit for bit.  Parity of the assembled system with the
oracle is test_gpu_mesh.py's test_elvals_suvals_match_oracle, which runs the
default (gather) path.
"""
import os

import numpy as np
import pytest

from cpfem_code_amd import capi
from cpfem_code_amd.problem import (LT_BRICK_1GP, Control, bcc_iron, fcc_plane_strain, fcc_tension,
                                    with_second_phase)
from gpbatch import mesh_state

pytestmark = pytest.mark.gpu


def _solver(prob, precision, mode):
    if mode == "scatter":
        os.environ["CPFEM_ASSEMBLY"] = "scatter"
    try:
        return capi.Solver(prob, precision=precision)
    finally:
        os.environ.pop("CPFEM_ASSEMBLY", None)


# residual stopping rule: the one the FP32 paths (precision 1, 2) take
CTL = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=20000)
CASES = {
    "tension 6^3, 8 points": lambda: fcc_tension(6, ninc=4, seed=5, control=CTL),
    "plane strain 5^3, 8 points": lambda: fcc_plane_strain(5, ninc=4, seed=2, control=CTL),
    "tension 5^3, 1 point, FCC+BCC": lambda: with_second_phase(
        fcc_tension(5, ninc=4, seed=4, lt=LT_BRICK_1GP, control=CTL), bcc_iron(), fraction=0.4, seed=1),
}


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("case", list(CASES))
def test_gather_system_matches_scatter(case, precision):
    prob = CASES[case]()
    print(f"assembly: {case}, precision {precision}", flush=True)
    out = {}
    for mode in ("gather", "scatter"):
        sv = _solver(prob, precision, mode)
        sv.set_state(mesh_state(prob, seed=11))
        sv.elvals(1)
        sv.suvals(1)
        m = sv.matrix()
        out[mode] = (m.indptr, m.indices, m.data, sv.system()[0].copy(), sv.get_state()["fc"].copy())
        sv.close()
    g, s = out["gather"], out["scatter"]
    for k, name in enumerate(("indptr", "indices", "vals", "rhs", "fc")):
        assert g[k].dtype == s[k].dtype and g[k].shape == s[k].shape, name
        assert g[k].tobytes() == s[k].tobytes(), name   # bit for bit (signed zeros too)
    assert np.abs(g[2]).max() > 0 and np.abs(g[4]).max() > 0


@pytest.mark.parametrize("precision", [0, 2])
def test_gather_increments_match_scatter(precision):
    """Increments 1-3 of plastic tension: every state array equal bit for bit."""
    prob = fcc_tension(6, ninc=4, seed=5, control=CTL)
    states = {}
    for mode in ("gather", "scatter"):
        sv = _solver(prob, precision, mode)
        codes = [sv.increment(inc)["code"] for inc in (1, 2, 3)]
        states[mode] = (codes, sv.get_state())
        sv.close()
    assert states["gather"][0] == states["scatter"][0]
    for k, v in states["gather"][1].items():
        assert v.tobytes() == states["scatter"][1][k].tobytes(), k

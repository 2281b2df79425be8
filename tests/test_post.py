"""CSVMKR-equivalent tables (cpfem_code_amd.post) against the compiled CSVMKR.

tests/golden/csvmkr.npz holds two .frs files and the point (mode 2) and element
(mode 1) tables the reference CSVMKR (src/csvmkr) wrote for them: the reference
driver's state after increment 2 of a 2^3 one-point FCC tension run, and a
synthetic state on 20-node HCP bricks with tie elements placed so that the
reader's compaction (data_in.f90:44-62) also drops the last brick.

CSVMKR computes in REAL(4) and prints G13.5 (five significant digits), so
values are compared to 1e-4 relative plus 1e-5 of the column's largest value.
Orientation triples are compared as rotations: near Phi = 0 the Bunge angles
phi1 and phi2 are not separately defined, and REAL(4) acos resolves angles
near zero only to about sqrt(2 * 2^-24) = 3.5e-4 rad (the element mean of
the small orientation increments sits there), so the bar is a misorientation
of 5e-4 rad.  The other functions are pinned by closed forms below.
"""
import math
import os

import numpy as np
import pytest

from cpfem_code_amd import fsd, post
from cpfem_code_amd.problem import BRICK20, cubic_compliance, hexagonal_compliance

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "csvmkr.npz")
ANGLES = {"gp": [(35, 38), (38, 41)], "el": [(35, 38), (38, 41)]}   # parsed columns (label first)


def _parse(text):
    lines = text.splitlines()
    return lines[:2], np.array([[float(t) for t in ln.split(",")] for ln in lines[2:] if ln.strip()])


def _rot_angle(a, b):
    qa, qb = post.quaternion(*a), post.quaternion(*b)
    return 2.0 * math.acos(min(1.0, abs(float(qa @ qb))))


@pytest.fixture(scope="module")
def golden():
    return np.load(FIX)


@pytest.mark.parametrize("case", ["fcc12", "b20"])
@pytest.mark.parametrize("mode", ["gp", "el"])
def test_tables_match_csvmkr(golden, tmp_path, case, mode):
    frs = tmp_path / "state.frs"
    frs.write_bytes(golden[f"{case}_frs"].tobytes())
    d = fsd.read_frs(str(frs))
    element = mode == "el"
    rows = post.elstate(d) if element else post.gpstate(d)
    out = tmp_path / "mine.csv"
    post.write_csv(str(out), rows, element=element)
    ref_head, ref = _parse(str(golden[f"{case}_{mode}"]))
    head, mine = _parse(out.read_text())
    assert head == ref_head
    assert mine.shape == ref.shape
    np.testing.assert_array_equal(mine[:, 0], ref[:, 0])          # point / element numbers
    if element:
        np.testing.assert_array_equal(mine[:, -1], ref[:, -1])    # material code
    skip = set()
    for lo, hi in ANGLES[mode]:
        for r in range(len(ref)):
            assert _rot_angle(mine[r, lo:hi], ref[r, lo:hi]) < 5e-4, (r, lo)
        skip.update(range(lo, hi))
    cols = [c for c in range(1, 49) if c not in skip]
    scale = np.abs(ref[:, cols]).max(0)
    tol = 1e-4 * np.abs(ref[:, cols]) + 1e-5 * scale
    bad = np.argwhere(np.abs(mine[:, cols] - ref[:, cols]) > tol)
    assert len(bad) == 0, [(int(r), cols[c], ref[r, cols[c]], mine[r, cols[c]]) for r, c in bad[:8]]


def test_g13_5_reproduces_reference_text(golden):
    """Formatting each printed value again gives the reference's text."""
    n = 0
    for key in ("fcc12_gp", "fcc12_el", "b20_gp", "b20_el"):
        for line in str(golden[key]).splitlines()[2:]:
            toks = line.split(",")
            vals = toks[1:49]
            for t in vals:
                assert post.g13_5(float(t)) == t, (key, t)
                n += 1
    assert n == 48 * (8 + 8 + 56 + 7)         # every value of the four tables


def test_kept_elements_compaction():
    # a tie that ends up second to last also drops the element after it
    assert post.kept_elements([8, 8, 8, 0, 8, 8, 8, 8, 0, 8]) == [0, 1, 2, 4, 5, 6, 7]
    assert post.kept_elements([8, 8, 0, 8, 0]) == [0, 1, 3]
    assert post.kept_elements([8, 0, 0]) == [0]
    assert post.kept_elements([0, 0]) == []


def _affine_element(nodes_local, rng):
    a = np.eye(3) + rng.normal(scale=0.2, size=(3, 3))
    b = rng.normal(size=3)
    return np.asarray(nodes_local, float) @ a.T + b, 8.0 * np.linalg.det(a), b


@pytest.mark.parametrize("lt", [12, 13, 22])
def test_point_weights_integrate_affine_volume(lt):
    """Point weights of an affine image of the reference cube sum to its volume
    (8 det A) and the weighted point coordinates to volume times the centre."""
    rng = np.random.default_rng(lt)
    local = BRICK20 if lt == 22 else [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                      (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    xl, vol, centre = _affine_element(local, rng)
    ng = 1 if lt == 12 else 8
    w = [post.point_weight(lt, ig, xl) for ig in range(ng)]
    tol = 2e-5 if lt == 22 else 1e-12          # SHAPE_22 tables carry 6 decimals
    assert abs(sum(v for v, _ in w) - vol) < tol * abs(vol)
    np.testing.assert_allclose(sum(v * x for v, x in w), vol * centre, rtol=tol, atol=tol)


@pytest.mark.parametrize("lt", [21, 23])
def test_tet_wedge_weights_integrate_affine_cell(lt):
    """The 6 tetrahedra (or 2 pentahedra) of a cell of the mesh generator,
    mapped affinely: their point weights sum to the mapped cell's volume and
    the weighted point coordinates to volume times its centre (tables of 6
    decimals, the weight constants 0.0416667 and 0.16666667)."""
    from cpfem_code_amd.problem import _tet10_cells, _wedge15_cells
    rng = np.random.default_rng(lt)
    a = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    vol = 8.0 * np.linalg.det(a)                  # cell of side 2
    centre = np.ones(3) @ a.T + b
    tot, mom = 0.0, np.zeros(3)
    for cell in (_tet10_cells() if lt == 21 else _wedge15_cells()):
        xl = np.asarray(cell, float) @ a.T + b
        for ig in range(4 if lt == 21 else 6):
            v, x = post.point_weight(lt, ig, xl)
            assert v > 0
            tot += v
            mom += v * x
    assert abs(tot - vol) < 3e-6 * vol
    np.testing.assert_allclose(mom, vol * centre, rtol=1e-5, atol=1e-5)


def _bunge(ph1, phi, ph2):
    """Crystal-to-specimen rotation of Bunge angles (columns are crystal axes)."""
    def rz(a):
        return np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])

    def rx(a):
        return np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    return rz(ph1) @ rx(phi) @ rz(ph2)


_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (2, 0), (0, 1)]
_ENG = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])


def _rotated_compliance(e, g):
    """Specimen-axes compliance (engineering shears) of crystal compliances e(1:9)
    by the fourth-order tensor rotation S'_ijkl = g_im g_jn g_kp g_lq S_mnpq."""
    cc = np.zeros((6, 6))
    cc[0, 0], cc[1, 1], cc[2, 2] = e[0], e[1], e[2]
    cc[1, 2] = cc[2, 1] = e[3]
    cc[2, 0] = cc[0, 2] = e[4]
    cc[0, 1] = cc[1, 0] = e[5]
    cc[3, 3], cc[4, 4], cc[5, 5] = e[6], e[7], e[8]
    s = np.zeros((3, 3, 3, 3))
    for a, (i, j) in enumerate(_PAIRS):
        for b, (k, l) in enumerate(_PAIRS):
            for p, q in {(i, j), (j, i)}:
                for r, t in {(k, l), (l, k)}:
                    s[p, q, r, t] = cc[a, b] / (_ENG[a] * _ENG[b])
    sp = np.einsum("im,jn,kp,lq,mnpq->ijkl", g, g, g, g, s)
    return np.array([[sp[i, j, k, l] * _ENG[a] * _ENG[b] for b, (k, l) in enumerate(_PAIRS)]
                     for a, (i, j) in enumerate(_PAIRS)])


@pytest.mark.parametrize("which", ["cubic", "hex"])
def test_elastic_strain_against_rotated_compliance(which):
    """ELASEPS against the tensor rotation of the crystal compliance.

    In the crystal frame (and any frame where the rotated compliance has no
    normal-shear or shear-shear coupling, as a 90 degree turn about a cube axis)
    the two agree.  In general the reference's normal-normal block is the
    rotated compliance and its normal-shear terms are half of it: ELASEPS
    applies the engineering factor 2 on the strain side only and keeps two of
    the four index orders of each shear compliance term (data_out.f90:297-346).
    Its shear-shear block is pinned by the CSVMKR tables alone."""
    e = cubic_compliance(108.0, 62.0, 28.3) if which == "cubic" else hexagonal_compliance(59.7, 26.2, 21.7, 61.7, 16.4)
    e = np.asarray(e, float)

    def matrix(ang):
        return np.stack([post.elastic_strain(e, *ang, np.eye(6)[k]) for k in range(6)], 1)

    for ang in [(0.0, 0.0, 0.0), (math.pi / 2, 0.0, 0.0), (0.0, math.pi / 2, 0.0)]:
        if which == "hex" and ang[1] != 0.0:
            continue            # a turn of the c axis couples shears for the hexagonal crystal
        np.testing.assert_allclose(matrix(ang), _rotated_compliance(e, _bunge(*ang)), rtol=1e-9, atol=1e-15)
    rng = np.random.default_rng(3)
    for _ in range(20):
        ang = (rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * math.pi))
        m, t = matrix(ang), _rotated_compliance(e, _bunge(*ang))
        tol = dict(rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(m[:3, :3], t[:3, :3], **tol)
        np.testing.assert_allclose(m[:3, 3:], 0.5 * t[:3, 3:], **tol)
        np.testing.assert_allclose(m[3:, :3], 0.5 * t[3:, :3], **tol)
        np.testing.assert_allclose(m, m.T, rtol=0, atol=1e-18)


def test_quaternion_euler_round_trip():
    rng = np.random.default_rng(5)
    for _ in range(200):
        ang = (rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * math.pi))
        back = post.euler(post.quaternion(*ang))
        np.testing.assert_allclose(_bunge(*back), _bunge(*ang), atol=1e-9)
    # the quaternion is the rotation's: angle of q equals the rotation angle of g
    ang = (0.3, 0.7, -1.1)
    q = post.quaternion(*ang)
    assert abs(2 * math.acos(q[0]) - math.acos((np.trace(_bunge(*ang)) - 1) / 2)) < 1e-12


def test_von_mises_closed_forms():
    assert post.von_mises([100, 0, 0, 0, 0, 0]) == pytest.approx(100.0)
    assert post.von_mises([0, 0, 0, 0, 0, 10]) == pytest.approx(10 * math.sqrt(3))
    assert post.von_mises([5, 5, 5, 0, 0, 0]) == pytest.approx(0.0)


def test_state_tables_matches_frs_reader(golden, tmp_path):
    """Tables from a Problem plus a state dict equal those from the .frs file."""
    from cpfem_code_amd.problem import LT_BRICK_1GP, Control, fcc_tension
    frs = tmp_path / "state.frs"
    frs.write_bytes(golden["fcc12_frs"].tobytes())
    d = fsd.read_frs(str(frs))
    p = fcc_tension(2, ninc=3, lt=LT_BRICK_1GP, seed=7, control=Control(maxit=20))
    t = post.state_tables(p, d)
    for fn in (post.gpstate, post.elstate):
        for a, b in zip(fn(d), fn(t)):
            assert a[0] == b[0]
            np.testing.assert_array_equal(a[1], b[1])

"""The Fortran driver (cpfem_code_amd/lib/fasolt_gpu) on a generated .fsd.

The driver reads the reference's inputs (dirname, fafcdf.dat, <name>.fsd,
nsave/isave on stdin), runs the increments on the GPU through the C ABI and
writes .frs/.slp/fafrep.txt; its results must equal the Python binding's run
of the same C ABI calls and agree with the FP64 oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from cpfem_code_amd import build, capi, fsd
from cpfem_code_amd.problem import Control, fcc_tension
from oracle import Oracle, OracleModel

pytestmark = pytest.mark.gpu


def test_fortran_host_matches_oracle(tmp_path):
    ctl = Control(sor_kind=1, sor_tol=1e-12, sor_omega=1.6, sor_maxit=50000)
    prob = fcc_tension(3, ninc=3, seed=11, control=ctl)
    d = str(tmp_path)
    fsd.write_fsd(prob, os.path.join(d, "case.fsd"))
    fsd.write_control_files(d, ctl)
    r = subprocess.run([build.HOST_EXE], input="case\n1\n2\n", text=True, cwd=d, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert os.path.exists(os.path.join(d, "case-2.frs")) and os.path.exists(os.path.join(d, "case-3.frs"))
    rep = open(os.path.join(d, "fafrep.txt")).read()
    assert rep.count("Increment number") == 3 and "Power:" in rep, rep[:3000]
    out = fsd.read_frs(os.path.join(d, "case-3.frs"))
    om = OracleModel(Oracle(64), prob)
    for inc in (1, 2, 3):
        om.increment(inc)
    # .frs holds REAL(4) values
    assert np.max(np.abs(out["x"] - om.x)) <= 2e-7 * np.max(np.abs(om.x))
    # point stresses: bounded by the constitutive tolerance (toler = 1e-3 on erepmd)
    e = np.abs(out["st"] - om.st) / np.max(np.abs(om.st))
    assert np.median(e) < 1e-6 and e.max() < 2e-3
    # the final dump is written after the increment loop, whose index is then ninc+1,
    # as in the reference program (Fasolt3.f90:38-65)
    assert out["inc"] == 4


def test_fortran_host_calibration_matches_python(tmp_path):
    """calib.dat makes the Fortran driver record the path and write the replay
    loss and its adjoint gradient (calib.txt); the same through the Python
    binding on the same problem gives the same numbers."""
    from cpfem_code_amd import calibrate
    ctl = Control(sor_kind=1, sor_tol=1e-12, sor_omega=1.6, sor_maxit=50000)
    prob = fcc_tension(3, ninc=4, seed=11, control=ctl)
    d = str(tmp_path)
    fsd.write_fsd(prob, os.path.join(d, "case.fsd"))
    fsd.write_control_files(d, ctl)
    p0 = calibrate.material_params(prob.materials[0]).astype(np.float32).astype(np.float64)
    target = np.array([30.0, 55.0, 60.0, 62.0])
    with open(os.path.join(d, "calib.dat"), "w") as f:
        f.write(" ".join(repr(float(v)) for v in p0) + "\n3\n" + " ".join(repr(float(v)) for v in target) + "\n")
    r = subprocess.run([build.HOST_EXE], input="case\n0\n", text=True, cwd=d, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = open(os.path.join(d, "calib.txt")).read().splitlines()
    nx = int(lines[0].split()[-1])
    loss = float(lines[1].split()[-1])
    grad = np.array([float(v) for v in lines[2].split()[1:]])
    # the same calls from Python, with fafcdf.dat's settings as the driver holds
    # them (REAL(4)); the problem's values are float32-representable already
    import dataclasses
    f4 = lambda v: float(np.float32(v))
    ctl4 = dataclasses.replace(ctl, toler=f4(ctl.toler), alpha=f4(ctl.alpha), afactor=f4(ctl.afactor),
                               efix=f4(ctl.efix), smalp=f4(ctl.smalp), smalf=f4(ctl.smalf), smald=f4(ctl.smald),
                               smale=f4(ctl.smale), smalr=f4(ctl.smalr))
    p = calibrate.record_path(dataclasses.replace(prob, control=ctl4), ninc=4, ckpt_every=10)
    try:
        assert p.solver.calib_exclude(p0) == nx
        _, l2, g2, _ = p.solver.calib_run(p0, target, comp=2)
    finally:
        p.solver.close()
    print(f"host: loss {loss:.6e} python {l2:.6e}; gradient {grad} python {g2}", flush=True)
    assert abs(loss - l2) <= 1e-9 * abs(l2)
    assert np.allclose(grad, g2, rtol=1e-8, atol=1e-12 * np.abs(g2).max())

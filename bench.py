"""Benchmark: Gauss-point updates/s and s/increment of FASOLT3's hot path at 64^3 FCC.

Workload (default): FCC aluminium, 64x64x64 8-node bricks with 2x2x2 points
(262,144 elements, 2,097,152 points, 823,875 DOFs), seeded random texture,
plane-strain compression -1e-3 per increment, SOR to a relative (Jacobi-scaled)
residual of 1e-8.  One step = one load increment: UPDATE_NODE, the damped
Newton loop (point updates, assembly, BCs, SOR solve, update) and the commit,
with all state resident in HBM.

Legs, in this order (each bounded by wall time, see --budget):
  1. FP64 path, SOR relaxation factor and sweep launch shape chosen by the
     library's autotuner at the first solve (in the untimed warm-up): the
     headline ``value`` (point updates per second of whole-increment wall time).
  2. Mixed path (FP32 SOR sweeps on a rounded copy of the FP64 matrix, FP64
     residuals and element products; precision 2), same settings.
  3. The same two paths at the configured omega 1.6, one timed increment each.
  4. One Newton system of the workload solved by each path against a tight
     FP64 solve: the displacement error of each at the 1e-8 residual.
  5. Graph mode: FCC tension at 16^3, increments 2-6 call by call and with
     each Newton iteration as one graph launch; time per increment of each
     and whether the final states are bit for bit equal.
  6. cpu_baseline: the FP64 oracle (the CPU restatement of the same algorithm,
     one core) on whole increments of the CPU-runnable 8^3 uniaxial-tension
     case, next to the GPU on the same increments.
Timed steps stop early when a leg's time is spent; ``steps`` is the number of
increments actually timed on the headline path.

``roofline`` describes the kernel with the largest share of the headline
increments: the constitutive kernels (FP64 vector ALU: algorithmic FLOPs of
the point updates from the kernels' own counters over their event-timed
duration per Newton iteration) or the FP64 colour-sweep kernel (algorithmic
bytes per launch over the event-timed launch duration); ``traffic`` comes from
the committed counter summary (profiles/pmc_traffic.json).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 64] [--budget S]
    python bench.py --sweep-probe 64     # fixed-length SOR run for profiler passes
    python bench.py --const-probe 262144 # constitutive kernels alone for counter passes
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s
FP64_PEAK = 78.6e12        # FLOP/s, vector
FP32_PEAK = 157.3e12
METRIC = "Gauss-pt updates/s + s/increment at 64³ FCC, FP64 vs mixed-prec SOR"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sor_bytes(prob, nnzb, s_a, s_x):
    """Algorithmic bytes of one SOR sweep: nnz*s_A + 4*nnzb + 4*(nnodes+1) + 3*ndof*s_x."""
    return 9 * nnzb * s_a + 4 * nnzb + 4 * (prob.nnod + 1) + 3 * prob.ndof * s_x


def nnzb_structured(dims):
    nx, ny, nz = dims

    def f(n):
        return 3 * (n + 1) - 2
    return f(nx) * f(ny) * f(nz)


def load_traffic():
    """HBM bytes per launch of the profiled kernels from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {k: v["hbm_bytes_per_launch"] for k, v in d.get("kernels", {}).items()}


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def traced_increment(sv, inc, ctl):
    """One increment through the per-step entry points with one progress line
    per Newton iteration (same state and stopping test as cpfem_increment)."""
    before = sv.stats()
    t0 = time.perf_counter()
    sv.update_node()
    nit, code = 0, 0
    while True:
        nit += 1
        c1, _, _ = sv.elvals(inc)
        if c1 < 0 or (c1 > 500 and c1 not in (987, 910)):
            code = max(code, c1)
            break
        sfc = sv.suvals(inc)
        c2, sweeps, err = sv.facsls()
        sdx, alp2 = sv.nr_update()
        code = max(code, c1, c2)
        log(f"[bench]   inc {inc} iteration {nit}: sdx {sdx:.3e} sfc {sfc:.3e} alp2 {alp2:.3f} sweeps {sweeps} "
            f"residual {err:.2e} code {max(c1, c2)}")
        if not (sdx >= ctl.toler and nit < ctl.maxit):
            break
    sv.cumslip(inc)
    sv.update_gptv()
    sv.reorient()
    after = sv.stats()
    keep = ("sor_omega", "sor_wpb", "pts_987_last")
    rec = {k: (after[k] - before[k]) if not isinstance(after[k], list) else
           [a - b for a, b in zip(after[k], before[k])]
           for k in after if k not in ("nit", "sor_sweeps", "err", "sdx", "sfc", "sumcpiter", "alp2", "sor_error")
           + keep}
    rec.update({k: after[k] for k in keep})
    rec.update(nit=nit, code=code, t_total_ms=1e3 * (time.perf_counter() - t0))
    return rec


def run_path(prob, precision, warmup, steps, warm_s, timed_s, capi, trace=False, label="", tune_points=False):
    """Warm-up increments (at least one, at most `warmup`, stopping once warm_s
    is spent), then timed increments (at least one, at most `steps`, stopping
    once timed_s is spent).  Returns the timed records, their wall time, the
    warm-up records and the solver's final state (on request)."""
    t0 = time.time()
    sv = capi.Solver(prob, precision=precision)
    log(f"[bench] {label}: precision={precision} setup {time.time() - t0:.1f} s")
    step = (lambda i: traced_increment(sv, i, prob.control)) if trace else sv.increment
    inc = 0
    warm = []
    t_w = time.perf_counter()
    while inc < max(warmup, 1):
        inc += 1
        st = step(inc)
        warm.append(st)
        log(f"[bench] {label} warmup inc {inc}: nit {st['nit']} sweeps {st['sor_sweeps_total']} "
            f"{st['t_total_ms']:.1f} ms code {st['code']} 987-points {st['pts_987_last']} omega "
            f"{st['sor_omega']:.3g} waves/block {st['sor_wpb']} tune {st['t_tune_ms']:.1f} ms")
        if time.perf_counter() - t_w > warm_s:
            break
    shapes = None
    if tune_points:
        t_tp = time.perf_counter()
        shapes = sv.autotune_points(inc + 1)
        log(f"[bench] {label}: point-kernel shapes tuned on increment {inc + 1}: first-pass budget {shapes[0]}, "
            f"elements per workgroup {shapes[1]} ({time.perf_counter() - t_tp:.1f} s)")
    recs = []
    t0 = time.perf_counter()
    while len(recs) < max(steps, 1) and inc < prob.ninc1:
        inc += 1
        st = step(inc)
        recs.append(st)
        log(f"[bench] {label} inc {inc}: nit {st['nit']} sweeps {st['sor_sweeps_total']} total "
            f"{st['t_total_ms']:.1f} ms const {st['t_const_ms']:.1f} asm {st['t_asm_ms']:.1f} sor "
            f"{st['t_sor_ms']:.1f} code {st['code']} 987-points {st['pts_987_last']} (updates {st['pts_987']}) "
            f"pts {st['gp_updates']} continued {st['pts_continued']} wide {st['pts_wide']} omega {st['sor_omega']:.4g}")
        if time.perf_counter() - t0 > timed_s:
            break
    wall = time.perf_counter() - t0
    sv.close()
    if recs:
        recs[0]["point_shapes"] = shapes
    return recs, wall, warm, inc


def leg_summary(recs, wall, warm, first_inc):
    n = len(recs)
    s = {k: sum(r[k] for r in recs) for k in ("nit", "gp_updates", "sor_sweeps_total", "flops_const",
                                               "transc_const", "pts_987", "sweep_launches")}
    t = {k: sum(r[k] for r in recs) * 1e-3 for k in ("t_const_ms", "t_asm_ms", "t_sor_ms", "t_suvals_ms",
                                                      "t_update_ms", "t_sweep_ms")}
    return {
        "increments_timed": list(range(first_inc, first_inc + n)),
        "s_per_increment": wall / n,
        "s_per_increment_each": [r["t_total_ms"] * 1e-3 for r in recs],
        "gp_updates_per_s": s["gp_updates"] / wall,
        "gp_updates_per_s_constitutive": s["gp_updates"] / t["t_const_ms"],
        "nr_iterations_per_increment": s["nit"] / n,
        "nr_iterations_each": [r["nit"] for r in recs],
        "sor_sweeps_per_increment": s["sor_sweeps_total"] / n,
        "constitutive_s_per_increment": t["t_const_ms"] / n,
        "assembly_s_per_increment": t["t_asm_ms"] / n,
        "sor_s_per_increment": t["t_sor_ms"] / n,
        "bcs_s_per_increment": t["t_suvals_ms"] / n,
        "update_s_per_increment": t["t_update_ms"] / n,
        "constitutive_pass_s_per_increment": [sum(r["t_const_pass_ms"][q] for r in recs) * 1e-3 / n
                                              for q in range(4)],
        "constitutive_flops_per_point_update": s["flops_const"] / max(s["gp_updates"], 1),
        "constitutive_transcendentals_per_point_update": s["transc_const"] / max(s["gp_updates"], 1),
        "sor_omega": recs[-1]["sor_omega"],
        "sor_waves_per_workgroup": recs[-1]["sor_wpb"],
        "omega_tuning_s_once": sum(r["t_tune_ms"] for r in warm + recs) * 1e-3,
        "faferr_codes_each": [r["code"] for r in recs],
        "points_987_final_iteration_each": [r["pts_987_last"] for r in recs],
        "point_updates_987_each": [r["pts_987"] for r in recs],
        "warmup_increments": len(warm),
        "point_kernel_shapes": recs[0].get("point_shapes"),
        "warmup_s_each": [r["t_total_ms"] * 1e-3 for r in warm],
        "warmup_nr_iterations_each": [r["nit"] for r in warm],
        "warmup_sweeps_each": [r["sor_sweeps_total"] for r in warm],
        "warmup_constitutive_s_each": [r["t_const_ms"] * 1e-3 for r in warm],
        "warmup_points_987_each": [r["pts_987_last"] for r in warm],
    }


def roofline_of(prob, recs, traffic, key_sor, s_a, asm_t="double"):
    """Roofline objects of the constitutive phase and the sweep kernel of a leg."""
    nnzb = nnzb_structured(prob.dims)
    launches = sum(r["sweep_launches"] for r in recs)
    t_sw = sum(r["t_sweep_ms"] for r in recs) * 1e-3
    nsw = sum(r["sor_sweeps_total"] for r in recs)
    per_launch = sor_bytes(prob, nnzb, s_a, s_a) * nsw / launches
    dur = t_sw / launches
    sor = {"bound": "hbm", "achieved": per_launch / dur / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": per_launch / dur / HBM_PEAK, "traffic": traffic.get(key_sor) if traffic else None,
           "bytes_per_launch": per_launch, "launch_us": dur * 1e6, "launches": launches,
           "kernel": f"{key_sor} (one colour of one SOR sweep)"}
    nit = sum(r["nit"] for r in recs)
    flops = sum(r["flops_const"] for r in recs)
    t_c = sum(r["t_const_ms"] for r in recs) * 1e-3
    const = {"bound": "fp64-valu", "achieved": flops / t_c / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
             "frac": flops / t_c / FP64_PEAK, "traffic": traffic.get("k_const") if traffic else None,
             "flops_per_launch": flops / nit, "launch_us": t_c / nit * 1e6, "launches": nit,
             "kernel": "k_const<12,*> (the constitutive kernels of one Newton iteration: every point's DSTRSS; "
                       "one 'launch' = the constitutive phase of one ELVALS, event-timed on the library stream)"}
    asm_flops = 75968 * prob.nlmnt * nit
    t_a = sum(r["t_asm_ms"] for r in recs) * 1e-3
    asm_keys = (f"k_assemble<{asm_t}, true>", f"k_gather_blocks<{asm_t}>", "k_gather_forces")
    asm_traffic = sum(traffic[k] for k in asm_keys) if traffic and all(k in traffic for k in asm_keys) else None
    asm = {"bound": "fp64-valu", "achieved": asm_flops / t_a / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
           "frac": asm_flops / t_a / FP64_PEAK, "traffic": asm_traffic,
           "launch_us": t_a / nit * 1e6, "launches": nit,
           "kernel": f"{' + '.join(asm_keys)} (element blocks of every element, then the gathers of the matrix "
                     "blocks and node forces: the assembly of one Newton iteration; traffic summed over the three)"}
    return {"sor": sor, "constitutive": const, "assembly": asm,
            "share_s": {"sor": sum(r["t_sor_ms"] for r in recs) * 1e-3, "constitutive": t_c, "assembly": t_a}}


def mixed_accuracy(prob, capi):
    """One Newton system (the first of increment 1) solved by each path to the
    1e-8 scaled residual, against a 1e-12 FP64 solve of the same system."""
    import dataclasses
    base = dataclasses.replace(prob.control, sor_omega=1.94, sor_maxit=400000)
    out = {}

    def solve(precision, tol):
        sv = capi.Solver(prob, precision=precision, control=dataclasses.replace(base, sor_tol=tol))
        sv.update_node()
        sv.elvals(1)
        sv.suvals(1)
        code, sweeps, err = sv.facsls()
        _, ddx = sv.system()
        sv.close()
        return ddx, sweeps, err

    exact, _, _ = solve(0, 1e-12)
    scale = np.abs(exact).max()
    for name, prec in (("fp64", 0), ("mixed", 2), ("mixed_fp32_matrix", 1)):
        d, sw, err = solve(prec, prob.control.sor_tol)
        out[name] = {"precision": prec, "sweeps": sw, "scaled_residual": err,
                     "displacement_rel_error_vs_tight_fp64": float(np.abs(d - exact).max() / scale)}
    out["note"] = ("max-norm error of the solution of one Newton system relative to its largest component, each "
                   "path stopped at the 1e-8 Jacobi-scaled residual (omega 1.94), against a 1e-12 FP64 solve of "
                   "the same system")
    return out


def cpu_baseline(capi, seconds):
    """The FP64 oracle (the CPU restatement, one core) on whole increments of the
    CPU-runnable case: FCC uniaxial tension along z, 8^3 bricks x 8 points,
    SOR omega 1.6 to the 1e-8 scaled residual; then the GPU on the same
    increments.  Both continue past FAFERR 987 (the reference would stop)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, OracleModel
    from cpfem_code_amd.problem import Control, fcc_tension
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=200000)
    prob = fcc_tension(8, ninc=50, seed=0, control=ctl)
    om = OracleModel(Oracle(64), prob)
    om.ctl.keep_going = 1
    upd, t_cpu, inc = 0, 0.0, 0
    while t_cpu < seconds and inc < 50:
        inc += 1
        t0 = time.perf_counter()
        om.increment(inc)
        t_cpu += time.perf_counter() - t0
        upd += prob.ngps * om.m.nit
    # the GPU on the same increments (FP64 path, same settings)
    sv = capi.Solver(prob, precision=0)
    sv.increment(1)
    sv.close()
    sv = capi.Solver(prob, precision=0)
    g_upd, t_gpu = 0, 0.0
    for i in range(1, inc + 1):
        t0 = time.perf_counter()
        st = sv.increment(i)
        t_gpu += time.perf_counter() - t0
        g_upd += st["gp_updates"]
    sv.close()
    return dict(value=upd / t_cpu, unit="GP-updates/s", cores=1, kind="port",
                sample=f"FP64 oracle (CPU restatement of the reference path, gcc -O2, one thread) on increments "
                       f"1-{inc} of FCC uniaxial tension, 8^3 bricks x 8 points ({prob.ngps} points), SOR omega 1.6 "
                       f"to the 1e-8 scaled residual: {upd} point updates in {t_cpu:.1f} s",
                host_cpu=cpu_model(), host_nproc=os.cpu_count(),
                gpu_same_increments={"gp_updates_per_s": g_upd / t_gpu, "s": t_gpu,
                                     "ratio_gpu_over_cpu": (g_upd / t_gpu) / (upd / t_cpu)})


def sweep_probe(n, sweeps, copy_mib):
    """Fixed-length SOR run on the first Newton system of the n^3 workload:
    exactly `sweeps` colour sweeps in the FP64 path and in both mixed paths,
    each timed with HIP events, then a device copy of copy_mib MiB whose
    FETCH_SIZE / WRITE_SIZE readings calibrate the counters under the same
    profiler."""
    from cpfem_code_amd import capi
    from cpfem_code_amd.problem import Control, fcc_plane_strain
    if sweeps % 16:
        raise SystemExit("--sweep-probe needs a multiple of the 16-sweep batch")
    ctl = Control(sor_kind=1, sor_tol=1e-30, sor_omega=1.6, sor_maxit=sweeps)
    prob = fcc_plane_strain(n, ninc=3, seed=2, control=ctl)
    print(f"sweep probe n={n} sweeps={sweeps} nodes={prob.nnod}", flush=True)
    for prec in (0, 2, 1):
        sv = capi.Solver(prob, precision=prec)
        sv.update_node()
        sv.elvals(1)
        sv.suvals(1)
        t0 = time.perf_counter()
        code, nsw, err = sv.facsls()
        wall = time.perf_counter() - t0
        st = sv.stats()
        per = 1e3 * st["t_sweep_ms"] / max(st["sweep_launches"], 1)
        print(f"precision {prec}: {nsw} sweeps, {st['sweep_launches']} colour launches, {per:.2f} us per launch "
              f"(events), facsls wall {1e3 * wall:.1f} ms, residual {err:.3e}", flush=True)
        sv.close()
    device_copy(copy_mib << 20)


def device_copy(nbytes):
    """A device-to-device copy of nbytes (the runtime's copy kernel), whose
    FETCH_SIZE / WRITE_SIZE readings calibrate the counters.  Called through
    the HIP runtime directly: no torch in a profiled process (torch's own
    kernel launches end the process with SIGSEGV under rocprofv3 here)."""
    import ctypes as C
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    a, b = C.c_void_p(), C.c_void_p()
    for p_ in (a, b):
        if hip.hipMalloc(C.byref(p_), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(f"calibration copy: {nbytes} bytes read and {nbytes} bytes written", flush=True)


def const_probe(n):
    """Constitutive kernels alone, for counter passes: cpfem_dstrss_batch (the
    multi-pass schedule over the 6-, 8- and 12-wide tiers) on n seeded FCC
    points near yield (random orientations, start stresses of 40 MPa, strain
    increments of 1e-3, 5 % elastic), twice (the first is a warm-up)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from gpbatch import make_gp_batch
    from cpfem_code_amd import capi
    from cpfem_code_amd.problem import Control, fcc_aluminium
    mat = fcc_aluminium()
    b = make_gp_batch(mat, n, seed=101, elastic_frac=0.05, stress=40.0)
    print(f"constitutive probe: {n} FCC points", flush=True)
    for rep in range(2):
        t0 = time.perf_counter()
        g = capi.dstrss_batch(mat, Control(), 1e-3, **b)
        wall = time.perf_counter() - t0
        print(f"call {rep}: {wall * 1e3:.1f} ms (host staging included), algorithmic FLOPs "
              f"{int(g['flops'].sum())}, plastic points {int((g['deff'] > 0).sum())}", flush=True)


def graph_mode_leg(capi, n, ninc):
    """The same increments of FCC uniaxial tension (n^3 bricks x 8 points,
    FP64, omega tuned at the first solve) run call by call and with each Newton
    iteration as one graph launch (cpfem_set_graph_mode): wall time of
    increments 2..ninc and whether the two final states are bit for bit equal."""
    import dataclasses
    from cpfem_code_amd.problem import Control, fcc_tension
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=0.0, sor_maxit=400000)
    prob = fcc_tension(n, ninc=ninc, seed=0, control=ctl)
    out, states = {}, {}
    for mode in ("calls", "graph"):
        sv = capi.Solver(prob, precision=0)
        if mode == "graph":
            sv.set_graph_mode(True)
        sv.increment(1)
        t0 = time.perf_counter()
        launches = 0
        nit = 0
        for inc in range(2, ninc + 1):
            st = sv.increment(inc)
            launches += st["graph_launches"]
            nit += st["nit"]
        wall = time.perf_counter() - t0
        states[mode] = sv.get_state()
        sv.close()
        out[mode] = {"s_per_increment": wall / (ninc - 1), "nr_iterations": nit, "graph_launches": launches}
        log(f"[bench] graph leg {mode}: {wall / (ninc - 1) * 1e3:.1f} ms per increment, {nit} Newton iterations, "
            f"{launches} graph launches")
    out["same_state_bitwise"] = all(np.array_equal(states["calls"][k], states["graph"][k])
                                    for k in ("x", "dx", "st", "sv"))
    out["workload"] = f"FCC uniaxial tension, {n}^3 bricks x 8 points, increments 2-{ninc}, FP64"
    return out


def calibration_leg(capi_mod, n, ctl):
    """Gradient of the macroscopic stress-strain curve (FCC tension, 50
    increments, FP64) by the adjoint, next to central differences."""
    import dataclasses
    from cpfem_code_amd import calibrate
    from cpfem_code_amd.problem import fcc_tension
    prob = fcc_tension(n, ninc=50, seed=4, control=dataclasses.replace(ctl, sor_omega=0.0))
    t0 = time.time()
    path = calibrate.record_path(prob, ninc=50, comp=2, ckpt_every=10, equilibrium=False)
    t_record = time.time() - t0
    p0 = calibrate.material_params(prob.materials[0])
    nx = calibrate.exclude_nonconverging(path, p0, rel=1e-6)
    target, _ = calibrate.seeded_target(path, p0, seed=4)
    calibrate.loss_and_grad(path, p0, target)
    t0 = time.time()
    loss, g, _, st = calibrate.loss_and_grad(path, p0, target)
    t_grad = time.time() - t0
    fd = calibrate.central_differences(path, p0, target, rel=1e-6)
    path.solver.close()
    return {"workload": f"FCC uniaxial tension, {n}^3 bricks x 8 points, 50 increments, FP64",
            "record_s": t_record, "points_excluded": nx, "loss_and_gradient_s": t_grad, "loss": loss,
            "gradient": g.tolist(), "central_differences": fd.tolist(),
            "max_rel_diff": float(np.max(np.abs(g - fd) / np.abs(fd)))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=None, help="mesh size (default: the workload's)")
    ap.add_argument("--workload", default="plane_strain", choices=["plane_strain", "tension", "hcp_compression"],
                    help="plane_strain: FCC 64^3 plane-strain compression (the headline); tension: FCC uniaxial "
                         "tension (32^3 FP64 drop-in case); hcp_compression: Mg-like HCP with twins (48^3)")
    ap.add_argument("--omega", default="auto",
                    help="SOR relaxation factor of legs 1-2: 'auto' (tuned with the launch shape at the first "
                         "solve, in the warm-up) or a number")
    ap.add_argument("--budget", type=float, default=360.0,
                    help="wall seconds for the timed work of all legs; each leg gets a share and stops early")
    ap.add_argument("--configured-omega", type=float, default=1.6,
                    help="omega of leg 3 (the configured factor); 0 skips the leg")
    ap.add_argument("--mixed-precision", type=int, default=2, choices=[1, 2],
                    help="mixed path: 2 = FP32 sweeps on a rounded copy of the FP64 matrix (default), 1 = FP32 "
                         "element products and matrix")
    ap.add_argument("--stall-tol", type=float, default=0.0,
                    help="> 0: CPSTRS also stops at a stationary stress (opt-in; the reference runs to maxit)")
    ap.add_argument("--accuracy", type=int, default=1, help="0 skips leg 4 (one-system accuracy check)")
    ap.add_argument("--tune-points", type=int, default=1,
                    help="1: tune the constitutive first-pass budget and the assembly's elements per workgroup "
                         "after the warm-up (cpfem_autotune_points, untimed); 0: the defaults (8, 4)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample; 0 skips it")
    ap.add_argument("--graph-leg", type=int, default=16,
                    help="mesh size of the graph-mode comparison leg (each Newton iteration as one graph launch "
                         "against the call-by-call path); 0 skips it")
    ap.add_argument("--calibration", type=int, default=0, help="mesh size of an opt-in calibration leg (0: none)")
    ap.add_argument("--trace-iterations", action="store_true",
                    help="drive each increment through the per-step calls and print every Newton iteration")
    ap.add_argument("--sweep-probe", type=int, default=0,
                    help="profiling mode: run exactly this many SOR sweeps per path on the first system and exit")
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--const-probe", type=int, default=0,
                    help="profiling mode: the constitutive batch kernels alone on this many seeded points, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU: the z-slab sharding over several GPUs is a "
                                   "design, not an implementation, so nothing is launched", "n_gpus": args.gpus}))
        sys.exit(2)
    t_start = time.time()
    from cpfem_code_amd import capi
    from cpfem_code_amd.problem import Control, fcc_plane_strain, fcc_tension, hcp_compression
    if args.sweep_probe:
        sweep_probe(args.n or 64, args.sweep_probe, args.copy_mib)
        return
    if args.const_probe:
        const_probe(args.const_probe)
        return
    name, cus = capi.device_info()
    import dataclasses
    ctl = Control(sor_kind=1, sor_tol=1e-8, sor_omega=1.6, sor_maxit=400000, stall_tol=args.stall_tol)
    ninc = 100
    if args.workload == "plane_strain":
        args.n = args.n or 64
        prob = fcc_plane_strain(args.n, ninc=ninc, seed=2, control=ctl)
        wl = f"FCC Al plane-strain compression, {args.n}^3 8-node bricks x 8 points, eps_zz=-1e-3/increment"
    elif args.workload == "tension":
        args.n = args.n or 32
        prob = fcc_tension(args.n, ninc=ninc, seed=0, control=ctl)
        wl = f"FCC Al uniaxial tension along z, {args.n}^3 8-node bricks x 8 points, eps_zz=1e-3/increment"
    else:
        args.n = args.n or 48
        prob = hcp_compression(args.n, ninc=ninc, seed=3, control=ctl)
        wl = (f"Mg-like HCP (12 slip + 6 twin systems, twinprob 1) compression along x, basal fibre, "
              f"{args.n}^3 8-node bricks x 8 points, eps_xx=-1e-3/increment")
    om_head = 0.0 if args.omega == "auto" else float(args.omega)
    wl += f", SOR to 1e-8 relative residual, omega {'tuned at the first solve' if om_head == 0 else om_head}"
    log(f"[bench] {name} {cus} CUs; problem {args.n}^3: {prob.nlmnt} elements, {prob.ngps} points, "
        f"{prob.ndof} DOFs")
    B = args.budget
    p_head = dataclasses.replace(prob, control=dataclasses.replace(ctl, sor_omega=om_head))
    legs = {}
    traffic = load_traffic()
    # 1-2: the headline paths
    recs64, wall64, warm64, _ = run_path(p_head, 0, args.warmup, args.steps, 0.08 * B, 0.22 * B, capi,
                                         trace=args.trace_iterations, label="fp64", tune_points=bool(args.tune_points))
    first = len(warm64) + 1
    legs["fp64"] = leg_summary(recs64, wall64, warm64, first)
    recsm, wallm, warmm, _ = run_path(p_head, args.mixed_precision, args.warmup, args.steps, 0.08 * B, 0.18 * B,
                                      capi, trace=args.trace_iterations, label="mixed", tune_points=bool(args.tune_points))
    legs["mixed"] = leg_summary(recsm, wallm, warmm, len(warmm) + 1)
    # 3: the configured relaxation factor, one timed increment each (increment 2)
    if args.configured_omega > 0 and time.time() - t_start < 0.75 * B:
        pc = dataclasses.replace(prob, control=dataclasses.replace(ctl, sor_omega=args.configured_omega))
        for lab, prec in (("fp64", 0), ("mixed", args.mixed_precision)):
            if time.time() - t_start > 0.9 * B:
                break
            r, w, wm, _ = run_path(pc, prec, 1, 1, 0.0, 0.0, capi, label=f"{lab} omega {args.configured_omega}")
            legs[f"{lab}_omega_{args.configured_omega}"] = leg_summary(r, w, wm, len(wm) + 1)
    acc = None
    if args.accuracy and args.workload == "plane_strain" and time.time() - t_start < B:
        acc = mixed_accuracy(prob, capi)
    graph = graph_mode_leg(capi, args.graph_leg, 6) if args.graph_leg > 0 else None
    cpu = cpu_baseline(capi, args.cpu_seconds) if args.cpu_seconds > 0 else None
    h = legs["fp64"]
    roof = roofline_of(prob, recs64, traffic, "k_sor_color<double>", 8)
    roof_m = roofline_of(prob, recsm, traffic, "k_sor_color<float>", 4,
                         "float" if args.mixed_precision == 1 else "double")
    dom = max(("sor", "constitutive", "assembly"), key=lambda k: roof["share_s"][k])
    line = {
        "metric": METRIC,
        "value": h["gp_updates_per_s"],
        "unit": "GP-updates/s",
        "n_gpus": 1,
        "steps": len(recs64),
        "warmup": len(warm64),
        "ms_per_step": 1e3 * h["s_per_increment"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh, texture and load path; no external data)",
        "config": {"workload": wl, "elements": prob.nlmnt, "points": prob.ngps, "dofs": prob.ndof},
        "s_per_increment_fp64": h["s_per_increment"],
        "s_per_increment_mixed": legs["mixed"]["s_per_increment"],
        "mixed_speedup_vs_fp64": h["s_per_increment"] / legs["mixed"]["s_per_increment"],
        "mixed_precision_mode": args.mixed_precision,
        "roofline": roof[dom],
        "roofline_all": {"fp64": {k: v for k, v in roof.items()}, "mixed": {"sor": roof_m["sor"],
                                                                            "constitutive": roof_m["constitutive"]}},
        "cpu_baseline": cpu,
        "legs": legs,
        "mixed_accuracy_one_system": acc,
        "graph_mode": graph,
        "device": f"{name} ({cus} CUs)",
        "note": ("value: point updates (points x Newton iterations) per second of whole-increment wall time on "
                 "the FP64 path over the timed increments (their indices in legs.fp64.increments_timed); the "
                 "relaxation factor is the library autotuner's (tuned once, in the warm-up; legs.*.sor_omega); "
                 "legs.*_omega_1.6 time increment 2 at the configured omega 1.6.  Both paths continue past "
                 "FAFERR 987 (points whose slip Newton ends above 0.5; counts per increment in legs.*), where "
                 "the reference program stops."),
        "bench_wall_s": time.time() - t_start,
    }
    if args.calibration > 0:
        line["calibration"] = calibration_leg(capi, args.calibration, ctl)
        line["bench_wall_s"] = time.time() - t_start
    print(json.dumps(line))


if __name__ == "__main__":
    main()

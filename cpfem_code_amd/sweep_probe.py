"""Fixed-length SOR run on the plane-strain benchmark system, for kernel profiles.

Builds the FCC plane-strain workload of bench.py (n^3 8-node bricks, seed 2),
forms the first Newton system of increment 1 (UPDATE_NODE, ELVALS, SUVALS) and
runs exactly ``--sweeps`` colour sweeps in the FP64 path and in the mixed path
(FP32 sweeps with FP64 residuals), timing each with HIP events.  A
device-to-device copy of ``--copy-mib`` MiB follows: a copy kernel of known
size whose FETCH_SIZE / WRITE_SIZE counter readings calibrate the counters
collected for the sweep kernels under the same profiler.

    python -m cpfem_code_amd.sweep_probe [--n 64] [--sweeps 64] [--copy-mib 512]
"""
import argparse
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--sweeps", type=int, default=64)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--paths", default="fp64,mixed")
    args = ap.parse_args()
    import torch
    from cpfem_code_amd import capi
    from cpfem_code_amd.problem import Control, fcc_plane_strain
    if args.sweeps % 16:
        raise SystemExit("--sweeps must be a multiple of the 16-sweep batch")
    ctl = Control(sor_kind=1, sor_tol=1e-30, sor_omega=1.6, sor_maxit=args.sweeps)
    prob = fcc_plane_strain(args.n, ninc=3, seed=2, control=ctl)
    print(f"sweep_probe n={args.n} sweeps={args.sweeps} nodes={prob.nnod}", flush=True)
    for path in args.paths.split(","):
        prec = 0 if path == "fp64" else 1
        sv = capi.Solver(prob, precision=prec)
        sv.update_node()
        sv.elvals(1)
        sv.suvals(1)
        t0 = time.perf_counter()
        code, sweeps, err = sv.facsls()
        wall = time.perf_counter() - t0
        st = sv.stats()
        per = 1e6 * st["t_sweep_ms"] * 1e-3 / max(st["sweep_launches"], 1)
        print(f"{path}: {sweeps} sweeps, {st['sweep_launches']} colour launches, {per:.2f} us per launch "
              f"(events), facsls wall {1e3 * wall:.1f} ms, residual {err:.3e}", flush=True)
        sv.close()
    nbytes = args.copy_mib << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration copy: {nbytes} bytes read and {nbytes} bytes written", flush=True)


if __name__ == "__main__":
    main()

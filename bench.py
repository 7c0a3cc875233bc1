"""Benchmark: Newton steps of the FCC crystal-plasticity BVP on one MI355X.

Workload (the configuration the metric is quoted on): 64^3 Q1 hex mesh of a unit cube (262,144
elements, 2,097,152 quadrature points, 823,875 dofs), 256 seeded Voronoi grains with uniformly
random orientations, FCC 12 {111}<110> systems (C11/C12/C44 = 170/124/75 GPa, s0 16 MPa, h0 180 MPa,
s_s 148 MPa, a 2.25, q 1.4), uniaxial tension in 20 steps of 5e-4 strain, 4 Newton iterations per
step, Jacobi-BiCG at 1e-10 relative residual.  Synthetic, seeded inputs (no data files).

A step is one Newton iteration: Dirichlet lines + assembly (qp update with tangent at every qp,
K_e/R_e, condensed scatter) + the Krylov solve + the displacement update; the post-increment pass
(stress-only qp pass, averages, reorientation, history commit) runs inside the timed region after
every 4th step.  value = qp updates per second = (qps per step) / (seconds per step).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "qp updates/s + Newton-step time, 64^3 FCC: mixed-precision vs plain FP64 GPU"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(args, quick):
    """Oracle (CPU FP64 restatement, test infrastructure) on a bounded sample of the same workload:
    a 16^3 mesh with the same material/grain statistics, 4 Newton steps of one increment past yield."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcase
    from plasticity_amd import synth
    n = 8 if quick else 16
    mat, pb, vox, rod, bcs, t = synth.fcc_workload(n, 16, 20, 5e-4, seed=2024)
    case = refcase.Case(t["m"] / np.linalg.norm(t["m"], axis=1)[:, None], t["n"] / np.linalg.norm(t["n"], axis=1)[:, None],
                        t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=n, ny=n, nz=n, delT=1.0,
                        total_time=20.0, bcs=bcs)
    case.elem_grain = vox.transpose(2, 1, 0).ravel().astype(np.uint32)  # element e = ex + n(ey + n ez)
    case.grain_rod = rod
    cores = os.cpu_count() or 1
    t0 = time.time()
    out = refcase.oracle_simulate(case, 2, nthreads=cores)
    dt = time.time() - t0
    steps = 2 * case.max_newton
    nq = 8 * n ** 3
    return {"value": nq * steps / dt, "unit": "qp updates/s", "cores": cores, "kind": "port",
            "sample": "oracle (CPU FP64 restatement) on %d^3 elements, 2 increments x 4 Newton steps (%d qp updates with tangent), %.1f s"
                      % (n, nq * steps, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--grains", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--solver", choices=["bicg", "cg"], default="bicg")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import plasticity_amd as pa
    from plasticity_amd import native, synth

    assert torch.cuda.is_available(), "no GPU"
    n = args.n
    mat, pb, vox, rod, bcs, _ = synth.fcc_workload(n, args.grains, 20, 5e-4, seed=2024)
    ctx = pa.Context(mat, pb)
    eg = ctx.element_grains(vox)
    ctx.set_grains(eg, rod)
    ctx.set_solver(native.PA_SOLVER_CG if args.solver == "cg" else native.PA_SOLVER_BICG,
                   native.PA_PREC_MIXED if args.precision == "mixed" else native.PA_PREC_FP64)
    nmax = pb.max_newton
    state = {"step": 0}
    log = []

    def newton_step():
        s = state["step"]
        inc, it = divmod(s, nmax)
        rn, its = ctx.newton_iteration(inc, it)
        asm_ms, kry_ms, _ = ctx.last_timings()
        rec = {"inc": inc, "it": it, "res": rn, "lin_iters": its, "asm_ms": asm_ms, "krylov_ms": kry_ms}
        if it == nmax - 1:
            rec["row"] = ctx.commit_increment().tolist()
            rec["post_ms"] = ctx.last_timings()[2]
        state["step"] += 1
        log.append(rec)
        print("step %d (increment %d, iteration %d): |R| %.3e, %d Krylov iterations, assembly %.1f ms, Krylov %.1f ms"
              % (s, inc, it, rn, its, asm_ms, kry_ms), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        newton_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        newton_step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timed = log[args.warmup:]
    ms_per_step = 1000.0 * dt / args.steps
    nq = ctx.nqp

    # roofline of the dominant kernel: the BiCG operator pair (K p and K^T p~), HBM-bound.
    # algorithmic bytes per launch: 2 reads of the nnz matrix values (8 B each) + Pr, Pl read and
    # Zr, Zl written once (8 B per dof each)
    reps = 20
    avg = ctypes.c_double()
    native.check(native.lib().pa_time_kernel(ctx._h, 0, reps, ctypes.byref(avg)))
    spmv_ms = avg.value
    bytes_spmv = 2 * ctx.nnz * 8 + 4 * ctx.ndof * 8
    achieved = bytes_spmv / (spmv_ms * 1e-3) / 1e9
    native.check(native.lib().pa_time_kernel(ctx._h, 1, 1, ctypes.byref(avg)))
    asm_sweep_ms = avg.value
    lin = [r["lin_iters"] for r in timed]
    out = {
        "metric": METRIC,
        "value": nq / (ms_per_step * 1e-3),
        "unit": "qp updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic (seeded Voronoi grains, random orientations)",
        "config": {"workload": "FCC uniaxial tension, %d^3 Q1 hex (%d qps, %d dofs), %d grains, 5e-4 strain/step, "
                               "4 Newton iterations/step, Jacobi-%s 1e-10" % (n, nq, ctx.ndof, args.grains, args.solver.upper()),
                   "form": "plain FP64" if args.precision == "fp64" else "mixed precision",
                   "mesh": [n, n, n], "nnz": ctx.nnz,
                   "krylov_iters_per_step": float(np.mean(lin)),
                   "assembly_ms_per_step": float(np.mean([r["asm_ms"] for r in timed])),
                   "krylov_ms_per_step": float(np.mean([r["krylov_ms"] for r in timed])),
                   "assembly_sweep_ms": asm_sweep_ms,
                   "parallelism": "single GPU"},
        "roofline": {"kernel": "k_bicg_spmv<double> (K p and K^T p~, 27-point node-block FP64)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": spmv_ms,
                     "algorithmic_bytes": bytes_spmv},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, quick=(n < 32))
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

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
SOLVER_NAMES = {"bicg": "Jacobi-BiCG", "cg": "Jacobi-PCG", "mgcg": "multigrid-PCG"}
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak
HBM_PEAK_GBS = 8000.0
# Algorithmic FP64 flops of the first qp pass per qp and Newton step, counted by the flop-counting
# build of the oracle on a 24^3 sample of this workload (tests/golden/make_flop_fixture.py): the
# flops of each qp up to its first numerically singular local system (those go to the
# pseudo-inverse pass that follows), all of them for the others.
FLOP_FIXTURE = os.path.join(ROOT, "tests", "golden", "qp_flops_fcc24.json")
# FP64 operations issued by one launch of the same pass at 64^3 (k_qp_tangent<12,float>), from
# rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (wave instructions x 64 lanes, FMA = 2; masked
# lanes count) and its HBM bytes (FETCH_SIZE x 2 + WRITE_SIZE KiB, the factors that make both
# counters read the known byte count of a plain copy): profiles/k_qp_tangent_mixed_64cube_fcc.md
# (means over the timed dispatches of the default window, --warmup 5 --steps 20)
PMC_QP_FLOPS = {("mixed", 64): (55784300 + 100712000 + 2 * 347436000 + 3884150) * 64}
PMC_QP_TRAFFIC_KIB = {("mixed", 64): 20712600 * 2 + 15814300}
# HBM bytes per launch of the Jacobi-PCG operator kernel k_cg_spmv at 64^3, measured the same way
# (profiles/k_cg_spmv_mixed_64cube_fcc.md)
PMC_TRAFFIC_KIB = {("cg", "mixed", 64): 142996 * 2 + 6373}


def first_pass_flops_per_qp(steps):
    """Mean algorithmic flops per qp of the first qp pass over the given Newton steps (indices from
    the start of the load path), from the flop fixture; None when it is absent."""
    if not os.path.exists(FLOP_FIXTURE):
        return None, None
    with open(FLOP_FIXTURE) as f:
        fx = json.load(f)
    rows = fx["assembly_pass"]
    per = [rows[s]["flops_to_first_singular"] / rows[s]["calls"] for s in steps if s < len(rows)]
    return (float(np.mean(per)) if per else None), fx["workload"]


def oracle_native_build():
    """The oracle compiled for this host (g++ -O3 -march=native -fopenmp) into a temporary
    directory, or the prebuilt portable one when that fails."""
    import shutil
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "oracle", "cp_oracle.cpp")
    cxx = shutil.which("g++")
    if cxx:
        out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so")
        flags = ["-O3", "-march=native", "-fopenmp", "-fPIC", "-std=c++17", "-shared"]
        r = subprocess.run([cxx] + flags + ["-o", out, src], capture_output=True, text=True, timeout=600)
        if r.returncode == 0:
            return out, "g++ " + " ".join(flags[:3])
    return os.path.join(ROOT, "oracle", "liboracle.so"), "prebuilt -O2 -march=x86-64-v2 (native build failed)"


def cpu_baseline(args, quick):
    """Oracle (CPU FP64 restatement, test infrastructure) built for the host, on a bounded sample of
    the same workload: a 32^3 mesh with the same material and 32 grains (1,024 elements per grain,
    as at 64^3 with 256), Jacobi-PCG (the solver the FP64 workload names), 3 increments x 4 Newton
    steps (about 10-15 s on 16 cores), with the assembly / solve / post-increment split."""
    import ctypes as ct
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcase
    from plasticity_amd import synth
    path, how = oracle_native_build()
    refcase.use_oracle_build(path)
    n, grains, incs = (8, 4, 2) if quick else (32, 32, 3)
    mat, pb, vox, rod, bcs, t = synth.fcc_workload(n, grains, 20, 5e-4, seed=2024)
    case = refcase.Case(t["m"] / np.linalg.norm(t["m"], axis=1)[:, None], t["n"] / np.linalg.norm(t["n"], axis=1)[:, None],
                        t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=n, ny=n, nz=n, delT=1.0,
                        total_time=20.0, bcs=bcs, solver=1)
    case.elem_grain = vox.transpose(2, 1, 0).ravel().astype(np.uint32)  # element e = ex + n(ey + n ez)
    case.grain_rod = rod
    cores = min(os.cpu_count() or 1, 16)
    t0 = time.time()
    out = refcase.oracle_simulate(case, incs, nthreads=cores)
    dt = time.time() - t0
    ns = (ct.c_longlong * 3)()
    refcase.oracle().or_last_timings(ns)
    steps = incs * case.max_newton
    nq = 8 * n ** 3
    return {"value": nq * steps / dt, "unit": "qp updates/s", "cores": cores, "kind": "port",
            "sample": "oracle (CPU FP64 restatement, %s) on %d^3 elements (%d qps, %d grains), Jacobi-PCG, %d increments "
                      "x 4 Newton steps: %.1f s" % (how, n, nq, grains, incs, dt),
            "assembly_s": ns[0] * 1e-9, "solve_s": ns[1] * 1e-9, "post_s": ns[2] * 1e-9,
            "krylov_iters_per_step": float(np.mean(out["lin_iters"])),
            "assembly_qp_updates_per_s": nq * steps / (ns[0] * 1e-9)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--grains", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--solver", choices=["bicg", "cg", "mgcg", "all"], default="all")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-warm", action="store_true", help="skip the learned warm-start measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-RVE workload")
    ap.add_argument("--no-ext", action="store_true", help="skip the periodic-BC and three-phase workloads")
    ap.add_argument("--only-ext", action="store_true", help="run the periodic-BC and three-phase workloads alone")
    ap.add_argument("--only-batch", action="store_true", help="run the batched-RVE workload alone and print its record")
    ap.add_argument("--full-path", type=int, default=1,
                    help="after the timed steps, run the multigrid-PCG forms to the end of the load path (untimed for value)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import plasticity_amd as pa
    from plasticity_amd import native, synth

    assert torch.cuda.is_available(), "no GPU"
    if args.only_batch:
        print(json.dumps({"batched_rves": batched_leg(pa, native, synth)}))
        return
    if args.only_ext:
        print(json.dumps({"periodic_and_phases": extensions_leg(pa, native, synth, {}, n=args.n, grains=args.grains)}))
        return
    n = args.n
    forms = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    solvers = ["bicg", "cg", "mgcg"] if args.solver == "all" else [args.solver]
    results = {}
    for solver in solvers:
        for form in forms:
            results[(form, solver)] = run_form(args, form, solver, pa, native, synth, torch)
    # headline: the last form (mixed when both ran) with the multigrid-preconditioned CG when it ran,
    # else Jacobi-PCG (the solver the FP64 workload names), else BiCG
    hs = "mgcg" if "mgcg" in solvers else ("cg" if "cg" in solvers else solvers[-1])
    head = results[(forms[-1], hs)]
    nq = head["nq"]
    out = {
        "metric": METRIC,
        "value": nq / (head["ms_per_step"] * 1e-3),
        "unit": "qp updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if forms[-1] == "fp64" else "f64 (return map, residual, Krylov) + f32 (tangents, matrix)",
        "data": "synthetic (seeded Voronoi grains, random orientations)",
        "config": {"workload": "FCC uniaxial tension, %d^3 Q1 hex (%d qps, %d dofs), %d grains, 5e-4 strain/step, "
                               "4 Newton iterations/step, %s 1e-10" % (n, nq, head["ndof"], args.grains, SOLVER_NAMES[hs]),
                   "form": forms[-1], "solver": hs, "mesh": [n, n, n], "nnz": head["nnz"], "parallelism": "single GPU",
                   "note": "Jacobi-PCG is the solver the FP64 workload names and Jacobi-BiCG the reference code's; "
                           "multigrid-PCG reaches the same stopping test (||D^-1 r|| <= 1e-10 ||b||) in ~30 instead of "
                           "~1,450 iterations (all three timed)"},
        "roofline": head["roofline"],
    }
    keys = ("ms_per_step", "qp_updates_per_s", "krylov_iters_per_step", "assembly_ms_per_step", "krylov_ms_per_step",
            "post_ms_per_increment", "qp_tangent_ms", "element_ms", "operator_us", "update_us", "update_GBps", "p_us",
            "singular_local_solves_per_step", "rerun_lane_solves", "qp_main_pass_ms", "true_rel_residual", "full_load_path")
    out["config"]["roofline_krylov_operator"] = head["roofline_krylov"]
    for (form, solver), r in results.items():
        out["config"]["%s_%s" % (form, solver)] = {k: r[k] for k in keys}
    if len(forms) == 2:
        for solver in solvers:
            a, b = results[("fp64", solver)], results[("mixed", solver)]
            out["config"]["speedup_mixed_vs_fp64_%s" % solver] = a["ms_per_step"] / b["ms_per_step"]
            if solver != "cg" and ("fp64", "cg") in results:
                out["config"]["speedup_mixed_%s_vs_fp64_cg" % solver] = results[("fp64", "cg")]["ms_per_step"] / b["ms_per_step"]
            rows_a, rows_b = np.array(a["rows"]), np.array(b["rows"])
            err = {}
            for name, cols in (("avg_stress", slice(6, 12)), ("avg_strain", slice(0, 6)), ("slip_total", slice(14, 15))):
                err[name] = float(np.abs(rows_b[:, cols] - rows_a[:, cols]).max() / np.abs(rows_a[:, cols]).max())
            err["qp_cauchy_stress"] = float(np.abs(b["T"] - a["T"]).max() / np.abs(a["T"]).max())
            err["crss"] = float(np.abs(b["s"] - a["s"]).max() / np.abs(a["s"]).max())
            err["displacement"] = float(np.abs(b["u"] - a["u"]).max() / np.abs(a["u"]).max())
            out["config"]["mixed_vs_fp64_gpu_diff_%s" % solver] = err
    if not args.no_ext:
        out["config"]["periodic_and_phases"] = extensions_leg(pa, native, synth, results, n=n, grains=args.grains)
    if not args.no_warm:
        out["config"]["learned_warm_start"] = warm_start_leg(pa, native)
    if not args.no_batch:
        out["config"]["batched_rves"] = batched_leg(pa, native, synth)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, quick=(n < 32))
    print(json.dumps(out))


def batched_leg(pa, native, synth):
    """The batched-RVE workload: 64 RVEs with edge n ~ U{8..24} elements (2.54 M qps), alternately
    BCC with 48 slip systems and HCP with 24, 16 Voronoi grains each, tension in 10 steps of 2.5e-4
    (yield within the last steps; 4 Newton iterations per step, Jacobi-BiCG at 1e-10 per RVE), all
    RVEs in the same launches, Newton iterations replayed from HIP graphs after the autotuner's
    pick; FP64 and the mixed form (FP32 tangents and matrices).  Every Newton step of the load path
    is timed (host clock around each call, the post-increment pass included)."""
    eps = 2.5e-4
    mats, rve_mat, rve_nel, pb, egs, rods = synth.batch_workload(strain_per_step=eps)
    res = {}
    for form, prec in (("fp64", native.PA_PREC_FP64), ("mixed", native.PA_PREC_MIXED)):
        b = pa.Batch(mats, rve_mat, rve_nel, pb)
        for r in range(b.n_rve):
            b.set_grains(r, egs[r], rods[r])
        b.set_solver(native.PA_SOLVER_BICG, prec)
        tune = b.tune()
        steps, asm, kry, post, its, tot = 0, [], [], [], [], 0.0
        rows = None
        for inc in range(pb.total_increments):
            for it in range(pb.max_newton):
                t0 = time.perf_counter()
                rn, k = b.newton_iteration(inc, it)
                tot += time.perf_counter() - t0
                ms = b.last_timings()
                asm.append(ms[0]); kry.append(ms[1]); its.append(float(np.mean(k)))
                steps += 1
            t0 = time.perf_counter()
            rows = b.commit_increment()
            tot += time.perf_counter() - t0
            post.append(b.last_timings()[2])
        qp_ms = b.qp_pass_ms(1)
        nq = b.nqp
        b.close()
        ms_step = 1000.0 * tot / steps
        res[form] = {"ms_per_step": ms_step, "qp_updates_per_s": nq / (ms_step * 1e-3), "assembly_ms_per_step": float(np.mean(asm)),
                     "krylov_ms_per_step": float(np.mean(kry)), "post_ms_per_increment": float(np.mean(post)),
                     "bicg_iters_per_step_mean_over_rves": float(np.mean(its)), "qp_pass_ms_last_state": qp_ms, "tuner": tune,
                     "rows": rows}
    a, m = res["fp64"], res["mixed"]
    diff = float(np.abs(m["rows"][:, :12] - a["rows"][:, :12]).max() / np.abs(a["rows"][:, :12]).max())
    for r in res.values():
        del r["rows"]
    return {"workload": "64 RVEs, edge U{8..24} (seed 2024), %d qps, %d nodes; BCC 48 systems / HCP 24 systems alternating; "
                        "16 grains each; 10 x %.1e tension; Jacobi-BiCG 1e-10" % (nq, sum(int(np.prod(np.asarray(n) + 1)) for n in rve_nel), eps),
            "n_rve": int(len(rve_mat)), "qps": nq, "newton_steps": 4 * pb.total_increments, "fp64": a, "mixed": m,
            "speedup_mixed_vs_fp64": a["ms_per_step"] / m["ms_per_step"], "mixed_vs_fp64_final_rows_rel_diff": diff,
            "note": "ms_per_step includes the post-increment passes (spread over the 4 steps of their increment)"}


def extensions_leg(pa, native, synth, results, n=64, grains=256, warmup=4, steps=8):
    """The 64^3 FCC tension workload under periodic boundary conditions (the reference's fcc/periodicBCs
    constraint lines and uniaxial-tension vertex rows, 5e-4 per step on the corner V100) and with
    three phases (FCC; HCP with 24 systems; FCC with kinematic hardening: grain g in phase g % 3 + 1),
    mixed form, Jacobi-PCG (multigrid does not take periodic lines) and multigrid-PCG for the phases:
    ms per Newton step over `steps` steps after `warmup`, against the same form and solver with face
    rules (the main legs)."""
    from plasticity_amd import model
    out = {}

    def timed(ctx, nmax=4):
        rows, its = [], []
        for s in range(warmup + steps):
            if s == warmup:
                t0 = time.perf_counter()
            inc, it = divmod(s, nmax)
            _, k = ctx.newton_iteration(inc, it)
            its.append(k)
            if it == nmax - 1:
                rows.append(ctx.commit_increment())
        dt = time.perf_counter() - t0
        return 1000.0 * dt / steps, float(np.mean(its[warmup:])), rows

    # periodic: the constraint lines of the reference's input, the vertex rows of its uniaxial tension
    d = os.path.join(ROOT, "tests", "golden", "fcc_periodicBCs")
    mat, pb, vox, rod, _, _ = synth.fcc_workload(n, grains, 20, 5e-4, seed=2024)
    lines = model.read_periodic_constraints(os.path.join(d, "PeriodicBCsConstraints.txt"), 12, 27, 9)
    known = [1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0]
    total = [0.0] * 24
    for r in (3, 12, 15, 21):  # u_x of V100, V110, V101, V111
        total[r] = 5e-4 * pb.total_increments
    ctx = pa.Context(mat, pb)
    ctx.set_grains(ctx.element_grains(vox), rod)
    ctx.set_solver(native.PA_SOLVER_CG, native.PA_PREC_MIXED)
    ctx.set_periodic(model.make_periodic(lines, known, total))
    ms, its, rows = timed(ctx)
    ctx.close()
    ref = results.get(("mixed", "cg"))
    out["periodic"] = {"workload": "%d^3 FCC, %d grains, periodic uniaxial tension 5e-4/step (fcc/periodicBCs lines), mixed, "
                                   "Jacobi-PCG 1e-10" % (n, grains),
                       "ms_per_step": ms, "qp_updates_per_s": 8 * n ** 3 / (ms * 1e-3), "krylov_iters_per_step": its,
                       "face_rule_ms_per_step": ref["ms_per_step"] if ref else None,
                       "last_row_Txx_Tyy_Tzz": [float(v) for v in rows[-1][6:9]]}
    # three phases
    t = synth.fcc_tables()
    h = synth.hcp24_tables()
    hcp = pa.make_material(h["m"], h["n"], h["q"], h["C"], h["s0"], h["h0"], h["ss"], h["pw"], h["tol"])
    kin = pa.make_material(t["m"], t["n"], t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], t["tol"],
                           C1=[10.0] * 12, C2=[1.0] * 12)
    ctx = pa.Context([mat, hcp, kin], pb)
    ctx.set_grain_phases(np.arange(len(rod)) % 3 + 1)
    ctx.set_grains(ctx.element_grains(vox), rod)
    ctx.set_solver(native.PA_SOLVER_MGCG, native.PA_PREC_MIXED)
    ms, its, rows = timed(ctx)
    qp_ms = ctx.last_qp_pass_ms()
    ctx.close()
    ref = results.get(("mixed", "mgcg"))
    out["phases"] = {"workload": "%d^3, %d grains in 3 phases (FCC 12, HCP 24, FCC 12 with kinematic hardening), tension "
                                 "5e-4/step, mixed, multigrid-PCG 1e-10" % (n, grains),
                     "ms_per_step": ms, "qp_updates_per_s": 8 * n ** 3 / (ms * 1e-3), "krylov_iters_per_step": its,
                     "qp_main_pass_ms_last": qp_ms, "one_phase_ms_per_step": ref["ms_per_step"] if ref else None}
    return out


def warm_start_leg(pa, native):
    """The learned warm start of the converged local loop (Max Solver Iterations 50): the network is
    trained on the GPU on the trial states of every Newton iteration of an 8^3 tension run (8
    grains, 5 increments of 1e-3) and judged on the trial states of a 16^3 run with other grains:
    mean local solves with and without it, and the largest stress difference between the two."""
    from plasticity_amd import synth, warmstart as ws

    def run(nel, grains, incs, seed):
        mat, pb, vox, rod, _, _ = synth.fcc_workload(nel, grains, 10, 1e-3, seed=seed)
        mat.max_solver_iter = 50
        ctx = pa.Context(mat, pb)
        eg = ctx.element_grains(vox)
        ctx.set_grains(eg, rod)
        st = ws.collect_trial_states(ctx, (nel, nel, nel), (1.0, 1.0, 1.0), eg, incs, pb.max_newton)
        return ctx, st

    t0 = time.perf_counter()
    ctx, st = run(8, 8, 5, 7)
    plain = ws.qp_update_warm(ctx, st["F"], st["Fp"], st["Fe"], st["s"], st["grain"], None)
    ctx.close()
    act = plain["iters"] > 0
    X, Y = plain["feat"][act], plain["slip_inc"][act]
    net = ws.WarmStartNet(ws.init_blob(X, Y, seed=0))
    epochs = 60
    t1 = time.perf_counter()
    losses = net.train(X, Y, epochs=epochs, batch=256, lr=2e-3, seed=1)
    train_s = time.perf_counter() - t1
    ctx, st = run(16, 16, 3, 8)
    p = ws.qp_update_warm(ctx, st["F"], st["Fp"], st["Fe"], st["s"], st["grain"], None)
    w = ws.qp_update_warm(ctx, st["F"], st["Fp"], st["Fe"], st["s"], st["grain"], net)
    ctx.close()
    a = p["iters"] > 0
    pred, Yt = np.maximum(w["x0"][a], 0), p["slip_inc"][a]
    err = np.linalg.norm(pred - Yt, axis=1) / np.maximum(np.linalg.norm(Yt, axis=1), 1e-12)
    dT = np.abs(w["T"] - p["T"]).max(axis=1) / np.abs(p["T"]).max(axis=1)
    net.close()
    return {"train_pairs": int(len(X)), "epochs": epochs, "train_s": train_s,
            "train_pairs_per_s": len(X) * epochs / train_s, "loss_first_last_epoch": [float(losses[0]), float(losses[-1])],
            "heldout_plastic_qps": int(a.sum()), "local_solves_plain": float(p["iters"][a].mean()),
            "local_solves_warm": float(w["iters"][a].mean()),
            "qps_with_fewer_solves": float((w["iters"][a] < p["iters"][a]).mean()),
            "prediction_rel_err_median": float(np.median(err)), "max_stress_diff_warm_vs_plain": float(dT.max()),
            "leg_s": time.perf_counter() - t0}


def run_form(args, form, solver, pa, native, synth, torch):
    n = args.n
    mat, pb, vox, rod, bcs, _ = synth.fcc_workload(n, args.grains, 20, 5e-4, seed=2024)
    ctx = pa.Context(mat, pb)
    eg = ctx.element_grains(vox)
    ctx.set_grains(eg, rod)
    ctx.set_solver({"bicg": native.PA_SOLVER_BICG, "cg": native.PA_SOLVER_CG, "mgcg": native.PA_SOLVER_MGCG}[solver],
                   native.PA_PREC_MIXED if form == "mixed" else native.PA_PREC_FP64)
    nmax = pb.max_newton
    state = {"step": 0}
    log = []

    def newton_step(check=True):
        """One Newton iteration; check: also the true relative residual ||b - K du|| / ||b|| of its
        Krylov solve (an extra operator product, so never inside the timed region)."""
        s = state["step"]
        inc, it = divmod(s, nmax)
        rn, its = ctx.newton_iteration(inc, it)
        asm_ms, kry_ms, _ = ctx.last_timings()
        rec = {"inc": inc, "it": it, "res": rn, "lin_iters": its, "asm_ms": asm_ms, "krylov_ms": kry_ms,
               "deferred": int(ctx.last_counts()[0]), "qp_pass_ms": ctx.last_qp_pass_ms()}
        if check:
            rec["true_rel_res"] = float(ctx.last_residual()[0])
        if it == nmax - 1:
            rec["row"] = ctx.commit_increment().tolist()
            rec["post_ms"] = ctx.last_timings()[2]
            rec["post_deferred"] = int(ctx.last_counts()[1])
        state["step"] += 1
        log.append(rec)
        print("[%s %s] step %d (increment %d, iteration %d): |R| %.3e, %d Krylov iterations, assembly %.1f ms, Krylov %.1f ms, "
              "%d singular local solves" % (form, solver, s, inc, it, rn, its, asm_ms, kry_ms, rec["deferred"]), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        newton_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        newton_step(check=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timed = log[args.warmup:]
    ms_per_step = 1000.0 * dt / args.steps
    vb = 4 if form == "mixed" else 8
    # kernel timings at a plastic state: one more Newton iteration (the next increment's first),
    # then time the kernels on that state (after a commit every qp sits on its yield surface)
    newton_step()
    newton_step()
    avg = ctypes.c_double()
    # the Krylov operator kernel: BiCG's pair (K p and K^T p~: two reads of the matrix values,
    # p, p~ read, z, z~ written) or CG's product (p read, q written); algorithmic bytes = matrix
    # values read once per product + FP64 vectors read or written once
    which, nmat, nvec = (0, 2, 4) if solver == "bicg" else (2, 1, 2)
    native.check(native.lib().pa_time_kernel(ctx._h, which, 20, ctypes.byref(avg)))
    spmv_ms = avg.value
    bytes_spmv = nmat * ctx.nnz * vb + nvec * ctx.ndof * 8
    achieved = bytes_spmv / (spmv_ms * 1e-3) / 1e9
    upd_us = p_us = None
    if solver == "cg":  # update kernel: x, p, r, q, D^-1 read; x, r, z written.  p kernel: z, p read, p written
        native.check(native.lib().pa_time_kernel(ctx._h, 5, 20, ctypes.byref(avg)))
        upd_us = avg.value * 1e3
        native.check(native.lib().pa_time_kernel(ctx._h, 6, 20, ctypes.byref(avg)))
        p_us = avg.value * 1e3
    native.check(native.lib().pa_time_kernel(ctx._h, 3, 1, ctypes.byref(avg)))
    qp_ms = avg.value
    native.check(native.lib().pa_time_kernel(ctx._h, 4, 1, ctypes.byref(avg)))
    el_ms = avg.value
    # the first qp pass alone (the dominant kernel): return map + dP/dF of every qp
    native.check(native.lib().pa_time_kernel(ctx._h, 7, 3, ctypes.byref(avg)))
    qp_main_ms = avg.value
    full = None
    if args.full_path and solver == "mgcg":
        # the rest of the load path (all increments), untimed for `value`: phase times from the
        # library's HIP events, plus the true residual of every solve
        while state["step"] < pb.total_increments * nmax:
            newton_step()
        steps_all = [r for r in log]
        full = {"steps": len(steps_all),
                "ms_per_step_events": float(np.mean([r["asm_ms"] + r["krylov_ms"] for r in steps_all]) +
                                      np.sum([r.get("post_ms", 0.0) for r in steps_all]) / len(steps_all)),
                "assembly_ms_per_step": float(np.mean([r["asm_ms"] for r in steps_all])),
                "krylov_ms_per_step": float(np.mean([r["krylov_ms"] for r in steps_all])),
                "krylov_iters_per_step": float(np.mean([r["lin_iters"] for r in steps_all])),
                "first_qp_pass_ms": float(np.mean([r["qp_pass_ms"] for r in steps_all]))}
    checked = [r["true_rel_res"] for r in log if "true_rel_res" in r]
    posts = [r["post_ms"] for r in timed if "post_ms" in r]
    kib = PMC_TRAFFIC_KIB.get(("cg" if solver == "mgcg" else solver, form, n)) if args.grains == 256 else None
    traffic = kib * 1024 if kib else None
    issued = PMC_QP_FLOPS.get((form, n)) if args.grains == 256 else None
    qp_kib = PMC_QP_TRAFFIC_KIB.get((form, n)) if args.grains == 256 else None
    # the first qp pass over the timed steps: HIP-event durations of each launch, and the flop
    # fixture's algorithmic flops per qp for the same Newton steps of the load path
    live_ms = float(np.mean([r["qp_pass_ms"] for r in timed]))
    fpq, fx_workload = first_pass_flops_per_qp(range(args.warmup, args.warmup + args.steps))
    qp_flops = fpq * ctx.nqp if fpq else None
    qp_tflops = qp_flops / (live_ms * 1e-3) / 1e12 if qp_flops else None
    res = {
        "nq": ctx.nqp, "ndof": ctx.ndof, "nnz": ctx.nnz,
        "ms_per_step": ms_per_step, "qp_updates_per_s": ctx.nqp / (ms_per_step * 1e-3),
        "krylov_iters_per_step": float(np.mean([r["lin_iters"] for r in timed])),
        "assembly_ms_per_step": float(np.mean([r["asm_ms"] for r in timed])),
        "krylov_ms_per_step": float(np.mean([r["krylov_ms"] for r in timed])),
        "post_ms_per_increment": float(np.mean(posts)) if posts else None,
        "singular_local_solves_per_step": float(np.mean([r["deferred"] for r in timed])),
        "qp_tangent_ms": qp_ms, "element_ms": el_ms, "operator_us": spmv_ms * 1e3, "update_us": upd_us, "p_us": p_us,
        "update_GBps": (8 * ctx.ndof * 8 / (upd_us * 1e-6) / 1e9) if upd_us else None,
        "qp_main_pass_ms": qp_main_ms,
        "rows": [r["row"] for r in log if "row" in r],
        "roofline": {"kernel": "k_qp_tangent<%d, %s> (first qp pass: return map + dP/dF of all %d qps, singular ones queued)"
                               % (12, "float" if form == "mixed" else "double", ctx.nqp),
                     "bound": "fp64-valu", "achieved": qp_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": qp_tflops / FP64_PEAK_TFLOPS if qp_tflops else None,
                     "traffic": qp_kib * 1024 if qp_kib else None, "avg_launch_ms": live_ms,
                     "avg_launch_source": "HIP events around each first-pass launch on the context stream, mean over the "
                                          "%d timed steps" % args.steps,
                     "flops_per_launch": qp_flops, "flops_per_qp": fpq,
                     "flops_source": ("algorithmic FP64 flops per qp of the same Newton steps, counted by the flop-counting "
                                      "oracle build on the %s (tests/golden/qp_flops_fcc24.json), x %d qps" % (fx_workload, ctx.nqp))
                                     if fpq else "flop fixture absent",
                     "issued_flops_per_launch": issued,
                     "issued_source": "rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 x 64 lanes, FMA = 2, masked lanes "
                                      "count (profiles/k_qp_tangent_mixed_64cube_fcc.md)" if issued else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE (profiles/k_qp_tangent_mixed_64cube_fcc.md)"
                                       if qp_kib else None},
        "roofline_krylov": {"kernel": "%s<%s> (27-point node-block %s)" % ("k_bicg_spmv" if solver == "bicg" else "k_cg_spmv",
                                                                      "float" if form == "mixed" else "double",
                                                                      "K p and K^T p~" if solver == "bicg" else "K p"),
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_ms": spmv_ms,
                     "algorithmic_bytes": bytes_spmv,
                     "traffic_source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE (profiles/k_cg_spmv_mixed_64cube_fcc.md)"
                                       if traffic else "not measured for this kernel/size"},
    }
    res["true_rel_residual"] = {"max": float(np.max(checked)) if checked else None,
                                "median": float(np.median(checked)) if checked else None,
                                "solves_checked": len(checked),
                                "note": "||b - K du|| / ||b|| after each Krylov solve outside the timed steps (the stopping "
                                        "test is ||D^-1 r|| <= 1e-10 ||b||, on the Jacobi-preconditioned residual)"}
    res["full_load_path"] = full
    res["rerun_lane_solves"] = int(ctx.last_counts()[2])
    res["T"] = ctx.get_state("T")
    res["s"] = ctx.get_state("s")
    res["u"] = ctx.get_state("u")
    ctx.close()
    return res


if __name__ == "__main__":
    main()

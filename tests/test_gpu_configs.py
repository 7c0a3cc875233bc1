"""The benchmark workloads at full size on the GPU, through the C ABI.

The workloads (uniaxial tension of a unit cube, FCC, seeded Voronoi grains; plasticity_amd.synth):
  C1  8^3,  8 grains, 10 steps of 1e-3, FP64, Jacobi-BiCG 1e-10 (the reference's solver);
  C2  32^3, 64 grains, 20 steps of 5e-4, FP64, Jacobi-PCG 1e-10;
  C3  64^3, 256 grains, 20 steps of 5e-4, mixed form (FP32 tangents and matrix), multigrid-PCG 1e-10.

Their plastic steps are large for the reference's fixed 4 Newton iterations: |R| stalls after the
first correction (190 -> ~3 on C1), so the state at the end of a step depends on round-off in the
solves and the oracle differs from itself by 1e-3..2e-2 when only its Krylov tolerance changes.  Those
workloads are therefore compared with the oracle in units of the oracle's own sensitivity (the
same run at the workload's tolerance and at a 3x tighter one), and checked at every increment for
what the method guarantees regardless of trajectory:
  * det Fp = 1 (Fp = exp(traceless) Fp_t, calculatePlasticity.cc:359);
  * slip resistances never decrease and never pass the saturation stress (:401-429);
  * accumulated slips never decrease (x_beta_old >= 0 after inactive-slip removal);
  * the Krylov solve met its stopping test ||D^-1 (B - K X)|| <= tol ||B|| (recomputed);
  * exactly 4 Newton iterations per step, all rows finite.
Where the reference's Newton converges (the fcc/simpleTension inputs through increment 4, |R| ->
5e-12 within 6 iterations) the mixed form is held to 1e-8 of the FP64 oracle, per quantity.
"""
import ctypes
import os

import numpy as np
import pytest

import plasticity_amd as pa
from plasticity_amd import native, synth
import refcase

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SS = 148.0  # saturation stress of the FCC workload (MPa)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _workload(n, grains, steps, eps, lin_tol, solver, seed=2024):
    _, _, vox, rod, bcs, t = synth.fcc_workload(n, grains, steps, eps, seed=seed)
    case = refcase.Case(t["m"] / np.linalg.norm(t["m"], axis=1)[:, None], t["n"] / np.linalg.norm(t["n"], axis=1)[:, None],
                        t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=n, ny=n, nz=n, delT=1.0,
                        total_time=float(steps), bcs=bcs, lin_tol=lin_tol, solver=solver)
    case.elem_grain = vox.transpose(2, 1, 0).ravel().astype(np.uint32)
    case.grain_rod = rod
    return case


def _ctx(case, method, precision):
    mat = pa.make_material(case.m, case.n, case.q, case.C, case.s0, case.h0, case.ss, case.pw, case.tol,
                           case.max_slip_search, case.max_solver_iter)
    pb = pa.make_problem((case.nx, case.ny, case.nz), case.span, case.total_increments, case.delT, case.bcs,
                         case.max_newton, case.lin_tol, case.lin_maxit)
    ctx = pa.Context(mat, pb)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    ctx.set_solver(method, precision)
    return ctx


def _run(ctx, case, n_inc, check_every=1):
    """n_inc increments, Newton iteration by Newton iteration; checks the stopping test of every
    Krylov solve and the history properties every `check_every` increments."""
    rows, prev = [], None
    for inc in range(n_inc):
        for it in range(case.max_newton):
            rn, n = ctx.newton_iteration(inc, it)
            assert np.isfinite(rn) and n > 0
            true_rel, stop_rel, solver_rel = ctx.last_residual()
            # the solver stopped on its own recurrence value <= tol; the recomputed value agrees
            assert solver_rel <= case.lin_tol * 1.0000001, (inc, it, solver_rel)
            assert stop_rel <= 10 * case.lin_tol, (inc, it, stop_rel, solver_rel)
        rows.append(ctx.commit_increment())
        assert np.all(np.isfinite(rows[-1]))
        if (inc + 1) % check_every == 0 or inc == n_inc - 1:
            prev = _history_properties(ctx, prev, inc)
    return np.array(rows)


def _history_properties(ctx, prev, inc):
    Fp = ctx.get_state("Fp").reshape(-1, 3, 3)
    s, slip = ctx.get_state("s"), ctx.get_state("slip")
    det = np.linalg.det(Fp)
    assert np.abs(det - 1.0).max() < 1e-8, (inc, np.abs(det - 1.0).max())
    assert s.max() <= SS * (1 + 1e-12), (inc, s.max())
    if prev is not None:
        assert (s >= prev["s"] - 1e-12 * SS).all(), inc
        assert (slip >= prev["slip"] - 1e-15).all(), inc
    print("increment %d: max |det Fp - 1| %.1e, s in [%.3f, %.3f], plastic qps %d" % (inc, np.abs(det - 1).max(), s.min(), s.max(),
                                                                                     int((slip.sum(axis=1) > 0).sum())), flush=True)
    return {"s": s, "slip": slip}


def _envelope_gate(rows_gpu, rows_ref, rows_ref2, factor=5.0):
    """Per-increment row errors of the GPU against the oracle, against the oracle's own sensitivity."""
    scale = np.abs(rows_ref[:, :12]).max(axis=1)
    e_gpu = np.abs(rows_gpu[:, :12] - rows_ref[:, :12]).max(axis=1) / scale
    e_self = np.abs(rows_ref2[:, :12] - rows_ref[:, :12]).max(axis=1) / scale
    print("row error vs oracle per increment", " ".join("%.1e" % v for v in e_gpu))
    print("oracle self-sensitivity         ", " ".join("%.1e" % v for v in e_self))
    assert e_gpu.max() <= factor * e_self.max(), (e_gpu.max(), e_self.max())
    assert np.median(e_gpu) <= factor * np.median(e_self), (np.median(e_gpu), np.median(e_self))
    return e_gpu, e_self


def test_c1_workload_tracks_the_oracle_within_its_own_sensitivity():
    case = _workload(8, 8, 10, 1e-3, 1e-10, 0)
    ref = refcase.oracle_simulate(case, 10, nthreads=8)
    case2 = _workload(8, 8, 10, 1e-3, 3.3e-11, 0)
    ref2 = refcase.oracle_simulate(case2, 10, nthreads=8)
    ctx = _ctx(case, native.PA_SOLVER_BICG, native.PA_PREC_FP64)
    rows = _run(ctx, case, 10)
    ctx.close()
    _envelope_gate(rows, ref["rows"], ref2["rows"])
    assert rows[-1, 14] > 0.02  # deeply plastic by the end (SlipTotal)
    # SlipTotal, the quantity most sensitive to active-set choices, to 5x the oracle's own spread
    d_gpu = np.abs(rows[:, 14] - ref["rows"][:, 14]).max()
    d_self = np.abs(ref2["rows"][:, 14] - ref["rows"][:, 14]).max()
    assert d_gpu <= 5 * d_self + 1e-12, (d_gpu, d_self)


def test_c2_workload_full_size():
    fx = np.load(os.path.join(HERE, "golden", "c2_fcc32_oracle.npz"))
    case = _workload(32, 64, 20, 5e-4, 1e-10, 1)
    ctx = _ctx(case, native.PA_SOLVER_CG, native.PA_PREC_FP64)
    rows = []
    for inc in range(2):
        rows.append(_run_one(ctx, case, inc))
    rows = np.array(rows)
    _envelope_gate(rows, fx["rows"][0], fx["rows"][1])
    # per-qp distributions after increment 1 against the oracle's (mean and 1/50/99th percentiles)
    T, s = ctx.get_state("T"), ctx.get_state("s")
    for name, a in (("T", T), ("s", s)):
        st = fx["stat_%s" % name]
        mine = np.stack([a.mean(0), a.std(0), np.percentile(a, 1, axis=0), np.percentile(a, 50, axis=0), np.percentile(a, 99, axis=0)])
        scale = np.abs(st[3:]).max()
        err = np.abs(mine - st).max() / scale
        print("%s distribution error %.2e" % (name, err))
        assert err < 2e-2, (name, err)
    prev = _history_properties(ctx, None, 1)
    for inc in range(2, 20):
        _run_one(ctx, case, inc)
        if inc % 6 == 5 or inc == 19:
            prev = _history_properties(ctx, prev, inc)
    ctx.close()


def _run_one(ctx, case, inc):
    for it in range(case.max_newton):
        rn, n = ctx.newton_iteration(inc, it)
        assert np.isfinite(rn) and n > 0
        _, stop_rel, solver_rel = ctx.last_residual()
        assert solver_rel <= case.lin_tol * 1.0000001 and stop_rel <= 10 * case.lin_tol, (inc, it, stop_rel, solver_rel)
    row = ctx.commit_increment()
    assert np.all(np.isfinite(row))
    return row


def test_c3_workload_full_size_mixed_multigrid():
    case = _workload(64, 256, 20, 5e-4, 1e-10, 1)
    ctx = _ctx(case, native.PA_SOLVER_MGCG, native.PA_PREC_MIXED)
    rows = _run(ctx, case, 20, check_every=10)
    ctx.close()
    assert rows[-1, 14] > 0 and rows[-1, 6] > rows[0, 6]
    assert np.all(np.diff(rows[:, 0]) > 0)  # Exx grows with the imposed stretch


def _converged_case():
    case = refcase.load_reference_case("fcc_simpleTension")
    case.lin_tol = 1e-14
    case.max_newton = 8
    return case


def test_mixed_form_converged_error_below_1e8():
    """fcc/simpleTension inputs, 5 increments with 8 Newton iterations each: the reference's Newton
    converges (|R| ~5e-12; 519 plastic qps in increment 4), so the mixed form (FP32 Jacobian, FP64
    residual) lands on the FP64 oracle's solution.  Per quantity <= 1e-8 relative; the FP64 form
    <= 1e-9."""
    case = _converged_case()
    n = 5
    ref = refcase.oracle_simulate(case, n, nthreads=8, want_state=True)
    assert ref["res_norms"][-1] < 1e-10
    st = ref["state"]
    for prec, bound in ((native.PA_PREC_MIXED, 1e-8), (native.PA_PREC_FP64, 1e-9)):
        ctx = _ctx(case, native.PA_SOLVER_BICG, prec)
        rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
        errs = {
            "avg_strain": _rel(rows[:, 0:6], ref["rows"][:, 0:6]),
            "avg_stress": _rel(rows[:, 6:12], ref["rows"][:, 6:12]),
            "qp_cauchy_stress": _rel(ctx.get_state("T"), st[:, 0:9]),
            "crss": _rel(ctx.get_state("s"), st[:, 27:39]),
            "Fp": _rel(ctx.get_state("Fp"), st[:, 9:18]),
            "displacement": _rel(ctx.get_state("u"), ref["u"]),
        }
        slip = ctx.get_state("slip")
        assert int((slip.sum(axis=1) > 0).sum()) > 100  # plastic
        print("precision", prec, {k: "%.1e" % v for k, v in errs.items()})
        for k, v in errs.items():
            assert v < bound, (prec, k, v)
        ctx.close()


@pytest.mark.parametrize("precision,bound,rbound", [(native.PA_PREC_FP64, 1e-9, 1e-9), (native.PA_PREC_MIXED, 1e-6, 1e-4)])
def test_plastic_jacobian_stagewise(precision, bound, rbound):
    """Condensed Jacobian and residual of increment 4, Newton iteration 1 (plastic qps present),
    after converged increments 0-3, against the oracle's (or_newton_system).  The mixed form's
    entries carry FP32 rounding (1e-6); its residual is the FP64 one at its own iterate, which differs
    from the FP64 iterate by the FP32 Jacobian's effect on the first correction (~1e-7 of du), seen
    ~70x larger against the iteration-1 residual (|R| 10 -> 0.14), hence 1e-4."""
    case = _converged_case()
    lib = refcase.oracle()
    pb = case.problem()
    nnz = lib.or_sparsity(ctypes.byref(pb), None, None)
    ndof = 3 * 9 ** 3
    K, R, du = np.zeros(nnz), np.zeros(ndof), np.zeros(ndof)
    mat = case.material(1)
    assert lib.or_newton_system(ctypes.byref(mat), ctypes.byref(pb), refcase.ptr(case.elem_grain), refcase.ptr(case.grain_rod),
                                ctypes.c_int(len(case.grain_rod)), ctypes.c_int(4), ctypes.c_int(1), refcase.ptr(K), refcase.ptr(R),
                                refcase.ptr(du)) == 0
    ctx = _ctx(case, native.PA_SOLVER_BICG, precision)
    for inc in range(4):
        ctx.run_increment(inc)
    ctx.newton_iteration(4, 0)
    ctx.assemble(4, 1)
    hardened = int((ctx.get_state("s_iter") > ctx.get_state("s") + 1e-12).any(axis=1).sum())
    assert hardened > 50  # qps on the plastic branch in this assembly
    _, _, gK = ctx.get_csr(values=True)
    gR = ctx.get_state("R")
    eK, eR = _rel(gK, K), _rel(gR, R)
    print("plastic Jacobian error %.1e, residual %.1e" % (eK, eR))
    assert eK < bound and eR < rbound, (eK, eR)
    ctx.close()

"""GPU parity of the HIP path against the oracle (through the C ABI).

Tolerances: per-qp stress, tangent, CRSS and Fp/Fe 1e-10 relative (to the largest entry of each
quantity); load-displacement rows 1e-9 relative; integer work (grain ids, CSR pattern) bit-exact.
"""
import numpy as np
import pytest

import plasticity_amd as pa
import refcase

pytestmark = pytest.mark.gpu


def _ctx(case, nel=None, n_grain=None):
    mat = pa.make_material(case.m, case.n, case.q, case.C, case.s0, case.h0, case.ss, case.pw, case.tol,
                           case.max_slip_search, case.max_solver_iter)
    pb = pa.make_problem(nel or (case.nx, case.ny, case.nz), case.span, case.total_increments, case.delT, case.bcs,
                         case.max_newton, case.lin_tol, case.lin_maxit)
    return pa.Context(mat, pb)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _random_states(case, n, rng, strain):
    F = np.eye(3) + rng.normal(scale=strain, size=(n, 3, 3))
    Fp = np.eye(3) + rng.normal(scale=strain / 5, size=(n, 3, 3))
    Fp /= np.cbrt(np.linalg.det(Fp))[:, None, None]
    s = rng.uniform(16.0, 60.0, size=(n, case.n_slip))
    return F, Fp, s


@pytest.mark.parametrize("strain", [2e-4, 2e-3, 1e-2])
def test_qp_update_matches_oracle(strain):
    case = refcase.fcc_case()
    rng = np.random.default_rng(11)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    n = 256
    F, Fp, s = _random_states(case, n, rng, strain)
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    ctx = _ctx(case, nel=(1, 1, 1))
    ctx.set_grains(np.zeros(1, np.uint32), rod)
    for flag in (0, 1):
        g = ctx.qp_update(F.reshape(n, 9), Fp.reshape(n, 9), s, grain, flag)
        nplastic = 0
        for i in range(n):
            o = refcase.oracle_qp(case, rod[grain[i]], F[i], Fp[i], np.append(s[i], 10e14), flag)
            assert g["n_active"][i] == o["n_active"], (i, g["n_active"][i], o["n_active"])
            nplastic += o["n_active"] > 0
            for k_g, k_o in (("P", "P"), ("T", "T"), ("Fe", "Fe"), ("Fp", "Fp")):
                assert _rel(g[k_g][i], o[k_o].ravel()) < 1e-10, (k_g, i, _rel(g[k_g][i], o[k_o].ravel()))
            assert _rel(g["s"][i], o["s"][:12]) < 1e-10
            if flag:
                assert _rel(g["dPdF"][i], o["dPdF"].ravel()) < 1e-10, (i, _rel(g["dPdF"][i], o["dPdF"].ravel()))
        if strain >= 2e-3:
            assert nplastic > n // 4  # the plastic branch is exercised
    ctx.close()


def test_element_grains_bit_exact():
    case = refcase.load_reference_case("fcc_simpleTension")
    ctx = _ctx(case)
    eg = ctx.element_grains(case.vox)
    assert np.array_equal(eg, case.elem_grain)
    ctx.close()
    case = refcase.load_reference_case("bcc_simpleTension")
    ctx = _ctx(case)
    assert np.array_equal(ctx.element_grains(case.vox), case.elem_grain)
    ctx.close()


@pytest.mark.parametrize("nel", [(1, 1, 1), (2, 3, 4), (8, 8, 8), (5, 8, 10)])
def test_csr_pattern_bit_exact(nel):
    import ctypes
    case = refcase.fcc_case(nx=nel[0], ny=nel[1], nz=nel[2])
    pb = case.problem()
    lib = refcase.oracle()
    nnz = lib.or_sparsity(ctypes.byref(pb), None, None)
    rp = np.zeros((nel[0] + 1) * (nel[1] + 1) * (nel[2] + 1) * 3 + 1, np.int64)
    ci = np.zeros(nnz, np.int32)
    lib.or_sparsity(ctypes.byref(pb), refcase.ptr(rp), refcase.ptr(ci))
    ctx = _ctx(case)
    assert ctx.nnz == nnz
    grp, gci = ctx.get_csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    ctx.close()


@pytest.mark.parametrize("lin_tol,row_tol,state_tol", [(1e-14, 1e-9, 1e-9), (1e-10, 1e-6, 1e-6)])
def test_newton_curve_matches_oracle_fcc_tension(lin_tol, row_tol, state_tol):
    """8^3 mesh with the fcc/simpleTension inputs: 8 increments through yield, 4 Newton iterations each.
    Yield (increment 5) brings hundreds of local solves whose A_PA has one singular value near
    1e-12 sigma_max, which both sides invert (compute_inverse_svd(0.0) semantics).

    At the reference's Krylov tolerance (1e-10 on the Jacobi-preconditioned residual,
    solveLinearSystem.cc:13) each Newton correction is only fixed to ~1e-4 and the BiCG stopping
    iteration depends on how the last residuals round, so two FP64 implementations agree to the
    tolerance-limited level.  With the linear solves driven to 1e-14 the comparison is well posed
    and the curve, the state and the Krylov iteration counts must match."""
    case = refcase.load_reference_case("fcc_simpleTension")
    case.lin_tol = lin_tol
    n = 8
    ref = refcase.oracle_simulate(case, n, nthreads=8, want_state=True)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows, its = [], []
    for inc in range(n):
        r, it, res = ctx.run_increment(inc)
        rows.append(r)
        its.append(it)
    rows = np.array(rows)
    its = np.array(its).ravel()
    print("lin iters gpu", its, "oracle", ref["lin_iters"])
    assert len(its) == len(ref["lin_iters"]) == 4 * n  # fixed Newton count (solveNonLinearSystem.cc:14)
    if lin_tol < 1e-12:
        # before yield (increments 0-4) the first two solves of each increment carry O(1)
        # right-hand sides and must take the same number of iterations; late Newton iterations
        # solve for a residual that is mostly round-off of cancelling element sums, and plastic
        # increments inherit 1e-11 state differences, so their counts only need to be close
        first = (np.arange(len(its)) % 4 < 2) & (np.arange(len(its)) < 20)
        assert np.abs(its - ref["lin_iters"])[first].max() <= 1
        assert np.abs(its - ref["lin_iters"]).max() <= 0.1 * ref["lin_iters"].max()
    # SlipTotal right after yield sums slips of systems that sit at the activation threshold; the
    # oracle itself moves it by 2e-6 (relative) between Krylov tolerances 1e-14 and 3e-15
    for cols, tol in ((slice(0, 6), row_tol), (slice(6, 12), row_tol), (slice(14, 15), max(row_tol, 1e-5))):
        scale = np.abs(ref["rows"][:, cols]).max()
        err = np.abs(rows[:, cols] - ref["rows"][:, cols]).max() / scale
        assert err < tol, (cols, err)
    st = ref["state"]
    nT = case.nT
    assert _rel(ctx.get_state("T"), st[:, 0:9]) < state_tol
    assert _rel(ctx.get_state("s"), st[:, 27:27 + 12]) < state_tol
    assert _rel(ctx.get_state("Fp"), st[:, 9:18]) < state_tol
    assert _rel(ctx.get_state("rotnew"), st[:, 27 + nT:30 + nT]) < state_tol
    assert _rel(ctx.get_state("u"), ref["u"]) < state_tol
    ctx.close()


@pytest.mark.parametrize("folder,n", [("fcc_simpleTension", 18), ("fcc_compression", 13), ("bcc_simpleTension", 30),
                                      ("fcc_FCC_RandomOrientationBlock", 12)])
def test_reference_golden_curve(folder, n):
    """The reference's own inputs (mesh, grain map, orientations, slip systems, BCs, 1e-10 BiCG) run
    through the GPU path reproduce its golden stress-strain rows (tests/compile_tests/*/
    correct_stressstrain.txt, 6 significant digits) to print precision, over the increments before
    rank-deficient active sets make the reference's own curve irreproducible."""
    case = refcase.load_reference_case(folder)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
    ctx.close()
    gold = case.golden[:n]
    for cols in (slice(0, 6), slice(6, 12)):
        scale = np.abs(gold[:, cols]).max(axis=1)
        err = (np.abs(rows[:, cols] - gold[:, cols]).max(axis=1) / scale).max()
        assert err < 1.2e-5, (folder, cols, err)
    slip = gold[:, 14] > 1e-7
    if slip.any():
        assert (np.abs(rows[slip, 14] - gold[slip, 14]) / gold[slip, 14]).max() < 5e-5
    assert (rows[~slip, 14] < 1e-6).all()


def test_first_newton_iteration_stagewise():
    """Condensed Jacobian, residual and Newton correction of increment 0 / iteration 0."""
    import ctypes
    case = refcase.fcc_case(nx=4, ny=3, nz=5, delT=1.0, total_time=10.0, lin_tol=1e-14)
    rng = np.random.default_rng(7)
    case.grain_rod = rng.uniform(-0.4, 0.4, size=(4, 3))
    case.elem_grain = rng.integers(0, 4, size=60).astype(np.uint32)
    lib = refcase.oracle()
    pb = case.problem()
    nnz = lib.or_sparsity(ctypes.byref(pb), None, None)
    ndof = 3 * 5 * 4 * 6
    K, R, du = np.zeros(nnz), np.zeros(ndof), np.zeros(ndof)
    mat = case.material(1)
    lib.or_first_newton(ctypes.byref(mat), ctypes.byref(pb), refcase.ptr(case.elem_grain), refcase.ptr(case.grain_rod),
                        ctypes.c_int(4), refcase.ptr(K), refcase.ptr(R), refcase.ptr(du))
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    ctx.assemble(0, 0)
    rp, ci, gK = ctx.get_csr(values=True)
    gR = ctx.get_state("R")
    assert _rel(gK, K) < 1e-12, _rel(gK, K)
    assert _rel(gR, R) < 1e-12, _rel(gR, R)
    ctx.linear_solve(0, 0)
    assert _rel(ctx.get_state("u"), du) < 1e-9
    ctx.close()


def test_mixed_precision_form_error_vs_oracle():
    """Mixed form (FP32 tangents and matrix values, FP64 return map, residual and Krylov
    recurrences): the Jacobian only steers Newton, so after the 4 Newton iterations of each
    increment the curve and the per-qp state stay close to the FP64 oracle."""
    case = refcase.load_reference_case("fcc_simpleTension")
    case.lin_tol = 1e-14
    n = 6
    ref = refcase.oracle_simulate(case, n, nthreads=8, want_state=True)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    ctx.set_solver(pa.native.PA_SOLVER_BICG, pa.native.PA_PREC_MIXED)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
    errs = {}
    for name, cols in (("strain", slice(0, 6)), ("stress", slice(6, 12)), ("slip", slice(14, 15))):
        errs[name] = np.abs(rows[:, cols] - ref["rows"][:, cols]).max() / np.abs(ref["rows"][:, cols]).max()
    errs["qp_stress"] = _rel(ctx.get_state("T"), ref["state"][:, 0:9])
    errs["crss"] = _rel(ctx.get_state("s"), ref["state"][:, 27:39])
    print("mixed-form errors vs FP64 oracle:", errs)
    for k, v in errs.items():
        assert v < 1e-5, (k, v)
    ctx.close()


def test_jacobi_pcg_option_matches_oracle():
    """The Jacobi-PCG solver option (the solver the benchmark workloads name) against the oracle's
    Jacobi-PCG, solves driven to 1e-14: same curve and Krylov counts for the O(1) right-hand sides."""
    case = refcase.load_reference_case("fcc_simpleTension")
    case.lin_tol = 1e-14
    n = 4
    ref = refcase.oracle_simulate(case, n, nthreads=8, want_state=True, solver=1)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    ctx.set_solver(pa.native.PA_SOLVER_CG, pa.native.PA_PREC_FP64)
    rows, its = [], []
    for inc in range(n):
        r, it, _ = ctx.run_increment(inc)
        rows.append(r)
        its.append(it)
    rows, its = np.array(rows), np.array(its).ravel()
    print("cg iters gpu", its, "oracle", ref["lin_iters"])
    first = np.arange(len(its)) % 4 < 2
    assert np.abs(its - ref["lin_iters"])[first].max() <= 1
    for cols in (slice(0, 6), slice(6, 12)):
        assert np.abs(rows[:, cols] - ref["rows"][:, cols]).max() / np.abs(ref["rows"][:, cols]).max() < 1e-9
    assert _rel(ctx.get_state("T"), ref["state"][:, 0:9]) < 1e-9
    ctx.close()

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
                           case.max_slip_search, case.max_solver_iter, case.C1, case.C2)
    pb = pa.make_problem(nel or (case.nx, case.ny, case.nz), case.span, case.total_increments, case.delT, case.bcs,
                         case.max_newton, case.lin_tol, case.lin_maxit, case.cyclic[1:] if case.cyclic[0] else None,
                         case.tabular, case.vel_grad)
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


@pytest.mark.parametrize("strain", [2e-3, 1e-2])
def test_qp_update_converged_local_loop_matches_oracle(strain):
    """Max Solver Iterations 50 (the local loop of calculatePlasticity.cc:322-441 run to the stress
    tolerance): stress, history and the tangent dP/dF of the last pass at 1e-10 of the oracle's."""
    case = refcase.fcc_case(max_solver_iter=50)
    rng = np.random.default_rng(23)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    n = 192
    F, Fp, s = _random_states(case, n, rng, strain)
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    ctx = _ctx(case, nel=(1, 1, 1))
    ctx.set_grains(np.zeros(1, np.uint32), rod)
    g = ctx.qp_update(F.reshape(n, 9), Fp.reshape(n, 9), s, grain, 1)
    nplastic = 0
    for i in range(n):
        o = refcase.oracle_qp(case, rod[grain[i]], F[i], Fp[i], np.append(s[i], 10e14), 1)
        assert g["n_active"][i] == o["n_active"], (i, g["n_active"][i], o["n_active"])
        nplastic += o["n_active"] > 0
        for k in ("P", "T", "Fe", "Fp"):
            assert _rel(g[k][i], o[k].ravel()) < 1e-10, (k, i)
        assert _rel(g["s"][i], o["s"][:12]) < 1e-10
        assert _rel(g["dPdF"][i], o["dPdF"].ravel()) < 1e-10, (i, _rel(g["dPdF"][i], o["dPdF"].ravel()))
    assert nplastic > n // 4
    ctx.close()


def test_cyclic_converged_local_loop_newton_matches_oracle():
    """The learned-warm-start workload's loading (FCC, cyclic +-1 % in steps of 1e-3, Max Solver
    Iterations 50) on 8^3 elements, 8 grains, through the Newton step's qp passes, 3 increments.
    The steps are large for 4 Newton iterations (|R| 775 -> 13), and the run is chaotic: the oracle
    itself moves |R| by 1.6e-5 at the second Newton iteration and the curve by 1e-2 after three
    increments when only its Krylov tolerance changes (1e-14 -> 3e-15).  So the first Newton
    iteration is held to 1e-10 and the rest to 5x the oracle's own spread."""
    from plasticity_amd import synth
    _, _, vox, rod, bcs, t = synth.fcc_workload(8, 8, 40, 1e-3, seed=2024)

    def case_at(lin_tol):
        c = refcase.Case(t["m"] / np.linalg.norm(t["m"], axis=1)[:, None], t["n"] / np.linalg.norm(t["n"], axis=1)[:, None],
                         t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=8, ny=8, nz=8, delT=1.0,
                         total_time=40.0, bcs=bcs, lin_tol=lin_tol, max_solver_iter=50)
        c.elem_grain = vox.transpose(2, 1, 0).ravel().astype(np.uint32)
        c.grain_rod = rod
        c.cyclic = (1, 2, 1, 10.0)
        return c

    case = case_at(1e-14)
    ref = refcase.oracle_simulate(case, 3, nthreads=8)
    ref2 = refcase.oracle_simulate(case_at(3e-15), 3, nthreads=8)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows, res = [], []
    for inc in range(3):
        for it in range(case.max_newton):
            rn, _ = ctx.newton_iteration(inc, it)
            assert np.isfinite(rn)
            res.append(rn)
        rows.append(ctx.commit_increment())
    ctx.close()
    res, rows = np.array(res), np.array(rows)
    r1, r2 = np.asarray(ref["res_norms"]).ravel(), np.asarray(ref2["res_norms"]).ravel()
    print("|R| GPU   ", " ".join("%.6g" % v for v in res))
    print("|R| oracle", " ".join("%.6g" % v for v in r1))
    assert abs(res[0] - r1[0]) <= 1e-10 * r1[0]
    assert np.abs(res - r1).max() / r1.max() <= 5 * np.abs(r2 - r1).max() / r1.max()
    d_gpu = np.abs(rows[:, :12] - ref["rows"][:, :12]).max()
    d_self = np.abs(ref2["rows"][:, :12] - ref["rows"][:, :12]).max()
    assert d_gpu <= 5 * d_self, (d_gpu, d_self)
    assert rows[-1, 14] > 0.02  # plastic (SlipTotal)


@pytest.mark.parametrize("strain", [2e-3, 1e-2])
def test_qp_update_with_kinematic_hardening_matches_oracle(strain):
    """Backstress (C_1 = 200, C_2 = 1 as in fcc/CyclicLoading) with random W_kh_conv of both signs."""
    case = refcase.fcc_case()
    case.C1, case.C2 = np.full(12, 200.0), np.full(12, 1.0)
    rng = np.random.default_rng(5)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    n = 192
    F, Fp, s = _random_states(case, n, rng, strain)
    W = rng.uniform(-20.0, 20.0, size=(n, 12))
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    ctx = _ctx(case, nel=(1, 1, 1))
    ctx.set_grains(np.zeros(1, np.uint32), rod)
    for flag in (0, 1):
        g = ctx.qp_update(F.reshape(n, 9), Fp.reshape(n, 9), s, grain, flag, W=W)
        nplastic = 0
        for i in range(n):
            o = refcase.oracle_qp(case, rod[grain[i]], F[i], Fp[i], np.append(s[i], 10e14), flag, W=W[i])
            assert g["n_active"][i] == o["n_active"], (i, g["n_active"][i], o["n_active"])
            nplastic += o["n_active"] > 0
            for k in ("P", "T", "Fe", "Fp"):
                assert _rel(g[k][i], o[k].ravel()) < 1e-10, (k, i)
            assert _rel(g["s"][i], o["s"][:12]) < 1e-10
            assert _rel(g["W"][i], o["W"][:12]) < 1e-10, (i, _rel(g["W"][i], o["W"][:12]))
            if flag:
                assert _rel(g["dPdF"][i], o["dPdF"].ravel()) < 1e-10, (i, _rel(g["dPdF"][i], o["dPdF"].ravel()))
        assert nplastic > n // 4
    ctx.close()


def _hcp_case():
    """hcp/simpleTension's 18 slip systems (basal, prismatic, pyramidal <a>, pyramidal <c+a>),
    hexagonal elastic constants and hardening, twinning off, loaded 4x faster to reach plasticity
    in 16 increments."""
    case = refcase.load_reference_case("hcp_simpleTension")
    case.bcs = [(5, 1, 0.0), (5, 2, 0.0), (5, 3, 0.0), (6, 3, 0.12)]
    return case


def test_qp_update_hcp_slip_matches_oracle():
    """18 HCP slip systems (the 24-system kernel variant) at 1e-10 against the oracle."""
    case = _hcp_case()
    rng = np.random.default_rng(3)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    n = 192
    F = np.eye(3) + rng.normal(scale=3e-3, size=(n, 3, 3))
    Fp = np.eye(3) + rng.normal(scale=6e-4, size=(n, 3, 3))
    Fp /= np.cbrt(np.linalg.det(Fp))[:, None, None]
    s = np.array([rng.uniform(1.0, 1.5) * case.s0 for _ in range(n)])
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
            for k in ("P", "T", "Fe", "Fp"):
                assert _rel(g[k][i], o[k].ravel()) < 1e-10, (k, i)
            assert _rel(g["s"][i], o["s"][:18]) < 1e-10
            if flag:
                assert _rel(g["dPdF"][i], o["dPdF"].ravel()) < 1e-10, (i, _rel(g["dPdF"][i], o["dPdF"].ravel()))
        assert nplastic > n // 4
    ctx.close()


def _bcc48_case(n=4, incs=10, eps=5e-4, lin_tol=1e-14):
    """BCC with 48 slip systems, the stiffness of the batched RVE workload (231/135/117 GPa) and
    bcc/simpleTension's hardening (s0 200, h0 1500, a 1, s_s 500 MPa), latent ratio 1.4 off-plane;
    n^3 elements, 4 grains, x-tension in `incs` steps of `eps`."""
    m, nrm = _bcc48_systems()
    m = m / np.linalg.norm(m, axis=1)[:, None]
    nrm = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    q = np.full((48, 48), 1.4)
    for a in range(48):
        for b in range(48):
            if np.allclose(nrm[a], nrm[b]):
                q[a, b] = 1.0
    C = np.zeros((6, 6))
    C[:3, :3] = 135e3
    np.fill_diagonal(C[:3, :3], 231e3)
    C[3, 3] = C[4, 4] = C[5, 5] = 117e3
    case = refcase.Case(m, nrm, q, C, [200.0] * 48, [1500.0] * 48, [500.0] * 48, [1.0] * 48, nx=n, ny=n, nz=n,
                        total_time=float(incs), bcs=((1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, eps * incs)), lin_tol=lin_tol)
    rng = np.random.default_rng(8)
    case.grain_rod = rng.uniform(-0.4, 0.4, size=(4, 3))
    case.elem_grain = rng.integers(0, 4, size=n ** 3).astype(np.uint32)
    return case


def test_bcc48_newton_steps_match_oracle():
    """Newton steps with 48 slip systems in the one-lane qp passes, whose active-set matrices live
    in a per-lane workspace of global memory (80 KB per lane, more than the private segment the
    runtime grants): 8 increments through yield on 4^3 elements against the oracle (Krylov solves
    at 1e-14; the oracle moves by <= 5e-11 when that tolerance changes to 3e-15), curve and
    per-qp state at 1e-9.  Two more increments leave the oracle's trajectory: with 48 systems
    A_PA is often rank-deficient with singular values at round-off, where the reference's own
    result depends on rounding (as in the per-qp test below), and the GPU's FMA rounding differs
    from the oracle's build; they are held to the method's invariants."""
    case = _bcc48_case()
    ref = refcase.oracle_simulate(case, 8, nthreads=8, want_state=True)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(8)])
    assert rows[-1, 14] > 0.02  # plastic (SlipTotal)
    assert _rel(rows[:, :12], ref["rows"][:, :12]) < 1e-9, _rel(rows[:, :12], ref["rows"][:, :12])
    assert _rel(rows[:, 14], ref["rows"][:, 14]) < 1e-8
    st = ref["state"]
    assert _rel(ctx.get_state("s"), st[:, 27:27 + 48]) < 1e-9
    assert _rel(ctx.get_state("T"), st[:, 0:9]) < 1e-9
    assert _rel(ctx.get_state("Fp"), st[:, 9:18]) < 1e-9
    s_prev, slip_prev = ctx.get_state("s"), ctx.get_state("slip")
    for inc in range(8, 10):
        row = ctx.run_increment(inc)[0]
        assert np.all(np.isfinite(row))
        Fp, s, slip = ctx.get_state("Fp").reshape(-1, 3, 3), ctx.get_state("s"), ctx.get_state("slip")
        assert np.abs(np.linalg.det(Fp) - 1).max() < 1e-8
        assert (s >= s_prev - 1e-12 * 500).all() and s.max() <= 500 * (1 + 1e-12)
        assert (slip >= slip_prev - 1e-15).all()
        s_prev, slip_prev = s, slip
    ctx.close()


def _bcc48_systems():
    """The 48 BCC systems {110}<111> (12), {112}<111> (12), {123}<111> (24): every plane normal of
    the family (up to sign) with every <111> direction (up to sign) in the plane."""
    import itertools

    def family(hkl):
        out = set()
        for perm in itertools.permutations(hkl):
            for sg in itertools.product((1, -1), repeat=3):
                v = tuple(a * b for a, b in zip(perm, sg))
                if v > tuple(-a for a in v):  # one of each +-pair
                    out.add(v)
        return sorted(out)
    dirs = family((1, 1, 1))
    m, n = [], []
    for fam in ((1, 1, 0), (1, 1, 2), (1, 2, 3)):
        for nv in family(fam):
            for d in dirs:
                if np.dot(nv, d) == 0:
                    n.append(nv)
                    m.append(d)
    return np.array(m, float), np.array(n, float)


def test_qp_update_bcc48_matches_oracle():
    """BCC with 48 slip systems (C11/C12/C44 = 231/135/117 GPa) through the batch entry point, whose
    cooperative kernel (a wavefront per qp, matrices in LDS) needs no private memory; per qp at
    1e-10 against the oracle, with and without the tangent.  The oracle takes unit slip vectors
    (as the reference normalises the file rows, init.cc:73-123); the library normalises its own."""
    m, nrm = _bcc48_systems()
    assert len(m) == 48 and np.allclose((m * nrm).sum(axis=1), 0)
    m = m / np.linalg.norm(m, axis=1)[:, None]
    nrm = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    q = np.full((48, 48), 1.4)
    for a in range(48):
        for b in range(48):
            if np.allclose(nrm[a], nrm[b]):
                q[a, b] = 1.0
    C = np.zeros((6, 6))
    C[:3, :3] = 135e3
    np.fill_diagonal(C[:3, :3], 231e3)
    C[3, 3] = C[4, 4] = C[5, 5] = 117e3
    case = refcase.Case(m, nrm, q, C, [30.0] * 48, [400.0] * 48, [150.0] * 48, [1.5] * 48)
    rng = np.random.default_rng(4)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    n = 128
    F = np.eye(3) + rng.normal(scale=3e-3, size=(n, 3, 3))
    Fp = np.eye(3) + rng.normal(scale=6e-4, size=(n, 3, 3))
    Fp /= np.cbrt(np.linalg.det(Fp))[:, None, None]
    s = np.array([rng.uniform(1.0, 1.5) * case.s0 for _ in range(n)])
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    ctx = _ctx(case, nel=(1, 1, 1))
    ctx.set_grains(np.zeros(1, np.uint32), rod)
    # states where the reference's own result depends on round-off (rank-deficient A_PA whose
    # smallest singular values sit at round-off): the oracle moves by > 1e-9 under 1e-14 relative
    # perturbations of F; no implementation can be held to 1e-10 there
    ill = {i for i in range(n) if _roundoff_dependent(case, rod[grain[i]], F[i], Fp[i], s[i])}
    assert len(ill) <= n // 20, sorted(ill)
    for flag in (0, 1):
        g = ctx.qp_update(F.reshape(n, 9), Fp.reshape(n, 9), s, grain, flag)
        nplastic = 0
        for i in range(n):
            o = refcase.oracle_qp(case, rod[grain[i]], F[i], Fp[i], np.append(s[i], 10e14), flag)
            nplastic += o["n_active"] > 0
            if i in ill:
                continue
            assert g["n_active"][i] == o["n_active"], (i, g["n_active"][i], o["n_active"])
            for k in ("P", "T", "Fe", "Fp"):
                assert _rel(g[k][i], o[k].ravel()) < 1e-10, (k, i)
            assert _rel(g["s"][i], o["s"][:48]) < 1e-10
            if flag:
                assert _rel(g["dPdF"][i], o["dPdF"].ravel()) < 1e-10, (i, _rel(g["dPdF"][i], o["dPdF"].ravel()))
        assert nplastic > n // 4
    ctx.close()


def _roundoff_dependent(case, rod, F, Fp, s, eps=1e-14, tries=6, tol=1e-9):
    """True when the oracle's stress moves by more than tol (or its active count changes) under
    `tries` random relative perturbations of F of size eps."""
    o = refcase.oracle_qp(case, rod, F, Fp, np.append(s, 10e14), 0)
    r = np.random.default_rng(12345)
    for _ in range(tries):
        o2 = refcase.oracle_qp(case, rod, F * (1 + eps * r.standard_normal((3, 3))), Fp, np.append(s, 10e14), 0)
        if _rel(o2["P"], o["P"]) > tol or o2["n_active"] != o["n_active"]:
            return True
    return False


def test_hcp_slip_curve_matches_oracle():
    """8^3 hcp/simpleTension mesh, grains and orientations, 18 slip systems, through yield (16
    increments, Krylov solves to 1e-14): curve to 1e-9, per-qp stress and CRSS to 1e-9."""
    case = _hcp_case()
    case.lin_tol = 1e-14
    n = 16
    ref = refcase.oracle_simulate(case, n, nthreads=8, want_state=True)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
    assert ref["rows"][-1, 14] > 1e-4  # plastic by the end
    for cols in (slice(0, 6), slice(6, 12)):
        err = np.abs(rows[:, cols] - ref["rows"][:, cols]).max() / np.abs(ref["rows"][:, cols]).max()
        assert err < 1e-9, (cols, err)
    assert abs(rows[-1, 14] - ref["rows"][-1, 14]) / ref["rows"][-1, 14] < 1e-6
    st = ref["state"]
    assert _rel(ctx.get_state("T"), st[:, 0:9]) < 1e-9
    assert _rel(ctx.get_state("s"), st[:, 27:27 + 18]) < 1e-9
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
                                      ("fcc_FCC_RandomOrientationBlock", 12), ("fcc_CyclicLoading", 13),
                                      ("fcc_tabularSimpleCompression", 16), ("fcc_DeformationGraidientBCs", 17)])
def test_reference_golden_curve(folder, n):
    """The reference's own inputs (mesh, grain map, orientations, slip systems, BCs, 1e-10 BiCG) run
    through the GPU path reproduce its golden stress-strain rows (tests/compile_tests/*/
    correct_stressstrain.txt, 6 significant digits) over the n increments (the oracle's horizon)
    before rank-deficient active sets make the reference's own curve irreproducible.

    Near the end of that horizon two FP64 implementations start to part: the Krylov solves stop at
    1e-10, so displacements differ at ~1e-9 between builds that round differently, and with dozens
    of singular local systems per iteration a borderline active-set choice eventually goes the
    other way.  Where that happens moves by a few increments with the compiled code (the per-qp
    results themselves agree with the oracle to 1e-11 on identical inputs), so the rows are held to
    print precision (1.2e-5) up to three increments before the horizon and to 5e-4 on the last three
    (observed noise there <= 1e-4; the reference's own harness gates at 1e-2)."""
    case = refcase.load_reference_case(folder)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
    ctx.close()
    gold = case.golden[:n]
    h1 = n - 3
    for cols in (slice(0, 6), slice(6, 12)):
        scale = np.abs(gold[:, cols]).max(axis=1)
        err = np.abs(rows[:, cols] - gold[:, cols]).max(axis=1) / scale
        assert err[:h1].max() < 1.2e-5, (folder, cols, err[:h1].max())
        assert err[h1:].max() < 5e-4, (folder, cols, err[h1:].max())
    slip = gold[:h1, 14] > 1e-7
    if slip.any():
        assert (np.abs(rows[:h1][slip, 14] - gold[:h1][slip, 14]) / gold[:h1][slip, 14]).max() < 5e-5
    assert (rows[:h1][~slip, 14] < 1e-6).all()
    late = gold[h1:, 14] > 1e-7
    if late.any():
        assert (np.abs(rows[h1:][late, 14] - gold[h1:][late, 14]) / gold[h1:][late, 14]).max() < 2e-3


def test_velocity_gradient_bc_affine_field_and_oracle():
    """Non-symmetric velocity gradient on a homogeneous grain: the GPU field is the affine (F_n - I) X at
    every node and matches the oracle's field and rows."""
    case = refcase.velgrad_case()
    n = 3
    ref = refcase.oracle_simulate(case, n, nthreads=4, want_state=True)
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
    u = ctx.get_state("u")
    ctx.close()
    assert _rel(u, refcase.velgrad_affine_u(case, n)) < 1e-9
    assert _rel(u, ref["u"]) < 1e-10
    assert _rel(rows[:, :12], ref["rows"][:, :12]) < 1e-9


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


def test_mixed_form_first_assembly_stagewise():
    """The mixed form's condensed Jacobian (FP32 element tangents, products and sums) equals the
    oracle's FP64 one to FP32 rounding (1e-6 of the largest entry), and so does its residual, whose
    iteration-0 condensation term -K_ij g_j uses the FP32 entries.  (The mixed form's plastic runs
    are checked end to end in test_mixed_precision_form_error_vs_oracle.)"""
    import ctypes
    case = refcase.fcc_case(nx=4, ny=3, nz=5, delT=1.0, total_time=10.0, lin_tol=1e-14)
    rng = np.random.default_rng(8)
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
    ctx.set_solver(pa.native.PA_SOLVER_BICG, pa.native.PA_PREC_MIXED)
    ctx.assemble(0, 0)
    rp, ci, gK = ctx.get_csr(values=True)
    assert _rel(gK, K) < 1e-6, _rel(gK, K)
    assert _rel(ctx.get_state("R"), R) < 1e-6
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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multigrid_pcg_matches_oracle(precision):
    """The multigrid-preconditioned CG (Galerkin levels 8^3 -> 4^3 -> 2^3, Chebyshev smoothing, dense
    coarsest solve) reaches the same solution as the oracle's Jacobi-BiCG: with the solves driven to
    1e-14, the curve and the per-qp stress match to 1e-9 (FP64 form) and the mixed form stays within
    its FP32-Jacobian level; and it needs several times fewer Krylov iterations than Jacobi-PCG."""
    case = refcase.load_reference_case("fcc_simpleTension")
    case.lin_tol = 1e-14
    n = 6
    ref = refcase.oracle_simulate(case, n, nthreads=8, want_state=True)
    its = {}
    for solver in (pa.native.PA_SOLVER_CG, pa.native.PA_SOLVER_MGCG):
        ctx = _ctx(case)
        ctx.set_grains(case.elem_grain, case.grain_rod)
        ctx.set_solver(solver, pa.native.PA_PREC_MIXED if precision == "mixed" else pa.native.PA_PREC_FP64)
        rows, it = [], []
        for inc in range(n):
            r, k, _ = ctx.run_increment(inc)
            rows.append(r)
            it.append(k)
        rows, its[solver] = np.array(rows), np.array(it).ravel()
        T = ctx.get_state("T")
        ctx.close()
        if solver != pa.native.PA_SOLVER_MGCG:
            continue
        tol = 1e-9 if precision == "fp64" else 1e-6
        for cols in (slice(0, 6), slice(6, 12)):
            err = np.abs(rows[:, cols] - ref["rows"][:, cols]).max() / np.abs(ref["rows"][:, cols]).max()
            assert err < tol, (cols, err)
        assert _rel(T, ref["state"][:, 0:9]) < (1e-8 if precision == "fp64" else 1e-5)
    print("krylov iterations: jacobi-pcg", its[pa.native.PA_SOLVER_CG], "mg-pcg", its[pa.native.PA_SOLVER_MGCG])
    assert (its[pa.native.PA_SOLVER_MGCG] > 0).all()
    assert its[pa.native.PA_SOLVER_MGCG].sum() * 3 < its[pa.native.PA_SOLVER_CG].sum()


def test_time_kernel_ids_used_by_bench():
    """Every kernel id bench.py times exists and runs on an assembled context (ids 0..7)."""
    import ctypes
    case = refcase.load_reference_case("fcc_simpleTension")
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    ctx.set_solver(pa.native.PA_SOLVER_MGCG, pa.native.PA_PREC_MIXED)
    for inc in range(6):
        ctx.run_increment(inc)
    ctx.newton_iteration(6, 0)
    avg = ctypes.c_double()
    for which in range(8):
        pa.native.check(pa.native.lib().pa_time_kernel(ctx._h, which, 1, ctypes.byref(avg)))
        assert avg.value > 0.0, which
    with pytest.raises(pa.native.PlasticityError):
        pa.native.check(pa.native.lib().pa_time_kernel(ctx._h, 8, 1, ctypes.byref(avg)))
    ctx.close()


def test_multigrid_rejects_meshes_that_do_not_halve():
    case = refcase.fcc_case(nx=4, ny=3, nz=5)
    ctx = _ctx(case)
    with pytest.raises(pa.native.PlasticityError, match="multigrid"):
        ctx.set_solver(pa.native.PA_SOLVER_MGCG, pa.native.PA_PREC_FP64)
    ctx.close()


def _von_mises(rows, c0):
    a = rows[:, c0:c0 + 6]
    return np.sqrt(0.5 * ((a[:, 0] - a[:, 1]) ** 2 + (a[:, 1] - a[:, 2]) ** 2 + (a[:, 2] - a[:, 0]) ** 2
                          + 6 * (a[:, 3] ** 2 + a[:, 4] ** 2 + a[:, 5] ** 2)))


def test_cyclic_kinematic_hardening_through_load_reversal():
    """fcc/CyclicLoading (kinematic hardening C_1 = 200, C_2 = 1; face 6 displaced in z, quarter cycle
    = 1000 increments) through the first load reversal, on the GPU.

    The reference's harness (tests/compile_tests/run_compile_tests.py:18,52-60) gates rows 0..998 at
    1e-2 on the row-wise maximum of the von Mises stress error.  Past increment 13 the reference's
    own curve carries round-off noise from rank-deficient active sets (row-to-row jitter up to 8e-3
    of the von Mises stress), and any other implementation follows its own realisation of that
    noise: the oracle, an independent FP64 implementation of the same algorithm, has median 1.65e-3,
    95th percentile 1.8e-2 and maximum 4.0e-2 against these rows.  So the gate applied
    here is: the GPU equals the oracle to 1e-6 over the first 12 increments (singular local
    systems appear from increment 6) and stays within 2e-3 of it up to increment 60 (past the
    first dozen increments of singular systems the two implementations take different
    borderline active-set choices, as the reference's own curve shows), and against the golden
    curve the median and 95th-percentile von Mises stress errors stay under 3e-3 and 2.5e-2, the von
    Mises strain and SlipTotal errors under 1e-2 on every row, and the unloading branch after the
    reversal (rows 999..1099) follows the golden Tzz within 1e-2 of its peak."""
    case = refcase.load_reference_case("fcc_CyclicLoading")
    n = 1100
    ctx = _ctx(case)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows = np.array([ctx.run_increment(inc)[0] for inc in range(n)])
    W = ctx.get_state("W")
    print("cyclic: reruns that solved in their lane", ctx.last_counts()[2])
    ctx.close()
    ref = refcase.oracle_simulate(case, 60, nthreads=8)["rows"]
    e_or = np.abs(rows[:60, :12] - ref[:, :12]).max(axis=1) / np.abs(ref[:, :12]).max(axis=1)
    print("cyclic: GPU vs oracle, first 12 increments %.2e, up to 60 %.2e" % (e_or[:12].max(), e_or.max()))
    assert e_or[:12].max() < 1e-6
    assert e_or.max() < 2e-3
    gold = case.golden[:n]
    k = 999
    with np.errstate(divide="ignore", invalid="ignore"):
        e_strain = np.nan_to_num(np.abs((_von_mises(gold, 0) - _von_mises(rows, 0)) / _von_mises(gold, 0)))[:k]
        e_stress = np.nan_to_num(np.abs((_von_mises(gold, 6) - _von_mises(rows, 6)) / _von_mises(gold, 6)))[:k]
        e_slip = np.nan_to_num(np.abs((gold[:, 14] - rows[:, 14]) / gold[:, 14]))[:k]
    print("cyclic vs golden rows 0..998: von Mises stress median %.2e p95 %.2e max %.2e; strain max %.2e; SlipTotal max %.2e"
          % (np.median(e_stress), np.percentile(e_stress, 95), e_stress.max(), e_strain.max(), e_slip.max()))
    assert np.median(e_stress) < 3e-3 and np.percentile(e_stress, 95) < 2.5e-2
    assert e_strain.max() < 1e-2 and e_slip.max() < 1e-2
    tz = np.abs(rows[k:, 8] - gold[k:, 8]).max() / np.abs(gold[:, 8]).max()
    print("cyclic: unloading Tzz %.2e of peak (rows 999..1099)" % tz)
    assert tz < 1e-2
    assert gold[1099, 8] > gold[998, 8] + 50  # the golden really unloads here
    assert np.abs(W).max() > 1.0  # the backstress evolved

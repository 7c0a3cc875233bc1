"""Pins of the oracle's building blocks by what the mathematics fixes (closed forms, invariants,
finite differences, brute force, scipy), chosen so that a dropped term, wrong sign or index, or a
transposed operand fails one of them."""
import ctypes

import numpy as np
import pytest
import scipy.linalg

import refcase
from refcase import ptr

L = refcase.oracle()


def odf(r):
    R = np.zeros(9)
    L.or_odfpoint(ptr(np.asarray(r, float)), ptr(R))
    return R.reshape(3, 3)


def expm(A):
    E = np.zeros(9)
    L.or_matrix_exponential(ptr(np.ascontiguousarray(A, float).ravel()), ptr(E))
    return E.reshape(3, 3)


def test_odfpoint_rotation_closed_form():
    rng = np.random.default_rng(0)
    for r in rng.normal(size=(20, 3)):
        R = odf(r)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14
        # Rodrigues vector = tan(theta/2) * axis, rotating counterclockwise about the axis
        th = 2 * np.arctan(np.linalg.norm(r))
        k = r / np.linalg.norm(r)
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        assert np.allclose(R, scipy.linalg.expm(th * K), atol=1e-13)
    assert np.array_equal(odf([0, 0, 0]), np.eye(3))


def test_matrix_exponential_closed_forms():
    N = np.array([[0, 2.5, 0], [0, 0, 0], [0, 0, 0.0]])  # nilpotent: exp = I + N exactly
    assert np.allclose(expm(N), np.eye(3) + N, atol=1e-15)
    th = 0.3
    W = np.array([[0, -th, 0], [th, 0, 0], [0, 0, 0.0]])
    assert np.allclose(expm(W), [[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]], atol=1e-11)
    D = np.diag([0.1, -0.2, 0.05])
    assert np.allclose(expm(D), np.diag(np.exp(np.diag(D))), atol=1e-11)
    rng = np.random.default_rng(1)
    A = rng.normal(scale=0.05, size=(3, 3))
    A -= np.trace(A) / 3 * np.eye(3)
    assert abs(np.linalg.det(expm(A)) - 1) < 1e-10  # traceless -> det 1


@pytest.mark.parametrize("n", [1, 3, 7, 12])
def test_pinv_solve_full_rank_and_rank_deficient(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + n * np.eye(n)
    b = rng.normal(size=n)
    for mode in (0, 1):
        x = np.zeros(n)
        L.or_pinv_solve(n, ptr(A), ptr(b), ptr(x), mode)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    if n >= 4:
        U, s, Vt = np.linalg.svd(A)
        # one singular value at 1e-12 sigma_max: LU fails its pivot test, and the fallback inverts
        # it like compute_inverse_svd(0.0): x carries the large component along that singular pair
        s1 = s.copy()
        s1[-1] = 1e-12 * s[0]
        A1 = np.ascontiguousarray((U * s1) @ Vt)
        x = np.zeros(n)
        L.or_pinv_solve(n, ptr(A1), ptr(b), ptr(x), 1)
        ref = Vt.T @ ((U.T @ b) / s1)
        assert abs(Vt[-1] @ x) > 1e6 * np.abs(b).max() / s[0]
        assert np.allclose(x, ref, rtol=1e-3, atol=1e-3 * np.abs(ref).max())
        # a single round-off singular value (exact rank n-1) with b off the range: the step along
        # its pair would be round-off amplified by 1e16, so it is dropped (minimum-norm solution)
        s0 = s.copy()
        s0[-1] = 0.0
        A0 = np.ascontiguousarray((U * s0) @ Vt)
        x = np.zeros(n)
        L.or_pinv_solve(n, ptr(A0), ptr(b), ptr(x), 1)
        assert np.allclose(x, np.linalg.pinv(A0, rcond=1e-10) @ b, rtol=1e-8, atol=1e-10)
        # a two-dimensional null space (round-off singular vectors) is dropped: minimum-norm solution
        s2 = s.copy()
        s2[-2:] = 0.0
        A2 = np.ascontiguousarray((U * s2) @ Vt)
        x = np.zeros(n)
        L.or_pinv_solve(n, ptr(A2), ptr(b), ptr(x), 1)
        assert np.allclose(x, np.linalg.pinv(A2, rcond=1e-10) @ b, rtol=1e-8, atol=1e-10)


def rotate_C(C6, R):
    """Rotate a Voigt stiffness through the full 4th-order tensor (independent of the Bond matrix)."""
    v = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    C = np.zeros((3, 3, 3, 3))
    for I, (i, j) in enumerate(v):
        for J, (k, l) in enumerate(v):
            for a, b in ((i, j), (j, i)):
                for c, d in ((k, l), (l, k)):
                    C[a, b, c, d] = C6[I, J]
    return np.einsum("ia,jb,kc,ld,abcd->ijkl", R, R, R, R, C)


def test_elastic_update_is_st_venant_kirchhoff_in_the_rotated_lattice():
    case = refcase.fcc_case()
    rng = np.random.default_rng(3)
    rot = rng.uniform(-0.4, 0.4, 3)
    F = np.eye(3) + rng.normal(scale=1e-4, size=(3, 3))
    o = refcase.oracle_qp(case, rot, F, flag=1)
    assert o["n_active"] == 0
    C = rotate_C(case.C, odf(rot))
    E = 0.5 * (F.T @ F - np.eye(3))
    S = np.einsum("ijkl,kl->ij", C, E)
    assert np.allclose(o["P"], F @ S, rtol=0, atol=1e-12 * np.abs(F @ S).max())
    T = F @ S @ F.T / np.linalg.det(F)
    assert np.allclose(o["T"], T, atol=1e-12 * np.abs(T).max())
    # tangent: central differences of P(F), and major symmetry of the elastic tangent
    D = o["dPdF"]
    h = 1e-7
    Dfd = np.zeros((9, 9))
    for c in range(9):
        dF = np.zeros(9)
        dF[c] = h
        Pp = refcase.oracle_qp(case, rot, F + dF.reshape(3, 3), flag=0)["P"].ravel()
        Pm = refcase.oracle_qp(case, rot, F - dF.reshape(3, 3), flag=0)["P"].ravel()
        Dfd[:, c] = (Pp - Pm) / (2 * h)
    # The reference's final tangent (calculatePlasticity.cc:899-913) differentiates P through
    # d(F^-1 Fe)^T with two terms that do not cancel when Fe = F: it equals the exact derivative
    # minus (Fe T*)(i,k) F^-1(j,l) + (Fe T*)(i,l) F^-1(j,k).  Pin that closed form.
    q3, Fi = F @ S, np.linalg.inv(F)
    extra = -(np.einsum("ik,jl->ijkl", q3, Fi) + np.einsum("il,jk->ijkl", q3, Fi)).reshape(9, 9)
    assert np.abs(D - (Dfd + extra)).max() < 1e-6 * np.abs(D).max()
    assert np.abs(extra).max() > 1e3 * np.abs(D - (Dfd + extra)).max()  # the extra term is resolved


def test_yield_onset_for_cube_oriented_crystal():
    """r = 0: a qp turns plastic exactly when max_a |tau_a| - s0 >= tol, tau_a = (Ce T*) : (m_a x n_a)."""
    case = refcase.fcc_case()
    m = case.m / np.linalg.norm(case.m, axis=1)[:, None]
    n = case.n / np.linalg.norm(case.n, axis=1)[:, None]
    C = rotate_C(case.C, np.eye(3))
    for eps in np.linspace(1e-4, 6e-4, 21):
        F = np.diag([1 + eps, 1 - 0.4 * eps, 1 - 0.4 * eps])
        E = 0.5 * (F.T @ F - np.eye(3))
        Ts = np.einsum("ijkl,kl->ij", C, E)
        tau = np.einsum("jk,aj,ak->a", F.T @ F @ Ts, m, n)
        o = refcase.oracle_qp(case, [0, 0, 0], F, flag=0)
        assert (o["n_active"] > 0) == ((np.abs(tau) - 16.0).max() >= 1e-9)


def test_plastic_update_closed_forms():
    """Single-pass update: s_a += sum_b q_ab h_b dgamma_b with h from the old s, and
    Fp = exp(sum_a sign(tau_a) dgamma_a S_a) Fp_t (calculatePlasticity.cc:255-270, 340-359, 401-429)."""
    case = refcase.fcc_case()
    rng = np.random.default_rng(5)
    rot = rng.uniform(-0.4, 0.4, 3)
    R = odf(rot)
    m = case.m / np.linalg.norm(case.m, axis=1)[:, None]
    n = case.n / np.linalg.norm(case.n, axis=1)[:, None]
    S = np.einsum("ij,ajk,lk->ail", R, np.einsum("aj,ak->ajk", m, n), R)
    hits = 0
    for trial in range(20):
        F = np.eye(3) + rng.normal(scale=3e-3, size=(3, 3))
        s_in = rng.uniform(16, 40, 12)
        o = refcase.oracle_qp(case, rot, F, s=np.append(s_in, 10e14), flag=0)
        if o["n_active"] == 0:
            continue
        hits += 1
        x = o["slip"]  # slip_conv = 0, so slip_iter = x_beta_old
        h = 180.0 * (1 - s_in / 148.0) ** 2.25
        s_ref = np.minimum(s_in + case.q @ (h * x), 148.0)
        assert np.allclose(o["s"][:12], s_ref, rtol=1e-13, atol=1e-11)
        Fe_tr = F  # Fp_t = I
        C = rotate_C(case.C, R)
        Ts = np.einsum("ijkl,kl->ij", C, 0.5 * (Fe_tr.T @ Fe_tr - np.eye(3)))
        tau = np.einsum("jk,ajk->a", Fe_tr.T @ Fe_tr @ Ts, S)
        X = np.einsum("a,ajk->jk", np.where(tau > 0, x, -x), S)
        assert np.allclose(o["Fp"], scipy.linalg.expm(X), atol=1e-10)
        assert (x >= -1e-15).all()
        assert np.allclose(o["Fe"], F @ np.linalg.inv(o["Fp"]), atol=1e-13)
    assert hits >= 10


def test_converged_return_map_lands_on_the_yield_surface_and_its_tangent_matches_fd():
    """With the local loop run to convergence the active systems satisfy |tau_a| = s_a and the
    consistent tangent approaches the derivative of the converged map."""
    case = refcase.fcc_case(max_solver_iter=200)
    rng = np.random.default_rng(9)
    rot = rng.uniform(-0.4, 0.4, 3)
    F = np.eye(3) + np.array([[8e-4, 1e-4, 0], [0, -3e-4, 2e-4], [1e-4, 0, -3e-4]])
    o = refcase.oracle_qp(case, rot, F, flag=1)
    assert o["n_active"] > 0
    C = rotate_C(case.C, odf(rot))
    Fe = o["Fe"]
    Ts = np.einsum("ijkl,kl->ij", C, 0.5 * (Fe.T @ Fe - np.eye(3)))
    R = odf(rot)
    m = case.m / np.linalg.norm(case.m, axis=1)[:, None]
    n = case.n / np.linalg.norm(case.n, axis=1)[:, None]
    S = np.einsum("ij,ajk,lk->ail", R, np.einsum("aj,ak->ajk", m, n), R)
    tau = np.einsum("jk,ajk->a", Fe.T @ Fe @ Ts, S)
    act = o["slip"] > 0
    assert np.allclose(np.abs(tau[act]), o["s"][:12][act], rtol=1e-8)
    D = o["dPdF"]
    h = 1e-8
    Dfd = np.zeros((9, 9))
    for c in range(9):
        dF = np.zeros(9)
        dF[c] = h
        Pp = refcase.oracle_qp(case, rot, F + dF.reshape(3, 3), flag=0)["P"].ravel()
        Pm = refcase.oracle_qp(case, rot, F - dF.reshape(3, 3), flag=0)["P"].ravel()
        Dfd[:, c] = (Pp - Pm) / (2 * h)
    err = np.abs(D - Dfd).max() / np.abs(D).max()
    assert err < 0.05, err  # the elastic tangent alone differs from Dfd by O(1)
    De = refcase.oracle_qp(refcase.fcc_case(s0=[1e6] * 12), rot, F, flag=1)["dPdF"]
    assert np.abs(De - Dfd).max() / np.abs(D).max() > 10 * err


def test_element_matrix_invariants():
    case = refcase.fcc_case()
    mat = case.material(1)
    rng = np.random.default_rng(4)
    rot = rng.uniform(-0.3, 0.3, 3)
    h = np.array([0.2, 0.125, 0.1])
    ue = rng.normal(scale=2e-6, size=24)  # strains ~2e-5: resolved shears ~1.5 MPa < s0 = 16, elastic
    Fp = np.tile(np.eye(3).ravel(), 8)
    s = np.tile(np.append(np.full(12, 16.0), 10e14), 8)
    Ke, Re = np.zeros(576), np.zeros(24)
    L.or_element_matrix(ctypes.byref(mat), ptr(rot), ptr(h), ptr(ue), ptr(Fp), ptr(s), ptr(Ke), ptr(Re))
    Ke = Ke.reshape(24, 24)
    # sum_v grad N_v = 0: nodal forces balance and rigid translations are in the null space
    assert np.abs(Re.reshape(8, 3).sum(axis=0)).max() < 1e-12 * np.abs(Re).max()
    for c in range(3):
        t = np.zeros(24)
        t[c::3] = 1.0
        assert np.abs(Ke @ t).max() < 1e-10 * np.abs(Ke).max()
    # consistency of K_e with R_e: -dR/du by central differences (elastic state)
    hh = 1e-9
    Kfd = np.zeros((24, 24))
    for d in range(24):
        up, um = ue.copy(), ue.copy()
        up[d] += hh
        um[d] -= hh
        Kp, Rp, Km, Rm = np.zeros(576), np.zeros(24), np.zeros(576), np.zeros(24)
        L.or_element_matrix(ctypes.byref(mat), ptr(rot), ptr(h), ptr(up), ptr(Fp), ptr(s), ptr(Kp), ptr(Rp))
        L.or_element_matrix(ctypes.byref(mat), ptr(rot), ptr(h), ptr(um), ptr(Fp), ptr(s), ptr(Km), ptr(Rm))
        Kfd[:, d] = -(Rp - Rm) / (2 * hh)
    # K_e = -dR_e/du up to the O(|T|/|C|) extra term of the reference tangent (see the elastic test)
    err = np.abs(Ke - Kfd).max() / np.abs(Ke).max()
    print("element FD error", err)
    assert 1e-7 < err < 1e-3, err


def test_polar_rotation_matches_scipy():
    rng = np.random.default_rng(8)
    for _ in range(20):
        Fe = np.eye(3) + rng.normal(scale=0.1, size=(3, 3))
        R = np.zeros(9)
        L.or_polar_rotation(ptr(np.ascontiguousarray(Fe).ravel()), ptr(R))
        Rs, _ = scipy.linalg.polar(Fe)
        assert np.allclose(R.reshape(3, 3), Rs, atol=1e-12)
    R = np.zeros(9)
    L.or_polar_rotation(ptr(np.eye(3).ravel()), ptr(R))
    assert np.allclose(R.reshape(3, 3), np.eye(3), atol=1e-15)


@pytest.mark.parametrize("nel", [(1, 1, 1), (2, 3, 4), (8, 8, 8)])
def test_sparsity_invariants(nel):
    case = refcase.fcc_case(nx=nel[0], ny=nel[1], nz=nel[2])
    pb = case.problem()
    nnz = L.or_sparsity(ctypes.byref(pb), None, None)
    if nel[0] == nel[1] == nel[2]:
        assert nnz == 9 * (3 * nel[0] + 1) ** 3
    ndof = 3 * np.prod(np.array(nel) + 1)
    rp = np.zeros(ndof + 1, np.int64)
    ci = np.zeros(nnz, np.int32)
    L.or_sparsity(ctypes.byref(pb), ptr(rp), ptr(ci))
    lens = np.diff(rp)
    assert set(np.unique(lens)) <= {81, 54, 36, 24, 27, 18, 12, 8 * 3 // 1}
    import scipy.sparse as sp
    A = sp.csr_matrix((np.ones(nnz), ci, rp), shape=(ndof, ndof))
    assert (A != A.T).nnz == 0  # symmetric pattern
    for r in range(ndof):
        row = ci[rp[r]:rp[r + 1]]
        assert (np.diff(row) > 0).all() and r in row


def test_grain_lookup_matches_brute_force_with_tie_rule():
    rng = np.random.default_rng(6)
    nv = np.array([5, 7, 4], np.int32)
    span = np.array([1.0, 2.0, 0.5])
    vox = rng.integers(1, 50, size=tuple(nv)).astype(np.uint32)
    st = span / nv
    pts = list(rng.uniform(0, 1, size=(200, 3)) * span)
    for _ in range(50):  # exact midpoints between voxel centres -> ties to the lower voxel
        i = rng.integers(0, nv - 1)
        pts.append((i + 1) * st)
    for p in pts:
        p = np.asarray(p, float)
        idx = []
        for d in range(3):
            c = np.arange(nv[d]) * st[d] + st[d] / 2
            dist = np.abs(c - p[d])
            idx.append(int(np.flatnonzero(dist == dist.min())[0]))  # lowest on ties
        got = L.or_material_id(ptr(vox), ptr(nv), ptr(span), ptr(p))
        assert got == vox[tuple(idx)]


def test_velocity_gradient_bc_drives_an_affine_field():
    """All-boundary velocity-gradient BCs on a homogeneous grain: Q1 with QGauss(2) reproduces the
    affine field, so interior nodes follow the boundary to round-off.  The non-symmetric L fails a
    transposed dF; non-unit element sizes fail node indices used for positions; a non-nilpotent L over
    three increments separates the implicit recursion (I - L dt)^-1 from F = I + L t or exp(L t)."""
    case = refcase.velgrad_case()
    n = 3
    out = refcase.oracle_simulate(case, n, nthreads=4, want_state=True)
    ua = refcase.velgrad_affine_u(case, n)
    err = np.abs(out["u"] - ua).max() / np.abs(ua).max()
    print("affine error", err)
    assert err < 1e-9
    # the alternatives named above really are distinguishable at this tolerance
    F_lin = np.eye(3) + case.vel_grad * n * case.delT
    F_imp = np.linalg.matrix_power(np.linalg.inv(np.eye(3) - case.vel_grad), n)
    assert np.abs(F_lin - F_imp).max() > 1e-4

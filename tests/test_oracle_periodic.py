"""Pins of the oracle's periodic boundary conditions (Use Periodic BCs: ellipticBVP periodicBCs.cc,
init.cc:66-173; the reference's fcc/periodicBCs inputs under tests/golden/fcc_periodicBCs) by what
the mathematics of periodic boundary conditions fixes:

* periodicity of the solution: at every increment, u on each positive face / edge / corner equals u
  at its periodic image plus the corner jump, u(x + L e_a) - u(x) = u(V_a) - u(V000);
* a homogeneous crystal under periodic uniaxial tension deforms homogeneously (F equal at every
  quadrature point) with no lateral stress (the free vertex dofs carry zero reaction);
* the condensed first Newton system (C^T K C, right-hand side C^T (R - K g), distribute) equals a
  Lagrange-multiplier solve of the unconstrained system with the constraints as side conditions;
* the tabular periodic table reverses the load at the table's time points.

There is no golden curve for these inputs in the reference (tests/compile_tests has none).
"""
import ctypes

import numpy as np
import pytest

import refcase
from refcase import ptr

L = refcase.oracle()
L.or_periodic_lines.restype = ctypes.c_longlong


def _case(prm="prm_UniaxialTension.prm"):
    return refcase.load_reference_case("fcc_periodicBCs", prm_name=prm)


def _lines(case):
    pb = case.problem(1)
    N = 3 * (case.nx + 1) * (case.ny + 1) * (case.nz + 1)
    con, rp, g = np.zeros(N, np.int32), np.zeros(N + 1, np.int64), np.zeros(N)
    nnz = L.or_periodic_lines(ctypes.byref(pb), ptr(con), ptr(rp), None, None, ptr(g), ctypes.c_longlong(0))
    assert nnz >= 0
    M, W = np.zeros(max(nnz, 1), np.int64), np.zeros(max(nnz, 1))
    L.or_periodic_lines(ctypes.byref(pb), ptr(con), ptr(rp), ptr(M), ptr(W), ptr(g), ctypes.c_longlong(nnz))
    return con.astype(bool), rp, M, W, g


def _node(n, i, j, k):
    return i + (n + 1) * (j + (n + 1) * k)


def _jumps(u, n):
    """Largest violation of u(x + L e_a) - u(x) = u(V_a) - u(V000) over all boundary node pairs."""
    U = u.reshape(n + 1, n + 1, n + 1, 3)  # [k][j][i][c]
    v000 = U[0, 0, 0]
    worst = 0.0
    for a, Va in ((0, U[0, 0, n]), (1, U[0, n, 0]), (2, U[n, 0, 0])):
        if a == 0:
            d = U[:, :, n] - U[:, :, 0]
        elif a == 1:
            d = U[:, n, :] - U[:, 0, :]
        else:
            d = U[n, :, :] - U[0, :, :]
        worst = max(worst, np.abs(d - (Va - v000)).max())
    return worst


@pytest.mark.parametrize("prm", ["prm_UniaxialTension.prm", "prm_simpleShear.prm", "prm_BiaxialTension.prm", "prm_pureShear.prm"])
def test_solution_is_periodic_with_the_corner_jumps(prm):
    case = _case(prm)
    for incs in (1, 6):
        out = refcase.oracle_simulate(case, incs, nthreads=4, want_state=True)
        u = out["u"]
        scale = np.abs(u).max()
        assert scale > 0
        assert _jumps(u, case.nx) < 1e-12 * scale, prm


def test_closed_lines_tie_each_dof_to_one_periodic_image():
    """Every constrained dof that is not a corner is one unconstrained dof of the same component
    with weight 1 (its periodic image) plus corner dofs; corners depend on corners only."""
    case = _case()
    con, rp, M, W, g = _lines(case)
    n = case.nx
    corners = {_node(n, i * n, j * n, k * n) for i in (0, 1) for j in (0, 1) for k in (0, 1)}
    assert con.sum() > 0
    for d in np.nonzero(con)[0]:
        ms, ws = M[rp[d]:rp[d + 1]], W[rp[d]:rp[d + 1]]
        assert not con[ms].any()  # closed: masters are unconstrained
        non_corner = [(m, w) for m, w in zip(ms, ws) if m // 3 not in corners]
        if d // 3 in corners:
            assert not non_corner
        else:
            assert len(non_corner) == 1 and non_corner[0][1] == 1.0 and non_corner[0][0] % 3 == d % 3


def test_homogeneous_crystal_deforms_homogeneously_with_free_lateral_faces():
    case = _case()
    case.elem_grain = np.zeros_like(case.elem_grain)
    case.grain_rod = np.array([[0.1, -0.2, 0.15]])
    out = refcase.oracle_simulate(case, 3, nthreads=4, want_state=True)
    st = out["state"]
    # per-qp Fe*Fp: recompute F from the stored Fe (cols 18..27) and Fp (9..18)
    Fe, Fp = st[:, 18:27].reshape(-1, 3, 3), st[:, 9:18].reshape(-1, 3, 3)
    F = Fe @ Fp
    assert np.abs(F - F[0]).max() < 1e-12
    row = out["rows"][-1]
    Txx, Tyy, Tzz = row[6], row[7], row[8]
    # V010_2 and V001_3 are free: the macroscopic Tyy, Tzz vanish with the Newton residual
    assert abs(Tyy) < 1e-9 * abs(Txx) and abs(Tzz) < 1e-9 * abs(Txx)
    # and the uniaxial stretch is the prescribed one: Exx = (1 + 5e-5 * incs)^2 / 2 - 1/2
    assert abs(row[0] - ((1 + 3 * 5e-5) ** 2 - 1) / 2) < 1e-15


def test_first_newton_system_matches_a_lagrange_multiplier_solve():
    case = _case()
    rng = np.random.default_rng(7)
    case.grain_rod = rng.uniform(-0.5, 0.5, size=(4, 3))
    case.elem_grain = rng.integers(0, 4, size=case.nx ** 3).astype(np.uint32)
    case.lin_tol, case.max_newton = 1e-15, 1
    n, N = case.nx, 3 * (case.nx + 1) ** 3
    out = refcase.oracle_simulate(case, 1, want_state=True)
    du = out["u"]
    mat = case.material()
    K = np.zeros((N, N))
    h = np.full(3, 1.0 / n)
    Fp, s = np.tile(np.eye(3).ravel(), 8), np.tile(np.append(case.s0, 10e14), 8)
    for e in range(n ** 3):
        ex, ey, ez = e % n, (e // n) % n, e // (n * n)
        nodes = [_node(n, ex + (v & 1), ey + ((v >> 1) & 1), ez + ((v >> 2) & 1)) for v in range(8)]
        Ke, Re = np.zeros(576), np.zeros(24)
        L.or_element_matrix(ctypes.byref(mat), ptr(np.ascontiguousarray(case.grain_rod[case.elem_grain[e]])), ptr(h),
                            ptr(np.zeros(24)), ptr(Fp), ptr(s), ptr(Ke), ptr(Re))
        dofs = np.array([3 * nd + c for nd in nodes for c in range(3)])
        K[np.ix_(dofs, dofs)] += Ke.reshape(24, 24)
        assert np.abs(Re).max() == 0.0  # F = I: no stress
    con, rp, M, W, g = _lines(case)
    rows = np.nonzero(con)[0]
    A = np.zeros((len(rows), N))
    for r, d in enumerate(rows):
        A[r, d] = 1.0
        for q in range(rp[d], rp[d + 1]):
            A[r, M[q]] -= W[q]
    sA = np.abs(K).max()  # constraint rows scaled to the stiffness for a balanced system
    KKT = np.block([[K, sA * A.T], [sA * A, np.zeros((len(rows), len(rows)))]])
    rhs = np.concatenate([np.zeros(N), sA * g[rows]])
    sol = np.linalg.solve(KKT, rhs)
    ref = sol[:N]
    assert np.abs(KKT @ sol - rhs).max() < 1e-12 * sA * np.abs(ref).max()
    print("condensed solve vs Lagrange multipliers: %.2e" % (np.abs(du - ref).max() / np.abs(ref).max()))
    assert np.abs(du - ref).max() < 1e-10 * np.abs(ref).max()


def test_tabular_periodic_table_reverses_the_load():
    """UniaxialCyclicTabular: periodic Tabular time 1 (200 increments of 0.005) and table
    0,1,3,.. with coefficients 1,-1,..: increments 0..199 load, 200.. unload (the counter moves
    when delT (inc + 1) - delT / 1000 passes the next table time)."""
    case = _case("prm_UniaxialCyclicTabular.prm")
    case.nx = case.ny = case.nz = 2
    case.elem_grain = refcase.oracle_element_grains(case)
    v100 = 3 * _node(2, 2, 0, 0)  # u_x of corner V100: the prescribed stretch
    out = refcase.oracle_simulate(case, 200, nthreads=4, want_state=True)
    assert abs(out["u"][v100] - 200 * 5e-5) < 1e-15
    out = refcase.oracle_simulate(case, 203, nthreads=4, want_state=True)
    assert abs(out["u"][v100] - 197 * 5e-5) < 1e-15
    dE = np.diff(out["rows"][:, 0])
    assert (dE[:199] > 0).all() and (dE[199:] < 0).all()


def test_unequal_edge_counts_are_refused():
    case = _case()
    case.nx = 4
    pb = case.problem(1)
    assert L.or_periodic_lines(ctypes.byref(pb), None, None, None, None, None, ctypes.c_longlong(0)) == -1

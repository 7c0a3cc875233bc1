"""The pseudo-inverse step of inactive-slip removal (inactiveSlipRemoval.cc:26,99) on the GPU.

The three device paths, the one-lane routine, the cooperative kernel that solves queued singular
systems, and the group routine the cooperative qp kernels run in place, must give bit-identical x
(a rerun of a quadrature point feeds one path's solution to a computation that elsewhere takes
another).  Against the oracle (or_pinv_solve, mode 1) and the
closed form V diag(1/sigma) U^T b: 1e-11 relative on full-rank systems, 1e-8 on exact rank
deficiency with b off the range (minimum-norm solution), 1e-3 where one singular value of
1e-11 sigma_max is inverted (its relative round-off is about eps / 1e-11).
"""
import ctypes

import numpy as np
import pytest

import plasticity_amd as pa
import refcase

pytestmark = pytest.mark.gpu

L = refcase.oracle()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _systems(n, count, kind, rng):
    A = np.zeros((count, n, n))
    b = rng.normal(size=(count, n))
    U = np.linalg.qr(rng.normal(size=(count, n, n)))[0]
    V = np.linalg.qr(rng.normal(size=(count, n, n)))[0]
    s = rng.uniform(1.0, 10.0, size=(count, n))
    s = -np.sort(-s, axis=1)
    if kind == "tiny":
        s[:, -1] = 1e-11 * s[:, 0]
    elif kind == "rank1def":
        s[:, -1] = 0.0
    elif kind == "rank2def":
        s[:, -2:] = 0.0
    if kind.startswith("rank"):
        # b carries a null-space component of at least |b|: the rule then drops the round-off
        # singular value whatever its computed size (with u_k.b near round-off the keep/drop
        # decision itself rests on round-off, as does the reference's inversion of it)
        for k in range(1 if kind == "rank1def" else 2):
            u = U[:, :, n - 1 - k]
            c = np.einsum("ei,ei->e", u, b)
            b += (np.where(c >= 0, 1.0, -1.0) * np.linalg.norm(b, axis=1))[:, None] * u
    A = np.einsum("eij,ej,ekj->eik", U, s, V)
    return np.ascontiguousarray(A), b, U, s, V


def _closed_form(b, U, s, V, rcond):
    inv = np.where(s > rcond * s[:, :1], 1.0 / np.where(s > 0, s, 1.0), 0.0)
    return np.einsum("eij,ej,ekj,ek->ei", V, inv, U, b)


@pytest.mark.parametrize("n", [3, 8, 12, 17, 24, 31, 48])
@pytest.mark.parametrize("kind", ["full", "tiny", "rank1def", "rank2def"])
def test_pinv_paths_bit_identical_and_match_oracle(n, kind):
    if kind == "rank2def" and n < 4:
        pytest.skip("needs n >= 4")
    rng = np.random.default_rng(1000 * n + len(kind))
    count = 96
    A, b, U, s, V = _systems(n, count, kind, rng)
    print("pinv", n, kind, count, flush=True)
    x0 = pa.model.pinv_solve(A, b, path=0)
    x1 = pa.model.pinv_solve(A, b, path=1)
    x2 = pa.model.pinv_solve(A, b, path=2)
    assert np.array_equal(x0.view(np.int64), x1.view(np.int64)), "lane and cooperative paths differ"
    assert np.array_equal(x0.view(np.int64), x2.view(np.int64)), "lane and in-place group paths differ"
    xo = np.zeros_like(x0)
    for e in range(count):
        Ae = np.ascontiguousarray(A[e])
        be = np.ascontiguousarray(b[e])
        xe = np.zeros(n)
        L.or_pinv_solve(n, _ptr(Ae), _ptr(be), _ptr(xe), 1)
        xo[e] = xe
    rcond = 0.0 if kind == "tiny" else 1e-10
    xc = _closed_form(b, U, s, V, rcond)
    tol = {"full": 1e-11, "tiny": 1e-3, "rank1def": 1e-8, "rank2def": 1e-8}[kind]
    scale = np.abs(xc).max(axis=1)
    assert (np.abs(x0 - xo).max(axis=1) <= tol * scale).all()
    assert (np.abs(x0 - xc).max(axis=1) <= tol * scale).all()
    if kind == "tiny":  # the tiny pair is inverted, not dropped
        assert (np.abs(np.einsum("ei,ei->e", V[:, :, -1], x0)) > 1e6 * np.abs(b).max(axis=1) / s[:, 0]).all()


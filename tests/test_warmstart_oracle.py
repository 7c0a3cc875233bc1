"""Pins of the warm-start oracle (CPU): the network's gradient against finite differences, the
bfloat16 rounding against hand-worked cases, the warm-started local loop against the plain one,
and the host's Gauss-point deformation gradients against an affine closed form."""
import os
import sys

import numpy as np

import refcase

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

from plasticity_amd import warmstart as ws  # noqa: E402


def _blob(rng, n=64):
    X = rng.normal(size=(n, mo.NIN))
    Y = np.abs(rng.normal(size=(n, mo.NOUT))) * 1e-3
    blob = ws.init_blob(X, Y, seed=3).astype(np.float64)
    # non-zero biases so that their gradients are exercised
    for l in range(4):
        nout = mo.NOUT if l == 3 else 64
        blob[mo.OFF_B[l]:mo.OFF_B[l] + nout] = rng.normal(scale=0.1, size=nout)
    return blob, X, Y


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {one: one, one + 2 ** -8: one, one + 3 * 2 ** -8: one + 2 ** -6, one + 2 ** -7: one + 2 ** -7,
             one + 2 ** -8 + 2 ** -20: one + 2 ** -7, -(one + 3 * 2 ** -8): -(one + 2 ** -6), 0.0: 0.0}
    for x, want in cases.items():
        assert mo.bf16(np.float32(x)) == np.float32(want), x


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    blob, X, Y = _blob(rng)
    loss, g = mo.loss_grad(blob, X, Y, quant=False)
    # padding carries no gradient
    for l in range(4):
        WT = g[mo.OFF_W[l]:mo.OFF_W[l] + mo.LN[l] * mo.LK[l]].reshape(mo.LN[l], mo.LK[l])
        kin = mo.NIN if l == 0 else 64
        nout = mo.NOUT if l == 3 else 64
        assert not WT[:, kin:].any() and not WT[nout:, :].any()
    picks = []
    for l in range(4):
        kin = mo.NIN if l == 0 else 64
        nout = mo.NOUT if l == 3 else 64
        for _ in range(12):
            picks.append(mo.OFF_W[l] + rng.integers(nout) * mo.LK[l] + rng.integers(kin))
        for _ in range(4):
            picks.append(mo.OFF_B[l] + rng.integers(nout))
    for p in picks:
        h = 1e-6 * max(1.0, abs(blob[p]))
        bp, bm = blob.copy(), blob.copy()
        bp[p] += h
        bm[p] -= h
        fd = (mo.loss_grad(bp, X, Y, quant=False)[0] - mo.loss_grad(bm, X, Y, quant=False)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (p, fd, g[p])


def test_mlp_quantised_gradient_close_to_exact():
    """bf16 operands move the gradient by about 2^-8 relative, not more (a layout slip in the
    quantised path would move it by O(1))."""
    rng = np.random.default_rng(1)
    blob, X, Y = _blob(rng, 256)
    l0, g0 = mo.loss_grad(blob, X, Y, quant=False)
    l1, g1 = mo.loss_grad(blob, X, Y, quant=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 3e-2 * np.linalg.norm(g0)


def _plastic_state(rng):
    rot = rng.uniform(-0.5, 0.5, 3)
    F = np.eye(3) + rng.normal(size=(3, 3)) * 1.5e-3
    return rot, F, np.eye(3), np.full(12, 16.0)


def test_oracle_warm_start_none_is_plain():
    case = refcase.fcc_case(max_solver_iter=50)
    rng = np.random.default_rng(2)
    for _ in range(20):
        rot, F, Fp, s = _plastic_state(rng)
        a = refcase.oracle_qp(case, rot, F, Fp, np.append(s, 1e15), flag=1)
        b = refcase.oracle_qp_warm(case, rot, F, Fp, s, None, flag=1)
        for k in ("P", "T", "dPdF", "Fp", "slip"):
            assert np.array_equal(a[k], b[k]), k


def test_oracle_warm_start_converges_to_the_same_state():
    """Every start ends on the plain loop's stress: from the converged slip the loop stops after
    the one solve that certifies the set; from a perturbed slip it runs on; from a prediction of
    the wrong set it continues from its first solve, exactly as without one."""
    case = refcase.fcc_case(max_solver_iter=50)
    rng = np.random.default_rng(3)
    n, one = 0, 0
    for _ in range(40):
        rot, F, Fp, s = _plastic_state(rng)
        a = refcase.oracle_qp_warm(case, rot, F, Fp, s, None)
        if a["n_active"] == 0:
            continue
        x = a["slip"].copy()
        starts = [x, x * (1 + 0.05 * rng.normal(size=12)), np.where(rng.random(12) < 0.5, x, 0.0), rng.random(12) * 1e-3]
        for x0 in starts:
            r = refcase.oracle_qp_warm(case, rot, F, Fp, s, x0)
            assert np.abs(r["T"] - a["T"]).max() <= 1e-9 * np.abs(a["T"]).max()
            assert r["iters"] >= 1
        n += 1
        one += refcase.oracle_qp_warm(case, rot, F, Fp, s, x)["iters"] == 1
    assert n >= 30 and one >= 0.9 * n


def test_qp_deformation_gradients_affine_closed_form():
    nel, span = (3, 4, 2), (1.0, 0.7, 1.3)
    G = np.array([[0.01, -0.02, 0.003], [0.004, 0.02, -0.01], [-0.005, 0.007, 0.015]])
    nx, ny, nz = nel
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    X = np.stack([i.ravel() * span[0] / nx, j.ravel() * span[1] / ny, k.ravel() * span[2] / nz], axis=1)
    u = (X @ G.T).ravel()
    F = ws.qp_deformation_gradients(u, nel, span)
    assert F.shape == (8 * nx * ny * nz, 9)
    assert np.abs(F - (np.eye(3) + G).ravel()).max() < 1e-14
    # a field linear in x only varies its gradient by element, not by qp; a bilinear one by qp
    u2 = np.zeros((len(X), 3))
    u2[:, 0] = 0.01 * X[:, 1] * X[:, 2]
    F2 = ws.qp_deformation_gradients(u2.ravel(), nel, span).reshape(-1, 8, 3, 3)
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    e = 1 + nx * (2 + ny * 1)  # ex=1, ey=2, ez=1
    for q in range(8):
        y = (2 + g[(q >> 1) & 1]) * span[1] / ny
        z = (1 + g[(q >> 2) & 1]) * span[2] / nz
        assert abs(F2[e, q, 0, 1] - 0.01 * z) < 1e-14 and abs(F2[e, q, 0, 2] - 0.01 * y) < 1e-14


def test_qp_deformation_gradients_element_subset():
    rng = np.random.default_rng(5)
    nel, span = (4, 3, 5), (1.0, 0.8, 1.2)
    u = rng.normal(size=3 * 5 * 4 * 6) * 1e-3
    full = ws.qp_deformation_gradients(u, nel, span).reshape(-1, 8, 9)
    elems = np.array([7, 0, 59, 23])
    sub = ws.qp_deformation_gradients(u, nel, span, elems).reshape(-1, 8, 9)
    assert np.array_equal(sub, full[elems])

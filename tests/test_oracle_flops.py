"""The flop-counting build of the oracle (oracle/liboracle_flops.so) and the flop fixture it wrote
(tests/golden/qp_flops_fcc24.json, the per-qp FP64 flops behind the qp pass's roofline).

* The count is exact where it can be counted by hand: LU with partial pivoting on a dense n x n
  system with one right-hand side costs sum_j j(j+1) + sum_j j add/sub, 1 + sum_j j(j+2) +
  sum_j (j+1) mul and 2n div (j = 0..n-1; the 1 mul is the pivot threshold).
* The counting build runs the oracle's algorithm unchanged: per-qp results bit-identical to the
  plain oracle's on plastic states.
* The fixture holds what the counting build gives today: an elastic qp's count equals the count of
  fixture step 0 (every qp elastic there), and every step counts each of the 24^3 x 8 qps once.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import refcase

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLOPS = os.path.join(ROOT, "oracle", "liboracle_flops.so")
FIXTURE = os.path.join(ROOT, "tests", "golden", "qp_flops_fcc24.json")


@pytest.fixture(scope="module")
def flop_lib():
    if not os.path.exists(FLOPS):
        pytest.skip("oracle/liboracle_flops.so is not built (make -C oracle)")
    return ctypes.CDLL(FLOPS)


def _totals(lib):
    t = (ctypes.c_longlong * 5)()
    lib.or_flop_total(t)
    return list(t)


@pytest.mark.parametrize("n", [1, 5, 12])
def test_lu_count_equals_the_hand_count(flop_lib, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(1.0, 2.0, (n, n)) + n * np.eye(n)  # no zero multipliers, no pivot swaps needed
    b = rng.uniform(1.0, 2.0, n)
    x = np.zeros(n)
    flop_lib.or_flop_reset()
    flop_lib.or_pinv_solve(ctypes.c_int(n), refcase.ptr(A), refcase.ptr(b), refcase.ptr(x), ctypes.c_int(1))
    add, mul, div, sq, fun = _totals(flop_lib)
    j = np.arange(n)
    assert (add, mul, div, sq, fun) == (int((j * (j + 1)).sum() + j.sum()), int(1 + (j * (j + 2)).sum() + (j + 1).sum()), 2 * n, 0, 0)
    assert np.abs(A @ x - b).max() < 1e-12


def test_counting_build_reproduces_the_oracle_bit_for_bit(flop_lib):
    case = refcase.fcc_case()
    rng = np.random.default_rng(3)
    plastic = 0
    for k in range(40):
        rot = rng.uniform(-0.5, 0.5, 3)
        E = rng.standard_normal((3, 3)) * 4e-3
        F = np.eye(3) + E
        refcase.use_oracle_build(None)
        a = refcase.oracle_qp(case, rot, F)
        refcase.use_oracle_build(FLOPS)
        try:
            b = refcase.oracle_qp(case, rot, F)
        finally:
            refcase.use_oracle_build(None)
        plastic += a["n_active"] > 0
        for key in ("P", "T", "dPdF", "Fp", "s"):
            assert np.array_equal(a[key], b[key]), (k, key)
    assert plastic > 10


def test_fixture_matches_the_counting_build(flop_lib):
    with open(FIXTURE) as f:
        fx = json.load(f)
    nq = 8 * 24 ** 3
    steps = fx["assembly_pass"]
    assert len(steps) == 80 and all(s["calls"] == nq for s in steps)
    assert all(s["singular_calls"] <= s["calls"] and s["flops_to_first_singular"] <= sum(s[k] for k in ("add", "mul", "div", "sqrt", "fun"))
               for s in steps)
    s0 = steps[0]
    assert s0["singular_calls"] == 0
    per_qp0 = sum(s0[k] for k in ("add", "mul", "div", "sqrt", "fun")) / nq
    # one elastic qp (strain far below yield) through the counting build
    case = refcase.fcc_case()
    refcase.use_oracle_build(FLOPS)
    try:
        flop_lib.or_flop_reset()
        out = refcase.oracle_qp(case, np.array([0.1, -0.2, 0.3]), np.eye(3) + 1e-6 * np.arange(9).reshape(3, 3))
        c = (ctypes.c_longlong * 8)()
        assert flop_lib.or_flop_counts(ctypes.c_int(0), ctypes.c_int(1), c) == 0  # qp_update alone (no grain set-up)
        assert c[5] == 1
        elastic = sum(c[:5])
    finally:
        refcase.use_oracle_build(None)
    assert out["n_active"] == 0
    assert elastic == per_qp0, (elastic, per_qp0)
    # plastic steps cost more than the elastic one, and the post pass (stress only) less than the
    # assembly pass of the same step
    assert min(sum(s[k] for k in ("add", "mul", "div", "sqrt", "fun")) for s in steps[1:]) > per_qp0 * nq
    post = fx["post_pass"][5]
    assert post["calls"] == nq
    assert sum(post[k] for k in ("add", "mul", "div", "sqrt", "fun")) < sum(steps[23][k] for k in ("add", "mul", "div", "sqrt", "fun"))

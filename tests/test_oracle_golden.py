"""The oracle against the reference's own golden stress-strain curves (tests/golden, 6 significant digits).

The reference gates these curves at 1e-2 relative on von Mises stress/strain and SlipTotal
(tests/compile_tests/run_compile_tests.py:18,55-60).  Here the oracle is held to print precision
(1e-5 of the row's largest component) for as long as the reference's own result is reproducible,
and to the harness's metric beyond that: once rank-deficient active sets occur (increment 19 of
fcc/simpleTension) the reference's pseudo-inverse (compute_inverse_svd(0.0),
inactiveSlipRemoval.cc:26,99) inverts round-off-level singular values and its curve carries
~1e-3 noise that no other implementation can reproduce.
"""
import ctypes

import numpy as np
import pytest

import refcase


def _row_err(rows, gold, cols):
    scale = np.abs(gold[:, cols]).max(axis=1)
    return (np.abs(rows[:, cols] - gold[:, cols]).max(axis=1) / scale).max()


def _vm_stress(a):
    return np.sqrt(0.5 * ((a[:, 6] - a[:, 7]) ** 2 + (a[:, 7] - a[:, 8]) ** 2 + (a[:, 8] - a[:, 6]) ** 2
                          + 6 * (a[:, 9] ** 2 + a[:, 10] ** 2 + a[:, 11] ** 2)))


# increments before the first rank-deficient active set perturbs the reference's own curve
@pytest.mark.parametrize("folder,n", [("fcc_simpleTension", 18), ("fcc_compression", 13), ("bcc_simpleTension", 30),
                                      ("fcc_FCC_RandomOrientationBlock", 12), ("fcc_CyclicLoading", 13),
                                      ("fcc_tabularSimpleCompression", 16), ("fcc_DeformationGraidientBCs", 17)])
def test_oracle_matches_golden_to_print_precision(folder, n):
    case = refcase.load_reference_case(folder)
    out = refcase.oracle_simulate(case, n, nthreads=8)
    rows, gold = out["rows"], case.golden[:n]
    assert _row_err(rows, gold, slice(6, 12)) < 1.2e-5
    assert _row_err(rows, gold, slice(0, 6)) < 1.2e-5
    slip = gold[:, 14] > 1e-7
    if slip.any():
        assert (np.abs(rows[slip, 14] - gold[slip, 14]) / gold[slip, 14]).max() < 5e-5
    assert (rows[~slip, 14] < 1e-6).all()
    # Newton iteration count is fixed (solveNonLinearSystem.cc:14) and every linear solve converged
    assert len(out["lin_iters"]) == 4 * n and (out["lin_iters"] < case.lin_maxit).all()


def test_oracle_follows_golden_within_harness_gate_after_singular_active_sets():
    case = refcase.load_reference_case("fcc_simpleTension")
    n = 40
    lib = refcase.oracle()
    c, s = ctypes.c_longlong(), ctypes.c_longlong()
    lib.or_pinv_stats(ctypes.byref(c), ctypes.byref(s))
    out = refcase.oracle_simulate(case, n, nthreads=1)
    lib.or_pinv_stats(ctypes.byref(c), ctypes.byref(s))
    assert s.value > 0  # rank-deficient local solves do occur in this regime
    vm, vg = _vm_stress(out["rows"]), _vm_stress(case.golden[:n])
    assert (np.abs(vm - vg) / vg).max() < 3e-3  # reference harness gate: 1e-2


def test_pinv_semantics_agree_before_the_singular_regime():
    """deal.II's SVD pseudo-inverse (mode 0) and the LU-with-SVD-fallback the kernels use (mode 1)
    give the same curve while every A_PA is nonsingular (increments 1-17 of fcc/simpleTension).
    The two solvers differ by ~1e-16 per solve; ill-conditioned A_PA amplify that to ~1e-10 by
    increment 12, the same floor any two FP64 implementations of this path share."""
    case = refcase.load_reference_case("fcc_simpleTension")
    a = refcase.oracle_simulate(case, 12, pinv_mode=0, nthreads=8)["rows"]
    b = refcase.oracle_simulate(case, 12, pinv_mode=1, nthreads=8)["rows"]
    assert np.abs(a[:, 6:12] - b[:, 6:12]).max() / np.abs(b[:, 6:12]).max() < 1e-9

"""Error paths of the Newton step on the GPU, forced from the host (every kernel still runs to its
normal end; a NaN input only makes them compute NaN).

* A NaN/Inf residual or Krylov norm ends the step with PA_ERR_NONFINITE at once, as the reference's
  PETSc solve stops on KSP_DIVERGED_NANORINF (solveLinearSystem.cc:13-31); the Krylov loop exits at
  its first test instead of running to Maximum linear solver iterations, and u is not updated.
* Changing the arithmetic form between assembly and solve is refused (PA_ERR_STATE): the Jacobian
  in memory belongs to the other form.
"""
import ctypes
import time

import numpy as np
import pytest

import plasticity_amd as pa
from plasticity_amd import native
import refcase

pytestmark = pytest.mark.gpu


def _ctx(case, C=None):
    mat = pa.make_material(case.m, case.n, case.q, case.C if C is None else C, case.s0, case.h0, case.ss, case.pw, case.tol)
    pb = pa.make_problem((4, 4, 4), case.span, case.total_increments, case.delT, case.bcs, lin_maxit=5000)
    ctx = pa.Context(mat, pb)
    ctx.set_grains(np.zeros(64, np.uint32), np.array([[0.1, -0.2, 0.3]]))
    return ctx


def test_nonfinite_residual_and_krylov_norm_stop_the_step():
    case = refcase.fcc_case()
    C = case.C.copy()
    C[0, 0] = np.nan
    ctx = _ctx(case, C)
    print("nonfinite: assembly with a NaN stiffness", flush=True)
    with pytest.raises(pa.PlasticityError, match="error 6"):
        ctx.newton_iteration(0, 0)
    # the solve alone: assemble without asking for the norm, then the Krylov loop must stop at its
    # first test (iteration 0) and leave u untouched
    lib = native.lib()
    native.check(lib.pa_assemble(ctx._h, ctypes.c_int(0), ctypes.c_int(0), None))
    its = ctypes.c_int(-1)
    t0 = time.perf_counter()
    rc = lib.pa_linear_solve(ctx._h, ctypes.c_int(0), ctypes.c_int(0), ctypes.byref(its))
    dt = time.perf_counter() - t0
    assert rc == native.PA_ERR_NONFINITE, (rc, lib.pa_last_error())
    assert its.value == 0 and dt < 2.0, (its.value, dt)
    assert np.all(ctx.get_state("u") == 0.0)
    ctx.close()


def test_solve_after_a_change_of_arithmetic_form_is_refused():
    case = refcase.fcc_case()
    ctx = _ctx(case)
    ctx.set_solver(native.PA_SOLVER_CG, native.PA_PREC_FP64)
    ctx.assemble(0, 0)
    ctx.set_solver(native.PA_SOLVER_CG, native.PA_PREC_MIXED)
    its = ctypes.c_int(-1)
    rc = native.lib().pa_linear_solve(ctx._h, ctypes.c_int(0), ctypes.c_int(0), ctypes.byref(its))
    assert rc == native.PA_ERR_STATE, rc
    ctx.assemble(0, 0)  # a fresh assembly in the new form makes the solve valid again
    rn, n = ctx.newton_iteration(0, 0)
    assert np.isfinite(rn) and n > 0
    ctx.close()


def test_cooperative_newton_passes_are_refused():
    """The cooperative Newton-step qp passes are switched off (a GPU memory fault with no cause
    found); selecting them is refused and the one-lane form stays in effect."""
    ctx = _ctx(refcase.fcc_case())
    with pytest.raises(pa.PlasticityError, match="switched off"):
        ctx.set_qp_form(native.PA_QP_COOP)
    ctx.set_qp_form(native.PA_QP_LANE)
    rn, its = ctx.newton_iteration(0, 0)
    assert np.isfinite(rn) and its > 0
    ctx.close()

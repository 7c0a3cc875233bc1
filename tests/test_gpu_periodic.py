"""Periodic boundary conditions on the GPU (pa_set_periodic: the constraint lines applied around the
node-block SpMV, cp_pbc.h) against the oracle's condensed assembly of the same lines, on the
reference's fcc/periodicBCs inputs (3^3 mesh, its 32^3-voxel grain file) and on larger meshes:

* stress-strain rows against the oracle at 1e-9 (Krylov solves at 1e-14 in both) from the first
  increment up to the first one where the two runs part, which must lie at least three increments
  past first yield: after yield a qp at a round-off active-set decision makes both results depend
  on rounding (the oracle itself moves by 1e-5..1e-3 when its Krylov tolerance changes);
* at every increment, every qp's committed state (T, s, Fp) equals the oracle's qp update on the
  GPU's own inputs (F from the GPU's displacement, the GPU's history) at 1e-10, up to qps on the
  other side of a round-off decision (shown by a perturbation of F within 1e-12);
* the solution is periodic with the corner jumps at every increment (a property of the constraint
  set, exact to rounding), and a homogeneous crystal stays homogeneous with zero lateral stress;
* the tabular periodic table unloads at its time point;
* the mixed form equals FP64 to 1e-8 on the elastic increments;
* multigrid-PCG is refused with periodic BCs, and the prm reader drives the same run.
"""
import os

import numpy as np
import pytest

import plasticity_amd as pa
from plasticity_amd import model, native
import refcase

pytestmark = pytest.mark.gpu

D = os.path.join(refcase.GOLDEN, "fcc_periodicBCs")


def _per(case):
    p = case.periodic
    tab = (p["tabular_time"], p["table_time"], p["table_coef"]) if p.get("tabular") else None
    return model.make_periodic(p, p["known"], p["total"], tab)


def _context(case, solver=native.PA_SOLVER_BICG, precision=native.PA_PREC_FP64):
    mat = pa.make_material(case.m, case.n, case.q, case.C, case.s0, case.h0, case.ss, case.pw, case.tol)
    pb = pa.make_problem((case.nx, case.ny, case.nz), case.span, case.total_increments, case.delT, [],
                         max_newton=case.max_newton, lin_tol=case.lin_tol, lin_maxit=case.lin_maxit)
    ctx = pa.Context(mat, pb)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    ctx.set_solver(solver, precision)
    ctx.set_periodic(_per(case))
    return ctx


def _case(prm, n=None, lin_tol=1e-14):
    case = refcase.load_reference_case("fcc_periodicBCs", prm_name=prm)
    if n is not None:
        case.nx = case.ny = case.nz = n
        case.elem_grain = refcase.oracle_element_grains(case)
    case.lin_tol = lin_tol
    return case


def _run(ctx, n_inc):
    rows = []
    for i in range(n_inc):
        rows.append(ctx.run_increment(i)[0])
    return np.array(rows)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _run_checked(ctx, case, ref, n_inc):
    """n_inc increments on the GPU; after each, the jumps and every qp's committed state against the
    oracle's qp update on the GPU's own inputs.  Returns the rows and the first increment where
    the GPU and oracle rows part by more than 1e-9 (n_inc if never)."""
    rows, rounded = [], 0
    for inc in range(n_inc):
        Fp0, s0 = ctx.get_state("Fp"), ctx.get_state("s")
        rows.append(ctx.run_increment(inc)[0])
        u = ctx.get_state("u")
        assert _jumps(u, case.nx) < 1e-12 * np.abs(u).max(), inc
        _, _, k = refcase.check_committed_state(case, u, ctx.get_state("T"), ctx.get_state("s"), ctx.get_state("Fp"), Fp0, s0)
        rounded += k
    rows = np.array(rows)
    e = np.abs(rows[:, :12] - ref["rows"][:, :12]).max(axis=1) / np.abs(ref["rows"][:, :12]).max(axis=1)
    parted = np.nonzero(e > 1e-9)[0]
    n_same = int(parted[0]) if len(parted) else n_inc
    plastic = np.nonzero(ref["rows"][:, 14] > 0)[0]
    first_plastic = int(plastic[0]) if len(plastic) else n_inc
    print("  equal to the oracle (%.1e) through increment %d (first yield at %d), %.1e after; %d qp updates across a "
          "round-off decision" % (e[:n_same].max() if n_same else 0.0, n_same - 1, first_plastic,
                                  e[n_same:].max() if n_same < n_inc else 0.0, rounded), flush=True)
    assert n_same >= min(n_inc, first_plastic + 3)
    assert np.isfinite(rows).all()
    return rows, n_same


@pytest.mark.parametrize("solver", [0, 1], ids=["bicg", "pcg"])
@pytest.mark.parametrize("prm", ["prm_UniaxialTension.prm", "prm_simpleShear.prm", "prm_BiaxialTension.prm", "prm_pureShear.prm"])
def test_periodic_runs_match_the_oracle(prm, solver):
    n_inc = 40
    case = _case(prm)
    case.solver = solver
    print("case %s, solver %d, %d increments" % (prm, solver, n_inc), flush=True)
    ref = refcase.oracle_simulate(case, n_inc, want_state=True)
    ctx = _context(case, native.PA_SOLVER_CG if solver else native.PA_SOLVER_BICG)
    rows, n_same = _run_checked(ctx, case, ref, n_inc)
    assert rows[-1, 14] > 0  # plastic by the end
    if n_same == n_inc:
        st = ref["state"]
        assert _rel(ctx.get_state("u"), ref["u"]) < 1e-9
        assert _rel(ctx.get_state("T"), st[:, 0:9]) < 1e-9 and _rel(ctx.get_state("s"), st[:, 27:27 + case.n_slip]) < 1e-9
    ctx.close()


def _jumps(u, n):
    U = u.reshape(n + 1, n + 1, n + 1, 3)
    v0 = U[0, 0, 0]
    return max(np.abs(U[:, :, n] - U[:, :, 0] - (U[0, 0, n] - v0)).max(), np.abs(U[:, n, :] - U[:, 0, :] - (U[0, n, 0] - v0)).max(),
               np.abs(U[n, :, :] - U[0, :, :] - (U[n, 0, 0] - v0)).max())


def test_16cube_polycrystal_is_periodic_and_follows_the_oracle():
    """16^3 elements, the reference's grain file sampled at the cell centres, the prm's 5e-5 steps up
    to first yield: the GPU field is periodic with the corner jumps and equals the oracle's over the
    elastic increments (no active set, hence no round-off decision)."""
    case = _case("prm_UniaxialTension.prm", n=16)
    n_inc = 10
    ref = refcase.oracle_simulate(case, n_inc, nthreads=8, want_state=True)
    ctx = _context(case)
    rows = _run(ctx, n_inc)
    u = ctx.get_state("u")
    elastic = ref["rows"][:, 14] == 0
    e = np.abs(rows[:, :12] - ref["rows"][:, :12]).max(axis=1) / np.abs(ref["rows"][:, :12]).max(axis=1)
    print("jumps %.2e, %d elastic increments, curve error there %.2e, slip at the end %.3e"
          % (_jumps(u, 16) / np.abs(u).max(), elastic.sum(), e[elastic].max(), rows[-1, 14]), flush=True)
    assert _jumps(u, 16) < 1e-12 * np.abs(u).max()
    assert elastic.sum() >= 5
    assert e[elastic].max() < 1e-9
    assert np.isfinite(rows).all()
    ctx.close()


def test_homogeneous_crystal_stays_homogeneous():
    case = _case("prm_UniaxialTension.prm", n=6)
    case.elem_grain = np.zeros_like(case.elem_grain)
    case.grain_rod = np.array([[0.1, -0.2, 0.15]])
    ctx = _context(case)
    rows = _run(ctx, 3)
    Fe, Fp = ctx.get_state("Fe").reshape(-1, 3, 3), ctx.get_state("Fp").reshape(-1, 3, 3)
    F = Fe @ Fp
    assert np.abs(F - F[0]).max() < 1e-12
    Txx, Tyy, Tzz = rows[-1, 6:9]
    assert abs(Tyy) < 1e-9 * abs(Txx) and abs(Tzz) < 1e-9 * abs(Txx)
    ctx.close()


def test_tabular_periodic_unloading_matches_the_oracle():
    case = _case("prm_UniaxialCyclicTabular.prm", n=2)
    n_inc = 206
    ref = refcase.oracle_simulate(case, n_inc, nthreads=4, want_state=True)
    ctx = _context(case)
    rows, _ = _run_checked(ctx, case, ref, n_inc)
    dE = np.diff(rows[:, 0])
    assert (dE[:199] > 0).all() and (dE[199:] < 0).all()
    u = ctx.get_state("u")
    assert abs(u[3 * 2] - 194 * 5e-5) < 1e-15  # corner V100 follows the table: 200 steps up, 6 down
    ctx.close()


def test_mixed_form_follows_fp64():
    """FP32 tangents and matrix, FP64 residual and Newton: the elastic increments (4 Newton
    iterations converge there) equal the FP64 run to 1e-8; the run stays finite and yields."""
    case = _case("prm_UniaxialTension.prm", n=8)
    n_inc = 30
    out = []
    for prec in (native.PA_PREC_FP64, native.PA_PREC_MIXED):
        ctx = _context(case, precision=prec)
        out.append(_run(ctx, n_inc))
        ctx.close()
    elastic = out[0][:, 14] == 0
    err = np.abs(out[1][elastic, :12] - out[0][elastic, :12]).max() / np.abs(out[0][elastic, :12]).max()
    print("mixed vs FP64 on %d elastic increments: %.2e (slip at the end %.3e)" % (elastic.sum(), err, out[0][-1, 14]), flush=True)
    assert elastic.sum() >= 5 and out[0][-1, 14] > 0
    assert err < 1e-8
    assert np.isfinite(out[1]).all()


def test_multigrid_is_refused_with_periodic_bcs():
    case = _case("prm_UniaxialTension.prm", n=4)
    ctx = _context(case)
    with pytest.raises(native.PlasticityError, match="periodic"):
        ctx.set_solver(native.PA_SOLVER_MGCG)
    ctx.close()


def test_prm_driven_run_equals_the_context_run(tmp_path):
    """model.main's path (userInputParameters -> crystalPlasticity -> run) on prm_simpleShear.prm,
    first 8 increments, against the Context run of the same case."""
    ui = model.userInputParameters(os.path.join(D, "prm_simpleShear.prm"))
    ui.outputDirectory = str(tmp_path)
    cp = model.crystalPlasticity(ui)
    cp.orientations.loadOrientations(ui.file(ui.grainIDFile), ui.headerLinesGrainIDFile, ui.file(ui.grainOrientationsFile),
                                     ui.numPts, ui.span)
    cp.orientations.loadOrientationVector(ui.file(ui.grainOrientationsFile))
    rows = cp.run(8, write=False)
    case = _case("prm_simpleShear.prm", lin_tol=1e-11)
    ctx = _context(case)
    rows2 = _run(ctx, 8)
    ctx.close()
    assert _rel(rows, rows2) < 1e-12

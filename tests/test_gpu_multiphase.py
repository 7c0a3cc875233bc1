"""Several phases on the GPU (pa_create_phases / pa_set_grain_phases; multiphaseInit.cc:68-361):
each grain row's phase selects its material (slip systems, elastic moduli, hardening, latent
ratios, kinematic hardening) against the oracle's several-phase run (or_simulate_phases), which
tests/test_oracle_multiphase.py pins (one phase = the single-phase run bit for bit; the phase
selects the material).

* rows against the oracle at 1e-9 (Krylov at 1e-14) from the first increment up to the first one
  where the two runs part, at least two increments past first yield (after yield round-off
  active-set decisions make both depend on rounding);
* at every increment every qp's committed state (T, s, Fp, backstress) equals the oracle's qp
  update of that qp's phase on the GPU's own inputs at 1e-10, up to qps across a round-off decision;
* a two-phase context whose grains all take phase 1 equals the one-phase context bit for bit in
  every stress and strain column and writes TwinRealVF, TwinMade and SlipTotal as 0, as the
  reference does with Enable Multiphase (updateAfterIncrement.cc:718-725);
* the reference's multiphase example (FCC, HCP with twinning, FCC with kinematic hardening) runs as
  the oracle's several-phase run does;
* phases that disagree on the global local-loop keys, grains set before their phases and phase ids
  out of range are refused.
"""
import numpy as np
import pytest

import plasticity_amd as pa
from plasticity_amd import native, synth
import refcase

pytestmark = pytest.mark.gpu


def _phase_cases(n=4, strain=1e-3, n_inc=8):
    bcs = ((1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, strain * n_inc))
    fcc = refcase.fcc_case(nx=n, ny=n, nz=n, delT=1.0, total_time=float(n_inc), lin_tol=1e-14, bcs=bcs)
    t = synth.hcp24_tables()
    hcp = refcase.Case(t["m"], t["n"], t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=n, ny=n, nz=n,
                       delT=1.0, total_time=float(n_inc), bcs=bcs, lin_tol=1e-14)
    kin = refcase.fcc_case(nx=n, ny=n, nz=n, delT=1.0, total_time=float(n_inc), lin_tol=1e-14, bcs=bcs)
    kin.C1, kin.C2 = np.full(12, 10.0), np.full(12, 1.0)
    kin.h0 = kin.h0 * 0.8
    case = refcase.fcc_case(nx=n, ny=n, nz=n, delT=1.0, total_time=float(n_inc), lin_tol=1e-14, bcs=bcs)
    rng = np.random.default_rng(11)
    ng = 6
    case.grain_rod = rng.uniform(-0.5, 0.5, size=(ng, 3))
    case.elem_grain = rng.integers(0, ng, size=n ** 3).astype(np.uint32)
    case.phases = [fcc, hcp, kin]
    case.grain_phase = np.array([1, 2, 3, 1, 2, 3], np.int32)
    return case


def _materials(case):
    mats = []
    for c in case.phases:
        q = c.twin["q"] if c.twin else c.q
        mats.append(pa.make_material(c.m, c.n, q, c.C, c.s0, c.h0, c.ss, c.pw, c.tol, max_solver_iter=c.max_solver_iter,
                                     C1=c.C1, C2=c.C2, twin=c.twin))
    return mats


def _problem(case):
    return pa.make_problem((case.nx, case.ny, case.nz), case.span, case.total_increments, case.delT, case.bcs,
                           lin_tol=case.lin_tol)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _check_qps(case, ctx, u, Fp0, s0, W0):
    """Per qp: the committed state against the oracle's update of the qp's phase on the GPU's inputs."""
    F = refcase.qp_F(u, case.nx, case.ny, case.nz, tuple(case.span[d] / (case.nx, case.ny, case.nz)[d] for d in range(3)))
    T1, s1, Fp1, W1 = ctx.get_state("T"), ctx.get_state("s"), ctx.get_state("Fp"), ctx.get_state("W")
    rounded = 0
    for Q in range(len(F)):
        g = case.elem_grain[Q // 8]
        pc = case.phases[case.grain_phase[g] - 1]
        ns = pc.nT if pc.twin else pc.n_slip  # twin systems follow the slip systems; else the inert one
        sq = s0[Q][:ns] if pc.twin else np.append(s0[Q][:ns], 10e14)
        rod = case.grain_rod[g]
        W = W0[Q][:ns] if pc.C1 is not None else None
        o = refcase.oracle_qp(pc, rod, F[Q], Fp0[Q].reshape(3, 3), sq, 0, W=W)
        if _rel(T1[Q], o["T"].ravel()) < 1e-10:
            assert _rel(s1[Q][:ns], o["s"][:ns]) < 1e-10 and _rel(Fp1[Q], o["Fp"].ravel()) < 1e-10, Q
            if pc.C1 is not None:
                assert np.abs(W1[Q][:ns] - o["W"][:ns]).max() <= 1e-10 * max(np.abs(o["W"][:ns]).max(), 1e-3), Q
            assert not np.any(s1[Q][ns:]), Q  # the systems the phase does not have stay untouched
            continue
        r = np.random.default_rng(1000 + Q)
        near = any(_rel(T1[Q], refcase.oracle_qp(pc, rod, F[Q] * (1 + 1e-12 * r.standard_normal((3, 3))), Fp0[Q].reshape(3, 3),
                                                  sq, 0, W=W)["T"].ravel()) < 1e-6 for _ in range(400))
        assert near, Q
        rounded += 1
    assert rounded <= max(1, len(F) // 100)
    return rounded


def test_phases_match_the_oracle():
    n_inc = 8
    case = _phase_cases(n_inc=n_inc)
    ref = refcase.oracle_simulate(case, n_inc, want_state=True)
    ctx = pa.Context(_materials(case), _problem(case))
    ctx.set_grain_phases(case.grain_phase)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows, rounded, slipped = [], 0, []
    for inc in range(n_inc):
        Fp0, s0, W0 = ctx.get_state("Fp"), ctx.get_state("s"), ctx.get_state("W")
        rows.append(ctx.run_increment(inc)[0])
        rounded += _check_qps(case, ctx, ctx.get_state("u"), Fp0, s0, W0)
        slipped.append(ctx.get_state("slip").max() > 0)
    rows = np.array(rows)
    e = np.abs(rows[:, :12] - ref["rows"][:, :12]).max(axis=1) / np.abs(ref["rows"][:, :12]).max(axis=1)
    parted = np.nonzero(e > 1e-9)[0]
    n_same = int(parted[0]) if len(parted) else n_inc
    # SlipTotal reads 0 with several phases: first yield from the slip state
    first_plastic = int(np.argmax(slipped)) if any(slipped) else n_inc
    print("equal to the oracle (%.1e) through increment %d (first yield at %d); %d qp updates across a round-off decision"
          % (e[:n_same].max(), n_same - 1, first_plastic, rounded), flush=True)
    assert slipped[-1]
    assert not rows[:, 12:15].any() and not ref["rows"][:, 12:15].any()
    assert n_same >= min(n_inc, first_plastic + 2)
    if n_same == n_inc:
        st = ref["state"]
        assert _rel(ctx.get_state("u"), ref["u"]) < 1e-9 and _rel(ctx.get_state("T"), st[:, 0:9]) < 1e-9
    ctx.close()


def test_two_phase_context_with_one_phase_in_use_equals_the_one_phase_context():
    """Two FCC phases (the second with other hardening; same system count and no kinematic
    hardening, so both contexts run the same kernel instantiation), every grain in phase 1."""
    case = _phase_cases(n_inc=4)
    mats = _materials(case)
    f = case.phases[0]
    other = pa.make_material(f.m, f.n, f.q, f.C, f.s0, f.h0 * 1.2, f.ss, f.pw, f.tol)
    pb = _problem(case)
    out = []
    for multi in (False, True):
        ctx = pa.Context([mats[0], other] if multi else mats[0], pb)
        if multi:
            ctx.set_grain_phases(np.ones(len(case.grain_rod), np.int32))
        ctx.set_grains(case.elem_grain, case.grain_rod)
        out.append((np.array([ctx.run_increment(i)[0] for i in range(4)]), ctx.get_state("T"), ctx.get_state("s")[:, :12]))
        ctx.close()
    (r1, T1, s1), (r2, T2, s2) = out
    assert r1[-1, 14] > 0
    assert np.array_equal(r1[:, :12], r2[:, :12]) and np.array_equal(T1, T2) and np.array_equal(s1, s2)
    assert not r2[:, 12:15].any()  # Enable Multiphase zeroes them (updateAfterIncrement.cc:718-725)


def test_bad_phase_inputs_are_refused():
    case = _phase_cases(n_inc=2)
    mats = _materials(case)
    pb = _problem(case)
    f = case.phases[0]
    other = pa.make_material(f.m, f.n, f.q, f.C, f.s0, f.h0, f.ss, f.pw, f.tol, max_solver_iter=4)
    with pytest.raises(native.PlasticityError, match="share"):
        pa.Context([mats[0], other], pb)
    ctx = pa.Context(mats, pb)
    with pytest.raises(native.PlasticityError, match="pa_set_grain_phases"):
        ctx.set_grains(case.elem_grain, case.grain_rod)
    with pytest.raises(native.PlasticityError, match="out of range"):
        ctx.set_grain_phases(np.array([1, 2, 4, 1, 2, 3], np.int32))
    ctx.close()


def test_reference_multiphase_example_matches_the_oracle():
    """fcc/simpleCompression_Multiphase (4 x 4 x 12 elements; grains 1-3 along z in phases 1-3: FCC,
    HCP with 18 slip and 6 twin systems, FCC with kinematic hardening; Max Solver Iterations 4),
    40 increments of compression through first yield: rows against the oracle's several-phase run,
    and the last increment per qp."""
    case = refcase.load_reference_case("fcc_simpleCompression_Multiphase", twinning=True)
    case.lin_tol = 1e-14
    n_inc = 40
    ref = refcase.oracle_simulate(case, n_inc, nthreads=8, want_state=True)
    ctx = pa.Context(_materials(case), _problem(case))
    ctx.set_grain_phases(case.grain_phase)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows, slipped = [], []
    for i in range(n_inc):
        if i == n_inc - 1:
            Fp0, s0, W0 = ctx.get_state("Fp"), ctx.get_state("s"), ctx.get_state("W")
        rows.append(ctx.run_increment(i)[0])
        slipped.append(ctx.get_state("slip").max() > 0)
    rows = np.array(rows)
    e = np.abs(rows[:, :12] - ref["rows"][:, :12]).max(axis=1) / np.abs(ref["rows"][:, :12]).max(axis=1)
    parted = np.nonzero(e > 1e-9)[0]
    n_same = int(parted[0]) if len(parted) else n_inc
    first_plastic = int(np.argmax(slipped)) if any(slipped) else n_inc
    print("equal to the oracle (%.1e) through increment %d (first yield at %d); largest twin fraction %.3e"
          % (e[:n_same].max(), n_same - 1, first_plastic, ctx.get_state("twin_fraction").max()), flush=True)
    assert first_plastic < n_inc and n_same >= min(n_inc, first_plastic + 2)
    # TwinRealVF, TwinMade and SlipTotal read 0 with Enable Multiphase, as in the golden's 1,000 rows
    assert not rows[:, 12:15].any() and not ref["rows"][:, 12:15].any()
    _check_qps(case, ctx, ctx.get_state("u"), Fp0, s0, W0)
    ctx.close()


def test_prm_driven_multiphase_run_equals_the_context_run(tmp_path):
    """model.main's path on fcc/simpleCompression_Multiphase's prm.prm (the phases from the prm keys
    with suffixes, the grain phases from the orientation file's fifth column), 10 increments, against
    the Context run built from the oracle's reader of the same files."""
    import os
    from plasticity_amd import model
    d = os.path.join(refcase.GOLDEN, "fcc_simpleCompression_Multiphase")
    ui = model.userInputParameters(os.path.join(d, "prm.prm"))
    cp = model.crystalPlasticity(ui)
    cp.orientations.loadOrientations(ui.file(ui.grainIDFile), ui.headerLinesGrainIDFile, ui.file(ui.grainOrientationsFile),
                                     ui.numPts, ui.span)
    cp.orientations.loadOrientationVector(ui.file(ui.grainOrientationsFile), True)
    rows = cp.run(10, write=False)
    case = refcase.load_reference_case("fcc_simpleCompression_Multiphase", twinning=True)
    ctx = pa.Context(_materials(case), _problem(case))
    ctx.set_grain_phases(case.grain_phase)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    rows2 = np.array([ctx.run_increment(i)[0] for i in range(10)])
    ctx.close()
    assert _rel(rows, rows2) < 1e-12



def test_no_reorientation_by_twinning_with_several_phases():
    """hcp/simpleTension with a Twin Threshold Fraction of 1e-6, 42 increments: the one-phase
    context reorients qps by twinning; the same material as a two-phase context (every grain in
    phase 1) reorients none, as the reference does with Enable Multiphase (F_r stays zero,
    updateAfterIncrement.cc:718-725), and the two runs are equal until the one-phase run's first
    reorientation."""
    from types import SimpleNamespace
    case = refcase.load_reference_case("hcp_simpleTension", twinning=True)
    case.twin = dict(case.twin)
    case.twin["threshold"], case.twin["saturation"] = 1e-6, 0.0
    n = 42
    out = []
    for phases in ([case], [case, case]):
        mats = _materials(SimpleNamespace(phases=phases))
        ctx = pa.Context(mats if len(phases) > 1 else mats[0], _problem(case))
        if len(phases) > 1:
            ctx.set_grain_phases(np.ones(len(case.grain_rod), np.int32))
        ctx.set_grains(case.elem_grain, case.grain_rod)
        rows = np.array([ctx.run_increment(i)[0] for i in range(n)])
        out.append((rows, int(ctx.get_state("twinned").sum()), float(ctx.get_state("twin_fraction").max())))
        ctx.close()
    (r1, tw1, _), (r2, tw2, f2) = out
    e = np.abs(r1[:, :12] - r2[:, :12]).max(axis=1) / np.abs(r1[:, :12]).max(axis=1)
    parted = np.nonzero(e > 1e-12)[0]
    print("one phase: %d qps reoriented; two phases: %d (largest twin fraction %.2e); rows part at increment %s"
          % (tw1, tw2, f2, parted[:1]), flush=True)
    assert tw1 > 0 and tw2 == 0 and f2 > 1e-4
    assert not r2[:, 12:15].any()
    assert len(parted) and parted[0] >= 38  # equal until the one-phase run's first reorientation

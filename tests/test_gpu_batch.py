"""Batched RVEs on the GPU (pa_batch_*): RVEs of uneven size and different materials (FCC 12, HCP 24
and BCC 48 slip systems) advanced in the same launches, each checked against the oracle run on that
RVE alone, with the Krylov solves (Jacobi-BiCG, the reference's, and Jacobi-PCG) at 1e-14.

* curves against each RVE's own oracle run at 1e-9 where the oracle's result does not move with
  its Krylov tolerance, and every RVE's committed per-qp state against the oracle's qp update on
  the batch's own inputs at 1e-10;
* graph replay and direct launches give bitwise the same rows, Krylov counts and state at every
  Newton iteration (each iteration replays the captured graphs again);
* the autotuner's launch choices leave every result bitwise unchanged.
"""
import numpy as np
import pytest

import plasticity_amd as pa
from plasticity_amd import native, synth
import refcase

pytestmark = pytest.mark.gpu

N_INC = 5


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _tables():
    fcc = synth.fcc_tables()
    fcc["m"] = fcc["m"] / np.linalg.norm(fcc["m"], axis=1)[:, None]
    fcc["n"] = fcc["n"] / np.linalg.norm(fcc["n"], axis=1)[:, None]
    return [fcc, synth.hcp24_tables(), synth.bcc48_tables()]


# (material, nx, ny, nz, grains)
RVES = [(0, 4, 4, 4, 4), (1, 3, 4, 5, 3), (2, 3, 3, 3, 2), (0, 5, 3, 4, 5), (1, 4, 4, 4, 4)]


def _setup(lin_tol=1e-14, solver=1):
    tabs = _tables()
    bcs = [(1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, 1e-3 * N_INC)]
    cases, grains = [], []
    rng = np.random.default_rng(31)
    for k, nx, ny, nz, ng in RVES:
        t = tabs[k]
        c = refcase.Case(t["m"], t["n"], t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=nx, ny=ny, nz=nz,
                         total_time=float(N_INC), bcs=bcs, lin_tol=lin_tol, solver=solver)
        c.grain_rod = rng.uniform(-0.5, 0.5, size=(ng, 3))
        c.elem_grain = rng.integers(0, ng, size=nx * ny * nz).astype(np.uint32)
        cases.append(c)
    mats = [pa.make_material(t["m"], t["n"], t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], t["tol"]) for t in tabs]
    pb = pa.make_problem((1, 1, 1), (1.0, 1.0, 1.0), N_INC, 1.0, bcs, lin_tol=lin_tol)
    return mats, pb, cases


def _batch(mats, pb, cases, graphs=True, solver=1):
    b = pa.Batch(mats, [r[0] for r in RVES], [r[1:4] for r in RVES], pb)
    for i, c in enumerate(cases):
        b.set_grains(i, c.elem_grain, c.grain_rod)
    b.set_graphs(graphs)
    b.set_solver(native.PA_SOLVER_CG if solver == 1 else native.PA_SOLVER_BICG, native.PA_PREC_FP64)
    return b


def _run(b, n_inc):
    rows, its = [], []
    for inc in range(n_inc):
        row, it, _ = b.run_increment(inc)
        rows.append(row)
        its.append(it)
    return np.array(rows), np.array(its)


@pytest.mark.parametrize("solver", [0, 1], ids=["bicg", "pcg"])
def test_batch_rves_match_their_own_oracle_runs(solver):
    """Every RVE against the oracle on that RVE alone.
    * Curves: the two HCP RVEs do not move when the oracle's Krylov tolerance goes from 1e-14 to
      3e-15, 1e-12 or 1e-11 (spread <= 1e-11) and are held to 1e-9 over all increments, with their
      per-qp state.  The FCC and BCC-48 RVEs take borderline active-set decisions at their first
      plastic Newton iterations and move by 1e-2..1 under those changes (the GPU's solves stop on
      the same test but sum in another order); their curves are checked through the per-qp pass
      below instead.
    * Per qp, every RVE: after increment 1 the batch's committed state (T, s, Fp) at every qp equals
      the oracle's calculatePlasticity(.,.,0) on the batch's own inputs at that qp (F from the
      batch's u, Fp and s committed after increment 0) at 1e-10.  A qp whose numerically singular
      active set leaves one singular value at round-off may differ: the keep/drop decision of
      its pseudo-inverse rests on rounding (one FCC qp of this run: the oracle drops it, the kernel
      code keeps it, on the GPU and compiled for the host alike, and the oracle itself keeps it
      under 6 % of random 1e-12 relative perturbations of F).  Such a qp passes only if the GPU's
      state is the oracle's at an input within 1e-12 of the GPU's (found among 400 seeded
      perturbations), and at most 1 % of qps may need it.
    """
    mats, pb, cases = _setup(solver=solver)
    others = [_setup(lin_tol=t, solver=solver)[2] for t in (3e-15, 1e-12, 1e-11)]
    b = _batch(mats, pb, cases, solver=solver)
    print("batch: %d RVEs, %d qps" % (b.n_rve, b.nqp), flush=True)
    rows = [b.run_increment(0)[0]]
    before = [(b.get_state(i, "Fp"), b.get_state(i, "s")) for i in range(len(cases))]
    rows.append(b.run_increment(1)[0])
    after = [(b.get_state(i, "T"), b.get_state(i, "s"), b.get_state(i, "Fp"), b.get_state(i, "u")) for i in range(len(cases))]
    for inc in range(2, N_INC):
        rows.append(b.run_increment(inc)[0])
    rows = np.array(rows)
    stable = 0
    for i, c in enumerate(cases):
        ref = refcase.oracle_simulate(c, N_INC, want_state=True)
        scale = np.abs(ref["rows"][:, :12]).max(axis=1)
        e = np.abs(rows[:, i, :12] - ref["rows"][:, :12]).max(axis=1) / scale
        spread = np.zeros(N_INC)
        for o in others:
            r2 = refcase.oracle_simulate(o[i], N_INC)["rows"]
            spread = np.maximum(spread, np.abs(r2[:, :12] - ref["rows"][:, :12]).max(axis=1) / scale)
        print("RVE %d (%d systems): curve error per increment %s, oracle spread %s" % (
            i, c.n_slip, " ".join("%.1e" % v for v in e), " ".join("%.1e" % v for v in spread)), flush=True)
        assert np.all(np.isfinite(rows[:, i]))
        assert rows[-1, i, 14] > 0  # every RVE has gone plastic
        if spread.max() < 1e-11:
            stable += 1
            assert e.max() < 1e-9
            assert _rel(rows[:, i, 14], ref["rows"][:, 14]) < 1e-8
            st = ref["state"]
            assert _rel(b.get_state(i, "T"), st[:, 0:9]) < 1e-9
            assert _rel(b.get_state(i, "s"), st[:, 27:27 + c.n_slip]) < 1e-9
            assert _rel(b.get_state(i, "u"), ref["u"]) < 1e-9
        # per qp, increment 1, on the batch's own inputs
        Fp0, s0 = before[i]
        T1, s1, Fp1, u1 = after[i]
        F = refcase.qp_F(u1, c.nx, c.ny, c.nz, (1.0 / c.nx, 1.0 / c.ny, 1.0 / c.nz))
        nq, skipped, plastic = len(F), 0, 0
        for Q in range(nq):
            rod = c.grain_rod[c.elem_grain[Q // 8]]
            Fpq = Fp0[Q].reshape(3, 3)
            sq = np.append(s0[Q], 10e14)
            o = refcase.oracle_qp(c, rod, F[Q], Fpq, sq, 0)
            plastic += o["n_active"] > 0
            if _rel(T1[Q], o["T"].ravel()) < 1e-10:
                assert _rel(s1[Q], o["s"][:c.n_slip]) < 1e-10, (i, Q)
                assert _rel(Fp1[Q], o["Fp"].ravel()) < 1e-10, (i, Q)
                continue
            r = np.random.default_rng(1000 + Q)
            near = False
            for _ in range(400):
                o2 = refcase.oracle_qp(c, rod, F[Q] * (1 + 1e-12 * r.standard_normal((3, 3))), Fpq, sq, 0)
                if _rel(T1[Q], o2["T"].ravel()) < 1e-6 and _rel(s1[Q], o2["s"][:c.n_slip]) < 1e-6:
                    near = True
                    break
            assert near, (i, Q, _rel(T1[Q], o["T"].ravel()))
            skipped += 1
        print("  per-qp pass of increment 1: %d qps, %d plastic, %d on the other side of a round-off decision" % (nq, plastic, skipped),
              flush=True)
        assert skipped <= max(1, nq // 100)
    assert stable >= 2
    b.close()


@pytest.mark.parametrize("solver", [0, 1], ids=["bicg", "pcg"])
def test_batch_graph_replay_equals_direct_launches_bitwise(solver):
    mats, pb, cases = _setup(lin_tol=1e-10)
    out = []
    for graphs in (True, False):
        b = _batch(mats, pb, cases, graphs, solver)
        rows, its = _run(b, 3)
        out.append((rows, its, [b.get_state(i, "s") for i in range(len(RVES))]))
        b.close()
    (r1, i1, s1), (r2, i2, s2) = out
    assert np.array_equal(r1, r2) and np.array_equal(i1, i2)
    for a, c in zip(s1, s2):
        assert np.array_equal(a, c)


def test_batch_autotuner_choices_leave_results_unchanged():
    mats, pb, cases = _setup(lin_tol=1e-10)
    b = _batch(mats, pb, cases)
    rep = b.tune()
    print("tuner:", rep, flush=True)
    assert rep["grid"] in (256, 512, 768, 1024) and rep["chunk"] in (8, 16, 32)
    assert all(v > 0 for v in rep["qp_ms"].values())
    rows_t, its_t = _run(b, 3)
    b.close()
    b = _batch(mats, pb, cases)
    rows, its = _run(b, 3)
    b.close()
    assert np.array_equal(rows, rows_t) and np.array_equal(its, its_t)

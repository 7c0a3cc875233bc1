"""GPU tests of the grain-averaged output (updateAfterIncrement.cc:611-707): per-grain volumes and
volume-averaged Cauchy stresses against the oracle's per-qp state (1e-9 relative to the largest
stress, volumes to rounding), and the GrainAverage<increment>.csv files the host writes from a prm.
"""
import os
import shutil

import numpy as np
import pytest

import plasticity_amd as pa
import refcase

pytestmark = pytest.mark.gpu


def test_grain_averages_match_oracle_state():
    """Two increments of a 4^3 FCC tension case whose GPU state matches the oracle's to ~1e-12 (the
    smoke case, grain ids shifted to start at 1, one more id without elements)."""
    case = refcase.fcc_case(nx=4, ny=4, nz=4, delT=1.0, total_time=10.0, lin_tol=1e-14,
                            bcs=((1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, 0.01)))
    rng = np.random.default_rng(7)
    rod = rng.uniform(-0.4, 0.4, size=(4, 3))
    eg = rng.integers(0, 4, size=64).astype(np.uint32)
    case.grain_rod = np.vstack([np.zeros(3), rod, np.zeros(3)])  # rows by grain id 0..5
    case.elem_grain = eg + 1                                      # ids 1..4; grain 5 has no element
    n_inc = 2
    ref = refcase.oracle_simulate(case, n_inc, want_state=True)
    T_or = ref["state"][:, :9]
    mat = pa.make_material(case.m, case.n, case.q, case.C, case.s0, case.h0, case.ss, case.pw, case.tol)
    pb = pa.make_problem((4, 4, 4), case.span, case.total_increments, case.delT, case.bcs, lin_tol=case.lin_tol)
    ctx = pa.Context(mat, pb)
    ctx.set_grains(case.elem_grain, case.grain_rod)
    for i in range(n_inc):
        ctx.run_increment(i)
    ga = ctx.grain_averages(5)
    T = ctx.get_state("T")
    ctx.close()
    scale = np.abs(T_or).max()
    assert np.abs(T - T_or).max() <= 1e-9 * scale  # the states the averages are taken over
    jxw = (1.0 / 4) ** 3 / 8
    qgrain = np.repeat(case.elem_grain, 8)
    comps = [0, 4, 8, 1, 2, 5]  # Txx Tyy Tzz Txy Txz Tyz
    for g in range(1, 6):
        sel = qgrain == g
        vol = sel.sum() * jxw
        assert abs(ga[g - 1, 0] - vol) <= 1e-15
        if vol > 0:
            assert np.abs(ga[g - 1, 1:] - (T[sel][:, comps] * jxw).sum(axis=0) / vol).max() <= 1e-13 * scale, g
            assert np.abs(ga[g - 1, 1:] - (T_or[sel][:, comps] * jxw).sum(axis=0) / vol).max() <= 1e-9 * scale, g
        else:
            assert not ga[g - 1].any()
    assert abs(ga[:, 0].sum() - 1.0) < 1e-14  # the unit cube
    # volume-weighted mean of the grain averages = the body average
    body = (ga[:, 1:] * ga[:, :1]).sum(axis=0)
    assert np.abs(body - T[:, comps].mean(axis=0)).max() <= 1e-12 * scale


def test_grain_average_files_from_prm(tmp_path):
    src = os.path.join(refcase.GOLDEN, "fcc_simpleTension")
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    with open(tmp_path / "prm.prm", "a") as f:
        f.write("\nset Write Grain Averaged Output = true\nset Skip Grain Averaged Output Steps = 2\n"
                "set Number of Grain Averaged Output Variables = 7\nset Output Directory = out\n")
    ui = pa.userInputParameters(str(tmp_path / "prm.prm"))
    problem = pa.crystalPlasticity(ui)
    problem.orientations.loadOrientations(ui.file(ui.grainIDFile), ui.headerLinesGrainIDFile,
                                          ui.file(ui.grainOrientationsFile), ui.numPts, ui.span)
    problem.orientations.loadOrientationVector(ui.file(ui.grainOrientationsFile))
    problem.run(n_increments=4)
    ga = problem.ctx.grain_averages(problem.orientations.numberOfGrains)
    out = tmp_path / "out"
    assert sorted(p for p in os.listdir(out) if p.startswith("GrainAverage")) == ["GrainAverage1.csv", "GrainAverage3.csv"]
    rows = [line.split(",") for line in open(out / "GrainAverage3.csv").read().splitlines()]
    assert len(rows) == problem.orientations.numberOfGrains and all(len(r) == 7 for r in rows)
    got = np.array([[float(v) for v in r] for r in rows])
    assert np.allclose(got, ga, rtol=1e-8, atol=0)  # setprecision(9)
    assert abs(got[:, 0].sum() - np.prod(ui.span)) <= 1e-8 * np.prod(ui.span)

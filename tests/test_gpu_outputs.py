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


def test_quadrature_output_file_matches_state(tmp_path):
    """QuadratureOutputs<inc>.csv of the fcc/simpleTension inputs (8^3 by refine factor 3, Krylov
    solves to 1e-14): 127 columns, one row per qp in deal.II cell order; stresses, Fe, Fp,
    orientation and slip equal the oracle's per-qp state to print precision (%8.5e)."""
    src = os.path.join(refcase.GOLDEN, "fcc_simpleTension")
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    with open(tmp_path / "prm.prm", "a") as f:
        f.write("\nset Write Quadrature Output = true\nset Skip Quadrature Output Steps = 2\nset Output Directory = out\n"
                "set Relative linear solver tolerance = 1.0e-14\n")
    ui = pa.userInputParameters(str(tmp_path / "prm.prm"))
    problem = pa.crystalPlasticity(ui)
    problem.orientations.loadOrientations(ui.file(ui.grainIDFile), ui.headerLinesGrainIDFile,
                                          ui.file(ui.grainOrientationsFile), ui.numPts, ui.span)
    problem.orientations.loadOrientationVector(ui.file(ui.grainOrientationsFile))
    problem.run(n_increments=2)
    out = tmp_path / "out"
    assert sorted(p for p in os.listdir(out) if p.startswith("Quadrature")) == ["QuadratureOutputs1.csv"]
    text = open(out / "QuadratureOutputs1.csv").read().splitlines()
    rows = np.array([[float(v) for v in line.split(",")[:-1]] for line in text])
    nel = (problem.problem.nx, problem.problem.ny, problem.problem.nz)
    nq = 8 * int(np.prod(nel))
    assert rows.shape == (nq, 127) and all(line.endswith(",") for line in text)
    case = refcase.load_reference_case("fcc_simpleTension")
    case.lin_tol = 1e-14
    st = refcase.oracle_simulate(case, 2, want_state=True)["state"]
    order = pa.model.dealii_cell_order(ui.subdivisions, ui.meshRefineFactor)
    Q = (8 * order[:, None] + np.arange(8)[None, :]).ravel()
    mix = [0, 4, 8, 1, 2, 3, 5, 6, 7]
    nT = case.nT
    want = {"grain": np.repeat(case.elem_grain[order], 8), "T": st[Q][:, mix], "Fp": st[Q][:, 9 + np.array(mix)],
            "Fe": st[Q][:, 18 + np.array(mix)], "rot": st[Q][:, 27 + nT:30 + nT], "slip": st[Q][:, 30 + nT:30 + nT + 12]}
    close = lambda a, b: np.allclose(a, b, rtol=2e-5, atol=1e-12 * max(1.0, np.abs(b).max()))
    assert np.array_equal(rows[:, 0], want["grain"])
    assert (rows[:, 1] == 1).all() and (rows[:, 3] == 0).all()
    assert close(rows[:, 2], np.prod(np.array(ui.span) / np.array(nel)) / 8)
    assert close(rows[:, 7:10], want["rot"])
    assert close(rows[:, 10:19], want["Fe"]) and close(rows[:, 19:28], want["Fp"])
    assert np.allclose(rows[:, 28:37], want["T"], rtol=2e-5, atol=1e-6 * np.abs(want["T"]).max())
    assert np.allclose(rows[:, 37:49], want["slip"], rtol=2e-5, atol=1e-9)
    assert not rows[:, 49:].any()
    # qp coordinates lie inside their cells, cell by cell in the given order
    h = np.array(ui.span) / np.array(nel)
    e = np.repeat(order, 8)
    cell = np.stack([e % nel[0], (e // nel[0]) % nel[1], e // (nel[0] * nel[1])], axis=1)
    assert (np.floor(rows[:, 4:7] / h + 1e-9).astype(int) == cell).all()

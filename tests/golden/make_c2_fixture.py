"""Writes tests/golden/c2_fcc32_oracle.npz: the oracle on the 32^3 FCC tension workload (32,768
elements, 262,144 qps, 64 seeded Voronoi grains, 5e-4 strain per step, FP64, Jacobi-PCG at 1e-10
relative on the preconditioned residual, 4 Newton iterations per step) for its first two increments,
at the workload's Krylov tolerance and at a 3x tighter one.  The plastic steps make the 4-iteration
Newton stall, so the two runs differ by the oracle's own sensitivity to round-off in the solves;
the GPU test scales its gate by that difference.

Stored (all float64 / int arrays, no pickles): rows [tol][inc][15], lin_iters and res_norms
[tol][inc*4], the per-qp state of a fixed sample of qps (every 4099th: T, Fp, s) and per-component
statistics over all qps (mean, standard deviation, 1/50/99th percentiles of T and s).

Run from the repository root: python tests/golden/make_c2_fixture.py  (about 10 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import refcase  # noqa: E402
from plasticity_amd import synth  # noqa: E402

N, GRAINS, STEPS, EPS, SEED = 32, 64, 20, 5e-4, 2024
N_INC = 2
TOLS = (1e-10, 3.3e-11)
SAMPLE_STRIDE = 4099


def c2_case(lin_tol):
    _, _, vox, rod, bcs, t = synth.fcc_workload(N, GRAINS, STEPS, EPS, seed=SEED)
    case = refcase.Case(t["m"] / np.linalg.norm(t["m"], axis=1)[:, None], t["n"] / np.linalg.norm(t["n"], axis=1)[:, None],
                        t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=N, ny=N, nz=N, delT=1.0,
                        total_time=float(STEPS), bcs=bcs, lin_tol=lin_tol, solver=1)
    case.elem_grain = vox.transpose(2, 1, 0).ravel().astype(np.uint32)
    case.grain_rod = rod
    return case


def main():
    out = {"n": N, "grains": GRAINS, "steps": STEPS, "eps": EPS, "seed": SEED, "tols": np.array(TOLS)}
    rows, its, res = [], [], []
    for k, tol in enumerate(TOLS):
        t0 = time.time()
        o = refcase.oracle_simulate(c2_case(tol), N_INC, nthreads=os.cpu_count() or 1, want_state=True)
        print("tol %.1e: %.0f s" % (tol, time.time() - t0), flush=True)
        rows.append(o["rows"])
        its.append(o["lin_iters"])
        res.append(o["res_norms"])
        if k == 0:
            st = o["state"]
            nT = 13
            T, Fp, s = st[:, 0:9], st[:, 9:18], st[:, 27:27 + 12]
            sample = np.arange(0, len(st), SAMPLE_STRIDE)
            out["sample_qp"] = sample
            out["sample_T"], out["sample_Fp"], out["sample_s"] = T[sample], Fp[sample], s[sample]
            for name, a in (("T", T), ("s", s)):
                out["stat_%s" % name] = np.stack([a.mean(0), a.std(0), np.percentile(a, 1, axis=0), np.percentile(a, 50, axis=0),
                                                  np.percentile(a, 99, axis=0)])
            out["u_norm"] = np.array([np.linalg.norm(o["u"])])
    out["rows"], out["lin_iters"], out["res_norms"] = np.array(rows), np.array(its), np.array(res)
    np.savez_compressed(os.path.join(HERE, "c2_fcc32_oracle.npz"), **out)
    print("rows at %.1e:" % TOLS[0], out["rows"][0])
    print("row difference between tolerances:", np.abs(out["rows"][0] - out["rows"][1]).max(axis=1) / np.abs(out["rows"][0]).max(axis=1))


if __name__ == "__main__":
    main()

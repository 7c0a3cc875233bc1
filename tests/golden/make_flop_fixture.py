"""Writes tests/golden/qp_flops_fcc24.json: the algorithmic FP64 flops of the qp update, counted by
the flop-counting build of the oracle (oracle/liboracle_flops.so: the oracle's source over a scalar
that counts + - * / sqrt and functions) on a 24^3 sample of the 64^3 FCC tension workload (the same
material, 16 seeded Voronoi grains for ~860 elements per grain against ~1,020 at 64^3 with 256,
random orientations, 5e-4 strain per step, 4 Newton iterations per step, Jacobi-PCG 1e-10), all 20
increments, single-threaded.

Per Newton step (increment * 4 + iteration) of the assembly pass (return map + consistent tangent
at every qp): qp calls, add/sub, mul, div, sqrt, function counts, the flops up to the first
numerically singular local system (what a pass that hands singular systems to a second kernel
computes) and the calls that met one.  Per increment: the same for the post-increment pass
(stress only).

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_flop_fixture.py       (about 10 minutes)
"""
import ctypes
import json
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

N, GRAINS, STEPS, EPS, SEED = 24, 16, 20, 5e-4, 2024
KEYS = ["calls", "add", "mul", "div", "sqrt", "fun", "flops_to_first_singular", "singular_calls"]


def flop_lib():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle_flops.so"))
    lib.or_simulate.restype = ctypes.c_int
    lib.or_flop_counts.restype = ctypes.c_int
    return lib


def counts(lib, step, pass_):
    c = (ctypes.c_longlong * 8)()
    assert lib.or_flop_counts(ctypes.c_int(step), ctypes.c_int(pass_), c) == 0
    c = list(c)
    return dict(zip(KEYS, [c[5], c[0], c[1], c[2], c[3], c[4], c[6], c[7]]))


def main():
    _, _, vox, rod, bcs, t = synth.fcc_workload(N, GRAINS, STEPS, EPS, seed=SEED)
    case = refcase.Case(t["m"] / np.linalg.norm(t["m"], axis=1)[:, None], t["n"] / np.linalg.norm(t["n"], axis=1)[:, None],
                        t["q"], t["C"], t["s0"], t["h0"], t["ss"], t["pw"], tol=t["tol"], nx=N, ny=N, nz=N, delT=1.0,
                        total_time=float(STEPS), bcs=bcs, solver=1)
    eg = np.ascontiguousarray(vox.transpose(2, 1, 0).ravel(), dtype=np.uint32)
    rod = np.ascontiguousarray(rod, float)
    mat, pb = case.material(1), case.problem(STEPS, 1)
    rows = np.zeros((STEPS, 15))
    its = np.zeros(STEPS * case.max_newton, np.int32)
    res = np.zeros(STEPS * case.max_newton)
    lib = flop_lib()
    lib.or_flop_reset()
    t0 = time.time()
    P = refcase.ptr
    rc = lib.or_simulate(ctypes.byref(mat), ctypes.byref(pb), P(eg), P(rod), ctypes.c_int(len(rod)), P(rows), P(its), P(res),
                         None, None)
    assert rc == 0, rc
    nsteps = STEPS * case.max_newton
    out = {
        "workload": "FCC uniaxial tension sample: %d^3 Q1 hex (%d qps), %d grains, seed %d, %d x %.0e strain, %d Newton "
                    "iterations/step, Jacobi-PCG 1e-10" % (N, 8 * N ** 3, GRAINS, SEED, STEPS, EPS, case.max_newton),
        "counting": "+ - * / 1 flop each, sqrt 1, exp/log/pow/tgamma 1 each ('fun'); comparisons, fabs, copies 0",
        "max_newton": case.max_newton,
        "assembly_pass": [counts(lib, s, 1) for s in range(nsteps)],
        "post_pass": [counts(lib, inc * case.max_newton + case.max_newton - 1, 0) for inc in range(STEPS)],
        "krylov_iterations": its.tolist(),
        "seconds": time.time() - t0,
    }
    with open(os.path.join(HERE, "qp_flops_fcc24.json"), "w") as f:
        json.dump(out, f, indent=0)
    for s, c in enumerate(out["assembly_pass"]):
        tot = c["add"] + c["mul"] + c["div"] + c["sqrt"] + c["fun"]
        print("step %2d: %.0f flops/qp, %.0f to the first singular system, %d singular" %
              (s, tot / c["calls"], c["flops_to_first_singular"] / c["calls"], c["singular_calls"]))
    print("%.0f s" % out["seconds"])


if __name__ == "__main__":
    main()

"""Test infrastructure: drives the CPU emulation of the wavefront-cooperative qp kernels
(libemu_qp.so, built with AddressSanitizer) and compares it with the oracle.  Run as a child process
with the ASan runtime preloaded (tests/test_emu_coop.py does this); prints one JSON line per check.

  python run_emu.py batch <case> <strain> <n> <seed> [kin]   per-qp update of random states, flag 0 and 1
  python run_emu.py pass  <case> <strain> <nx> <seed> [kin]  the assembly and post-increment passes on an
                                                             nx^3 mesh with a random displacement field
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import plasticity_amd as pa  # noqa: E402
import refcase  # noqa: E402


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    L = ctypes.CDLL(os.path.join(HERE, "libemu_qp.so"))
    L.emu_qp_batch.restype = ctypes.c_int
    L.emu_qp_pass.restype = ctypes.c_int
    return L


def make_case(name):
    if name == "fcc":
        return refcase.fcc_case()
    if name == "hcp":
        case = refcase.load_reference_case("hcp_simpleTension")
        return case
    if name == "bcc48":
        return bcc48_case()
    raise ValueError(name)


def bcc48_case():
    """BCC with the 48 {110}, {112}, {123} <111> systems (C11/C12/C44 = 231/135/117 GPa), latent ratio
    1 on coplanar systems and 1.4 otherwise (s0 30, h0 400, s_s 150, a 1.5 MPa)."""
    dirs, norms = [], []
    b111 = [np.array(v, float) for v in ((1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1))]
    fams = [(1, 1, 0), (1, 1, 2), (1, 2, 3)]
    import itertools
    for fam in fams:
        planes = set()
        for p in itertools.permutations(fam):
            for sg in itertools.product((1, -1), repeat=3):
                v = tuple(a * b for a, b in zip(p, sg))
                if tuple(-x for x in v) not in planes:
                    planes.add(v)
        for pl in sorted(planes):
            for d in b111:
                if abs(np.dot(pl, d)) < 1e-12:
                    dirs.append(d)
                    norms.append(np.array(pl, float))
    m, n = np.array(dirs), np.array(norms)
    ns = len(m)
    q = np.array([[1.0 if np.allclose(n[a], n[b]) else 1.4 for b in range(ns)] for a in range(ns)])
    C = np.zeros((6, 6))
    C[:3, :3] = 135e3
    np.fill_diagonal(C[:3, :3], 231e3)
    C[3, 3] = C[4, 4] = C[5, 5] = 117e3
    return refcase.Case(m, n, q, C, [30.0] * ns, [400.0] * ns, [150.0] * ns, [1.5] * ns)


def material(case):
    return pa.make_material(case.m, case.n, case.q, case.C, case.s0, case.h0, case.ss, case.pw, case.tol,
                            case.max_slip_search, case.max_solver_iter, case.C1, case.C2)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def random_states(case, n, rng, strain):
    F = np.eye(3) + rng.normal(scale=strain, size=(n, 3, 3))
    Fp = np.eye(3) + rng.normal(scale=strain / 5, size=(n, 3, 3))
    Fp /= np.cbrt(np.linalg.det(Fp))[:, None, None]
    s = np.array([rng.uniform(1.0, 3.5) * case.s0 for _ in range(n)])
    s = np.minimum(s, 0.95 * case.ss)  # below saturation (the hardening law's base stays positive)
    return F, Fp, s


def check_batch(case, strain, n, seed, kin):
    L = lib()
    rng = np.random.default_rng(seed)
    ns = case.n_slip
    if kin:
        case.C1, case.C2 = np.full(ns, 200.0), np.full(ns, 1.0)
    mat = material(case)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    F, Fp, s = random_states(case, n, rng, strain)
    W = rng.uniform(-20.0, 20.0, size=(n, ns)) if kin else None
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    out = {"check": "batch", "n": n, "ns": ns, "strain": strain, "kin": kin}
    for flag in (0, 1):
        P, T, Fe, Fpo = (np.zeros((n, 9)) for _ in range(4))
        D = np.zeros((n, 81))
        so, xo = np.zeros((n, ns)), np.zeros((n, ns))
        na = np.zeros(n, np.int32)
        Wio = np.ascontiguousarray(W.copy()) if kin else None
        rc = L.emu_qp_batch(ctypes.byref(mat), ptr(np.ascontiguousarray(rod)), ctypes.c_int(8), ctypes.c_longlong(n),
                            ptr(np.ascontiguousarray(F.reshape(n, 9))), ptr(np.ascontiguousarray(Fp.reshape(n, 9))),
                            ptr(np.ascontiguousarray(s)), ptr(grain), ctypes.c_int(flag), ptr(P), ptr(T), ptr(D), ptr(Fe),
                            ptr(Fpo), ptr(so), ptr(xo), ptr(na), ptr(Wio) if kin else None)
        assert rc == 0
        err = {"P": 0.0, "T": 0.0, "Fe": 0.0, "Fp": 0.0, "s": 0.0, "dPdF": 0.0, "W": 0.0}
        nact_mismatch, nplastic = 0, 0
        for i in range(n):
            o = refcase.oracle_qp(case, rod[grain[i]], F[i], Fp[i], np.append(s[i], 10e14), flag, W=None if W is None else W[i])
            nact_mismatch += int(na[i] != o["n_active"])
            nplastic += int(o["n_active"] > 0)
            for k, g in (("P", P), ("T", T), ("Fe", Fe), ("Fp", Fpo)):
                err[k] = max(err[k], rel(g[i], o[k].ravel()))
            err["s"] = max(err["s"], rel(so[i], o["s"][:ns]))
            if kin:
                err["W"] = max(err["W"], rel(Wio[i], o["W"][:ns]))
            if flag:
                err["dPdF"] = max(err["dPdF"], rel(D[i], o["dPdF"].ravel()))
        out["flag%d" % flag] = {"err": err, "n_active_mismatch": nact_mismatch, "plastic": nplastic}
    print(json.dumps(out), flush=True)


def qp_F(nel, h, u):
    """F at every qp Q = 8 e + q (getElementalValues.cc:129-138), host restatement."""
    nx, ny, nz = nel
    g0, g1 = 0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)
    nq = 8 * nx * ny * nz
    F = np.zeros((nq, 3, 3))
    for e in range(nx * ny * nz):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        for q in range(8):
            G = np.eye(3)
            for v in range(8):
                nd = (ex + (v & 1)) + (nx + 1) * ((ey + ((v >> 1) & 1)) + (ny + 1) * (ez + ((v >> 2) & 1)))
                N, dN = [], []
                for d in range(3):
                    xi = g1 if (q >> d) & 1 else g0
                    b = (v >> d) & 1
                    N.append(xi if b else 1 - xi)
                    dN.append((1.0 if b else -1.0) / h[d])
                g = np.array([dN[0] * N[1] * N[2], N[0] * dN[1] * N[2], N[0] * N[1] * dN[2]])
                G += np.outer(u[3 * nd:3 * nd + 3], g)
            F[8 * e + q] = G
    return F


def check_pass(case, strain, nx, seed, kin, mixed=0):
    L = lib()
    rng = np.random.default_rng(seed)
    ns = case.n_slip
    if kin:
        case.C1, case.C2 = np.full(ns, 200.0), np.full(ns, 1.0)
    mat = material(case)
    nel = (nx, max(1, nx - 1), nx)
    nelem = nel[0] * nel[1] * nel[2]
    nq = 8 * nelem
    nnode = (nel[0] + 1) * (nel[1] + 1) * (nel[2] + 1)
    h = np.array([1.0 / nel[0], 1.0 / nel[1], 1.0 / nel[2]])
    rod = rng.uniform(-0.5, 0.5, size=(4, 3))
    egrain = rng.integers(0, 4, size=nelem).astype(np.uint32)
    # displacement: a uniform stretch plus noise at the nodes
    X = np.array([[i * h[0], j * h[1], k * h[2]] for k in range(nel[2] + 1) for j in range(nel[1] + 1) for i in range(nel[0] + 1)])
    Fm = np.eye(3) + rng.normal(scale=strain, size=(3, 3))
    u = ((X @ (Fm - np.eye(3)).T) + rng.normal(scale=strain * h[0] / 2, size=X.shape)).ravel()
    _, Fp, s = random_states(case, nq, rng, strain)
    slip = rng.uniform(0, 1e-3, size=(nq, ns))
    W = rng.uniform(-20.0, 20.0, size=(nq, ns)) if kin else np.zeros((nq, ns))
    cm = lambda a: np.ascontiguousarray(a.reshape(nq, -1).T)  # component-major [c][nq]
    F = qp_F(nel, h, u)
    out = {"check": "pass", "nel": nel, "ns": ns, "strain": strain, "kin": kin, "mixed": mixed}
    for flag in (1, 0):
        Fe_i, Fp_i, s_i, sl_i = np.zeros((9, nq)), np.zeros((9, nq)), np.zeros((ns, nq)), np.zeros((ns, nq))
        W_i = np.zeros((ns, nq))
        D, P, T = np.zeros((81, nq)), np.zeros((9, nq)), np.zeros((9, nq))
        nsing = ctypes.c_uint(0)
        nel_c = (ctypes.c_int * 3)(*nel)
        rc = L.emu_qp_pass(ctypes.byref(mat), ptr(np.ascontiguousarray(rod)), ctypes.c_int(4), nel_c, ptr(h), ptr(u), ptr(egrain),
                           ptr(cm(Fp)), ptr(cm(s)), ptr(cm(slip)), ptr(cm(W)), ctypes.c_int(flag), ctypes.c_int(mixed),
                           ctypes.c_int(3), ptr(Fe_i), ptr(Fp_i), ptr(s_i), ptr(sl_i), ptr(W_i), ptr(D), ptr(P), ptr(T),
                           ctypes.byref(nsing))
        assert rc == 0
        jxw = h.prod() / 8
        err = {"Fe": 0.0, "Fp": 0.0, "s": 0.0, "slip": 0.0, "P": 0.0, "T": 0.0, "dPdF": 0.0, "W": 0.0}
        for Q in range(nq):
            o = refcase.oracle_qp(case, rod[egrain[Q // 8]], F[Q], Fp[Q], np.append(s[Q], 10e14), flag, W=W[Q] if kin else None,
                                  slip=slip[Q])
            err["Fe"] = max(err["Fe"], rel(Fe_i[:, Q], o["Fe"].ravel()))
            err["Fp"] = max(err["Fp"], rel(Fp_i[:, Q], o["Fp"].ravel()))
            err["s"] = max(err["s"], rel(s_i[:, Q], o["s"][:ns]))
            err["slip"] = max(err["slip"], rel(sl_i[:, Q], o["slip"]))
            if kin:
                err["W"] = max(err["W"], rel(W_i[:, Q], o["W"][:ns]))
            if flag:
                err["P"] = max(err["P"], rel(P[:, Q], jxw * o["P"].ravel()))
                err["dPdF"] = max(err["dPdF"], rel(D[:, Q], jxw * o["dPdF"].ravel()))
            else:
                err["T"] = max(err["T"], rel(T[:, Q], o["T"].ravel()))
        out["flag%d" % flag] = {"err": err, "nsing": nsing.value}
    print(json.dumps(out), flush=True)


def check_golden(folder, k):
    """The two qp passes on a reference case's mesh, grains and load, at the state the oracle reaches:
    the history committed after k increments, the displacement at the end of increment k + 1."""
    L = lib()
    case = refcase.load_reference_case(folder)
    ns, nT = case.n_slip, case.nT
    nel = (case.nx, case.ny, case.nz)
    nq = 8 * case.nx * case.ny * case.nz
    if k > 0:
        st = refcase.oracle_simulate(case, k, want_state=True)["state"]
        Fp, s, slip = st[:, 9:18], st[:, 27:27 + ns], st[:, 27 + nT + 3:27 + nT + 3 + ns]
    else:
        Fp, s, slip = np.tile(np.eye(3).ravel(), (nq, 1)), np.tile(case.s0, (nq, 1)), np.zeros((nq, ns))
    u = refcase.oracle_simulate(case, k + 1, want_state=True)["u"]
    h = np.array([case.span[0] / nel[0], case.span[1] / nel[1], case.span[2] / nel[2]])
    mat = material(case)
    rod = np.ascontiguousarray(case.grain_rod, float)
    eg = np.ascontiguousarray(case.elem_grain, np.uint32)
    F = qp_F(nel, h, u)
    cm = lambda a: np.ascontiguousarray(np.asarray(a).reshape(nq, -1).T)
    out = {"check": "golden", "folder": folder, "k": k}
    for flag in (1, 0):
        Fe_i, Fp_i, s_i, sl_i, W_i = np.zeros((9, nq)), np.zeros((9, nq)), np.zeros((ns, nq)), np.zeros((ns, nq)), np.zeros((ns, nq))
        D, P, T = np.zeros((81, nq)), np.zeros((9, nq)), np.zeros((9, nq))
        nsing = ctypes.c_uint(0)
        rc = L.emu_qp_pass(ctypes.byref(mat), ptr(rod), ctypes.c_int(len(rod)), (ctypes.c_int * 3)(*nel), ptr(h), ptr(u), ptr(eg),
                           ptr(cm(Fp)), ptr(cm(s)), ptr(cm(slip)), ptr(np.zeros((ns, nq))), ctypes.c_int(flag), ctypes.c_int(0),
                           ctypes.c_int(4), ptr(Fe_i), ptr(Fp_i), ptr(s_i), ptr(sl_i), ptr(W_i), ptr(D), ptr(P), ptr(T),
                           ctypes.byref(nsing))
        assert rc == 0
        err = {"Fp": 0.0, "s": 0.0, "T": 0.0, "dPdF": 0.0}
        for Q in range(0, nq, 7):
            o = refcase.oracle_qp(case, rod[eg[Q // 8]], F[Q], Fp[Q].reshape(3, 3), np.append(s[Q], 10e14), flag, slip=slip[Q])
            err["Fp"] = max(err["Fp"], rel(Fp_i[:, Q], o["Fp"].ravel()))
            err["s"] = max(err["s"], rel(s_i[:, Q], o["s"][:ns]))
            if flag:
                err["dPdF"] = max(err["dPdF"], rel(D[:, Q], h.prod() / 8 * o["dPdF"].ravel()))
            else:
                err["T"] = max(err["T"], rel(T[:, Q], o["T"].ravel()))
        out["flag%d" % flag] = {"err": err, "nsing": nsing.value}
    print(json.dumps(out), flush=True)


if __name__ == "__main__" and sys.argv[1] not in ("lane", "twin"):
    if sys.argv[1] == "golden":
        check_golden(sys.argv[2], int(sys.argv[3]))
        sys.exit(0)
    mode, name, strain, n, seed = sys.argv[1], sys.argv[2], float(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5])
    kin = len(sys.argv) > 6 and sys.argv[6] == "kin"
    case = make_case(name)
    if mode == "batch":
        check_batch(case, strain, n, seed, kin)
    else:
        check_pass(case, strain, n, seed, kin, mixed=int(len(sys.argv) > 7 and sys.argv[7] == "mixed"))


def run_form(L, fn, case, mat, rod, F, Fp, s, grain, flag, W):
    n, ns = len(F), case.n_slip
    P, T, Fe, Fpo = (np.zeros((n, 9)) for _ in range(4))
    D = np.zeros((n, 81))
    so, xo = np.zeros((n, ns)), np.zeros((n, ns))
    na = np.zeros(n, np.int32)
    Wio = None if W is None else np.ascontiguousarray(W.copy())
    rc = getattr(L, fn)(ctypes.byref(mat), ptr(np.ascontiguousarray(rod)), ctypes.c_int(len(rod)), ctypes.c_longlong(n),
                        ptr(np.ascontiguousarray(F.reshape(n, 9))), ptr(np.ascontiguousarray(Fp.reshape(n, 9))),
                        ptr(np.ascontiguousarray(s)), ptr(grain), ctypes.c_int(flag), ptr(P), ptr(T), ptr(D), ptr(Fe),
                        ptr(Fpo), ptr(so), ptr(xo), ptr(na), None if Wio is None else ptr(Wio))
    assert rc == 0
    return {"P": P, "T": T, "D": D, "Fe": Fe, "Fp": Fpo, "s": so, "x": xo, "na": na, "W": Wio}


def check_lane(case, strain, n, seed):
    """The cooperative form against the one-lane routine, both on the host: same arithmetic."""
    L = lib()
    rng = np.random.default_rng(seed)
    mat = material(case)
    rod = rng.uniform(-0.5, 0.5, size=(8, 3))
    F, Fp, s = random_states(case, n, rng, strain)
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    out = {"check": "lane", "n": n, "ns": case.n_slip}
    a = run_form(L, "emu_qp_batch", case, mat, rod, F, Fp, s, grain, 1, None)
    b = run_form(L, "emu_qp_lane", case, mat, rod, F, Fp, s, grain, 1, None)
    out["coop_vs_lane"] = {k: rel(a[k], b[k]) for k in ("P", "D", "s", "x")}
    out["na_mismatch"] = int((a["na"] != b["na"]).sum())
    errs = []
    for i in range(n):
        o = refcase.oracle_qp(case, rod[grain[i]], F[i], Fp[i], np.append(s[i], 10e14), 1)
        errs.append(rel(b["P"][i], o["P"].ravel()))
    out["lane_vs_oracle_P_max"] = max(errs)
    out["lane_vs_oracle_P_bad"] = int(sum(e > 1e-8 for e in errs))
    print(json.dumps(out), flush=True)


if __name__ == "__main__" and sys.argv[1] == "lane":
    check_lane(make_case(sys.argv[2]), float(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5]))


def check_twin(n, seed):
    """hcp/simpleTension's 18 slip + 6 twin systems: the cooperative form against the one-lane
    routine (bit for bit) and against the oracle's twin return map, on the random states of the GPU
    test (tests/test_gpu_twinning.py::test_qp_update_with_twin_systems_matches_oracle)."""
    L = lib()
    case = refcase.load_reference_case("hcp_simpleTension", twinning=True)
    tw = case.twin
    ns = case.n_slip
    # the emulator's material carries all 24 systems; the twin ones are marked by emu_set_nslip
    sysc = refcase.Case(np.vstack([case.m, tw["m"]]), np.vstack([case.n, tw["n"]]), tw["q"], case.C,
                        np.concatenate([case.s0, tw["s0"]]), np.concatenate([case.h0, tw["h0"]]),
                        np.concatenate([case.ss, tw["ss"]]), np.concatenate([case.pw, tw["pw"]]), tol=case.tol)
    mat = material(sysc)
    L.emu_set_nslip(ctypes.c_int(ns))
    rng = np.random.default_rng(seed)
    rod = rng.uniform(-0.6, 0.6, size=(8, 3))
    F = np.eye(3) + rng.normal(scale=2.5e-3, size=(n, 3, 3))
    Fp = np.eye(3) + rng.normal(scale=4e-4, size=(n, 3, 3))
    Fp /= np.cbrt(np.linalg.det(Fp))[:, None, None]
    s0 = np.concatenate([case.s0, tw["s0"]])
    s = np.array([rng.uniform(1.0, 1.3) * s0 for _ in range(n)])
    grain = rng.integers(0, 8, size=n).astype(np.uint32)
    out = {"check": "twin", "n": n}
    for flag in (0, 1):
        a = run_form(L, "emu_qp_batch", sysc, mat, rod, F, Fp, s, grain, flag, None)
        b = run_form(L, "emu_qp_lane", sysc, mat, rod, F, Fp, s, grain, flag, None)
        bad_lane, bad_coop, twins = [], [], 0
        for i in range(n):
            o = refcase.oracle_qp_twin(case, rod[grain[i]], F[i], Fp[i], s[i], flag)
            twins += int((o["x"][ns:] > 0).any())
            if rel(b["P"][i], o["P"].ravel()) > 1e-10 or rel(b["s"][i], o["s"]) > 1e-10:
                bad_lane.append(i)
            if rel(a["P"][i], o["P"].ravel()) > 1e-10 or rel(a["s"][i], o["s"]) > 1e-10:
                bad_coop.append(i)
        out["flag%d" % flag] = {"coop_vs_lane": {k: rel(a[k], b[k]) for k in ("P", "s", "x")}, "twin_qps": twins,
                                "lane_vs_oracle_bad": bad_lane[:10], "coop_vs_oracle_bad": bad_coop[:10]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__" and sys.argv[1] == "twin":
    check_twin(int(sys.argv[2]), int(sys.argv[3]))

"""The library's closed periodic constraint lines (pa_periodic_lines: the host build that
pa_set_periodic uploads) against the oracle's (or_periodic_lines), for every periodicBCs input of
the reference and several mesh sizes: the same constrained dofs, masters, weights and
inhomogeneities.  Host only, no device call."""
import ctypes

import numpy as np
import pytest

import refcase
from refcase import ptr
from plasticity_amd import native, model

PRMS = ["prm_UniaxialTension.prm", "prm_simpleShear.prm", "prm_BiaxialTension.prm", "prm_pureShear.prm",
        "prm_UniaxialTensionTabular.prm", "prm_UniaxialCyclicTabular.prm"]


def _product_lines(case):
    L = native.lib()
    pb = model.make_problem((case.nx, case.ny, case.nz), case.span, case.total_increments, case.delT, [],
                            lin_tol=case.lin_tol)
    per = refcase.periodic_struct(case.periodic)
    N = 3 * (case.nx + 1) * (case.ny + 1) * (case.nz + 1)
    con, rp, g = np.zeros(N, np.int32), np.zeros(N + 1, np.int64), np.zeros(N)
    n = ctypes.c_longlong()
    rc = L.pa_periodic_lines(ctypes.byref(pb), ctypes.byref(per), ptr(con), ptr(rp), None, None, ptr(g), ctypes.c_longlong(0),
                             ctypes.byref(n))
    if rc:
        return rc, None
    M, W = np.zeros(max(n.value, 1), np.int64), np.zeros(max(n.value, 1))
    native.check(L.pa_periodic_lines(ctypes.byref(pb), ctypes.byref(per), ptr(con), ptr(rp), ptr(M), ptr(W), ptr(g),
                                     ctypes.c_longlong(n.value), ctypes.byref(n)))
    return 0, (con, rp, M[:n.value], W[:n.value], g)


def _oracle_lines(case):
    L = refcase.oracle()
    L.or_periodic_lines.restype = ctypes.c_longlong
    pb = case.problem(1)
    N = 3 * (case.nx + 1) * (case.ny + 1) * (case.nz + 1)
    con, rp, g = np.zeros(N, np.int32), np.zeros(N + 1, np.int64), np.zeros(N)
    nnz = L.or_periodic_lines(ctypes.byref(pb), ptr(con), ptr(rp), None, None, ptr(g), ctypes.c_longlong(0))
    if nnz < 0:
        return -1, None
    M, W = np.zeros(max(nnz, 1), np.int64), np.zeros(max(nnz, 1))
    L.or_periodic_lines(ctypes.byref(pb), ptr(con), ptr(rp), ptr(M), ptr(W), ptr(g), ctypes.c_longlong(nnz))
    return 0, (con, rp, M[:nnz], W[:nnz], g)


@pytest.mark.parametrize("prm", PRMS)
@pytest.mark.parametrize("n", [2, 3, 5])
def test_library_lines_equal_the_oracle_lines(prm, n):
    case = refcase.load_reference_case("fcc_periodicBCs", prm_name=prm)
    case.nx = case.ny = case.nz = n
    rc_p, P = _product_lines(case)
    rc_o, O = _oracle_lines(case)
    assert rc_p == 0 and rc_o == 0
    con_p, rp_p, M_p, W_p, g_p = P
    con_o, rp_o, M_o, W_o, g_o = O
    assert np.array_equal(con_p, con_o)
    assert np.array_equal(g_p, g_o)
    for d in np.nonzero(con_o)[0]:
        a = {int(m): w for m, w in zip(M_p[rp_p[d]:rp_p[d + 1]], W_p[rp_p[d]:rp_p[d + 1]]) if w != 0.0}
        b = {int(m): w for m, w in zip(M_o[rp_o[d]:rp_o[d + 1]], W_o[rp_o[d]:rp_o[d + 1]]) if w != 0.0}
        assert a == b, d


def test_unequal_element_counts_are_refused_by_both():
    case = refcase.load_reference_case("fcc_periodicBCs", prm_name=PRMS[0])
    case.nx = 4
    assert _product_lines(case)[0] == native.PA_ERR_UNSUPPORTED
    assert _oracle_lines(case)[0] == -1

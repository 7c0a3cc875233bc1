"""The oracle's twinning (calculatePlasticity.cc:421-452, 962-1019; updateAfterIncrement.cc:176,
222-231, 287-305, 719-720, 780-821) against what pins it:

* the reference's own golden curve of hcp/simpleTension (tests/compile_tests/hcp/simpleTension/
  correct_stressstrain.txt, 6 significant digits), whose run twins from increment 38: strain and
  stress rows, TwinRealVF (the volume-averaged twin fraction) and SlipTotal over 45 increments;
* the reorientation by twinning as a rotation: the quaternion product of the orientation with
  (0, n) is R_new = R_old (2 n n^T - I), a half turn about the twin normal in the crystal frame;
* the reorientation test itself on hand-set fractions: the largest fraction (first of equals)
  against F_T = min(1, threshold + saturation F_e / F_r), only with F_r > 0 and only once.
"""
import ctypes

import numpy as np
import pytest

import refcase


def _twin_case():
    return refcase.load_reference_case("hcp_simpleTension", twinning=True)


def test_oracle_twinning_reproduces_the_hcp_tension_golden():
    case = _twin_case()
    assert case.twin is not None and case.nT == 24
    n = 45
    rows = refcase.oracle_simulate(case, n, nthreads=4)["rows"]
    gold = case.golden[:n]
    tw = gold[:, 12] > 0
    assert tw.any() and not tw[:38].any() and tw[38]  # the golden run twins from increment 38
    for cols in (slice(0, 6), slice(6, 12)):
        err = np.abs(rows[:, cols] - gold[:, cols]).max(axis=1) / np.abs(gold[:, cols]).max(axis=1)
        print("rows vs golden, before / after twinning starts: %.1e %.1e" % (err[~tw].max(), err[tw].max()))
        assert err[~tw].max() < 5e-6 and err[tw].max() < 5e-5, (cols, err)  # observed <= 3e-6, 2.6e-5
    assert (rows[~tw, 12] == 0).all()
    rel = np.abs(rows[tw, 12] - gold[tw, 12]) / gold[tw, 12]
    assert rel.max() < 1.5e-2, rel  # observed 7e-3 at the first twinned increment, 2e-3 after
    sl = gold[:, 14] > 1e-7
    assert (np.abs(rows[sl, 14] - gold[sl, 14]) / gold[sl, 14]).max() < 1e-3
    assert (rows[:, 13] == 0).all() and (gold[:, 13] == 0).all()  # no reorientation this early


def _odf(r):
    R = np.zeros(9)
    refcase.oracle().or_odfpoint(refcase.ptr(np.ascontiguousarray(r, float)), refcase.ptr(R))
    return R.reshape(3, 3)


def test_twin_reorientation_is_a_half_turn_about_the_twin_normal():
    case = _twin_case()
    mat = case.material()
    lib = refcase.oracle()
    rng = np.random.default_rng(5)
    ntw, ns = case.n_twin, case.n_slip
    for trial in range(12):
        rot = rng.uniform(-0.6, 0.6, 3)
        k = trial % ntw
        tw = np.full(ntw, 0.1)
        tw[k] = 0.9  # above the threshold 0.7
        rot_iter = np.zeros(3)
        twin_iter = ctypes.c_int(0)
        s = np.zeros(case.nT)
        pos = lib.or_twin_event(ctypes.byref(mat), ctypes.c_double(0.05), ctypes.c_double(0.0), refcase.ptr(tw), ctypes.c_int(0),
                                refcase.ptr(rot), refcase.ptr(rot_iter), ctypes.byref(twin_iter), refcase.ptr(s))
        assert pos == k and twin_iter.value == 1
        assert (s[ns:] == 1e5).all() and (s[:ns] == 0).all()
        nrm = case.twin["n"][k]
        H = 2 * np.outer(nrm, nrm) - np.eye(3)
        assert np.abs(_odf(rot_iter) - _odf(rot) @ H).max() < 1e-12


@pytest.mark.parametrize("F_r,F_e,twin_conv,fracs,expect", [
    (0.05, 0.0, 0, [0.2, 0.71, 0.3, 0.71, 0.0, 0.1], 1),   # first of two equal maxima
    (0.05, 0.0, 0, [0.2, 0.69, 0.3, 0.5, 0.0, 0.1], -1),   # below the threshold 0.7
    (0.0, 0.0, 0, [0.2, 0.95, 0.3, 0.5, 0.0, 0.1], -1),    # F_r = 0: no reorientation (:992)
    (0.05, 0.0, 1, [0.2, 0.95, 0.3, 0.5, 0.0, 0.1], -1),   # already twinned
    (0.05, 0.5, 0, [0.2, 0.95, 0.3, 0.5, 0.0, 0.1], 1),    # saturation factor 0 (the prm): F_T = 0.7
])
def test_twin_event_rule(F_r, F_e, twin_conv, fracs, expect):
    case = _twin_case()
    mat = case.material()
    rot, rot_iter, s = np.array([0.1, 0.2, -0.1]), np.full(3, 7.0), np.zeros(case.nT)
    twin_iter = ctypes.c_int(twin_conv)
    pos = refcase.oracle().or_twin_event(ctypes.byref(mat), ctypes.c_double(F_r), ctypes.c_double(F_e),
                                         refcase.ptr(np.array(fracs, float)), ctypes.c_int(twin_conv), refcase.ptr(rot),
                                         refcase.ptr(rot_iter), ctypes.byref(twin_iter), refcase.ptr(s))
    assert pos == expect
    if expect < 0:
        assert (rot_iter == 7.0).all() and (s == 0).all() and twin_iter.value == twin_conv


def test_twin_threshold_with_saturation_factor_is_capped_at_one():
    """F_T = threshold + saturation F_e / F_r, at most 1 (:970-985): with saturation 0.2, F_e 0.05
    and F_r 0.01, F_T would be 1.7 -> 1, so a fraction of 0.99 does not reorient."""
    case = _twin_case()
    case.twin = dict(case.twin, saturation=0.2)
    mat = case.material()
    rot, rot_iter, s = np.zeros(3), np.zeros(3), np.zeros(case.nT)
    tw = np.array([0.99, 0, 0, 0, 0, 0.0])
    ti = ctypes.c_int(0)
    args = (ctypes.byref(mat), ctypes.c_double(0.01), ctypes.c_double(0.05), refcase.ptr(tw), ctypes.c_int(0), refcase.ptr(rot),
            refcase.ptr(rot_iter), ctypes.byref(ti), refcase.ptr(s))
    assert refcase.oracle().or_twin_event(*args) == -1
    tw[0] = 1.0 + 1e-9
    assert refcase.oracle().or_twin_event(*args) == 0

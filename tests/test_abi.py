"""The C-ABI library loads and exports every symbol include/plasticity_amd.h declares (no GPU calls)."""
import os
import re

import plasticity_amd.native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "plasticity_amd.h")).read()
    return sorted(set(re.findall(r"\b(pa_[a-z_]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    L = nat.lib()
    names = declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert sorted(nat.EXPORTS) == names


def test_error_path_without_gpu_is_reported():
    import ctypes
    import plasticity_amd as pa
    mat = nat.PaMaterial()
    mat.n_slip = 0  # invalid: rejected before any device call
    pb = nat.PaProblem()
    h = ctypes.c_void_p()
    rc = nat.lib().pa_create(ctypes.byref(mat), ctypes.byref(pb), ctypes.byref(h))
    assert rc == 1
    assert b"n_slip" in nat.lib().pa_last_error()


def test_struct_layout_matches_header():
    """ctypes mirrors of pa_material / pa_problem have the header's sizes (field order = header)."""
    import ctypes
    src = open(os.path.join(ROOT, "include", "plasticity_amd.h")).read()
    assert "double C1[PA_MAX_SLIP]" in src and "double C2[PA_MAX_SLIP]" in src
    ms, d = nat.MAX_SLIP, ctypes.sizeof(ctypes.c_double)
    # n_slip, m, n, q, C, s0 h0 ss pw, stress_tol, 4 ints, C1, C2 (with the int -> double padding)
    body = 4 + 4 + (3 * ms * 2 + ms * ms + 36 + 4 * ms + 1) * d + 4 * 4
    body += (-body) % d
    assert ctypes.sizeof(nat.PaMaterial) == body + 2 * ms * d


def test_prm_with_kinematic_hardening_and_cyclic_loading_maps_onto_the_abi():
    """fcc/CyclicLoading's prm.prm: kinematic hardening and the cyclic BC reach pa_material / pa_problem
    (userInputParameters.cc:274-279; init.cc:327-348)."""
    from plasticity_amd import model
    ui = model.userInputParameters(os.path.join(ROOT, "tests", "golden", "fcc_CyclicLoading", "prm.prm"))
    cp = model.crystalPlasticity(ui)
    assert cp.material.kinematic_hardening == 1
    assert list(cp.material.C1[:12]) == [200.0] * 12 and list(cp.material.C2[:12]) == [1.0] * 12
    pb = cp.problem
    assert (pb.cyclic, pb.cyclic_face, pb.cyclic_dof, pb.quarter_cycle_time) == (1, 6, 3, 5.0)
    assert cp.totalIncrements == 4000
    assert pb.face_dof_constrained[5][2] == 1 and pb.face_total_u[5][2] == -0.05

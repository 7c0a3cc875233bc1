"""The C-ABI library loads and exports every symbol include/plasticity_amd.h declares (no GPU calls)."""
import os
import re

import numpy as np

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


def test_pinv_solve_rejects_bad_arguments_before_any_device_call():
    import numpy as np
    import pytest
    import plasticity_amd as pa
    with pytest.raises(pa.PlasticityError):
        pa.model.pinv_solve(np.eye(3)[None], np.ones((1, 3)), path=2)
    with pytest.raises(pa.PlasticityError):
        pa.model.pinv_solve(np.zeros((1, 49, 49)), np.zeros((1, 49)))
    assert pa.model.pinv_solve(np.zeros((0, 4, 4)), np.zeros((0, 4))).shape == (0, 4)

def test_warm_start_and_output_entry_points_reject_bad_arguments_before_any_device_call():
    """PA_ERR_ARG (1) for null handles and out-of-range sizes, checked before the device is touched."""
    import ctypes
    import numpy as np
    from plasticity_amd import native
    L = native.lib()
    X = np.zeros((4, 18), np.float32)
    Y = np.zeros((4, 12), np.float32)
    g = np.zeros(11472, np.float32)  # PA_MLP_NPARAM
    loss = ctypes.c_double()
    assert L.pa_mlp_create(None) == 1
    assert L.pa_mlp_set_params(None, native.ptr(g)) == 1
    assert L.pa_mlp_get_params(None, native.ptr(g)) == 1
    assert L.pa_mlp_predict(None, ctypes.c_longlong(4), native.ptr(X), native.ptr(Y)) == 1
    assert L.pa_mlp_grad(None, ctypes.c_longlong(4), native.ptr(X), native.ptr(Y), native.ptr(g), ctypes.byref(loss)) == 1
    assert L.pa_mlp_train(None, ctypes.c_longlong(4), native.ptr(X), native.ptr(Y), 1, 4, ctypes.c_float(1e-3),
                          ctypes.c_ulonglong(0), None) == 1
    assert L.pa_grain_averages(None, 4, native.ptr(np.zeros(28))) == 1
    assert L.pa_qp_update_warm(None, None, ctypes.c_longlong(0), *([None] * 17)) == 1
    assert L.pa_mlp_destroy(None) == 0


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of pa_material / pa_problem have the sizes and field offsets the C compiler
    gives the header's structs."""
    import ctypes
    import shutil
    import subprocess
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "plasticity_amd.h"\n'
                   'int main(void) { printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(pa_material), sizeof(pa_problem),'
                   ' offsetof(pa_material, C1), offsetof(pa_material, C2), offsetof(pa_problem, table_time),'
                   ' offsetof(pa_problem, table_u)); printf("%zu %zu\\n", offsetof(pa_problem, velgrad),'
                   ' offsetof(pa_problem, vel_grad)); return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(nat.PaMaterial), ctypes.sizeof(nat.PaProblem), nat.PaMaterial.C1.offset, nat.PaMaterial.C2.offset,
            nat.PaProblem.table_time.offset, nat.PaProblem.table_u.offset, nat.PaProblem.velgrad.offset,
            nat.PaProblem.vel_grad.offset]
    assert got == want


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


def test_prm_with_tabular_bcs_maps_onto_the_abi():
    """fcc/tabularSimpleCompression's prm.prm: the table reaches pa_problem (init.cc:389-415)."""
    from plasticity_amd import model
    ui = model.userInputParameters(os.path.join(ROOT, "tests", "golden", "fcc_tabularSimpleCompression", "prm.prm"))
    pb = model.crystalPlasticity(ui).problem
    assert (pb.tabular, pb.n_table) == (1, 3)
    assert list(pb.table_time[:3]) == [0.0, 20.0, 100.0]
    assert list(pb.table_u[1][0][:3]) == [0.0, -0.2, -0.2]
    assert [pb.face_dof_constrained[0][d] for d in range(3)] == [1, 1, 1] and pb.face_dof_constrained[1][0] == 1


def test_prm_with_velocity_gradient_bc_maps_onto_the_abi():
    """fcc/DeformationGraidientBCs's prm.prm: L reaches pa_problem (userInputParameters.cc:78-83) and no
    face rule is set (Use Simple BCs = false)."""
    from plasticity_amd import model
    ui = model.userInputParameters(os.path.join(ROOT, "tests", "golden", "fcc_DeformationGraidientBCs", "prm.prm"))
    cp = model.crystalPlasticity(ui)
    pb = cp.problem
    assert pb.velgrad == 1
    assert [[pb.vel_grad[a][b] for b in range(3)] for a in range(3)] == [[-0.005, 0, 0], [0, -0.005, 0], [0, 0, 0.01]]
    assert all(pb.face_dof_constrained[f][d] == 0 for f in range(6) for d in range(3))
    assert cp.totalIncrements == 1000 and (pb.nx, pb.ny, pb.nz) == (8, 8, 8)


def test_periodic_struct_layout_matches_header(tmp_path):
    """ctypes mirror of pa_periodic: size and field offsets as the C compiler gives them."""
    import ctypes
    import shutil
    import subprocess
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = ["vertex_dofs", "vertex_coef", "n_edge_lines", "edge_coef", "n_face_lines", "face_coef", "vertex_known",
              "vertex_total", "tabular", "tabular_time", "n_table", "table_time", "table_coef"]
    src = tmp_path / "layout.c"
    body = "".join(' printf("%%zu\\n", offsetof(pa_periodic, %s));' % f for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "plasticity_amd.h"\n'
                   'int main(void) { printf("%zu\\n", sizeof(pa_periodic));' + body + ' return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(nat.PaPeriodic)] + [getattr(nat.PaPeriodic, f).offset for f in fields]
    assert got == want


def test_prm_with_periodic_bcs_maps_onto_the_abi():
    """fcc/periodicBCs prm_UniaxialCyclicTabular.prm: the constraint file's 12 vertex, 27 edge and 9
    face lines, the vertex rows and the tabular periodic table reach pa_periodic (ellipticBVP
    init.cc:66-131, userInputParameters.cc:126-136), as the oracle's reader of the same file has them."""
    import refcase
    from plasticity_amd import model
    d = os.path.join(ROOT, "tests", "golden", "fcc_periodicBCs")
    ui = model.userInputParameters(os.path.join(d, "prm_UniaxialCyclicTabular.prm"))
    cp = model.crystalPlasticity(ui)
    P = cp.periodic
    assert (P.n_vertex_lines, P.n_edge_lines, P.n_face_lines) == (12, 27, 9)
    assert list(P.vertex_dofs[6]) == [4, 22, 4, 7, 1] and list(P.vertex_coef[6]) == [4.0, 1.0, 4.0, -1.0, 1.0]
    assert list(P.edge_dofs[0]) == [3, 9, 18, 9] and list(P.face_dofs[8]) == [17, 14, 11, 2]
    assert list(P.vertex_known[:6]) == [1, 1, 1, 1, 1, 1] and P.vertex_total[3] == 0.01
    assert (P.tabular, P.tabular_time, P.n_table) == (1, 1.0, 7)
    assert list(P.table_coef[:7]) == [1, -1, 1, -1, 1, -1, 1]
    assert all(cp.problem.face_dof_constrained[f][k] == 0 for f in range(6) for k in range(3))
    ref = refcase.periodic_struct(refcase.load_reference_case("fcc_periodicBCs", prm_name="prm_UniaxialCyclicTabular.prm").periodic)
    assert bytes(P) == bytes(ref)


def test_prm_with_several_phases_maps_onto_the_abi():
    """fcc/simpleCompression_Multiphase's prm.prm: three phases (userInputParameters.cc:286-545) reach
    three pa_material (phase 2 with 18 slip and 6 twin systems, phase 3 with kinematic hardening) as
    the oracle's reader has them, and the orientation file's fifth column gives the grain phases."""
    import refcase
    from plasticity_amd import model
    d = os.path.join(ROOT, "tests", "golden", "fcc_simpleCompression_Multiphase")
    ui = model.userInputParameters(os.path.join(d, "prm.prm"))
    cp = model.crystalPlasticity(ui)
    assert len(cp.materials) == 3
    m1, m2, m3 = cp.materials
    assert (m1.n_slip, m1.twinning, m1.kinematic_hardening) == (12, 0, 0)
    assert (m2.n_slip, m2.twinning, m2.n_twin, m2.twin_shear, m2.twin_threshold) == (18, 1, 6, 0.126, 0.7)
    assert (m3.n_slip, m3.kinematic_hardening, m3.C1[0], m3.C2[0]) == (12, 1, 10.0, 1.0)
    assert m2.C[2][2] == 61.5e3 and m2.C[5][5] == 16.8e3 and m1.max_solver_iter == 4 == m2.max_solver_iter
    ref = refcase.load_reference_case("fcc_simpleCompression_Multiphase", twinning=True)
    for mat, case in zip(cp.materials, ref.phases):
        nsys = case.nT if case.twin else case.n_slip
        assert np.allclose([mat.s0[a] for a in range(nsys)], list(case.s0) + (list(case.twin["s0"]) if case.twin else []))
    cp.orientations.loadOrientationVector(os.path.join(d, "orientations.txt"), True)
    assert list(cp.orientations.phase[1:4]) == [1, 2, 3]

"""ctypes binding of the C ABI in include/plasticity_amd.h (libplasticity_amd.so, built in-tree).

The library is required: every entry point here calls the HIP implementation and raises when the
shared object is missing or a call fails.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplasticity_amd.so")
MAX_SLIP = 48
MAX_TABLE = 32

PA_SOLVER_BICG, PA_SOLVER_CG, PA_SOLVER_MGCG = 0, 1, 2
PA_PREC_FP64, PA_PREC_MIXED = 0, 1
PA_QP_LANE, PA_QP_COOP = 0, 1
PA_ERR_ARG, PA_ERR_UNSUPPORTED, PA_ERR_HIP, PA_ERR_IO, PA_ERR_STATE, PA_ERR_NONFINITE = 1, 2, 3, 4, 5, 6

# every symbol include/plasticity_amd.h declares
EXPORTS = [
    "pa_last_error", "pa_create", "pa_destroy", "pa_element_grains", "pa_set_grains", "pa_set_solver", "pa_set_qp_form", "pa_last_residual",
    "pa_newton_iteration", "pa_assemble", "pa_linear_solve", "pa_commit_increment", "pa_run_increment",
    "pa_get_state", "pa_get_csr", "pa_sizes", "pa_qp_update", "pa_last_timings", "pa_last_qp_pass_ms", "pa_time_kernel",
    "pa_last_counts", "pa_pinv_solve", "pa_pinv_solve_noalloc", "pa_mlp_create", "pa_mlp_destroy", "pa_mlp_set_params", "pa_mlp_get_params",
    "pa_mlp_predict", "pa_mlp_grad", "pa_mlp_train", "pa_qp_update_warm", "pa_grain_averages",
    "pa_batch_create", "pa_batch_destroy", "pa_batch_set_grains", "pa_batch_set_solver", "pa_batch_set_graphs",
    "pa_batch_newton_iteration", "pa_batch_commit_increment", "pa_batch_tune", "pa_batch_qp_pass_ms", "pa_batch_last_timings",
    "pa_batch_get_state", "pa_batch_sizes", "pa_set_periodic", "pa_periodic_lines", "pa_create_phases", "pa_set_grain_phases", "pa_project_fields",
]


class PaMaterial(ctypes.Structure):
    _fields_ = [
        ("n_slip", ctypes.c_int),
        ("m", ctypes.c_double * 3 * MAX_SLIP), ("n", ctypes.c_double * 3 * MAX_SLIP),
        ("q", ctypes.c_double * MAX_SLIP * MAX_SLIP), ("C", ctypes.c_double * 6 * 6),
        ("s0", ctypes.c_double * MAX_SLIP), ("h0", ctypes.c_double * MAX_SLIP),
        ("ss", ctypes.c_double * MAX_SLIP), ("pw", ctypes.c_double * MAX_SLIP),
        ("stress_tol", ctypes.c_double), ("max_slip_search", ctypes.c_int), ("max_solver_iter", ctypes.c_int),
        ("kinematic_hardening", ctypes.c_int), ("twinning", ctypes.c_int),
        ("C1", ctypes.c_double * MAX_SLIP), ("C2", ctypes.c_double * MAX_SLIP),
        ("n_twin", ctypes.c_int), ("twin_shear", ctypes.c_double), ("twin_threshold", ctypes.c_double),
        ("twin_saturation", ctypes.c_double),
    ]


class PaProblem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("span", ctypes.c_double * 3),
        ("total_increments", ctypes.c_int), ("delT", ctypes.c_double), ("max_newton", ctypes.c_int),
        ("lin_tol", ctypes.c_double), ("lin_maxit", ctypes.c_int),
        ("face_dof_constrained", ctypes.c_int * 3 * 6), ("face_total_u", ctypes.c_double * 3 * 6),
        ("cyclic", ctypes.c_int), ("cyclic_face", ctypes.c_int), ("cyclic_dof", ctypes.c_int),
        ("quarter_cycle_time", ctypes.c_double),
        ("tabular", ctypes.c_int), ("n_table", ctypes.c_int), ("table_time", ctypes.c_double * MAX_TABLE),
        ("table_u", ctypes.c_double * MAX_TABLE * 3 * 6),
        ("velgrad", ctypes.c_int), ("vel_grad", ctypes.c_double * 3 * 3),
    ]


MAX_PBC_LINES = 64


class PaPeriodic(ctypes.Structure):
    """pa_periodic: the periodic constraint file's lines, the prm's vertex rows and the tabular table."""
    _fields_ = [
        ("n_vertex_lines", ctypes.c_int), ("vertex_dofs", ctypes.c_int * 5 * MAX_PBC_LINES),
        ("vertex_coef", ctypes.c_double * 5 * MAX_PBC_LINES),
        ("n_edge_lines", ctypes.c_int), ("edge_dofs", ctypes.c_int * 4 * MAX_PBC_LINES), ("edge_coef", ctypes.c_double * 4 * MAX_PBC_LINES),
        ("n_face_lines", ctypes.c_int), ("face_dofs", ctypes.c_int * 4 * MAX_PBC_LINES), ("face_coef", ctypes.c_double * 4 * MAX_PBC_LINES),
        ("vertex_known", ctypes.c_int * 24), ("vertex_total", ctypes.c_double * 24),
        ("tabular", ctypes.c_int), ("tabular_time", ctypes.c_double), ("n_table", ctypes.c_int),
        ("table_time", ctypes.c_double * MAX_TABLE), ("table_coef", ctypes.c_int * MAX_TABLE),
    ]


class PlasticityError(RuntimeError):
    pass


_LIB = None


def lib():
    """Load libplasticity_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise PlasticityError("%s is missing: build it with __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.pa_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise PlasticityError("plasticity_amd error %d: %s" % (rc, lib().pa_last_error().decode()))


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)

"""Readers for the reference's input formats and a ctypes driver for the oracle.

Test infrastructure only.  File formats follow the reference readers:
  prm.prm            deal.II ParameterHandler "set Key = value" (src/userInputParameters/userInputParameters.cc:17-545)
  slip/twin files    one vector per line, normalised on read (src/materialModels/crystalPlasticity/init.cc:73-123)
  latent hardening   n_slip x n_slip table (init.cc:51-69)
  grain-ID voxels    header lines, then for x: for y: one line of z values (src/utilityObjects/crystalOrientationsIO.cc:85-105)
  orientations       1 header line, then "id r1 r2 r3" (crystalOrientationsIO.cc:24-53)
  BC file            header lines, then "face dof totalU" (src/ellipticBVP/init.cc:327-348)
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
MAXT = 49


def parse_prm(path):
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line.startswith("set "):
                continue
            k, v = line[4:].split("=", 1)
            out[" ".join(k.split())] = v.strip()
    return out


def _floats(s):
    return [float(x) for x in s.replace(",", " ").split()]


def read_vectors(path, n):
    rows = []
    with open(path) as f:
        for line in f:
            vals = line.split()
            if len(vals) < 3:
                continue
            v = np.array([float(x) for x in vals[:3]])
            rows.append(v / np.sqrt((v * v).sum()))
            if len(rows) == n:
                break
    return np.array(rows)


def read_latent(path, n):
    rows = []
    with open(path) as f:
        for line in f:
            vals = line.split()
            if not vals:
                continue
            rows.append([float(x) for x in vals[:n]])
            if len(rows) == n:
                break
    return np.array(rows)


def read_voxels(path, header, nv):
    with open(path) as f:
        lines = f.read().split("\n")[header:]
    vox = np.zeros(nv, dtype=np.uint32)
    li = 0
    for x in range(nv[0]):
        for y in range(nv[1]):
            vals = lines[li].split()
            li += 1
            vox[x, y, :] = [int(v) for v in vals[: nv[2]]]
    return vox


def read_orientations(path, with_phase=False):
    """id -> Rodrigues rows; with_phase: also the fifth column, the grain's phase (multiphase
    orientation files, crystalOrientationsIO.cc:30-50)."""
    table, phase = {}, {}
    with open(path) as f:
        f.readline()
        for line in f:
            vals = line.split()
            if len(vals) >= 4:
                table[int(vals[0])] = [float(v) for v in vals[1:4]]
                if with_phase and len(vals) >= 5:
                    phase[int(vals[0])] = int(float(vals[4]))
    n = max(table) + 1
    rod = np.zeros((n, 3))
    for k, v in table.items():
        rod[k] = v
    if not with_phase:
        return rod
    ph = np.ones(n, np.int32)
    for k, v in phase.items():
        ph[k] = v
    return rod, ph


def read_bcs(path, header, count):
    with open(path) as f:
        lines = [l for l in f.read().split("\n")]
    rules = []
    for line in lines[header: header + count]:
        v = line.split()
        rules.append((int(v[0]), int(v[1]), float(v[2])))
    return rules


class Material(ctypes.Structure):
    _fields_ = [
        ("n_slip", ctypes.c_int), ("n_twin", ctypes.c_int), ("nT", ctypes.c_int),
        ("m", ctypes.c_double * 3 * MAXT), ("n", ctypes.c_double * 3 * MAXT),
        ("q", ctypes.c_double * MAXT * MAXT), ("C", ctypes.c_double * 6 * 6),
        ("s0", ctypes.c_double * MAXT), ("h0", ctypes.c_double * MAXT),
        ("ss", ctypes.c_double * MAXT), ("pw", ctypes.c_double * MAXT),
        ("C1", ctypes.c_double * MAXT), ("C2", ctypes.c_double * MAXT),
        ("tol", ctypes.c_double), ("max_slip_search", ctypes.c_int),
        ("max_solver_iter", ctypes.c_int), ("pinv_mode", ctypes.c_int),
        ("twinning", ctypes.c_int), ("twin_shear", ctypes.c_double),
        ("twin_threshold", ctypes.c_double), ("twin_saturation", ctypes.c_double),
    ]


class Problem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("span", ctypes.c_double * 3), ("n_increments", ctypes.c_int),
        ("total_increments", ctypes.c_int), ("delT", ctypes.c_double),
        ("max_newton", ctypes.c_int), ("lin_tol", ctypes.c_double),
        ("lin_maxit", ctypes.c_int), ("solver", ctypes.c_int),
        ("face_dof_constrained", ctypes.c_int * 3 * 6),
        ("face_total_u", ctypes.c_double * 3 * 6),
        ("cyclic", ctypes.c_int), ("cyclic_face", ctypes.c_int), ("cyclic_dof", ctypes.c_int),
        ("quarter_cycle_time", ctypes.c_double), ("nthreads", ctypes.c_int),
        ("tabular", ctypes.c_int), ("n_table", ctypes.c_int), ("table_time", ctypes.c_double * 32),
        ("table_u", ctypes.c_double * 32 * 3 * 6),
        ("velgrad", ctypes.c_int), ("vel_grad", ctypes.c_double * 3 * 3),
        ("periodic", ctypes.c_void_p),
    ]


MAX_PBC = 64


class Periodic(ctypes.Structure):
    """or_periodic: the constraint file's vertex / edge / face lines, the vertex rows of the prm and
    the tabular periodic table (ellipticBVP init.cc:66-131, userInputParameters.cc:126-136)."""
    _fields_ = [
        ("n_vertex_lines", ctypes.c_int), ("vertex_dofs", ctypes.c_int * 5 * MAX_PBC), ("vertex_coef", ctypes.c_double * 5 * MAX_PBC),
        ("n_edge_lines", ctypes.c_int), ("edge_dofs", ctypes.c_int * 4 * MAX_PBC), ("edge_coef", ctypes.c_double * 4 * MAX_PBC),
        ("n_face_lines", ctypes.c_int), ("face_dofs", ctypes.c_int * 4 * MAX_PBC), ("face_coef", ctypes.c_double * 4 * MAX_PBC),
        ("vertex_known", ctypes.c_int * 24), ("vertex_total", ctypes.c_double * 24),
        ("tabular", ctypes.c_int), ("tabular_time", ctypes.c_double), ("n_table", ctypes.c_int),
        ("table_time", ctypes.c_double * 32), ("table_coef", ctypes.c_int * 32),
    ]


def read_periodic_file(path, n_vertex, n_edge, n_face):
    """PeriodicBCsConstraints.txt as ellipticBVP init.cc:78-126 reads it: 4 header lines, n_vertex
    pairs of lines (5 integers: entry count and vertex rows; 5 coefficients), 5 header lines,
    n_edge pairs (4 rows, 4 coefficients), 5 header lines, n_face pairs."""
    with open(path) as f:
        lines = f.read().split("\n")
    pos = [0]

    def take():
        line = lines[pos[0]]
        pos[0] += 1
        return line

    def block(n_header, n_lines, width):
        for _ in range(n_header):
            take()
        rows, coefs = [], []
        for _ in range(n_lines):
            rows.append([int(v) for v in take().split()[:width]])
            coefs.append([float(v) for v in take().split()[:width]])
        return rows, coefs
    vr, vc = block(4, n_vertex, 5)
    er, ec = block(5, n_edge, 4)
    fr, fc = block(5, n_face, 4)
    return dict(vertex=(vr, vc), edge=(er, ec), face=(fr, fc))


def periodic_struct(per):
    """or_periodic (and pa_periodic, the same layout) from a periodic dict."""
    P = Periodic()
    for name, (rows, coefs) in (("vertex", per["vertex"]), ("edge", per["edge"]), ("face", per["face"])):
        setattr(P, "n_%s_lines" % name, len(rows))
        R, Cf = getattr(P, "%s_dofs" % name), getattr(P, "%s_coef" % name)
        for i, (r, c) in enumerate(zip(rows, coefs)):
            for k in range(len(r)):
                R[i][k] = r[k]
                Cf[i][k] = c[k]
    for i in range(24):
        P.vertex_known[i] = int(per["known"][i])
        P.vertex_total[i] = float(per["total"][i])
    if per.get("tabular"):
        P.tabular, P.tabular_time = 1, per["tabular_time"]
        P.n_table = len(per["table_time"])
        for i, (t, c) in enumerate(zip(per["table_time"], per["table_coef"])):
            P.table_time[i], P.table_coef[i] = t, int(c)
    return P


class Case:
    """Material + problem + grains, from explicit arrays (the oracle's view of a prm case)."""

    def __init__(self, m, n, q, C, s0, h0, ss, pw, tol=1e-9, max_slip_search=1,
                 max_solver_iter=1, nx=8, ny=8, nz=8, span=(1.0, 1.0, 1.0), delT=1.0,
                 total_time=10.0, bcs=((1, 1, 0.0), (1, 2, 0.0), (1, 3, 0.0), (2, 1, 0.01)),
                 max_newton=4, lin_tol=1e-10, lin_maxit=50000, elem_grain=None, grain_rod=None,
                 solver=0):
        self.m = np.asarray(m, float)
        self.n = np.asarray(n, float)
        self.q = np.asarray(q, float)
        self.C = np.asarray(C, float)
        self.n_slip = len(self.m)
        self.s0, self.h0, self.ss, self.pw = [np.asarray(x, float) for x in (s0, h0, ss, pw)]
        self.tol, self.max_slip_search, self.max_solver_iter = tol, max_slip_search, max_solver_iter
        self.nx, self.ny, self.nz, self.span = nx, ny, nz, tuple(span)
        self.delT, self.total_time = delT, total_time
        self.total_increments = int(round(total_time / delT))
        self.bcs = list(bcs)
        self.max_newton, self.lin_tol, self.lin_maxit, self.solver = max_newton, lin_tol, lin_maxit, solver
        self.elem_grain = None if elem_grain is None else np.asarray(elem_grain, np.uint32)
        self.grain_rod = None if grain_rod is None else np.asarray(grain_rod, float)
        self.cyclic = (0, 1, 1, 1.0)
        self.C1 = self.C2 = None  # kinematic hardening off
        self.tabular = None       # (times, [(face, dof, displacements)]) for tabular BCs
        self.vel_grad = None      # 3x3 velocity gradient L for velocity-gradient BCs
        # twinning: dict(m, n (normalised), s0, h0, ss, pw per twin system, q (latent ratios over
        # slip + twin systems), shear, threshold, saturation); None = the inert twin system
        self.twin = None
        self.periodic = None      # periodic BCs: dict of read_periodic_file + known, total (and the tabular table)

    # --- reference defaults for the inert twin system (userInputParameters.cc:300-311)
    def material(self, pinv_mode=1):
        mat = Material()
        ns = self.n_slip
        tw = self.twin
        ntw = len(tw["m"]) if tw else 1
        mat.n_slip, mat.n_twin, mat.nT = ns, ntw, ns + ntw
        for a in range(ns):
            for j in range(3):
                mat.m[a][j] = self.m[a, j]
                mat.n[a][j] = self.n[a, j]
            mat.s0[a], mat.h0[a], mat.ss[a], mat.pw[a] = self.s0[a], self.h0[a], self.ss[a], self.pw[a]
            for b in range(ns):
                mat.q[a][b] = self.q[a, b]
        if tw:  # twin systems after the slip systems (init.cc:125-175, 250-262)
            for k in range(ntw):
                a = ns + k
                for j in range(3):
                    mat.m[a][j] = tw["m"][k][j]
                    mat.n[a][j] = tw["n"][k][j]
                mat.s0[a], mat.h0[a], mat.ss[a], mat.pw[a] = tw["s0"][k], tw["h0"][k], tw["ss"][k], tw["pw"][k]
            for a in range(ns + ntw):
                for b in range(ns + ntw):
                    mat.q[a][b] = tw["q"][a][b]
            mat.twinning, mat.twin_shear = 1, tw["shear"]
            mat.twin_threshold, mat.twin_saturation = tw["threshold"], tw["saturation"]
        else:
            # dummy twin: n=(1,0,0), m=(0,1,0), s0=10e14, h0=1, a=0, s_s=10e15 (init.cc:178-185)
            mat.n[ns][0], mat.m[ns][1] = 1.0, 1.0
            mat.s0[ns], mat.h0[ns], mat.pw[ns], mat.ss[ns] = 10e14, 1.0, 0.0, 10e15
        for i in range(6):
            for j in range(6):
                mat.C[i][j] = self.C[i, j]
        mat.tol, mat.max_slip_search, mat.max_solver_iter = self.tol, self.max_slip_search, self.max_solver_iter
        if self.C1 is not None:
            for a in range(ns):
                mat.C1[a], mat.C2[a] = self.C1[a], self.C2[a]
        mat.pinv_mode = pinv_mode
        return mat

    def problem(self, n_increments=None, nthreads=1):
        pb = Problem()
        pb.nx, pb.ny, pb.nz = self.nx, self.ny, self.nz
        for d in range(3):
            pb.span[d] = self.span[d]
        pb.total_increments = self.total_increments
        pb.n_increments = self.total_increments if n_increments is None else n_increments
        pb.delT = self.delT
        pb.max_newton, pb.lin_tol, pb.lin_maxit, pb.solver = self.max_newton, self.lin_tol, self.lin_maxit, self.solver
        for f, d, u in self.bcs:
            pb.face_dof_constrained[f - 1][d - 1] = 1
            pb.face_total_u[f - 1][d - 1] = u
        pb.cyclic, pb.cyclic_face, pb.cyclic_dof, pb.quarter_cycle_time = self.cyclic
        pb.nthreads = nthreads
        if self.tabular is not None:
            times, rules = self.tabular
            pb.tabular, pb.n_table = 1, len(times)
            for i, t in enumerate(times):
                pb.table_time[i] = t
            for f, d, us in rules:
                pb.face_dof_constrained[f - 1][d - 1] = 1
                for i, u in enumerate(us):
                    pb.table_u[f - 1][d - 1][i] = u
        if self.vel_grad is not None:
            pb.velgrad = 1
            for a in range(3):
                for b in range(3):
                    pb.vel_grad[a][b] = self.vel_grad[a][b]
        if getattr(self, "periodic", None):
            self._pbc = periodic_struct(self.periodic)  # kept alive with the case
            pb.periodic = ctypes.cast(ctypes.pointer(self._pbc), ctypes.c_void_p)
        return pb

    @property
    def n_twin(self):
        return len(self.twin["m"]) if self.twin else 1

    @property
    def nT(self):
        return self.n_slip + self.n_twin


def _true(prm, key):
    return prm.get(key, "false").strip().lower() == "true"


def _phase_case(prm, d, k, base):
    """Phase k's material (keys with the suffix " k" for k > 1, userInputParameters.cc:286-545) as a
    Case sharing the base case's mesh and loading."""
    sfx = "" if k == 1 else " %d" % k
    ns = int(prm["Number of Slip Systems" + sfx])
    C = np.array([_floats(prm["Elastic Stiffness%s row %d" % (sfx, i + 1)]) for i in range(6)])
    m = read_vectors(os.path.join(d, prm["Slip Directions File" + sfx]), ns)
    n = read_vectors(os.path.join(d, prm["Slip Normals File" + sfx]), ns)
    tw = _true(prm, "Twinning enabled" + sfx)
    ntw = int(prm["Number of Twin Systems" + sfx]) if tw else 0
    q = read_latent(os.path.join(d, prm["Latent Hardening Ratio filename" + sfx]), ns + ntw)
    c = Case(m, n, q[:ns, :ns], C, _floats(prm["Initial Slip Resistance" + sfx]), _floats(prm["Initial Hardening Modulus" + sfx]),
             _floats(prm["Saturation Stress" + sfx]), _floats(prm["Power Law Exponent" + sfx]), tol=base.tol,
             max_slip_search=base.max_slip_search, max_solver_iter=base.max_solver_iter)
    if tw:
        c.twin = dict(m=read_vectors(os.path.join(d, prm["Twin Directions File" + sfx]), ntw),
                      n=read_vectors(os.path.join(d, prm["Twin Normals File" + sfx]), ntw),
                      s0=_floats(prm["Initial Slip Resistance Twin" + sfx]), h0=_floats(prm["Initial Hardening Modulus Twin" + sfx]),
                      ss=_floats(prm["Saturation Stress Twin" + sfx]), pw=_floats(prm["Power Law Exponent Twin" + sfx]), q=q,
                      shear=float(prm["Characteristic Twin Shear" + sfx]), threshold=float(prm["Twin Threshold Fraction" + sfx]),
                      saturation=float(prm.get("Twin Saturation Factor" + sfx, "0")))
    if _true(prm, "Enable Kinematic Hardening" + sfx):
        c.C1 = np.array(_floats(prm["C_1 Slip Kinematic Hardening" + sfx]))
        c.C2 = np.array(_floats(prm["C_2 Slip Kinematic Hardening" + sfx]))
    return c


def load_reference_case(folder, twinning=False, prm_name="prm.prm"):
    """A Case built from one of the reference application folders copied under tests/golden; with
    twinning=True the prm's twin systems and parameters are read (userInputParameters.cc twin
    block), otherwise the case runs with the inert twin system."""
    d = os.path.join(GOLDEN, folder)
    prm = parse_prm(os.path.join(d, prm_name))
    ns = int(prm["Number of Slip Systems"])
    C = np.array([_floats(prm["Elastic Stiffness row %d" % (i + 1)]) for i in range(6)])
    m = read_vectors(os.path.join(d, prm["Slip Directions File"]), ns)
    n = read_vectors(os.path.join(d, prm["Slip Normals File"]), ns)
    q = read_latent(os.path.join(d, prm["Latent Hardening Ratio filename"]), ns)
    refine = int(prm.get("Refine factor", "0"))
    sub = [int(prm.get("Subdivisions %s" % a, "1")) for a in "XYZ"]
    nel = [s * 2 ** refine for s in sub]
    span = [float(prm["Domain size %s" % a]) for a in "XYZ"]
    nv = [int(float(prm["Voxels in %s direction" % a])) for a in "XYZ"]
    vox = read_voxels(os.path.join(d, prm["Grain ID file name"]), int(prm.get("Header Lines GrainID File", "0")), nv)
    rod = read_orientations(os.path.join(d, prm["Orientations file name"]))
    tabular = None
    if prm.get("Use Tabular BCs", "false").strip().lower() == "true":  # init.cc:389-415
        nt = int(prm["Number of time data for Tabular BCs"])
        rules = []
        with open(os.path.join(d, prm["Tabular Boundary condition filename"])) as f:
            for line in f:
                v = line.split()
                if len(v) >= 2 + nt and len(rules) < int(prm["Number of tabular boundary conditions"]):
                    rules.append((int(v[0]), int(v[1]), [float(x) for x in v[2:2 + nt]]))
        tabular = (_floats(prm["Tabular Time Table"])[:nt], rules)
        bcs = []
    elif _true(prm, "Use velocity gradient BC") or _true(prm, "Use Periodic BCs"):  # userInputParameters.cc:78, 126
        bcs = []
    else:
        bcs = read_bcs(os.path.join(d, prm["Boundary condition filename"]), int(prm["BC file number of header lines"]),
                       int(prm["Number of boundary conditions"]))
    case = Case(m, n, q, C, _floats(prm["Initial Slip Resistance"]), _floats(prm["Initial Hardening Modulus"]),
                _floats(prm["Saturation Stress"]), _floats(prm["Power Law Exponent"]),
                tol=float(prm["Stress Tolerance"]), max_slip_search=int(prm["Max Slip Search Iterations"]),
                max_solver_iter=int(prm["Max Solver Iterations"]), nx=nel[0], ny=nel[1], nz=nel[2], span=span,
                delT=float(prm["Time increments"]), total_time=float(prm["Total time"]), bcs=bcs,
                max_newton=int(prm["Maximum non linear iterations"]),
                lin_tol=float(prm["Relative linear solver tolerance"]),
                lin_maxit=int(prm["Maximum linear solver iterations"]))
    if twinning and _true(prm, "Twinning enabled"):
        ntw = int(prm["Number of Twin Systems"])
        case.twin = dict(m=read_vectors(os.path.join(d, prm["Twin Directions File"]), ntw),
                         n=read_vectors(os.path.join(d, prm["Twin Normals File"]), ntw),
                         s0=_floats(prm["Initial Slip Resistance Twin"]), h0=_floats(prm["Initial Hardening Modulus Twin"]),
                         ss=_floats(prm["Saturation Stress Twin"]), pw=_floats(prm["Power Law Exponent Twin"]),
                         q=read_latent(os.path.join(d, prm["Latent Hardening Ratio filename"]), ns + ntw),
                         shear=float(prm["Characteristic Twin Shear"]), threshold=float(prm["Twin Threshold Fraction"]),
                         saturation=float(prm["Twin Saturation Factor"]))
    if prm.get("Enable Kinematic Hardening", "false").strip().lower() == "true":  # userInputParameters.cc:274-279
        case.C1 = np.array(_floats(prm["C_1 Slip Kinematic Hardening"]))
        case.C2 = np.array(_floats(prm["C_2 Slip Kinematic Hardening"]))
    if prm.get("Enable cyclic loading", "false").strip().lower() == "true":
        case.cyclic = (1, int(prm["Cyclic loading face"]), int(prm["Cyclic loading direction"]), float(prm["Quarter cycle time"]))
    case.tabular = tabular
    if _true(prm, "Use velocity gradient BC"):
        case.vel_grad = np.array([_floats(prm["Velocity gradient row %d" % (i + 1)]) for i in range(3)])
    case.periodic = None
    if _true(prm, "Use Periodic BCs"):
        case.periodic = read_periodic_file(os.path.join(d, prm["Periodic Boundary condition Constraint filename"]),
                                           int(prm["Number of Vertices Constraints"]), int(prm["Number of Edges Constraints"]),
                                           int(prm["Number of Faces Constraints"]))
        case.periodic["known"] = [int(v) for v in prm["Vertices Periodic BCs row 1"].split(",")]
        case.periodic["total"] = _floats(prm["Vertices Periodic BCs row 2"])
        if _true(prm, "Use Tabular Periodic BCs"):
            case.periodic.update(tabular=True, tabular_time=float(prm["periodic Tabular time"]),
                                 table_time=_floats(prm["Tabular Periodic Time Table"]),
                                 table_coef=[int(v) for v in prm["Tabular Periodic Time Table Coefficient"].split(",")])
    case.vox, case.nv = vox, nv
    case.grain_rod = rod
    case.phases = None  # multiphase: one Case per phase (materials), grain_phase[grain row] in 1..n
    if _true(prm, "Enable Multiphase"):
        case.phases = [_phase_case(prm, d, k, case) for k in range(1, int(prm["Number of Phases"]) + 1)]
        case.grain_rod, case.grain_phase = read_orientations(os.path.join(d, prm["Orientations file name"]), with_phase=True)
    case.elem_grain = oracle_element_grains(case)
    golden = os.path.join(d, "correct_stressstrain.txt")
    case.golden = np.loadtxt(golden, skiprows=1) if os.path.exists(golden) else None
    return case


# ------------------------------------------------------------------ oracle
_ORACLE = None
_ORACLE_PATH = None


def use_oracle_build(path):
    """Load the oracle from another build of oracle/cp_oracle.cpp (same source, other flags)."""
    global _ORACLE, _ORACLE_PATH
    _ORACLE, _ORACLE_PATH = None, path


def oracle():
    global _ORACLE
    if _ORACLE is None:
        path = _ORACLE_PATH or os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (run __graft_entry__.build())" % path)
        lib = ctypes.CDLL(path)
        D, I, P = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
        lib.or_qp_update.restype = ctypes.c_int
        lib.or_simulate.restype = ctypes.c_int
        lib.or_sparsity.restype = ctypes.c_longlong
        lib.or_material_id.restype = ctypes.c_uint
        lib.or_set_pinv_rtol.argtypes = [ctypes.c_double]
        _ORACLE = lib
    return _ORACLE


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def oracle_element_grains(case):
    lib = oracle()
    nel = case.nx * case.ny * case.nz
    out = np.zeros(nel, np.uint32)
    vox = np.ascontiguousarray(case.vox, dtype=np.uint32)
    nv = np.array(case.nv, np.int32)
    pb = case.problem()
    lib.or_element_grains(ptr(vox), ptr(nv), ctypes.byref(pb), ptr(out))
    return out


def oracle_simulate(case, n_increments, pinv_mode=1, nthreads=1, want_state=False, solver=None):
    lib = oracle()
    if solver is not None:
        case.solver = solver
    mat = case.material(pinv_mode)
    pb = case.problem(n_increments, nthreads)
    rows = np.zeros((n_increments, 15))
    its = np.zeros(n_increments * case.max_newton, np.int32)
    res = np.zeros(n_increments * case.max_newton)
    nq = case.nx * case.ny * case.nz * 8
    if getattr(case, "phases", None):  # the largest phase's sizes
        tws = [ph.n_twin for ph in case.phases if ph.twin]
        W = 27 + max(ph.nT for ph in case.phases) + 3 + max(ph.n_slip for ph in case.phases) + (4 + max(tws) if tws else 0)
    else:
        W = 27 + case.nT + 3 + case.n_slip + (4 + case.n_twin if case.twin else 0)
    state = np.zeros((nq, W)) if want_state else None
    u = np.zeros(3 * (case.nx + 1) * (case.ny + 1) * (case.nz + 1)) if want_state else None
    rod = np.ascontiguousarray(case.grain_rod, float)
    eg = np.ascontiguousarray(case.elem_grain, np.uint32)
    if getattr(case, "phases", None):
        mats = (Material * len(case.phases))(*[ph.material(pinv_mode) for ph in case.phases])
        gp = np.ascontiguousarray(case.grain_phase, np.int32)
        rc = lib.or_simulate_phases(mats, ctypes.c_int(len(case.phases)), ptr(gp), ctypes.byref(pb), ptr(eg), ptr(rod),
                                    ctypes.c_int(len(rod)), ptr(rows), ptr(its), ptr(res), ptr(state) if want_state else None,
                                    ptr(u) if want_state else None)
    else:
        rc = lib.or_simulate(ctypes.byref(mat), ctypes.byref(pb), ptr(eg), ptr(rod), ctypes.c_int(len(rod)),
                             ptr(rows), ptr(its), ptr(res), ptr(state) if want_state else None,
                             ptr(u) if want_state else None)
    assert rc == 0, rc
    out = {"rows": rows, "lin_iters": its, "res_norms": res}
    if want_state:
        out["state"], out["u"] = state, u
    return out


def qp_F(u, nx, ny, nz, h):
    """F = I + sum_v u_v (x) grad N_v at every qp (Q = 8 e + q, element e with x fastest, Gauss
    point q with bit d selecting 1/2 +- 1/(2 sqrt 3) along axis d), from nodal u (x fastest)."""
    g = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))
    nnx, nny = nx + 1, ny + 1
    F = np.zeros((nx * ny * nz * 8, 3, 3))
    for e in range(nx * ny * nz):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        for q in range(8):
            Fq = np.eye(3)
            for v in range(8):
                nd = (ex + (v & 1)) + nnx * ((ey + ((v >> 1) & 1)) + nny * (ez + ((v >> 2) & 1)))
                N, dN = [], []
                for d in range(3):
                    xi = g[(q >> d) & 1]
                    b = (v >> d) & 1
                    N.append(xi if b else 1 - xi)
                    dN.append((1.0 if b else -1.0) / h[d])
                grad = np.array([dN[0] * N[1] * N[2], N[0] * dN[1] * N[2], N[0] * N[1] * dN[2]])
                Fq += np.outer(u[3 * nd:3 * nd + 3], grad)
            F[8 * e + q] = Fq
    return F


def check_committed_state(case, u1, T1, s1, Fp1, Fp0, s0, tol=1e-10, max_frac=0.01):
    """Every qp's committed state after an increment (T1, s1, Fp1, from the run under test) against
    the oracle's calculatePlasticity(.,.,0) on that run's own inputs: F from its displacement u1, Fp0
    and s0 committed before the increment.  A qp whose numerically singular active set leaves a
    singular value at round-off may take the other side of the pseudo-inverse's keep/drop decision;
    it passes only if the run's state is the oracle's at an input within 1e-12 of its own (400 seeded
    perturbations of F), and at most max_frac of the qps may need that.
    Returns (qps, plastic qps, qps on the other side of a round-off decision)."""
    n_s = case.n_slip
    F = qp_F(u1, case.nx, case.ny, case.nz, tuple(case.span[d] / (case.nx, case.ny, case.nz)[d] for d in range(3)))
    nq, skipped, plastic = len(F), 0, 0
    rel = lambda a, b: np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)
    for Q in range(nq):
        rod = case.grain_rod[case.elem_grain[Q // 8]]
        Fpq = np.asarray(Fp0[Q]).reshape(3, 3)
        sq = np.append(s0[Q][:n_s], 10e14)
        o = oracle_qp(case, rod, F[Q], Fpq, sq, 0)
        plastic += o["n_active"] > 0
        if rel(T1[Q], o["T"].ravel()) < tol:
            assert rel(s1[Q][:n_s], o["s"][:n_s]) < tol, Q
            assert rel(Fp1[Q], o["Fp"].ravel()) < tol, Q
            continue
        r = np.random.default_rng(1000 + Q)
        near = False
        for _ in range(400):
            o2 = oracle_qp(case, rod, F[Q] * (1 + 1e-12 * r.standard_normal((3, 3))), Fpq, sq, 0)
            if rel(T1[Q], o2["T"].ravel()) < 1e-6 and rel(s1[Q][:n_s], o2["s"][:n_s]) < 1e-6:
                near = True
                break
        assert near, (Q, rel(T1[Q], o["T"].ravel()))
        skipped += 1
    assert skipped <= max(1, int(max_frac * nq))
    return nq, plastic, skipped


def oracle_qp(case, rot, F, Fp=None, s=None, flag=1, pinv_mode=1, slip=None, W=None):
    lib = oracle()
    mat = case.material(pinv_mode)
    nT, ns = case.nT, case.n_slip
    Fp = np.eye(3) if Fp is None else np.asarray(Fp, float)
    s = np.array(list(case.s0) + [10e14]) if s is None else np.asarray(s, float)
    W = np.zeros(nT) if W is None else np.append(np.asarray(W, float)[:ns], np.zeros(nT - ns))
    slip_c = np.zeros(ns) if slip is None else np.asarray(slip, float)
    P, T, D, Fe, Fpo = np.zeros(9), np.zeros(9), np.zeros(81), np.zeros(9), np.zeros(9)
    so, Wo, sl = np.zeros(nT), np.zeros(nT), slip_c.copy()
    F = np.ascontiguousarray(np.asarray(F, float).reshape(9))
    Fp = np.ascontiguousarray(Fp.reshape(9))
    rot = np.ascontiguousarray(np.asarray(rot, float))
    na = lib.or_qp_update(ctypes.byref(mat), ptr(rot), ptr(F), ptr(Fp), ptr(s), ptr(W), ptr(slip_c), ctypes.c_int(flag),
                          ptr(P), ptr(T), ptr(D), ptr(Fe), ptr(Fpo), ptr(so), ptr(Wo), ptr(sl))
    return {"P": P.reshape(3, 3), "T": T.reshape(3, 3), "dPdF": D.reshape(9, 9), "Fe": Fe.reshape(3, 3),
            "Fp": Fpo.reshape(3, 3), "s": so, "slip": sl, "n_active": na, "W": Wo}


def oracle_qp_twin(case, rot, F, Fp, s, flag=1, pinv_mode=1):
    """or_qp_update_twin for a twinning case: s has n_slip + n_twin entries; returns the oracle's
    outputs with x = the slip increments of all systems (slip_iter and twin fraction times the
    characteristic twin shear, from zero histories)."""
    lib = oracle()
    mat = case.material(pinv_mode)
    nT, ns, ntw = case.nT, case.n_slip, case.n_twin
    P, T, D, Fe, Fpo = np.zeros(9), np.zeros(9), np.zeros(81), np.zeros(9), np.zeros(9)
    so, Wo, sl, tw = np.zeros(nT), np.zeros(nT), np.zeros(ns), np.zeros(ntw)
    na = lib.or_qp_update_twin(ctypes.byref(mat), ptr(np.ascontiguousarray(rot, float)),
                               ptr(np.ascontiguousarray(F, float).reshape(9)), ptr(np.ascontiguousarray(Fp, float).reshape(9)),
                               ptr(np.ascontiguousarray(s, float)), ptr(np.zeros(nT)), ptr(np.zeros(ns)), ptr(np.zeros(ntw)),
                               ctypes.c_int(flag), ptr(P), ptr(T), ptr(D), ptr(Fe), ptr(Fpo), ptr(so), ptr(Wo), ptr(sl), ptr(tw))
    return {"P": P.reshape(3, 3), "T": T.reshape(3, 3), "dPdF": D.reshape(9, 9), "Fe": Fe.reshape(3, 3),
            "Fp": Fpo.reshape(3, 3), "s": so, "x": np.concatenate([sl, tw * case.twin["shear"]]), "n_active": na}


def oracle_qp_warm(case, rot, F, Fp, s, x0=None, flag=0, pinv_mode=1):
    """or_qp_update_warm: the local loop started from x0 (n_slip values, None = from zero)."""
    lib = oracle()
    mat = case.material(pinv_mode)
    nT, ns = case.nT, case.n_slip
    s = np.append(np.asarray(s, float)[:ns], [10e14] * (nT - ns))
    W = np.zeros(nT)
    P, T, D, Fe, Fpo = np.zeros(9), np.zeros(9), np.zeros(81), np.zeros(9), np.zeros(9)
    so, Wo, sl = np.zeros(nT), np.zeros(nT), np.zeros(ns)
    it = ctypes.c_int(0)
    x0 = None if x0 is None else np.ascontiguousarray(x0, float)
    na = lib.or_qp_update_warm(ctypes.byref(mat), ptr(np.ascontiguousarray(rot, float)),
                               ptr(np.ascontiguousarray(F, float).reshape(9)), ptr(np.ascontiguousarray(Fp, float).reshape(9)),
                               ptr(s), ptr(W), ptr(np.zeros(ns)), ctypes.c_int(flag), ptr(x0) if x0 is not None else None,
                               ptr(P), ptr(T), ptr(D), ptr(Fe), ptr(Fpo), ptr(so), ptr(Wo), ptr(sl), ctypes.byref(it))
    return {"P": P.reshape(3, 3), "T": T.reshape(3, 3), "dPdF": D.reshape(9, 9), "Fe": Fe.reshape(3, 3),
            "Fp": Fpo.reshape(3, 3), "s": so[:ns], "slip": sl, "n_active": na, "iters": it.value}


def fcc_case(**kw):
    """The FCC material of the 8^3 tension workload (cubic 170/124/75 GPa, s0 16, h0 180, s_s 148, a 2.25, q 1.4)."""
    d = os.path.join(GOLDEN, "fcc_simpleTension")
    m = read_vectors(os.path.join(d, "slipDirections.txt"), 12)
    n = read_vectors(os.path.join(d, "slipNormals.txt"), 12)
    q = read_latent(os.path.join(d, "LatentHardeningRatio.txt"), 12)
    C = np.zeros((6, 6))
    C[:3, :3] = 124e3
    np.fill_diagonal(C[:3, :3], 170e3)
    C[3, 3] = C[4, 4] = C[5, 5] = 75e3
    args = dict(s0=[16.0] * 12, h0=[180.0] * 12, ss=[148.0] * 12, pw=[2.25] * 12)
    args.update(kw)
    return Case(m, n, q, C, **args)


def velgrad_case():
    """A single rotated FCC grain (elastic: s0 = 1e6) on a 4x3x5 mesh of span (1, 0.8, 1.2) driven by a
    non-symmetric, non-nilpotent velocity gradient.  Under all-boundary affine BCs a homogeneous body
    deforms affinely, so after n increments every node sits at u = (F_n - I) X."""
    case = fcc_case(s0=[1e6] * 12, nx=4, ny=3, nz=5, span=(1.0, 0.8, 1.2), delT=1.0, total_time=10.0,
                    bcs=(), lin_tol=1e-14)
    case.vel_grad = np.array([[0.012, 0.02, -0.004], [0.003, -0.006, 0.008], [-0.005, 0.001, 0.009]])
    case.grain_rod = np.array([[0.21, -0.13, 0.34]])
    case.elem_grain = np.zeros(case.nx * case.ny * case.nz, np.uint32)
    return case


def velgrad_affine_u(case, n):
    """(F_n - I) X at every node, F_(k+1) = (I - L delT)^-1 F_k from F_0 = I (updateBeforeIncrement.cc:14-31)."""
    M = np.linalg.inv(np.eye(3) - case.vel_grad * case.delT)
    F = np.linalg.matrix_power(M, n)
    h = np.array(case.span) / np.array([case.nx, case.ny, case.nz])
    k, j, i = np.meshgrid(np.arange(case.nz + 1), np.arange(case.ny + 1), np.arange(case.nx + 1), indexing="ij")
    X = np.stack([i.ravel() * h[0], j.ravel() * h[1], k.ravel() * h[2]], axis=1)  # node = i + nnx (j + nny k)
    return (X @ (F - np.eye(3)).T).ravel()

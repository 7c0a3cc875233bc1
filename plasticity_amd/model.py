"""Host-side mirror of the reference's crystalPlasticity<3> / ellipticBVP<3> / userInputParameters API
over the C ABI (include/plasticity_amd.h).

Reference surfaces mirrored (paths relative to the reference repository root):
  userInputParameters(file)                     src/userInputParameters/userInputParameters.cc:5-545
  crystalOrientationsIO.loadOrientations(...)   src/utilityObjects/crystalOrientationsIO.cc:66-111
  crystalOrientationsIO.loadOrientationVector   src/utilityObjects/crystalOrientationsIO.cc:11-62
  crystalPlasticity(userInputs).run()           include/crystalPlasticity.h:20, src/ellipticBVP/run.cc:9-30
  applications/crystalPlasticity/main.cc:10-68  -> plasticity_amd.model.main(prm)
"""
import ctypes
import os

import numpy as np

from . import native
from .native import PaMaterial, PaProblem, check, f64, ptr


# ------------------------------------------------------------------ prm input
class userInputParameters:
    """The prm keys this path reads, with the reference's declared defaults
    (userInputParameters.cc:546-968).  Unknown keys are kept in .raw and ignored."""

    def __init__(self, path):
        self.path = os.path.abspath(path)
        self.dir = os.path.dirname(self.path)
        raw = {}
        with open(path) as f:
            for line in f:
                line = line.split("#", 1)[0].strip()
                if not line.startswith("set "):
                    continue
                k, v = line[4:].split("=", 1)
                raw[" ".join(k.split())] = v.strip()
        self.raw = raw
        g = lambda k, d=None: raw.get(k, d)
        fl = lambda s: [float(x) for x in s.replace(",", " ").split()] if s else []
        b = lambda k, d: (g(k, d).strip().lower() == "true")
        self.dim = int(g("Number of dimensions", "3"))
        if self.dim != 3:
            raise native.PlasticityError("only 3-d problems are supported")
        self.feOrder = int(g("Order of finite elements", "-1"))
        self.quadOrder = int(g("Order of quadrature", "-1"))
        if self.feOrder != 1 or self.quadOrder != 2:
            raise native.PlasticityError("only Q1 elements with QGauss(2) are implemented")
        self.span = [float(g("Domain size %s" % a, "-1")) for a in "XYZ"]
        self.subdivisions = [int(g("Subdivisions %s" % a, "1")) for a in "XYZ"]
        self.meshRefineFactor = int(g("Refine factor", "-1"))
        self.readExternalMesh = b("Use external mesh", "false")
        self.externalMeshParameter = float(g("External mesh parameter", "0"))
        self.writeOutput = b("Write Output", "false")
        self.outputDirectory = g("Output Directory", "results")
        self.skipOutputSteps = int(g("Skip Output Steps", "-1"))  # userInputParameters.cc:55, 161-167
        self.output_Eqv_strain = b("Output Equivalent strain", "false")
        self.output_Eqv_stress = b("Output Equivalent stress", "false")
        self.output_Twin = b("Output Twin fractions", "false")
        self.delT = float(g("Time increments", "-1"))
        self.totalTime = float(g("Total time", "-1"))
        self.enableSimpleBCs = b("Use Simple BCs", "true")
        self.BCfilename = g("Boundary condition filename", "")
        self.BCheaderLines = int(g("BC file number of header lines", "0"))
        self.NumberofBCs = int(g("Number of boundary conditions", "0"))
        self.maxLinearSolverIterations = int(g("Maximum linear solver iterations", "-1"))
        self.relLinearSolverTolerance = float(g("Relative linear solver tolerance", "-1"))
        self.maxNonLinearIterations = int(g("Maximum non linear iterations", "-1"))
        self.elasticStiffness1 = [fl(g("Elastic Stiffness row %d" % (i + 1), "")) for i in range(6)]
        self.numSlipSystems1 = int(g("Number of Slip Systems", "0"))
        self.latentHardeningRatioFileName1 = g("Latent Hardening Ratio filename", "")
        self.initialSlipResistance1 = fl(g("Initial Slip Resistance", ""))
        self.initialHardeningModulus1 = fl(g("Initial Hardening Modulus", ""))
        self.powerLawExponent1 = fl(g("Power Law Exponent", ""))
        self.saturationStress1 = fl(g("Saturation Stress", ""))
        self.slipDirectionsFile1 = g("Slip Directions File", "")
        self.slipNormalsFile1 = g("Slip Normals File", "")
        self.enableKinematicHardening1 = b("Enable Kinematic Hardening", "false")  # userInputParameters.cc:274-279
        self.C_1_slip1 = fl(g("C_1 Slip Kinematic Hardening", "")) if self.enableKinematicHardening1 else []
        self.C_2_slip1 = fl(g("C_2 Slip Kinematic Hardening", "")) if self.enableKinematicHardening1 else []
        self.enableTwinning1 = b("Twinning enabled", "false")  # userInputParameters.cc twin block
        if self.enableTwinning1:
            self.numTwinSystems1 = int(g("Number of Twin Systems", "0"))
            self.initialSlipResistanceTwin1 = fl(g("Initial Slip Resistance Twin", ""))
            self.initialHardeningModulusTwin1 = fl(g("Initial Hardening Modulus Twin", ""))
            self.powerLawExponentTwin1 = fl(g("Power Law Exponent Twin", ""))
            self.saturationStressTwin1 = fl(g("Saturation Stress Twin", ""))
            self.twinSaturationFactor = float(g("Twin Saturation Factor", "0"))
            self.twinThresholdFraction = float(g("Twin Threshold Fraction", "0"))
            self.twinShear = float(g("Characteristic Twin Shear", "0"))
            self.twinDirectionsFile1 = g("Twin Directions File", "")
            self.twinNormalsFile1 = g("Twin Normals File", "")
        self.enableMultiphase = b("Enable Multiphase", "false")
        self.modelStressTolerance = float(g("Stress Tolerance", "-1"))
        self.modelMaxSlipSearchIterations = int(g("Max Slip Search Iterations", "-1"))
        self.modelMaxSolverIterations = int(g("Max Solver Iterations", "-1"))
        self.grainIDFile = g("Grain ID file name", "")
        self.numPts = [int(float(g("Voxels in %s direction" % a, "0"))) for a in "XYZ"]
        self.grainOrientationsFile = g("Orientations file name", "")
        self.headerLinesGrainIDFile = int(g("Header Lines GrainID File", "0"))
        self.enableCyclicLoading = b("Enable cyclic loading", "false")
        self.cyclicLoadingFace = int(g("Cyclic loading face", "1"))
        self.cyclicLoadingDOF = int(g("Cyclic loading direction", "1"))
        self.quarterCycleTime = float(g("Quarter cycle time", "1"))
        self.enableTabularBCs = b("Use Tabular BCs", "false")  # userInputParameters.cc:84-88
        self.Tabular_BCfilename = g("Tabular Boundary condition filename", "")
        self.tabularBCs_InputStepNumber = int(g("Number of time data for Tabular BCs", "0"))
        self.tabularNumberofBCs = int(g("Number of tabular boundary conditions", "0"))
        self.tabularTimeInput = fl(g("Tabular Time Table", ""))
        if self.enableTabularBCs and (self.enableSimpleBCs or self.enableCyclicLoading):
            raise native.PlasticityError("tabular boundary conditions together with simple or cyclic ones are not implemented")
        # grain-averaged output (userInputParameters.cc:52-60, 663-671)
        self.tabularOutput = b("Tabular Output", "false")
        self.tabularTimeOutput = fl(g("Tabular Time Output Table", ""))
        self.writeGrainAveragedOutput = b("Write Grain Averaged Output", "false")
        self.skipGrainAveragedOutputSteps = int(g("Skip Grain Averaged Output Steps", "-1"))
        self.numberOfGrainAverageDataOutput = int(g("Number of Grain Averaged Output Variables", "-1"))
        self.writeQuadratureOutput = b("Write Quadrature Output", "false")
        self.skipQuadratureOutputSteps = int(g("Skip Quadrature Output Steps", "-1"))
        for key in ("Enable Advanced Twin Model", "Enable Advanced Rate Dependent Model", "Enable User Material Model"):
            if b(key, "false"):
                raise native.PlasticityError("'%s' is outside the implemented path" % key)
        if self.writeGrainAveragedOutput and b("Output Buffer Layer Removal Feature", "false"):
            raise native.PlasticityError("the output buffer layer is outside the implemented path")
        self.useVelocityGrad = b("Use velocity gradient BC", "false")  # userInputParameters.cc:78-83
        self.targetVelGrad = [fl(g("Velocity gradient row %d" % (i + 1), "0, 0, 0")) for i in range(3)]
        if self.useVelocityGrad and any(len(r) != 3 for r in self.targetVelGrad):
            raise native.PlasticityError("'Velocity gradient row' needs three values")
        # periodic BCs (userInputParameters.cc:126-136)
        self.enablePeriodicBCs = b("Use Periodic BCs", "false")
        self.Periodic_BCfilename = g("Periodic Boundary condition Constraint filename", "PeriodicBCsConstraints.txt")
        self.numberVerticesConstraint = int(g("Number of Vertices Constraints", "12"))
        self.numberEdgesConstraint = int(g("Number of Edges Constraints", "27"))
        self.numberFacesConstraint = int(g("Number of Faces Constraints", "9"))
        self.periodicBCsInput = [int(float(x)) for x in fl(g("Vertices Periodic BCs row 1", ""))]
        self.periodicBCsInput2 = fl(g("Vertices Periodic BCs row 2", ""))
        self.enableTabularPeriodicBCs = b("Use Tabular Periodic BCs", "false")
        self.periodicTabularTime = float(g("periodic Tabular time", "-1"))
        self.tabularPeriodicTimeInput = fl(g("Tabular Periodic Time Table", ""))
        self.tabularPeriodicCoef = [int(float(x)) for x in fl(g("Tabular Periodic Time Table Coefficient", ""))]
        if self.enablePeriodicBCs and (len(self.periodicBCsInput) != 24 or len(self.periodicBCsInput2) != 24):
            raise native.PlasticityError("'Vertices Periodic BCs row 1/2' need 24 values")
        for key in ("Use Neumann BCs",
                    "Use Indentation BCs", "Use Torsion BCs", "Flag To Use Taylor Model", "Use Element Deletion"):
            if b(key, "false"):
                raise native.PlasticityError("'%s' is outside the implemented path" % key)
        if self.readExternalMesh:
            raise native.PlasticityError("external meshes are outside the implemented path")
        # several phases (userInputParameters.cc:286-545): phase k's keys carry the suffix " k"
        self.numberofPhases = int(g("Number of Phases", "1")) if self.enableMultiphase else 1
        if not 1 <= self.numberofPhases <= 4:
            raise native.PlasticityError("'Number of Phases' must be 1 to 4")
        self.phaseKeys = []
        for k in range(2, self.numberofPhases + 1):
            sfx = " %d" % k
            ph = dict(elasticStiffness=[fl(g("Elastic Stiffness%s row %d" % (sfx, i + 1), "")) for i in range(6)],
                      numSlipSystems=int(g("Number of Slip Systems" + sfx, "0")),
                      latentHardeningRatioFileName=g("Latent Hardening Ratio filename" + sfx, ""),
                      initialSlipResistance=fl(g("Initial Slip Resistance" + sfx, "")),
                      initialHardeningModulus=fl(g("Initial Hardening Modulus" + sfx, "")),
                      powerLawExponent=fl(g("Power Law Exponent" + sfx, "")), saturationStress=fl(g("Saturation Stress" + sfx, "")),
                      slipDirectionsFile=g("Slip Directions File" + sfx, ""), slipNormalsFile=g("Slip Normals File" + sfx, ""),
                      enableKinematicHardening=b("Enable Kinematic Hardening" + sfx, "false"),
                      C_1=fl(g("C_1 Slip Kinematic Hardening" + sfx, "")), C_2=fl(g("C_2 Slip Kinematic Hardening" + sfx, "")),
                      enableTwinning=b("Twinning enabled" + sfx, "false"))
            if ph["enableTwinning"]:
                ph.update(numTwinSystems=int(g("Number of Twin Systems" + sfx, "0")),
                          initialSlipResistanceTwin=fl(g("Initial Slip Resistance Twin" + sfx, "")),
                          initialHardeningModulusTwin=fl(g("Initial Hardening Modulus Twin" + sfx, "")),
                          powerLawExponentTwin=fl(g("Power Law Exponent Twin" + sfx, "")),
                          saturationStressTwin=fl(g("Saturation Stress Twin" + sfx, "")),
                          twinSaturationFactor=float(g("Twin Saturation Factor" + sfx, "0")),
                          twinThresholdFraction=float(g("Twin Threshold Fraction" + sfx, "0")),
                          twinShear=float(g("Characteristic Twin Shear" + sfx, "0")),
                          twinDirectionsFile=g("Twin Directions File" + sfx, ""), twinNormalsFile=g("Twin Normals File" + sfx, ""))
            self.phaseKeys.append(ph)

    def file(self, name):
        return name if os.path.isabs(name) else os.path.join(self.dir, name)


def read_vectors(path, n):
    """Slip/twin vector files (init.cc:73-123); normalisation happens in the library."""
    rows = []
    with open(path) as f:
        for line in f:
            v = line.split()
            if len(v) >= 3:
                rows.append([float(x) for x in v[:3]])
            if len(rows) == n:
                break
    if len(rows) < n:
        raise native.PlasticityError("%s: expected %d vectors" % (path, n))
    return np.array(rows)


def read_periodic_constraints(path, n_vertex, n_edge, n_face):
    """PeriodicBCsConstraints.txt (ellipticBVP init.cc:78-126): 4 header lines, then per vertex line a
    line of 5 integers (entry count, vertex rows) and a line of 5 coefficients; 5 header lines, per
    edge line 4 rows and 4 coefficients; 5 header lines, per face line 4 rows and 4 coefficients."""
    try:
        with open(path) as f:
            lines = f.read().split("\n")
    except OSError as e:
        raise native.PlasticityError("Unable to open Periodic BCfilename %s" % path) from e
    out, pos = {}, 0
    for name, head, count, width in (("vertex", 4, n_vertex, 5), ("edge", 5, n_edge, 4), ("face", 5, n_face, 4)):
        pos += head
        rows, coefs = [], []
        for _ in range(count):
            rows.append([int(v) for v in lines[pos].split()[:width]])
            coefs.append([float(v) for v in lines[pos + 1].split()[:width]])
            pos += 2
        out[name] = (rows, coefs)
    return out


def make_periodic(lines, known, total, tabular=None):
    """pa_periodic from the constraint file's lines, the vertex rows 1 and 2 of the prm and, for
    tabular periodic BCs, (periodic Tabular time, time table, coefficient table)."""
    P = native.PaPeriodic()
    for name in ("vertex", "edge", "face"):
        rows, coefs = lines[name]
        if len(rows) > native.MAX_PBC_LINES:
            raise native.PlasticityError("more than %d %s constraint lines" % (native.MAX_PBC_LINES, name))
        setattr(P, "n_%s_lines" % name, len(rows))
        R, C = getattr(P, "%s_dofs" % name), getattr(P, "%s_coef" % name)
        for i, (r, c) in enumerate(zip(rows, coefs)):
            for k in range(len(r)):
                R[i][k], C[i][k] = r[k], c[k]
    for i in range(24):
        P.vertex_known[i], P.vertex_total[i] = int(known[i]), float(total[i])
    if tabular is not None:
        t, times, coefs = tabular
        if len(times) > native.MAX_TABLE or len(coefs) < len(times):
            raise native.PlasticityError("tabular periodic table: at most %d times, one coefficient each" % native.MAX_TABLE)
        P.tabular, P.tabular_time, P.n_table = 1, float(t), len(times)
        for i, v in enumerate(times):
            P.table_time[i], P.table_coef[i] = v, int(coefs[i])
    return P


def read_table(path, n):
    rows = []
    with open(path) as f:
        for line in f:
            v = line.split()
            if v:
                rows.append([float(x) for x in v[:n]])
            if len(rows) == n:
                break
    if len(rows) < n:
        raise native.PlasticityError("%s: expected %d rows" % (path, n))
    return np.array(rows)


class crystalOrientationsIO:
    """Voxel grain map and id -> Rodrigues table (crystalOrientationsIO.cc)."""

    def __init__(self):
        self.vox = None
        self.numPts = None
        self.eulerAngles = None
        self.numberOfGrains = 0
        self.phase = None

    def loadOrientations(self, voxelFile, headerLines, orientationFile, numPts, span):
        with open(voxelFile) as f:
            lines = f.read().split("\n")[headerLines:]
        nv = [int(x) for x in numPts]
        vox = np.zeros(nv, dtype=np.uint32)
        li = 0
        for x in range(nv[0]):
            for y in range(nv[1]):
                vals = lines[li].split()
                li += 1
                vox[x, y, :] = [int(v) for v in vals[: nv[2]]]
        self.vox, self.numPts = vox, nv

    def loadOrientationVector(self, eulerFile, enableMultiphase=False, numVoxelData=0):
        """id -> Rodrigues vector; with several phases the fifth column is the grain's phase
        (crystalOrientationsIO.cc:11-62), kept in .phase (1-based; 1 for rows without an id)."""
        table, phase = {}, {}
        last = 0
        with open(eulerFile) as f:
            f.readline()
            for line in f:
                v = line.split()
                if len(v) >= 4:
                    table[int(v[0])] = [float(x) for x in v[1:4]]
                    if enableMultiphase:
                        if len(v) < 5:
                            raise native.PlasticityError("several phases: orientation row %s has no phase column" % v[0])
                        phase[int(v[0])] = int(float(v[4]))
                    last = int(v[0])
        self.numberOfGrains = last  # the last id read (crystalOrientationsIO.cc:61)
        rod = np.zeros((max(table) + 1, 3))
        ph = np.ones(max(table) + 1, dtype=np.int32)
        for k, v in table.items():
            rod[k] = v
            if enableMultiphase:
                ph[k] = phase[k]
        self.eulerAngles = rod
        self.phase = ph if enableMultiphase else None


# ------------------------------------------------------------------ context
class Context:
    """One device context (C-ABI handle)."""

    def __init__(self, mat, pb: PaProblem):
        """mat: one PaMaterial, or a list of 2..4 (one per phase: pa_create_phases; then
        set_grain_phases before set_grains)."""
        self._h = ctypes.c_void_p()
        L = native.lib()
        if isinstance(mat, (list, tuple)):
            arr = (PaMaterial * len(mat))(*mat)
            check(L.pa_create_phases(arr, ctypes.c_int(len(mat)), ctypes.byref(pb), ctypes.byref(self._h)))
            self.phases = list(mat)
            mat = max(mat, key=lambda m: m.n_slip)
        else:
            check(L.pa_create(ctypes.byref(mat), ctypes.byref(pb), ctypes.byref(self._h)))
            self.phases = [mat]
        self.mat, self.pb = mat, pb
        ne, nq, nd, nz = (ctypes.c_longlong() for _ in range(4))
        check(L.pa_sizes(self._h, ctypes.byref(ne), ctypes.byref(nq), ctypes.byref(nd), ctypes.byref(nz)))
        self.nelem, self.nqp, self.ndof, self.nnz = ne.value, nq.value, nd.value, nz.value
        self.n_slip = mat.n_slip
        self.n_twin = mat.n_twin if mat.twinning else 0
        self.n_sys = self.n_slip + self.n_twin  # systems of the return map (slip, then twin)

    def close(self):
        if self._h:
            native.lib().pa_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def element_grains(self, vox):
        vox = np.ascontiguousarray(vox, dtype=np.uint32)
        nv = np.array(vox.shape, dtype=np.int32)
        out = np.zeros(self.nelem, dtype=np.uint32)
        check(native.lib().pa_element_grains(self._h, ptr(vox), ptr(nv), ptr(out)))
        return out

    def set_grain_phases(self, grain_phase):
        """Phase (1-based) of each grain row, before set_grains (pa_set_grain_phases)."""
        gp = np.ascontiguousarray(grain_phase, dtype=np.int32)
        check(native.lib().pa_set_grain_phases(self._h, ptr(gp), ctypes.c_int(len(gp))))

    def set_grains(self, elem_grain, grain_rod):
        eg = np.ascontiguousarray(elem_grain, dtype=np.uint32)
        rod = f64(grain_rod)
        check(native.lib().pa_set_grains(self._h, ptr(eg), ptr(rod), ctypes.c_int(len(rod))))

    def set_solver(self, method=native.PA_SOLVER_BICG, precision=native.PA_PREC_FP64):
        check(native.lib().pa_set_solver(self._h, ctypes.c_int(method), ctypes.c_int(precision)))

    def set_periodic(self, per: native.PaPeriodic):
        """Periodic boundary conditions in place of the face rules (pa_set_periodic)."""
        check(native.lib().pa_set_periodic(self._h, ctypes.byref(per)))

    def last_residual(self):
        """(true ||B - K X|| / ||B||, ||D^-1(B - K X)|| / ||B||, the solver's own value) of the last solve."""
        out = np.zeros(3)
        check(native.lib().pa_last_residual(self._h, ptr(out)))
        return out

    def set_qp_form(self, form):
        """Layout of the Newton step's qp passes: native.PA_QP_LANE (the default and the only one
        accepted; native.PA_QP_COOP is refused while its GPU memory fault is unexplained)."""
        check(native.lib().pa_set_qp_form(self._h, ctypes.c_int(form)))

    def assemble(self, inc, it):
        rn = ctypes.c_double()
        check(native.lib().pa_assemble(self._h, ctypes.c_int(inc), ctypes.c_int(it), ctypes.byref(rn)))
        return rn.value

    def linear_solve(self, inc, it):
        n = ctypes.c_int()
        check(native.lib().pa_linear_solve(self._h, ctypes.c_int(inc), ctypes.c_int(it), ctypes.byref(n)))
        return n.value

    def newton_iteration(self, inc, it):
        rn, n = ctypes.c_double(), ctypes.c_int()
        check(native.lib().pa_newton_iteration(self._h, ctypes.c_int(inc), ctypes.c_int(it), ctypes.byref(rn), ctypes.byref(n)))
        return rn.value, n.value

    def commit_increment(self):
        row = np.zeros(15)
        check(native.lib().pa_commit_increment(self._h, ptr(row)))
        return row

    def run_increment(self, inc):
        row = np.zeros(15)
        its = np.zeros(self.pb.max_newton, dtype=np.int32)
        res = np.zeros(self.pb.max_newton)
        check(native.lib().pa_run_increment(self._h, ctypes.c_int(inc), ptr(row), ptr(its), ptr(res)))
        return row, its, res

    # name: (pa_get_state id, components per qp; "sys": slip + twin systems, "slip": slip systems,
    # "twin": twin systems, None: per node)
    STATE = {"T": (0, 9), "Fp": (1, 9), "Fe": (2, 9), "s": (3, "sys"), "rotnew": (4, 3), "slip": (5, "slip"),
             "u": (6, None), "R": (7, None), "Fp_iter": (8, 9), "Fe_iter": (9, 9), "s_iter": (10, "sys"),
             "W": (11, "sys"), "W_iter": (12, "sys"), "twin_fraction": (13, "twin"), "twinned": (14, 1), "rot": (15, 3)}

    def get_state(self, name):
        which, comps = self.STATE[name]
        if comps is None:
            out = np.zeros(self.ndof)
        else:
            comps = {"sys": self.n_sys, "slip": self.n_slip, "twin": self.n_twin}.get(comps, comps)
            out = np.zeros((self.nqp, comps))
        check(native.lib().pa_get_state(self._h, ctypes.c_int(which), ptr(out), ctypes.c_longlong(out.size)))
        return out

    def project_fields(self):
        """The L2 projections onto the nodes of von Mises stress, equivalent strain and the twinned
        flag of the last committed increment (pa_project_fields): array [3, nnode]."""
        nn = self.ndof // 3
        out = np.zeros((3, nn))
        its = np.zeros(3, dtype=np.int32)
        check(native.lib().pa_project_fields(self._h, ptr(out), ptr(its)))
        self.last_projection_iters = its
        return out

    def grain_averages(self, n_grains):
        """Grain-averaged output (updateAfterIncrement.cc:611-707): rows [volume, Txx, Tyy, Tzz,
        Txy, Txz, Tyz] for grain ids 1..n_grains, from the Cauchy stress of the last commit."""
        out = np.zeros((n_grains, 7))
        check(native.lib().pa_grain_averages(self._h, ctypes.c_int(n_grains), ptr(out)))
        return out

    def get_csr(self, values=False):
        rp = np.zeros(self.ndof + 1, dtype=np.int64)
        ci = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz) if values else None
        check(native.lib().pa_get_csr(self._h, ptr(rp), ptr(ci), ptr(val)))
        return (rp, ci, val) if values else (rp, ci)

    def last_timings(self):
        ms = np.zeros(3)
        check(native.lib().pa_last_timings(self._h, ptr(ms)))
        return ms

    def last_qp_pass_ms(self):
        """HIP-event duration of the first qp pass (return map + tangent at every qp) of the latest
        assembly, in ms."""
        ms = ctypes.c_double()
        check(native.lib().pa_last_qp_pass_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def last_counts(self):
        """Quadrature points that went through the pseudo-inverse pass in the last assembly and
        the last post-increment pass, and the reruns (since creation) that met a singular system
        other than the one solved for them."""
        c = np.zeros(3, np.int64)
        check(native.lib().pa_last_counts(self._h, ptr(c)))
        return c

    def qp_update(self, F, Fp, s, grain, flag=1, W=None):
        """calculatePlasticity over a batch; W: backstress W_kh_conv (kinematic hardening only).  The
        per-system arrays have n_sys = n_slip (+ n_twin with twinning) entries per qp."""
        F = f64(F).reshape(-1, 9)
        n = len(F)
        ns = self.n_sys
        Fp = f64(Fp).reshape(n, 9)
        s = f64(s).reshape(n, ns)
        grain = np.ascontiguousarray(grain, dtype=np.uint32).reshape(n)
        out = {k: np.zeros((n, 9)) for k in ("P", "T", "Fe", "Fp")}
        out["dPdF"] = np.zeros((n, 81)) if flag else None
        out["s"] = np.zeros((n, ns))
        out["slip_inc"] = np.zeros((n, ns))
        out["n_active"] = np.zeros(n, dtype=np.int32)
        out["W"] = np.zeros((n, ns))
        Win = None if W is None else f64(W).reshape(n, ns)
        check(native.lib().pa_qp_update(self._h, ctypes.c_longlong(n), ptr(F), ptr(Fp), ptr(s), ptr(grain), ctypes.c_int(flag),
                                        ptr(out["P"]), ptr(out["T"]), ptr(out["dPdF"]), ptr(out["Fe"]), ptr(out["Fp"]),
                                        ptr(out["s"]), ptr(out["slip_inc"]), ptr(out["n_active"]), ptr(Win), ptr(out["W"])))
        return out


class Batch:
    """Many independent RVEs (each the reference's boundary-value problem on its own box mesh)
    advanced together through the C ABI's batch entry points: one persistent qp pass over all of
    them per assembly, per-RVE Jacobi-BiCG (or PCG) in shared launches, Newton iterations replayed
    from HIP graphs.  mats: list of PaMaterial (at most 4); rve_mat[r]: material index of RVE r;
    rve_nel[r]: (nx, ny, nz) elements of RVE r; pb: PaProblem (span, boundary rules, increments,
    Newton count, Krylov tolerance) shared by all RVEs."""

    STATE = {"T": (0, 9), "Fp": (1, 9), "Fe": (2, 9), "s": (3, "sys"), "rotnew": (4, 3), "slip": (5, "sys"), "u": (6, None),
             "R": (7, None)}

    def __init__(self, mats, rve_mat, rve_nel, pb: PaProblem):
        L = native.lib()
        self.mats = list(mats)
        self.rve_mat = np.ascontiguousarray(rve_mat, dtype=np.int32)
        self.rve_nel = np.ascontiguousarray(np.asarray(rve_nel, dtype=np.int32).reshape(-1, 3))
        self.n_rve = len(self.rve_mat)
        arr = (PaMaterial * len(self.mats))(*self.mats)
        self._h = ctypes.c_void_p()
        check(L.pa_batch_create(arr, ctypes.c_int(len(self.mats)), ctypes.byref(pb), ctypes.c_int(self.n_rve), ptr(self.rve_mat),
                                ptr(self.rve_nel), ctypes.byref(self._h)))
        self.pb = pb
        sz = np.zeros(4, dtype=np.int64)
        check(L.pa_batch_sizes(self._h, ptr(sz)))
        self.nqp, self.nelem, self.nnode, self.ndof = (int(v) for v in sz)

    def close(self):
        if self._h:
            native.lib().pa_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grains(self, rve, elem_grain, grain_rod):
        eg = np.ascontiguousarray(elem_grain, dtype=np.uint32)
        rod = f64(grain_rod)
        check(native.lib().pa_batch_set_grains(self._h, ctypes.c_int(rve), ptr(eg), ptr(rod), ctypes.c_int(len(rod))))

    def set_solver(self, method=native.PA_SOLVER_BICG, precision=native.PA_PREC_FP64):
        check(native.lib().pa_batch_set_solver(self._h, ctypes.c_int(method), ctypes.c_int(precision)))

    def set_graphs(self, on):
        check(native.lib().pa_batch_set_graphs(self._h, ctypes.c_int(1 if on else 0)))

    def newton_iteration(self, inc, it):
        """|R| and CG iterations of every RVE."""
        rn = np.zeros(self.n_rve)
        its = np.zeros(self.n_rve, dtype=np.int32)
        check(native.lib().pa_batch_newton_iteration(self._h, ctypes.c_int(inc), ctypes.c_int(it), ptr(rn), ptr(its)))
        return rn, its

    def commit_increment(self):
        """stressstrain.txt rows of every RVE, shape (n_rve, 15)."""
        rows = np.zeros((self.n_rve, 15))
        check(native.lib().pa_batch_commit_increment(self._h, ptr(rows)))
        return rows

    def run_increment(self, inc):
        res, its = [], []
        for it in range(self.pb.max_newton):
            r, k = self.newton_iteration(inc, it)
            res.append(r)
            its.append(k)
        return self.commit_increment(), np.array(its).T, np.array(res).T

    def tune(self):
        """Autotuner report: dict(grid, chunk, qp_ms per grid candidate, krylov_ms per chunk candidate)."""
        rep = np.zeros(9)
        check(native.lib().pa_batch_tune(self._h, ptr(rep)))
        return {"grid": int(rep[0]), "chunk": int(rep[1]), "qp_ms": dict(zip((256, 512, 768, 1024), rep[2:6].tolist())),
                "krylov_ms": dict(zip((8, 16, 32), rep[6:9].tolist()))}

    def qp_pass_ms(self, flag=1):
        ms = ctypes.c_double()
        check(native.lib().pa_batch_qp_pass_ms(self._h, ctypes.c_int(flag), ctypes.byref(ms)))
        return ms.value

    def last_timings(self):
        ms = np.zeros(4)
        check(native.lib().pa_batch_last_timings(self._h, ptr(ms)))
        return ms

    def get_state(self, rve, name):
        which, comps = self.STATE[name]
        nel = int(np.prod(self.rve_nel[rve]))
        if comps is None:
            out = np.zeros(3 * int(np.prod(self.rve_nel[rve] + 1)))
        else:
            if comps == "sys":
                comps = self.mats[self.rve_mat[rve]].n_slip
            out = np.zeros((8 * nel, comps))
        check(native.lib().pa_batch_get_state(self._h, ctypes.c_int(rve), ctypes.c_int(which), ptr(out), ctypes.c_longlong(out.size)))
        return out


def pinv_solve(A, b, path=0):
    """x = pinv(A) b for a batch of n x n systems (A: (count, n, n), b: (count, n)) on the GPU:
    the pseudo-inverse step of inactiveSlipRemoval.cc:26,99; path 0 = one lane per system,
    1 = the cooperative kernel for queued systems, 2 = the group routine of the cooperative qp
    kernels (a group of lanes per system in both)."""
    A = f64(A)
    count, n = A.shape[0], A.shape[1]
    b = f64(b).reshape(count, n)
    x = np.zeros((count, n))
    check(native.lib().pa_pinv_solve(ctypes.c_int(path), ctypes.c_int(n), ctypes.c_longlong(count), ptr(A), ptr(b), ptr(x)))
    return x


def make_material(m, n, q, C, s0, h0, ss, pw, tol, max_slip_search=1, max_solver_iter=1, C1=None, C2=None, twin=None):
    """pa_material from the prm inputs; C1, C2: kinematic hardening (C_1 / C_2 Slip Kinematic
    Hardening), off when None; twin: dict(m, n, s0, h0, ss, pw, shear, threshold, saturation) of the
    twin systems (Twin Directions / Normals File rows and the Twin parameters), with q then the
    latent ratios over slip + twin systems; None = twinning off."""
    mat = PaMaterial()
    ns = len(m)
    ntw = len(twin["m"]) if twin else 0
    if ns + ntw > native.MAX_SLIP:
        raise native.PlasticityError("at most %d slip + twin systems" % native.MAX_SLIP)
    mat.n_slip = ns
    for a in range(ns):
        for j in range(3):
            mat.m[a][j] = m[a][j]
            mat.n[a][j] = n[a][j]
        mat.s0[a], mat.h0[a], mat.ss[a], mat.pw[a] = s0[a], h0[a], ss[a], pw[a]
    for k in range(ntw):
        a = ns + k
        for j in range(3):
            mat.m[a][j] = twin["m"][k][j]
            mat.n[a][j] = twin["n"][k][j]
        mat.s0[a], mat.h0[a], mat.ss[a], mat.pw[a] = twin["s0"][k], twin["h0"][k], twin["ss"][k], twin["pw"][k]
    for a in range(ns + ntw):
        for b in range(ns + ntw):
            mat.q[a][b] = q[a][b]
    if twin:
        mat.twinning, mat.n_twin = 1, ntw
        mat.twin_shear, mat.twin_threshold, mat.twin_saturation = twin["shear"], twin["threshold"], twin["saturation"]
    for i in range(6):
        for j in range(6):
            mat.C[i][j] = C[i][j]
    mat.stress_tol = tol
    mat.max_slip_search = max_slip_search
    mat.max_solver_iter = max_solver_iter
    if C1 is not None:
        mat.kinematic_hardening = 1
        for a in range(ns):
            mat.C1[a], mat.C2[a] = C1[a], C2[a]
    return mat


def make_problem(nel, span, total_increments, delT, bcs, max_newton=4, lin_tol=1e-10, lin_maxit=50000, cyclic=None, tabular=None,
                 vel_grad=None):
    """pa_problem; bcs: (face, dof, total displacement) rules; cyclic: (face, dof, quarter cycle time);
    tabular: (times, [(face, dof, displacements at the times)]); vel_grad: 3x3 velocity gradient L,
    which moves every boundary dof no face rule constrains by (F_(n+1) - F_n) X."""
    pb = PaProblem()
    pb.nx, pb.ny, pb.nz = nel
    for d in range(3):
        pb.span[d] = span[d]
    pb.total_increments, pb.delT = total_increments, delT
    pb.max_newton, pb.lin_tol, pb.lin_maxit = max_newton, lin_tol, lin_maxit
    for f, d, u in bcs:
        pb.face_dof_constrained[f - 1][d - 1] = 1
        pb.face_total_u[f - 1][d - 1] = u
    if cyclic:
        pb.cyclic = 1
        pb.cyclic_face, pb.cyclic_dof, pb.quarter_cycle_time = cyclic
    if tabular:
        times, rules = tabular
        if not 2 <= len(times) <= native.MAX_TABLE:
            raise native.PlasticityError("tabular boundary conditions need 2..%d time points" % native.MAX_TABLE)
        pb.tabular, pb.n_table = 1, len(times)
        for i, t in enumerate(times):
            pb.table_time[i] = t
        for f, d, us in rules:
            pb.face_dof_constrained[f - 1][d - 1] = 1
            for i, u in enumerate(us[:len(times)]):
                pb.table_u[f - 1][d - 1][i] = u
    if vel_grad is not None:
        pb.velgrad = 1
        for a in range(3):
            for b in range(3):
                pb.vel_grad[a][b] = vel_grad[a][b]
    return pb


class crystalPlasticity:
    """crystalPlasticity<3>(userInputs) (include/crystalPlasticity.h:13-264) over the device context."""

    def __init__(self, userInputs: userInputParameters):
        self.userInputs = ui = userInputs
        ns = ui.numSlipSystems1
        m = read_vectors(ui.file(ui.slipDirectionsFile1), ns)
        n = read_vectors(ui.file(ui.slipNormalsFile1), ns)
        twin = None
        if ui.enableTwinning1:  # init.cc:125-175: twin systems after the slip systems
            ntw = ui.numTwinSystems1
            twin = dict(m=read_vectors(ui.file(ui.twinDirectionsFile1), ntw), n=read_vectors(ui.file(ui.twinNormalsFile1), ntw),
                        s0=ui.initialSlipResistanceTwin1, h0=ui.initialHardeningModulusTwin1, ss=ui.saturationStressTwin1,
                        pw=ui.powerLawExponentTwin1, shear=ui.twinShear, threshold=ui.twinThresholdFraction,
                        saturation=ui.twinSaturationFactor)
        q = read_table(ui.file(ui.latentHardeningRatioFileName1), ns + (len(twin["m"]) if twin else 0))
        kin = ui.enableKinematicHardening1
        self.material = make_material(m, n, q, ui.elasticStiffness1, ui.initialSlipResistance1, ui.initialHardeningModulus1,
                                      ui.saturationStress1, ui.powerLawExponent1, ui.modelStressTolerance,
                                      ui.modelMaxSlipSearchIterations, ui.modelMaxSolverIterations,
                                      ui.C_1_slip1 if kin else None, ui.C_2_slip1 if kin else None, twin)
        self.materials = [self.material]
        for ph in ui.phaseKeys:  # phases 2.. (init.cc:200-960)
            ns2 = ph["numSlipSystems"]
            tw2 = None
            if ph["enableTwinning"]:
                ntw2 = ph["numTwinSystems"]
                tw2 = dict(m=read_vectors(ui.file(ph["twinDirectionsFile"]), ntw2), n=read_vectors(ui.file(ph["twinNormalsFile"]), ntw2),
                           s0=ph["initialSlipResistanceTwin"], h0=ph["initialHardeningModulusTwin"], ss=ph["saturationStressTwin"],
                           pw=ph["powerLawExponentTwin"], shear=ph["twinShear"], threshold=ph["twinThresholdFraction"],
                           saturation=ph["twinSaturationFactor"])
            q2 = read_table(ui.file(ph["latentHardeningRatioFileName"]), ns2 + (len(tw2["m"]) if tw2 else 0))
            k2 = ph["enableKinematicHardening"]
            self.materials.append(make_material(read_vectors(ui.file(ph["slipDirectionsFile"]), ns2),
                                                read_vectors(ui.file(ph["slipNormalsFile"]), ns2), q2, ph["elasticStiffness"],
                                                ph["initialSlipResistance"], ph["initialHardeningModulus"], ph["saturationStress"],
                                                ph["powerLawExponent"], ui.modelStressTolerance, ui.modelMaxSlipSearchIterations,
                                                ui.modelMaxSolverIterations, ph["C_1"] if k2 else None, ph["C_2"] if k2 else None, tw2))
        nel = [s * 2 ** ui.meshRefineFactor for s in ui.subdivisions]
        self.totalIncrements = int(round(ui.totalTime / ui.delT))  # ellipticBVP.cc:40
        bcs = []
        if ui.enableSimpleBCs or ui.enableCyclicLoading:  # init.cc:327-348
            with open(ui.file(ui.BCfilename)) as f:
                lines = f.read().split("\n")[ui.BCheaderLines: ui.BCheaderLines + ui.NumberofBCs]
            for line in lines:
                v = line.split()
                bcs.append((int(v[0]), int(v[1]), float(v[2])))
        cyc = (ui.cyclicLoadingFace, ui.cyclicLoadingDOF, ui.quarterCycleTime) if ui.enableCyclicLoading else None
        tab = None
        if ui.enableTabularBCs:  # init.cc:389-415: no header lines; face dof u(t_0) .. u(t_n-1) per line
            nt, rules = ui.tabularBCs_InputStepNumber, []
            with open(ui.file(ui.Tabular_BCfilename)) as f:
                for line in f:
                    v = line.split()
                    if len(v) >= 2 + nt and len(rules) < ui.tabularNumberofBCs:
                        rules.append((int(v[0]), int(v[1]), [float(x) for x in v[2:2 + nt]]))
            tab = (ui.tabularTimeInput[:nt], rules)
        self.problem = make_problem(nel, ui.span, self.totalIncrements, ui.delT, bcs, ui.maxNonLinearIterations,
                                    ui.relLinearSolverTolerance, ui.maxLinearSolverIterations, cyc, tab,
                                    ui.targetVelGrad if ui.useVelocityGrad else None)
        self.periodic = None
        if ui.enablePeriodicBCs:  # ellipticBVP init.cc:66-131
            lines = read_periodic_constraints(ui.file(ui.Periodic_BCfilename), ui.numberVerticesConstraint,
                                              ui.numberEdgesConstraint, ui.numberFacesConstraint)
            tabp = ((ui.periodicTabularTime, ui.tabularPeriodicTimeInput, ui.tabularPeriodicCoef)
                    if ui.enableTabularPeriodicBCs else None)
            self.periodic = make_periodic(lines, ui.periodicBCsInput, ui.periodicBCsInput2, tabp)
        self.orientations = crystalOrientationsIO()
        self.ctx = None

    def _init(self):
        multi = len(self.materials) > 1
        self.ctx = Context(self.materials if multi else self.material, self.problem)
        if self.periodic is not None:
            self.ctx.set_periodic(self.periodic)
        if multi:
            if self.orientations.phase is None:
                raise native.PlasticityError("several phases: load the orientations with enableMultiphase=True")
            self.ctx.set_grain_phases(self.orientations.phase)
        eg = self.ctx.element_grains(self.orientations.vox)
        self.ctx.set_grains(eg, self.orientations.eulerAngles)
        self.cellOrientationMap = eg

    def run(self, n_increments=None, write=True, log=None):
        """solve() (src/ellipticBVP/solve.cc:69-135); returns the stressstrain rows."""
        if self.ctx is None:
            self._init()
        n = self.totalIncrements if n_increments is None else n_increments
        rows = np.zeros((n, 15))
        self.lin_iters = np.zeros((n, self.problem.max_newton), dtype=np.int32)
        ui = self.userInputs
        outdir = ui.file(ui.outputDirectory)
        for inc in range(n):
            rows[inc], self.lin_iters[inc], res = self.ctx.run_increment(inc)
            if log:
                log("increment %d: residuals %s, linear iterations %s" % (inc, res, self.lin_iters[inc]))
            if write and ui.writeQuadratureOutput and self._output_due(inc, ui.skipQuadratureOutputSteps):
                os.makedirs(outdir, exist_ok=True)
                write_quadrature_output(os.path.join(outdir, "QuadratureOutputs%d.csv" % inc), self.ctx, self.cellOrientationMap,
                                        (self.problem.nx, self.problem.ny, self.problem.nz), ui.span,
                                        dealii_cell_order(ui.subdivisions, ui.meshRefineFactor))
            if write and ui.writeOutput and self._output_due(inc, ui.skipOutputSteps):  # solve.cc:122-124
                os.makedirs(outdir, exist_ok=True)
                write_vtu_output(outdir, inc, self.totalIncrements, self.ctx, self.cellOrientationMap,
                                 (self.problem.nx, self.problem.ny, self.problem.nz), ui.span,
                                 [name for name, on in (("Eqv_stress", ui.output_Eqv_stress), ("Eqv_strain", ui.output_Eqv_strain),
                                                        ("Twin", ui.output_Twin)) if on])
            if write and ui.writeGrainAveragedOutput and self._output_due(inc, ui.skipGrainAveragedOutputSteps):
                if ui.numberOfGrainAverageDataOutput != 7:  # updateAfterIncrement.cc:657-660
                    raise native.PlasticityError("The size of numberOfGrainAverageDataOutput defined in prm.prm is not correct")
                os.makedirs(outdir, exist_ok=True)
                write_grain_average(os.path.join(outdir, "GrainAverage%d.csv" % inc),
                                    self.ctx.grain_averages(self.orientations.numberOfGrains))
        if write:
            os.makedirs(outdir, exist_ok=True)
            write_stressstrain(os.path.join(outdir, "stressstrain.txt"), rows)
        return rows

    def _output_due(self, inc, skip):
        """updateAfterIncrement.cc:323-339, 344, 612: every skip-th increment, or the increments of
        the tabular output times (int(t / delT + 0.1))."""
        ui = self.userInputs
        if ui.tabularOutput:
            return sum(int(t / ui.delT + 0.1) == inc + 1 for t in ui.tabularTimeOutput) == 1
        return (inc + 1) % skip == 0


def dealii_cell_order(subdivisions, refine):
    """Element index e = ex + nx (ey + ny ez) of each active cell in deal.II's iteration order for
    subdivided_hyper_rectangle + refine_global (mesh.cc:32,101): coarse cells lexicographic (x
    fastest); each refinement replaces a cell by its 8 children in place, child c at offset
    (c & 1, c >> 1 & 1, c >> 2 & 1).  Returns an array of length nelem."""
    sx, sy, sz = subdivisions
    n = 2 ** refine
    nx, ny = sx * n, sy * n
    fine = np.zeros((n ** 3, 3), dtype=np.int64)  # offsets of the fine cells of one coarse cell, in order
    for k in range(n ** 3):
        x = y = z = 0
        for lev in range(refine):  # child digits from the coarsest level down
            c = (k >> (3 * (refine - 1 - lev))) & 7
            x, y, z = 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * z + ((c >> 2) & 1)
        fine[k] = (x, y, z)
    order = []
    for cz in range(sz):
        for cy in range(sy):
            for cx in range(sx):
                ex = cx * n + fine[:, 0]
                ey = cy * n + fine[:, 1]
                ez = cz * n + fine[:, 2]
                order.append(ex + nx * (ey + ny * ez))
    return np.concatenate(order)


def write_quadrature_output(path, ctx, elem_grain, nel, span, cell_order):
    """QuadratureOutputs<increment>.csv (updateAfterIncrement.cc:343-608, writeQuadratureOutput.cc:4-49):
    one row per quadrature point, cells in deal.II order, values as "%8.5e ,".  Columns: grain id,
    phase (1), JxW, twin flag (twin_ouput: 1 once reoriented by twinning), x, y, z, rotnew (3),
    Fe_conv, Fp_conv and Cauchy stress (each 11,22,33,12,13,21,23,31,32), slip fractions 0..83,
    twin fractions 0..5.  The reference reads the slip and twin fraction vectors up to index 83 and
    5 whatever their length (n_slip, and n_twin or 1), past their end; those entries are written
    as 0 here."""
    nx, ny, nz = nel
    h = np.asarray(span, float) / np.array([nx, ny, nz])
    jxw = h.prod() / 8.0
    nq = ctx.nqp
    rot = ctx.get_state("rotnew").reshape(nq, 3)
    Fe = ctx.get_state("Fe").reshape(nq, 9)
    Fp = ctx.get_state("Fp").reshape(nq, 9)
    T = ctx.get_state("T").reshape(nq, 9)
    slip = ctx.get_state("slip").reshape(nq, -1)
    twinned = ctx.get_state("twinned").reshape(nq) if ctx.n_twin else np.zeros(nq)
    twf = ctx.get_state("twin_fraction").reshape(nq, -1) if ctx.n_twin else np.zeros((nq, 0))
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    mix = [0, 4, 8, 1, 2, 3, 5, 6, 7]
    e = np.asarray(cell_order, np.int64)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    cols = []
    for q in range(8):
        Q = 8 * e + q
        X = np.stack([(ex + g[q & 1]) * h[0], (ey + g[(q >> 1) & 1]) * h[1], (ez + g[(q >> 2) & 1]) * h[2]], axis=1)
        sl = np.zeros((len(e), 84))
        sl[:, :min(84, slip.shape[1])] = slip[Q, :84]
        tw = np.zeros((len(e), 6))
        tw[:, :min(6, twf.shape[1])] = twf[Q, :6]
        block = np.hstack([np.asarray(elem_grain, float)[e][:, None], np.ones((len(e), 1)), np.full((len(e), 1), jxw),
                           twinned[Q][:, None], X, rot[Q], Fe[Q][:, mix], Fp[Q][:, mix], T[Q][:, mix], sl, tw])
        cols.append(block)
    rows = np.stack(cols, axis=1).reshape(len(e) * 8, -1)  # cell-major, qp-minor
    with open(path, "w") as f:
        for r in rows:
            f.write("".join("%8.5e ," % v for v in r) + "\n")


# VTK cell type of a hexahedron, and its vertex order as lexicographic Q1 vertex ids (x fastest)
_VTK_HEX = 12
_VTK_HEX_ORDER = (0, 1, 3, 2, 4, 5, 7, 6)


def _vtu_piece(path, nel, span, point_data, cell_data):
    """One VTK XML UnstructuredGrid file of the box mesh (nodes x fastest, undeformed, as DataOut writes
    the mesh) with ASCII point and cell arrays: point_data / cell_data are lists of (name, array
    [n] or [n, 3])."""
    nx, ny, nz = nel
    xs = [np.linspace(0.0, span[d], n + 1) for d, n in enumerate(nel)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.transpose(2, 1, 0).ravel(), ey.transpose(2, 1, 0).ravel(), ez.transpose(2, 1, 0).ravel()
    conn = np.stack([(ex + (v & 1)) + (nx + 1) * ((ey + ((v >> 1) & 1)) + (ny + 1) * (ez + ((v >> 2) & 1)))
                     for v in _VTK_HEX_ORDER], axis=1)
    nel_t = nx * ny * nz

    def arr(name, a):
        a = np.asarray(a)
        nc = 1 if a.ndim == 1 else a.shape[1]
        body = " ".join("%.9g" % v for v in a.ravel())
        return '<DataArray type="Float64" Name="%s" NumberOfComponents="%d" format="ascii">%s</DataArray>' % (name, nc, body)
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n'
                "<UnstructuredGrid>\n")
        f.write('<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(pts), nel_t))
        f.write("<Points>%s</Points>\n" % arr("Points", pts))
        f.write('<Cells><DataArray type="Int64" Name="connectivity" format="ascii">%s</DataArray>'
                % " ".join(str(v) for v in conn.ravel()))
        f.write('<DataArray type="Int64" Name="offsets" format="ascii">%s</DataArray>' % " ".join(str(8 * (e + 1)) for e in range(nel_t)))
        f.write('<DataArray type="UInt8" Name="types" format="ascii">%s</DataArray></Cells>\n' % " ".join([str(_VTK_HEX)] * nel_t))
        f.write("<PointData>%s</PointData>\n" % "".join(arr(n, a) for n, a in point_data))
        f.write("<CellData>%s</CellData>\n" % "".join(arr(n, a) for n, a in cell_data))
        f.write("</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def _pvtu(path, piece, point_data, cell_data):
    """The parallel record naming the one piece (DataOut::write_pvtu_record)."""
    def decl(items):
        return "".join('<PDataArray type="Float64" Name="%s" NumberOfComponents="%d"/>' % (n, 1 if np.asarray(a).ndim == 1 else 3)
                       for n, a in items)
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="LittleEndian">\n'
                '<PUnstructuredGrid GhostLevel="0">\n<PPointData>%s</PPointData>\n<PCellData>%s</PCellData>\n'
                '<PPoints><PDataArray type="Float64" NumberOfComponents="3"/></PPoints>\n<Piece Source="%s"/>\n'
                "</PUnstructuredGrid>\n</VTKFile>\n" % (decl(point_data), decl(cell_data), piece))


def write_vtu_output(outdir, inc, total_increments, ctx, elem_grain, nel, span, projected):
    """ellipticBVP::output (output.cc:4-270) on one process: solution-<inc>.0000.vtu with the nodal
    displacement "u" and the cell data meshGrain_ID and subdomain, and when fields are selected
    projectedFields-<inc>.0000.vtu with the L2-projected fields (pa_project_fields: Eqv_stress,
    Eqv_strain, Twin), subdomain and meshGrain_ID; each with its .pvtu record.  The increment is
    written with 4 digits below 10000 increments."""
    digits = 4 if total_increments < 10000 else int(np.ceil(np.log10(total_increments))) + 1
    tag = "%0*d" % (digits, inc)
    u = ctx.get_state("u").reshape(-1, 3)
    grain = np.asarray(elem_grain, dtype=np.float64)
    sub = np.zeros(len(grain))
    cells = [("meshGrain_ID", grain), ("subdomain", sub)]
    _vtu_piece(os.path.join(outdir, "solution-%s.0000.vtu" % tag), nel, span, [("u", u)], cells)
    _pvtu(os.path.join(outdir, "solution-%s.pvtu" % tag), "solution-%s.0000.vtu" % tag, [("u", u)], cells)
    if projected:
        fields = ctx.project_fields()
        names = ("Eqv_stress", "Eqv_strain", "Twin")
        pdata = [(n, fields[names.index(n)]) for n in projected]
        pcells = [("subdomain", sub), ("meshGrain_ID", grain)]
        _vtu_piece(os.path.join(outdir, "projectedFields-%s.0000.vtu" % tag), nel, span, pdata, pcells)
        _pvtu(os.path.join(outdir, "projectedFields-%s.pvtu" % tag), "projectedFields-%s.0000.vtu" % tag, pdata, pcells)


def write_grain_average(path, data):
    """GrainAverage<increment>.csv: one row per grain, comma separated, setprecision(9), appended
    (updateAfterIncrement.cc:686-706)."""
    with open(path, "a") as f:
        for row in data:
            f.write(",".join("%.9g" % v for v in row) + "\n")


def write_stressstrain(path, rows):
    """results/stressstrain.txt with the reference header and default 6-significant-digit
    ostream formatting (updateAfterIncrement.cc:735-771)."""
    head = "Exx\tEyy\tEzz\tEyz\tExz\tExy\tTxx\tTyy\tTzz\tTyz\tTxz\tTxy\tTwinRealVF\tTwinMade\tSlipTotal\n"
    with open(path, "w") as f:
        f.write(head)
        for r in rows:
            f.write("\t".join("%.6g" % v for v in r) + "\n")


def main(prm):
    """applications/crystalPlasticity/main.cc:10-68: read prm, load the microstructure, run."""
    ui = userInputParameters(prm)
    problem = crystalPlasticity(ui)
    problem.orientations.loadOrientations(ui.file(ui.grainIDFile), ui.headerLinesGrainIDFile, ui.file(ui.grainOrientationsFile),
                                          ui.numPts, ui.span)
    problem.orientations.loadOrientationVector(ui.file(ui.grainOrientationsFile), ui.enableMultiphase)
    return problem.run()

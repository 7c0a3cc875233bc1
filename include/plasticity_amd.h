/*
 * plasticity_amd: MI355X (gfx950) crystal-plasticity Newton step behind a C ABI.
 *
 * The reference (PRISMS-Plasticity 1.5.0) has no FFI for this path: its boundary is the C++
 * class pair ellipticBVP<dim> (include/ellipticBVP.h:26-226) and crystalPlasticity<dim>
 * (include/crystalPlasticity.h:13-264).  Each entry point below names the reference member it
 * stands in for.  Conventions:
 *   - every function returns 0 on success or a PA_ERR_* code; pa_last_error() describes the
 *     last failure of the calling thread; no C++ exception crosses this boundary;
 *   - host buffers are caller-owned and only read/written during the call; device memory is
 *     owned by the context; one context = one HIP stream on the current device;
 *   - matrices are row-major 3x3 (9 doubles); dP/dF is 81 doubles with
 *     dPdF[(3i+k)*9 + 3j+l] = dP_ik/dF_jl (the reference's dP_dF[i][k][j][l]);
 *   - quadrature points are numbered Q = 8*e + q with elements e = ex + nx*(ey + ny*ez) and
 *     q = qx + 2qy + 4qz of the 2x2x2 Gauss rule; nodes n = ix + (nx+1)*(iy + (ny+1)*iz);
 *     dof = 3n + component.
 */
#ifndef PLASTICITY_AMD_H
#define PLASTICITY_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define PA_MAX_SLIP 48
#define PA_MAX_TABLE 32

enum {
  PA_OK = 0,
  PA_ERR_ARG = 1,         /* invalid argument / size */
  PA_ERR_UNSUPPORTED = 2, /* configuration outside what the kernels implement */
  PA_ERR_HIP = 3,         /* HIP runtime error (no GPU, allocation, launch) */
  PA_ERR_IO = 4,          /* unreadable input file (the reference exits with status 1) */
  PA_ERR_STATE = 5,       /* call out of order (e.g. no grains set) */
  PA_ERR_NONFINITE = 6    /* a NaN/Inf residual, Krylov norm or averaged stress: the reference's PETSc
                           * solve stops on it (KSP_DIVERGED_NANORINF); no correction is applied */
};

enum { PA_SOLVER_BICG = 0, PA_SOLVER_CG = 1, PA_SOLVER_MGCG = 2 };
enum { PA_PREC_FP64 = 0, PA_PREC_MIXED = 1 };
enum { PA_QP_LANE = 0, PA_QP_COOP = 1 };

/* Single-phase material tables, as crystalPlasticity<dim>::init() builds them from the prm
 * inputs (src/materialModels/crystalPlasticity/init.cc:6-196, multiphaseInit.cc:4-67). */
typedef struct {
  int n_slip;                            /* Number of Slip Systems */
  double m[PA_MAX_SLIP][3];              /* Slip Directions File rows (normalised by the library) */
  double n[PA_MAX_SLIP][3];              /* Slip Normals File rows (normalised by the library) */
  double q[PA_MAX_SLIP][PA_MAX_SLIP];    /* Latent Hardening Ratio file */
  double C[6][6];                        /* Elastic Stiffness rows 1-6 (MPa, Voigt 11,22,33,23,13,12) */
  double s0[PA_MAX_SLIP];                /* Initial Slip Resistance */
  double h0[PA_MAX_SLIP];                /* Initial Hardening Modulus */
  double ss[PA_MAX_SLIP];                /* Saturation Stress */
  double pw[PA_MAX_SLIP];                /* Power Law Exponent */
  double stress_tol;                     /* Stress Tolerance */
  int max_slip_search;                   /* Max Slip Search Iterations (must be 1) */
  int max_solver_iter;                   /* Max Solver Iterations */
  int kinematic_hardening;               /* Enable Kinematic Hardening */
  int twinning;                          /* Twinning enabled */
  double C1[PA_MAX_SLIP];                /* C_1 Slip Kinematic Hardening (used when kinematic_hardening) */
  double C2[PA_MAX_SLIP];                /* C_2 Slip Kinematic Hardening */
  /* Twinning (calculatePlasticity.cc:421-452, 962-1019; used when twinning): the n_twin twin
   * systems follow the slip systems in m, n (Twin Directions / Twin Normals File rows), s0, h0,
   * ss, pw (the Twin parameters: Initial Slip Resistance / Initial Hardening Modulus / Saturation
   * Stress / Power Law Exponent Twin) and q (latent ratios over all n_slip + n_twin systems);
   * n_slip + n_twin <= PA_MAX_SLIP. */
  int n_twin;                            /* Number of Twin Systems */
  double twin_shear;                     /* Characteristic Twin Shear */
  double twin_threshold;                 /* Twin Threshold Fraction */
  double twin_saturation;                /* Twin Saturation Factor */
} pa_material;

/* Mesh, loading and solver inputs of ellipticBVP<dim> (userInputParameters.cc:17-235). */
typedef struct {
  int nx, ny, nz;                        /* elements per direction = Subdivisions * 2^Refine factor */
  double span[3];                        /* Domain size X/Y/Z */
  int total_increments;                  /* round(Total time / Time increments), ellipticBVP.cc:40 */
  double delT;                           /* Time increments */
  int max_newton;                        /* Maximum non linear iterations */
  double lin_tol;                        /* Relative linear solver tolerance */
  int lin_maxit;                         /* Maximum linear solver iterations */
  int face_dof_constrained[6][3];        /* BC file rules (face 1..6 = x-,x+,y-,y+,z-,z+; dof 1..3) */
  double face_total_u[6][3];             /* FinalDisplacement of each rule */
  int cyclic;                            /* Enable cyclic loading */
  int cyclic_face, cyclic_dof;           /* Cyclic loading face / direction (1-based) */
  double quarter_cycle_time;             /* Quarter cycle time */
  int tabular;                           /* Use Tabular BCs: the table drives the constrained (face,dof) rules */
  int n_table;                           /* Number of time data for Tabular BCs (2..PA_MAX_TABLE) */
  double table_time[PA_MAX_TABLE];       /* Tabular Time Table */
  double table_u[6][3][PA_MAX_TABLE];    /* Tabular BC file: displacement of each rule at the table times */
  int velgrad;                           /* Use velocity gradient BC: every boundary dof follows F */
  double vel_grad[3][3];                 /* Velocity gradient rows 1-3 (L) */
} pa_problem;

/* Periodic boundary conditions (Use Periodic BCs): the constraint file PeriodicBCsConstraints.txt
 * as ellipticBVP::init reads it (src/ellipticBVP/init.cc:66-131) and the prm's vertex rows and
 * tabular table (userInputParameters.cc:126-136).  Rows: vertex 0..23 = V000 V100 V010 V001 V110
 * V101 V011 V111 x 3 components; edge 0..35 = EX000 EX001 EX010 EX011 EY000 EY001 EY100 EY101
 * EZ000 EZ010 EZ100 EZ110 x 3 (EXab: y = a, z = b; EYab: x = a, z = b; EZab: x = a, y = b);
 * face 0..17 = FXN FXP FYN FYP FZN FZP x 3. */
#define PA_MAX_PBC_LINES 64
typedef struct {
  int n_vertex_lines;                        /* Number of Vertices Constraints */
  int vertex_dofs[PA_MAX_PBC_LINES][5];      /* entry count, then the line's vertex rows */
  double vertex_coef[PA_MAX_PBC_LINES][5];   /* (count), then their coefficients */
  int n_edge_lines;                          /* Number of Edges Constraints */
  int edge_dofs[PA_MAX_PBC_LINES][4];        /* constrained edge row, edge row, vertex row, vertex row */
  double edge_coef[PA_MAX_PBC_LINES][4];
  int n_face_lines;                          /* Number of Faces Constraints */
  int face_dofs[PA_MAX_PBC_LINES][4];        /* positive-face row (FXP/FYP/FZP), face row, vertex row, vertex row */
  double face_coef[PA_MAX_PBC_LINES][4];
  int vertex_known[24];                      /* Vertices Periodic BCs row 1 */
  double vertex_total[24];                   /* Vertices Periodic BCs row 2 (over the load path) */
  int tabular;                               /* Use Tabular Periodic BCs */
  double tabular_time;                       /* periodic Tabular time */
  int n_table;                               /* entries of the two tables below (<= PA_MAX_TABLE) */
  double table_time[PA_MAX_TABLE];           /* Tabular Periodic Time Table */
  int table_coef[PA_MAX_TABLE];              /* Tabular Periodic Time Table Coefficient: 1 load, -1 unload, 0 hold */
} pa_periodic;

typedef struct pa_context pa_context;

/* Last error message of the calling thread (never NULL). */
const char *pa_last_error(void);

/* Replaces crystalPlasticity<dim>::crystalPlasticity(userInputParameters&)
 * (src/materialModels/crystalPlasticity/crystalPlasticity.cc:4-76) plus ellipticBVP::mesh()/init()
 * (src/ellipticBVP/mesh.cc:7-124, init.cc:8-558): allocates per-qp state (Fp = I, s = s0),
 * builds the 27-point block sparsity on the device. */
int pa_create(const pa_material *mat, const pa_problem *pb, pa_context **out);
int pa_destroy(pa_context *ctx);

/* Several phases (Enable Multiphase; crystalPlasticity::init's multiphase tables, init.cc:200-960,
 * and multiphaseInit.cc:68-361): mats[k] is phase k + 1's material (1..4 phases, each with its own
 * systems, moduli, hardening, kinematic hardening and twinning; the phases share Max Solver
 * Iterations and Stress Tolerance, which the reference reads once).  Per-qp state is sized for the
 * largest phase.  Then pa_set_grain_phases (each grain row's phase, 1-based: the orientation
 * file's fifth column, crystalOrientationsIO.cc:42-45) before pa_set_grains, which builds each
 * grain's tables (and twin variants) from its phase and starts s at its phase's s0.  As the
 * reference does with Enable Multiphase: after each increment F_r, F_e and F_s are zero
 * (updateAfterIncrement.cc:718-725), so the row's TwinRealVF, TwinMade and SlipTotal read 0 and
 * no qp reorients by twinning (the reorientation needs F_r > 0, calculatePlasticity.cc:969-993);
 * twin fractions still grow and harden in a twinning phase.  A context made here with one phase
 * follows the same rule.  The batch qp entry points (pa_qp_update, pa_qp_update_warm) take one
 * phase. */
int pa_create_phases(const pa_material *mats, int n_phases, const pa_problem *pb, pa_context **out);
int pa_set_grain_phases(pa_context *ctx, const int *grain_phase, int n_grain_rows);

/* Replaces crystalOrientationsIO::getMaterialID over cell centres
 * (src/utilityObjects/crystalOrientationsIO.cc:115-194, loadOrientations.cc:4-33), on the GPU.
 * vox: [nv0][nv1][nv2] grain ids, z fastest (the grain-ID file order). */
int pa_element_grains(pa_context *ctx, const unsigned *vox, const int nv[3], unsigned *elem_grain);

/* Replaces crystalPlasticity::init's orientation load (init.cc:223-245) and
 * loadOrientationVector (crystalOrientationsIO.cc:11-62): grain id of each element and the
 * Rodrigues vector of each grain id (rows of 3, indexed by id). */
int pa_set_grains(pa_context *ctx, const unsigned *elem_grain, const double *grain_rod, int n_grain_rows);

/* Krylov method (BiCG = the reference's PETScWrappers::SolverBiCG, solveLinearSystem.cc:13-16;
 * CG = Jacobi-PCG) and arithmetic form (FP64, or mixed: FP32 element tangents / matrix values,
 * FP64 residual, return map and Krylov recurrences). */
int pa_set_solver(pa_context *ctx, int method, int precision);

/* Periodic boundary conditions in place of the face rules: replaces setPeriodicityConstraintsInit
 * and setPeriodicityConstraints (src/ellipticBVP/periodicBCs.cc:23-864, 1206-1310, with
 * setFace/Edge/NodeConstraints :866-1204) and the condensed assembly and distribute of their
 * affine constraints (assemble.cc:20-35, 78-82; solveLinearSystem.cc:32).  The dofs of the
 * positive faces and of nine edges follow their periodic images plus corner terms, the corners
 * follow the vertex lines and rows; the Newton iteration solves for the unconstrained dofs (the
 * Jacobian applied as C^T K C, Jacobi on its diagonal).  Needs equal element counts nx = ny = nz
 * >= 2 (the reference sizes every edge by Subdivisions Y - 1, periodicBCs.cc:49) and a Krylov
 * method other than PA_SOLVER_MGCG; the face rules of the problem are ignored (as
 * boundaryConditions.cc:455 does).  Call before the first Newton iteration. */
int pa_set_periodic(pa_context *ctx, const pa_periodic *per);

/* The closed constraint lines pa_set_periodic builds for the problem's mesh, on the host (no
 * device call; no counterpart in the reference, whose lines live in an AffineConstraints object):
 * constrained[d] = 1 for a constrained dof, whose line is u_d = sum weight[k] u_{master[k]} over
 * k in [rowptr[d], rowptr[d+1]) plus g[d] at the first Newton iteration of a loading increment.
 * rowptr has 3 (nx+1)(ny+1)(nz+1) + 1 entries; master and weight hold up to cap entries;
 * *n_entries receives the number of entries (call with cap 0 to size the arrays). */
int pa_periodic_lines(const pa_problem *pb, const pa_periodic *per, int *constrained, long long *rowptr, long long *master,
                      double *weight, double *g, long long cap, long long *n_entries);

/* Layout of the Newton step's quadrature-point passes (calculatePlasticity over every qp in
 * getElementalValues.cc:141 and updateAfterIncrement.cc:105).  No counterpart in the reference
 * (one CPU call per qp); the results agree to rounding in both forms.
 *   PA_QP_LANE (default): one lane per qp; numerically singular local systems are deferred to a
 *     cooperative pseudo-inverse pass and replayed.  With 48 slip systems the lane's active-set
 *     matrices live in a context-owned workspace of global memory (80 KB per lane, 2.6 GB).
 *   PA_QP_COOP: a group of lanes per qp (lane = slip system / tangent column, matrices in LDS),
 *     persistent wavefronts on a work queue, singular systems solved in place by the group.
 *     Refused (PA_ERR_UNSUPPORTED): these passes took a GPU memory fault whose cause is not
 *     found; the batch entry point pa_qp_update keeps the cooperative routine. */
int pa_set_qp_form(pa_context *ctx, int form);

/* One Newton iteration of solveNonLinearSystem (src/ellipticBVP/solveNonLinearSystem.cc:14-47):
 * applyDirichletBCs + assemble (boundaryConditions.cc:452-514, assemble.cc:6-104,
 * getElementalValues.cc:5-168) then solveLinearSystem (solveLinearSystem.cc:8-36).
 * res_norm: ||R||_2 of the condensed residual; lin_iters: Krylov iterations. */
int pa_newton_iteration(pa_context *ctx, int increment, int iteration, double *res_norm, int *lin_iters);

/* Split halves of pa_newton_iteration, for timing and parity.  pa_assemble returns
 * PA_ERR_NONFINITE when the residual norm it reports is NaN/Inf; pa_linear_solve returns it when the
 * Krylov solve stops on a NaN/Inf norm (u is then left unchanged) and PA_ERR_STATE when the
 * arithmetic form changed (pa_set_solver) since the last assembly. */
int pa_assemble(pa_context *ctx, int increment, int iteration, double *res_norm);
int pa_linear_solve(pa_context *ctx, int increment, int iteration, int *lin_iters);

/* updateAfterIncrement (src/materialModels/crystalPlasticity/updateAfterIncrement.cc:6-305,
 * 710-771): stress-only qp pass, volume averages, reorient() (reorient.cc:4-100), conv <- iter.
 * row: Exx Eyy Ezz Eyz Exz Exy Txx Tyy Tzz Tyz Txz Txy TwinRealVF TwinMade SlipTotal (TwinRealVF,
 * TwinMade: volume-averaged twin fraction and reoriented volume; 0 without twinning). */
int pa_commit_increment(pa_context *ctx, double row[15]);

/* solve() body for one increment (src/ellipticBVP/solve.cc:69-91): max_newton Newton iterations
 * then pa_commit_increment.  lin_iters / res_norms (optional) receive max_newton entries. */
int pa_run_increment(pa_context *ctx, int increment, double row[15], int *lin_iters, double *res_norms);

/* State readback for parity.  which: 0 T (9/qp), 1 Fp_conv (9), 2 Fe_conv (9), 3 s_conv (n_sys),
 * 4 rotnew (3), 5 slip_conv (n_slip), 6 u (3/node), 7 R (3/node), 8 Fp_iter, 9 Fe_iter, 10 s_iter,
 * 11 W_kh_conv (n_sys, backstress; kinematic hardening only), 12 W_kh_iter; with twinning
 * 13 twinfraction_conv (n_twin), 14 twin_conv (1: reoriented by twinning, else 0), 15 rot_conv (3,
 * the orientation of the constitutive update).  n_sys = n_slip (+ n_twin with twinning).
 * Output is qp-major (or node-major).  n = number of doubles the caller provides. */
int pa_get_state(pa_context *ctx, int which, double *host, long long n);

/* The post-processed fields of the last committed increment, L2-projected onto the nodes
 * (ellipticBVP::projection, projectionFunctions.cc:143-223, with the values of
 * updateAfterIncrement.cc:147-192): fields[0..nnode) von Mises stress sqrt(3/2)|dev T|
 * ("Eqv_stress"), [nnode..2 nnode) equivalent strain sqrt(2/3)|dev E| ("Eqv_strain"),
 * [2 nnode..3 nnode) the twinned flag ("Twin"); each the solution of M x = b (the consistent Q1
 * mass matrix, Jacobi-PCG at the problem's Krylov tolerance, solveLinearSystem.cc:43-65);
 * iters (optional, 3) the CG iterations.  Uses the Jacobian's storage: the next Newton step
 * assembles again. */
int pa_project_fields(pa_context *ctx, double *fields, int *iters);

/* Grain-averaged output of updateAfterIncrement (src/materialModels/crystalPlasticity/
 * updateAfterIncrement.cc:611-707) over the Cauchy stress of the last pa_commit_increment: row r
 * (grain id r + 1 of the element grain map, the reference's cellOrientationMap - 1) holds
 * [volume, Txx, Tyy, Tzz, Txy, Txz, Tyz], the stresses volume-averaged when the volume exceeds
 * 1e-15.  out: n_grains x 7 doubles.  Deterministic (fixed summation order) on the device. */
int pa_grain_averages(pa_context *ctx, int n_grains, double *out);

/* Sparsity of the Jacobian (DoFTools::make_sparsity_pattern over the full 27-node stencil,
 * src/ellipticBVP/init.cc:183-199), built on the GPU: row_ptr (ndof+1), col_idx (nnz); values
 * (optional, nnz) = the last assembled condensed Jacobian.  Sizes: pa_sizes. */
int pa_get_csr(pa_context *ctx, long long *row_ptr, int *col_idx, double *values);
int pa_sizes(pa_context *ctx, long long *nelem, long long *nqp, long long *ndof, long long *nnz);

/* calculatePlasticity(cellID, q, flag) over a batch of quadrature points (calculatePlasticity.cc:13-959),
 * host buffers in/out, with grain[i] selecting the orientation row set by pa_set_grains.
 * F, Fp_conv: 9/qp; s_conv: n_slip/qp.  Outputs P, T, Fe, Fp: 9/qp; dPdF: 81/qp (flag 1, may be
 * NULL for flag 0); s_out, slip_inc (x_beta_old): n_slip/qp; n_active: 1/qp.  With kinematic
 * hardening, W_conv (n_slip/qp, NULL = zero) is the backstress W_kh_conv and W_out (may be NULL)
 * receives W_kh_tau; both are ignored otherwise.  With twinning the per-system arrays (s_conv,
 * s_out, slip_inc, W) have n_slip + n_twin entries per qp (slip systems, then twin systems) and
 * each qp uses its grain's own orientation (no reorientation test: that belongs to the Newton
 * step's history). */
int pa_qp_update(pa_context *ctx, long long nqp, const double *F, const double *Fp_conv, const double *s_conv,
                 const unsigned *grain, int flag, double *P, double *T, double *dPdF, double *Fe, double *Fp,
                 double *s_out, double *slip_inc, int *n_active, const double *W_conv, double *W_out);

/* x = pinv(A) b for `count` dense n x n systems (row-major A: n*n per system; b, x: n per
 * system; n <= 48, count <= 2^20), host buffers, on the current device: the pseudo-inverse step
 * that inactiveSlipRemoval.cc:26,99 takes with FullMatrix::compute_inverse_svd(0.0) when the
 * active-set matrix A_PA is singular, with the same small-singular-value rule as the quadrature-
 * point kernels.  path 0 solves each system in one lane (the routine a quadrature point runs in
 * its own lane), path 1 with the cooperative kernel that solves queued singular systems (a group
 * of lanes per system), path 2 with the routine the PA_QP_COOP qp kernels run in place (a group of
 * lanes per system, rows in LDS); all three give bit-identical x.  No context needed. */
int pa_pinv_solve(int path, int n, long long count, const double *A, const double *b, double *x);

/* pa_pinv_solve without device allocation.  No counterpart in the reference: it is for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same arguments, kernels and results (bit for bit) as pa_pinv_solve, but
 * every device buffer the solve needs (the staged A, b, x, the padded store and its per-system
 * flags and queue length) lives in one file-scope device array of PA_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends; no hipMalloc/hipFree or other allocating call is made.
 * Capacity: count systems of size n (leading dimension NS = 12, 24 or 48 for n <= 12, 24, 48)
 * need 8 (n^2 + 2n + NS^2 + 2 NS) + 12 bytes each (each of ten segments rounded up to 256
 * bytes): about 12,400 systems at n = 12 and 870 at n = 48.  A larger count returns PA_ERR_ARG with the limit
 * in pa_last_error.  All calls on one GPU share its array: calls that overlap on one GPU (from
 * several host threads) race and must be serialised by the caller. */
#define PA_NOALLOC_SCRATCH_BYTES (32u << 20)
int pa_pinv_solve_noalloc(int path, int n, long long count, const double *A, const double *b, double *x);

/* Residuals of the last linear solve (no counterpart in the reference, whose PETSc solve reports
 * only its own stopping norm): with B the condensed residual and X the correction it solved for,
 * out[0] = ||B - K X|| / ||B|| (true residual, K the matrix in memory: FP32 values in the mixed
 * form), out[1] = ||D^-1 (B - K X)|| / ||B|| (recomputed stopping quantity), out[2] = the solver's own
 * recurrence value of that quantity / ||B||.  Valid until the next assembly. */
int pa_last_residual(pa_context *ctx, double out[3]);

/* Timing of the last pa_newton_iteration / pa_run_increment phases on the context stream (ms):
 * [0] assembly (qp + tangent + K_e + scatter), [1] Krylov, [2] post-increment pass. */
int pa_last_timings(pa_context *ctx, double ms[3]);

/* Instrumentation (no counterpart in the reference): duration (ms, HIP events on the context
 * stream) of the first qp pass of the latest assembly, the kernel that runs the return map and
 * the consistent tangent at every quadrature point (numerically singular local systems are handed
 * to the pseudo-inverse pass that follows it).  PA_ERR_STATE before the first assembly. */
int pa_last_qp_pass_ms(pa_context *ctx, double *ms);

/* Instrumentation (no counterpart in the reference): quadrature points whose local active-set
 * matrix was numerically singular and went through the pseudo-inverse pass, in the last assembly
 * (counts[0]) and the last post-increment pass (counts[1]); counts[2]: reruns, since the context
 * was created, that met a singular system other than the one solved for them and solved it in
 * their own lane. */
int pa_last_counts(pa_context *ctx, long long counts[3]);

/* Instrumentation (no counterpart in the reference): times `reps` launches of one kernel
 * between HIP events on the context stream, on the context's current data.  which:
 *   0 = BiCG operator pair  Zr = K Pr, Zl = K^T Pl  (the Krylov hot kernel)
 *   1 = one full assembly (qp-tangent launch + 8 colour launches of K_e + scatter)
 *   2 = Jacobi-PCG operator  Q = K P
 *   3 = the qp-tangent launch alone (qp update + dP/dF of every qp)
 *   4 = the 8 element colour launches alone (K_e, R_e, condensed scatter)
 *   5 = Jacobi-PCG update  x += a p, r -= a q, z = D^-1 r  (with the z.z, r.z partial sums)
 *   6 = Jacobi-PCG direction  p = z + b p
 *   7 = the first qp pass alone (every qp on the LU path; singular local systems only queued)
 * avg_ms receives the mean time of one launch (of one sweep for which = 1 and 4).  The matrix
 * is not re-zeroed between repetitions, and the Krylov vectors are overwritten (timing only). */
int pa_time_kernel(pa_context *ctx, int which, int reps, double *avg_ms);

/* ---- Learned warm start of the local return map (no counterpart in the reference: its local
 * loop, calculatePlasticity.cc:311-441, always starts from zero slip).  An MLP 18 -> 64 -> 64 -> 64
 * -> 12 (SiLU; BF16 weights and activations on the matrix cores, FP32 accumulation and master
 * weights) predicts the 12 slip increments of an FCC quadrature point from its 12 trial ratios
 * (tau_a - W_a)/s_a and the lattice-frame increment of its elastic Green-Lagrange strain (6, Voigt
 * 11,22,33,23,13,12).  The plain local loop then corrects the prediction and judges convergence by
 * the reference's test (||b_PA||_inf <= Stress Tolerance), so the network changes how many local
 * solves a qp takes, not what it converges to.
 *
 * Parameter blob (PA_MLP_NBLOB floats): per layer l = 0..3, W^T[N_l][K_l] row-major then b[N_l],
 * padded sizes (K, N) = (32, 64), (64, 64), (64, 64), (64, 16) (padding zero); then in_mean[18],
 * in_scale[18], out_scale[12].  Network input = (x - in_mean) * in_scale, prediction = y * out_scale. */
#define PA_MLP_NIN 18
#define PA_MLP_NOUT 12
#define PA_MLP_NPARAM 11472
#define PA_MLP_NBLOB 11520
typedef struct pa_mlp pa_mlp;
int pa_mlp_create(pa_mlp **out);
int pa_mlp_destroy(pa_mlp *net);
int pa_mlp_set_params(pa_mlp *net, const float *blob);   /* resets the Adam moments */
int pa_mlp_get_params(pa_mlp *net, float *blob);
/* Y[n][12] = prediction for X[n][18] (host buffers). */
int pa_mlp_predict(pa_mlp *net, long long n, const float *X, float *Y);
/* Gradient of the loss mean_{samples, 12 outputs} (y - Y/out_scale)^2 w.r.t. the PA_MLP_NPARAM
 * weights and biases, and the loss, over n <= 65536 samples as one batch. */
int pa_mlp_grad(pa_mlp *net, long long n, const float *X, const float *Y, float *grad, double *loss);
/* Adam (beta 0.9/0.999, eps 1e-8) on that loss, minibatches of `batch` samples in a seeded
 * per-epoch permutation, learning rate cosine-decayed from lr to lr/20; epoch_loss (optional,
 * `epochs` entries) receives the mean minibatch loss of each epoch.  Runs on the device. */
int pa_mlp_train(pa_mlp *net, long long n, const float *X, const float *Y, int epochs, int batch, float lr,
                 unsigned long long seed, double *epoch_loss);
/* pa_qp_update for FCC without kinematic hardening, with the local loop started from the network's
 * prediction (net may be NULL: the reference's start from zero).  Fe_conv (9/qp) enters the strain
 * increment input.  iters: local solves per qp; feat (optional, 18/qp): the network inputs;
 * x0 (optional, 12/qp, net only): the predictions.  Material Max Solver Iterations caps the loop. */
int pa_qp_update_warm(pa_context *ctx, pa_mlp *net, long long nqp, const double *F, const double *Fp_conv,
                      const double *Fe_conv, const double *s_conv, const unsigned *grain, int flag, double *P,
                      double *T, double *dPdF, double *Fe, double *Fp, double *s_out, double *slip_inc,
                      int *n_active, int *iters, float *feat, float *x0);

/* ---- Batched RVEs.  No counterpart in the reference, which runs one RVE per job; each RVE of a
 * batch is the reference's boundary-value problem (ellipticBVP<3> + crystalPlasticity<3> on its own
 * box mesh), and all of them advance together: one Newton iteration of every RVE (assemble.cc:6-104
 * per RVE, then a Jacobi-BiCG or Jacobi-PCG solve per RVE, solveLinearSystem.cc:8-36) runs in the
 * same launches, the qp work of all RVEs drained by one persistent kernel from a device work queue,
 * replayed from HIP graphs captured once per batch.
 *   mats: 1..4 materials (no twinning, no kinematic hardening); rve_mat[r]: material of RVE r;
 *   rve_nel[3r..3r+2]: its elements along x, y, z; pb: domain span, boundary rules (simple face
 *   rules only), increments, Newton count and Krylov tolerance, shared by every RVE (pb->nx/ny/nz
 *   and pb->solver are not used).  Every RVE needs pa_batch_set_grains before the first step. */
typedef struct pa_batch pa_batch;
int pa_batch_create(const pa_material *mats, int n_mats, const pa_problem *pb, int n_rve, const int *rve_mat,
                    const int *rve_nel, pa_batch **out);
void pa_batch_destroy(pa_batch *batch);
/* Grain of each element of RVE `rve` (row into grain_rod, n_grain_rows x 3 Rodrigues vectors), as
 * pa_set_grains for one context (crystalPlasticity::loadOrientations, loadOrientations.cc:4-33). */
int pa_batch_set_grains(pa_batch *batch, int rve, const unsigned *elem_grain, const double *grain_rod, int n_grain_rows);
/* Krylov method (PA_SOLVER_BICG, the reference's and the default, or PA_SOLVER_CG) and arithmetic
 * form (PA_PREC_FP64 or PA_PREC_MIXED: FP32 element tangents and matrices), as pa_set_solver. */
int pa_batch_set_solver(pa_batch *batch, int method, int precision);
/* Replay the Newton iteration from HIP graphs (1, the default) or launch its kernels directly (0);
 * the results are bitwise the same. */
int pa_batch_set_graphs(pa_batch *batch, int on);
/* One Newton iteration of every RVE (solveNonLinearSystem.cc:14-47 per RVE): res_norms[r] = |R| of
 * RVE r's assembly, lin_iters[r] its CG iterations (either may be NULL). */
int pa_batch_newton_iteration(pa_batch *batch, int inc, int it, double *res_norms, int *lin_iters);
/* updateAfterIncrement for every RVE (updateAfterIncrement.cc:6-305, 710-771): rows[15r..15r+14]
 * = RVE r's stressstrain.txt row (may be NULL). */
int pa_batch_commit_increment(pa_batch *batch, double *rows);
/* Autotuner: times the persistent qp pass at 256/512/768/1024 one-wave blocks and the Krylov chunk
 * at 8/16/32 iterations on the current state and keeps the fastest.  report (may be NULL, 9
 * doubles): grid, chunk, qp-pass ms per grid, solve ms per chunk.  Graphs are re-captured after. */
int pa_batch_tune(pa_batch *batch, double *report);
/* The persistent qp pass alone (flag 1: with the tangent) on the current state, HIP-event ms. */
int pa_batch_qp_pass_ms(pa_batch *batch, int flag, double *ms);
/* ms of the last assembly, Krylov loop, post-increment pass and timed qp pass. */
int pa_batch_last_timings(pa_batch *batch, double ms[4]);
/* State of RVE `rve` as pa_get_state ids 0-7 (T, Fp, Fe, s, rotnew, slip, u, R), per qp / node. */
int pa_batch_get_state(pa_batch *batch, int rve, int which, double *host, long long n);
/* Totals of the batch: quadrature points, elements, nodes, dofs. */
int pa_batch_sizes(pa_batch *batch, long long out[4]);

#ifdef __cplusplus
}
#endif
#endif

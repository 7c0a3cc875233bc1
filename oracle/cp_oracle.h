/*
 * CPU FP64 restatement of the PRISMS-Plasticity crystal-plasticity Newton step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: it is loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and by nothing else.  The product (plasticity_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm; the .cpp cites the
 * reference file:line each block follows (paths relative to the reference
 * repository root).
 */
#ifndef CP_ORACLE_H
#define CP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXT 49 /* BCC-48 + the inert twin system */
#define OR_MAX_TABLE 32 /* time points of tabular boundary conditions */

/* pinv_mode: how x = pinv(A_PA) b_PA is formed inside inactive_slip_removal
 *   0 = Jacobi-SVD pseudo-inverse with threshold 0 (deal.II
 *       LAPACKFullMatrix::compute_inverse_svd(0.0) semantics,
 *       src/materialModels/crystalPlasticity/inactiveSlipRemoval.cc:26,99)
 *   1 = LU with partial pivoting; when a pivot is <= 1e-10 max|A| falls back to
 *       the SVD pseudo-inverse.  A single singular value <= 1e-10 sigma_max is still
 *       inverted (the reference's behaviour) when the step along its singular pair is
 *       determined by the data; otherwise the small ones are dropped (minimum-norm
 *       solution).  The semantics the HIP kernel implements. */
typedef struct {
  int n_slip, n_twin, nT;          /* nT = n_slip + n_twin (n_twin=1 dummy when twinning is off) */
  double m[OR_MAXT][3], n[OR_MAXT][3];  /* normalised slip directions / normals */
  double q[OR_MAXT][OR_MAXT];       /* latent hardening ratios */
  double C[6][6];                   /* crystal-frame elastic stiffness, Voigt (11,22,33,23,13,12), MPa */
  double s0[OR_MAXT], h0[OR_MAXT], ss[OR_MAXT], pw[OR_MAXT];
  double C1[OR_MAXT], C2[OR_MAXT];  /* kinematic hardening (0 when disabled) */
  double tol;                       /* Stress Tolerance */
  int max_slip_search, max_solver_iter;
  int pinv_mode;
  /* Twinning (calculatePlasticity.cc:421-451, 962-1019): with twinning = 1 the n_twin systems after
   * the n_slip slip systems are real twin systems (their s0/h0/ss/pw, latent rows and normalised
   * m/n in the arrays above); twin_shear = Characteristic Twin Shear, twin_threshold = Twin
   * Threshold Fraction, twin_saturation = Twin Saturation Factor.  twinning = 0: one inert
   * system (n_twin = 1, s0 = 1e15) that never activates. */
  int twinning;
  double twin_shear, twin_threshold, twin_saturation;
} or_material;

/* Periodic boundary conditions (Use Periodic BCs; ellipticBVP init.cc:66-173, periodicBCs.cc).
 * Rows of the constraint file (PeriodicBCsConstraints.txt): vertex rows 0..23 = V000 V100 V010 V001
 * V110 V101 V011 V111 x 3 components; edge rows 0..35 = EX000 EX001 EX010 EX011 EY000 EY001 EY100
 * EY101 EZ000 EZ010 EZ100 EZ110 x 3; face rows 0..17 = FXN FXP FYN FYP FZN FZP x 3. */
#define OR_MAX_PBC_LINES 64
typedef struct {
  int n_vertex_lines;                            /* Number of Vertices Constraints */
  int vertex_dofs[OR_MAX_PBC_LINES][5];          /* entry count, then vertex rows */
  double vertex_coef[OR_MAX_PBC_LINES][5];       /* (count), then their coefficients */
  int n_edge_lines;                              /* Number of Edges Constraints */
  int edge_dofs[OR_MAX_PBC_LINES][4];            /* edge row, edge row, vertex row, vertex row */
  double edge_coef[OR_MAX_PBC_LINES][4];
  int n_face_lines;                              /* Number of Faces Constraints */
  int face_dofs[OR_MAX_PBC_LINES][4];            /* positive-face row, face row, vertex row, vertex row */
  double face_coef[OR_MAX_PBC_LINES][4];
  int vertex_known[24];                          /* Vertices Periodic BCs row 1 */
  double vertex_total[24];                       /* Vertices Periodic BCs row 2 */
  int tabular;                                   /* Use Tabular Periodic BCs */
  double tabular_time;                           /* periodic Tabular time */
  int n_table;
  double table_time[OR_MAX_TABLE];               /* Tabular Periodic Time Table */
  int table_coef[OR_MAX_TABLE];                  /* Tabular Periodic Time Table Coefficient (1, -1, 0) */
} or_periodic;

typedef struct {
  int nx, ny, nz;                   /* elements per direction (after refinement) */
  double span[3];
  int n_increments;                 /* increments to run (<= total) */
  int total_increments;             /* round(Total time / Time increments) */
  double delT;
  int max_newton;                   /* Maximum non linear iterations */
  double lin_tol;                   /* Relative linear solver tolerance */
  int lin_maxit;
  int solver;                       /* 0 = Jacobi-BiCG (reference), 1 = Jacobi-PCG */
  int face_dof_constrained[6][3];
  double face_total_u[6][3];        /* FinalDisplacement of each (face,dof) rule */
  int cyclic;                       /* Enable cyclic loading */
  int cyclic_face, cyclic_dof;      /* 1-based as in prm */
  double quarter_cycle_time;
  int nthreads;                     /* OpenMP threads for the CPU baseline (<=0: all) */
  int tabular;                      /* Use Tabular BCs: face_total_u is ignored, the table below drives */
  int n_table;                      /* Number of time data for Tabular BCs (<= OR_MAX_TABLE) */
  double table_time[OR_MAX_TABLE];  /* Tabular Time Table */
  double table_u[6][3][OR_MAX_TABLE];  /* displacement of each (face,dof) rule at the table times */
  int velgrad;                      /* Use velocity gradient BC */
  double vel_grad[3][3];            /* L */
  const or_periodic *periodic;      /* Use Periodic BCs (NULL: off); the face rules above must be empty */
} or_problem;

/* Reorientation by twinning at the end of calculatePlasticity (calculatePlasticity.cc:962-1019),
 * for one qp after its update: F_r, F_e = the volume-averaged twin fraction and reoriented volume
 * of the previous increment (updateAfterIncrement.cc:222-231, 719-720); tw_iter = the qp's twin
 * fractions (n_twin); twin_conv = 1 once the qp has twinned in an earlier increment; rot_conv
 * its orientation.  When the largest twin fraction exceeds F_T = min(1, threshold + saturation
 * F_e / F_r) (F_r > 0, not yet twinned), rot_iter receives the orientation rotated by 180 degrees
 * about that twin system's normal (quaternion product), *twin_iter = 1 and the twin systems'
 * s_iter entries 1e5; returns the twin system (0-based) or -1 (nothing written). */
int or_twin_event(const or_material *mat, double F_r, double F_e, const double *tw_iter, int twin_conv,
                  const double rot_conv[3], double rot_iter[3], int *twin_iter, double *s_iter);

/* One calculatePlasticity(cellID,q,flag) call.  In/out arrays follow the
 * reference member layout (row-major 3x3, dP_dF[i][k][j][l] flattened as
 * (3i+k)*9 + 3j+l).  slip_iter is only written on the plastic path
 * (calculatePlasticity.cc:454-456), so it is in/out.
 * Returns the number of active systems after inactive-slip removal (0 = elastic). */
int or_qp_update(const or_material *mat, const double rot[3], const double F[9],
                 const double Fp_conv[9], const double *s_conv, const double *W_conv,
                 const double *slip_conv, int flag,
                 double P[9], double T[9], double dPdF[81], double Fe_iter[9],
                 double Fp_iter[9], double *s_iter, double *W_iter, double *slip_iter);
/* or_qp_update with twin fractions: tw_conv (n_twin) in, tw_iter (n_twin) in/out (written on the
 * plastic path as tw_conv + slip increment / twin_shear, calculatePlasticity.cc:450-452). */
int or_qp_update_twin(const or_material *mat, const double rot[3], const double F[9], const double Fp_conv[9],
                      const double *s_conv, const double *W_conv, const double *slip_conv, const double *tw_conv, int flag,
                      double P[9], double T[9], double dPdF[81], double Fe_iter[9], double Fp_iter[9], double *s_iter,
                      double *W_iter, double *slip_iter, double *tw_iter);

/* Test-only: scale block `block` of the consistent tangent by (1 + rel) in every later
 * or_qp_update call (block 0 = none; ids in cp_oracle.cpp at g_pert_block).  Lets a test show
 * that its tangent checks detect an error of that size in each block. */
void or_tangent_perturbation(int block, double rel);

/* or_qp_update with a warm start (no counterpart in the reference): x0 (n_slip values, may be
 * NULL = the reference's start from zero) holds predicted slip increments; those of the
 * potentially-active systems that are positive are applied once before the first solve of the
 * local loop (calculatePlasticity.cc:322-441), which then runs unchanged.  iters receives the
 * number of local solves the loop took (0 on the elastic path). */
int or_qp_update_warm(const or_material *mat, const double rot[3], const double F[9], const double Fp_conv[9],
                      const double *s_conv, const double *W_conv, const double *slip_conv, int flag, const double *x0,
                      double P[9], double T[9], double dPdF[81], double Fe_iter[9], double Fp_iter[9], double *s_iter,
                      double *W_iter, double *slip_iter, int *iters);

/* Nearest-voxel grain lookup of a point (crystalOrientationsIO.cc:115-194).
 * vox is [nvx][nvy][nvz] (z fastest) as read from the grain-ID file. */
unsigned or_material_id(const unsigned *vox, const int nv[3], const double span[3], const double p[3]);

/* Grain id of every element (cell centre lookup, loadOrientations.cc:4-33),
 * elements ordered ex + nx*(ey + ny*ez). */
void or_element_grains(const unsigned *vox, const int nv[3], const or_problem *pb, unsigned *elem_grain);

/* Full-pattern CSR of the structured Q1^3 mesh (init.cc:183-199); row_ptr has
 * ndof+1 entries.  Returns nnz; col_idx may be NULL to query the size. */
long long or_sparsity(const or_problem *pb, long long *row_ptr, int *col_idx);

/* Run n_increments increments of the Newton loop (solve.cc:69-135) and write
 * the 15-column stressstrain rows (updateAfterIncrement.cc:735-771).
 * grain_rod is indexed by the grain id in elem_grain (rows of 3).
 * lin_iters (optional) receives n_increments*max_newton Krylov iteration counts,
 * res_norms (optional) the Newton residual norms.
 * state_out (optional) receives the final per-qp state, qp-major:
 *   [T 9 | Fp 9 | Fe 9 | s nT | rotnew 3 | slip n_slip] per qp, followed with twinning by
 *   [rot 3 | twinned 1 | twin fractions n_twin].
 * With twinning, rows[12] = TwinRealVF (volume-averaged twin fraction) and rows[13] = TwinMade
 * (reoriented volume fraction), updateAfterIncrement.cc:719-720, 755.
 * u_out (optional) receives the final displacement (3 per node). */
int or_simulate(const or_material *mat, const or_problem *pb, const unsigned *elem_grain,
                const double *grain_rod, int n_grain_rows,
                double *rows, int *lin_iters, double *res_norms, double *state_out,
                double *u_out);

/* or_simulate with several phases (multiphaseInit.cc:68-361, Enable Multiphase): mats[k] is phase
 * k + 1's material (n_phases <= 4), grain_phase[g] (1-based) the phase of grain row g.  Per-qp state
 * has the largest phase's size (state_out rows as or_simulate with those sizes); TwinRealVF and
 * SlipTotal sum the first numTwinSystems1 / numSlipSystems1 fractions of every qp (phase 1's counts,
 * numTwinSystems1 = 1 without its twinning; updateAfterIncrement.cc:222-236) and F_T is each
 * phase's threshold fraction alone (calculatePlasticity.cc:969-980). */
int or_simulate_phases(const or_material *mats, int n_phases, const int *grain_phase, const or_problem *pb,
                       const unsigned *elem_grain, const double *grain_rod, int n_grain_rows, double *rows, int *lin_iters,
                       double *res_norms, double *state_out, double *u_out);

/* The closed periodic constraints of pb (pb->periodic set) at unit load scale (the increment-0
 * inhomogeneities of setPeriodicityConstraintsInc0): for every dof d, constrained[d] = 1 if the
 * dof is a constrained line, whose entries are master[rowptr[d]..rowptr[d+1]) with weight[...],
 * and g[d] its inhomogeneity.  rowptr has ndof + 1 entries.  Returns the number of entries (call
 * with cap 0 to size master/weight) or -1 if the input is refused. */
long long or_periodic_lines(const or_problem *pb, int *constrained, long long *rowptr, long long *master, double *weight,
                            double *g, long long cap);

/* Increment 0, Newton iteration 0 only: the condensed CSR values (pattern of or_sparsity),
 * the condensed residual and the Newton correction du (after constraints.distribute). */
int or_first_newton(const or_material *mat, const or_problem *pb, const unsigned *elem_grain,
                    const double *grain_rod, int n_grain_rows, double *K_vals, double *R, double *du);

/* Increment inc, Newton iteration it (after inc full increments and `it` Newton iterations of
 * increment inc): the condensed CSR values, residual and Newton correction, as or_first_newton. */
int or_newton_system(const or_material *mat, const or_problem *pb, const unsigned *elem_grain,
                     const double *grain_rod, int n_grain_rows, int inc, int it, double *K_vals, double *R,
                     double *du);

/* Jacobi-BiCG (method 0) / Jacobi-PCG (1) on a CSR system with absolute tolerance tol on the
 * preconditioned residual norm; returns the iteration count. */
int or_krylov(int method, long long n, const long long *rp, const int *ci, const double *val, const double *b,
              double *x, double tol, int maxit);

/* Counts of inactive-slip-removal solves and of those whose A_PA was numerically
 * singular (LU pivot <= 1e-10 max|A|) since the last call; resets the counters. */
void or_pinv_stats(long long *calls, long long *singular);
/* The first A_PA (n x n, row-major) and b_PA that took the fallback since the process started;
 * returns n (0 if none). */
int or_pinv_sample(double *A, double *b);
/* Diagnostics: with enable = 1 every solve adds its sigma_min / sigma_max to a decade histogram
 * (bin k: [1e-(k+1), 1e-k), bin 23: exactly singular); hist receives and resets the counts. */
void or_pinv_histogram(int enable, long long hist[24]);
/* With the histogram enabled, every solve with sigma_min < 1e-8 sigma_max is also recorded as
 * (n, sigma_min/sigma_max, second smallest/sigma_max, |u_min . b|/|b|); returns and clears them. */
int or_pinv_records(double *out, int max_records);
/* Overrides the relative pivot / singular-value level of pinv_mode 1 (experiments only). */
void or_set_pinv_rtol(double rtol);
/* Wall-clock split of the latest or_simulate (ns): [0] assembly (qp passes, element matrices,
 * scatter), [1] Krylov solves and updates, [2] post-increment passes. */
void or_last_timings(long long ns[3]);

/* Single increment helpers exposed for property tests. */
void or_odfpoint(const double r[3], double R[9]);
void or_matrix_exponential(const double A[9], double E[9]);
void or_polar_rotation(const double Fe[9], double R[9]);
int  or_pinv_solve(int n, const double *A, const double *b, double *x, int mode);
int  or_element_matrix(const or_material *mat, const double rot[3], const double h[3],
                       const double ue[24], const double *Fp_conv, const double *s_conv,
                       double Ke[576], double Re[24]);

#ifdef __cplusplus
}
#endif
#endif

// Quadrature-point passes of the Newton step: mesh helpers, the argument blocks of assembly
// stage 1 (getElementalValues.cc:129-141 with calculatePlasticity(.,.,1)) and of the post-increment
// pass (updateAfterIncrement.cc:77-126 with calculatePlasticity(.,.,0)), and the bodies of their
// cooperative form and of the batch entry point, one group of lanes per qp (coop::qp_update).  The
// bodies take the block's LDS as an argument; the __global__ wrappers that declare it live in
// cp_api.hip.  Included by cp_api.hip inside its anonymous namespace.
#pragma once

struct MeshDims {
  int nx, ny, nz, nnx, nny, nnz;
  long long nnode, nelem;
};

// Q1 gradient of vertex v at Gauss point q of a box cell with sides h
__device__ __forceinline__ void shape_grad(int q, int v, const double *h, double g[3]) {
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  double N[3], dN[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xi = ((q >> d) & 1) ? g1 : g0;
    const int b = (v >> d) & 1;
    N[d] = b ? xi : 1.0 - xi;
    dN[d] = (b ? 1.0 : -1.0) / h[d];
  }
  g[0] = dN[0] * N[1] * N[2];
  g[1] = N[0] * dN[1] * N[2];
  g[2] = N[0] * N[1] * dN[2];
}

// F = I + sum_v u_v (x) grad N_v at qp Q = 8 e + q (getElementalValues.cc:129-138)
__device__ __forceinline__ void qp_F(const MeshDims &m, const double *h, const double *u, long long Q, double F[9]) {
  const long long e = Q >> 3;
  const int q = (int)(Q & 7);
  const int ex = (int)(e % m.nx), ey = (int)((e / m.nx) % m.ny), ez = (int)(e / ((long long)m.nx * m.ny));
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const long long nd = (ex + (v & 1)) + (long long)m.nnx * ((ey + ((v >> 1) & 1)) + (long long)m.nny * (ez + ((v >> 2) & 1)));
    double g[3];
    shape_grad(q, v, h, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double uc = u[3 * nd + c];
#pragma unroll
      for (int j = 0; j < 3; ++j) F[3 * c + j] += uc * g[j];
    }
  }
  F[0] += 1; F[4] += 1; F[8] += 1;
}

// Store of the numerically singular local systems that the one-lane passes defer to the
// cooperative pseudo-inverse kernel (cp_api.hip, k_sing_svd)
constexpr int SING_ROUNDS = 1;
struct SingStore {
  unsigned cap;  // entries with storage
  double *A;     // [cap][NS*NS] A_PA, leading dimension NS
  double *b;     // [cap][NS]
  int *n;        // [cap]
  double *X;     // [cap][SING_ROUNDS][NS] solutions in the order the qp meets its singular systems
  unsigned long long *key;  // [cap][SING_ROUNDS] hash of the bits of each solved system
  int *have;     // [cap] solutions stored
  int *done;     // [cap] 1 once the qp is finished
  unsigned *miss;  // replays whose system differed from the stored one (solved in their lane)
};

// qps a wavefront claims from a work queue at a time (a multiple of the groups per wavefront)
constexpr int QP_CHUNK = 16;

// Persistent wavefronts: each claims QP_CHUNK consecutive qps from the work queue (one counter,
// zeroed before the launch), one qp per group at a time, until the queue is empty.  Active sets
// and local iteration counts vary from qp to qp, so the queue (not a fixed grid partition)
// balances the work; a qp's result does not depend on which group computes it.
template <int G, typename F>
__device__ __forceinline__ void qp_queue(unsigned *work, long long nq, F &&one) {
  constexpr int GPW = 64 / G;
  const int gi = (int)(threadIdx.x & 63) / G;
  for (;;) {
    unsigned base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(work, (unsigned)QP_CHUNK);
    base = (unsigned)__shfl((int)base, 0);
    if ((long long)base >= nq) break;
    for (int k = gi; k < QP_CHUNK; k += GPW) {
      const long long Q = (long long)base + k;
      if (Q < nq) one(Q);
    }
  }
}

// ---- twinning (calculatePlasticity.cc:421-452, 962-1019).  The twin systems follow the slip
// systems in the system arrays (s, x); a qp's lattice orientation is one of 1 + ntw variants of
// its grain's (0: the grain's, 1 + k: turned 180 degrees about twin system k's normal, at most one
// reorientation per qp), each with its own orientation tables (row grain * (1 + ntw) + variant).
struct TwinArgs {
  int on;                 // twinning enabled (everything below unused otherwise)
  int nslip, ntw;         // slip systems; twin systems
  double shear;           // Characteristic Twin Shear
  double F_T;             // min(1, Twin Threshold Fraction + Twin Saturation Factor F_e / F_r) of this increment
  int armed;              // F_r > 0 (calculatePlasticity.cc:992): a qp may reorient in this increment
  const double *tw_conv;  // [ntw][nq] twin fractions
  double *tw_iter;
  const int *var_conv;    // [nq] orientation variant
  int *var_iter;
  const double *vrod;     // [grain][nvar][3] Rodrigues vectors of the variants
  int nvar;               // orientation-table rows per grain: 1 + the largest twin-system count
};

// After a qp update: the twin fractions (written on the plastic path only, as the reference writes
// twinfraction_iter, :450-452; otherwise the last call's stand) and the reorientation test of
// :986-1016 (largest fraction, first of equals, above F_T; once per qp): the variant, and twin
// slip resistances of 1e5.
__device__ __forceinline__ void twin_epilogue(const TwinArgs &t, long long Q, long long nq, bool plastic, const double *x,
                                              double *s_iter) {
  double fmax = 0.0;
  int pos = 0;
  for (int k = 0; k < t.ntw; ++k) {
    double f;
    if (plastic) {
      f = t.tw_conv[k * nq + Q] + x[t.nslip + k] / t.shear;
      t.tw_iter[k * nq + Q] = f;
    } else {
      f = t.tw_iter[k * nq + Q];
    }
    if (k == 0 || f > fmax) { fmax = f; pos = k; }
  }
  if (t.var_conv[Q] == 0 && t.armed && fmax > t.F_T) {
    t.var_iter[Q] = 1 + pos;
    for (int k = 0; k < t.ntw; ++k) s_iter[(t.nslip + k) * nq + Q] = 100000.0;
  }
}
// Several phases (multiphaseInit.cc:68-361): each grain row's phase selects its material block
// (hardening, latent ratios, kinematic hardening, tolerance) and its system count; the grain
// tables already hold the phase's rotated Schmid tensors and elastic moduli.
struct PhaseArgs {
  const int *gphase;  // phase (0-based) of each grain row; null: one phase
  int mstride;        // doubles per material block
  int ns[4];          // systems of each phase (slip, then twin)
  int nslip[4], ntw[4];
  double shear[4];    // Characteristic Twin Shear of each phase
  double F_T[4];      // min(1, Twin Threshold Fraction): with several phases the saturation term is not
                      // used (calculatePlasticity.cc:969-985)
};
// the twinning arguments of a qp's phase (the tables and fraction arrays are shared)
__device__ __forceinline__ void phase_twin(const PhaseArgs &ph, unsigned grain, TwinArgs &tw) {
  const int p = ph.gphase[grain];
  tw.nslip = ph.nslip[p];
  tw.ntw = ph.ntw[p];
  tw.shear = ph.shear[p];
  tw.F_T = ph.F_T[p];
}
__device__ __forceinline__ void qp_phase(const PhaseArgs &ph, const double *mat0, int ns0, unsigned grain, const double *&mat, int &ns) {
  if (ph.gphase) {
    const int p = ph.gphase[grain];
    mat = mat0 + (long long)p * ph.mstride;
    ns = ph.ns[p];
  } else {
    mat = mat0;
    ns = ns0;
  }
}

// orientation-table row of qp Q in element e
__device__ __forceinline__ long long table_row(const TwinArgs &t, const unsigned *egrain, long long Q) {
  return t.on ? (long long)egrain[Q >> 3] * t.nvar + t.var_conv[Q] : (long long)egrain[Q >> 3];
}

// ---- assembly, stage 1: F from u, calculatePlasticity(.,.,1); writes the iter history,
// JxW * dP/dF (as DT) and JxW * P.
struct QpArgs {
  MeshDims m;
  double h[3], JxW;
  const double *u;
  const unsigned *egrain;
  const double *gtab;
  int gstride;
  const double *mat;
  int ns, max_solver_iter;
  long long nq;
  const double *Fp_conv, *s_conv, *slip_conv, *W_conv;
  double *Fe_iter, *Fp_iter, *s_iter, *slip_iter, *W_iter;  // W_*: backstress (kinematic hardening only)
  void *D;          // [81][nq] of DT
  double *P;        // [9][nq]
  // one-lane form: qps deferred to the pseudo-inverse pass and their singular systems
  unsigned *queue, *qcount;
  SingStore sing;
  // cooperative form
  unsigned *work;   // work queue: the next unclaimed qp (zeroed before each launch)
  unsigned *nsing;  // pseudo-inverse solves (numerically singular A_PA) of the pass
  TwinArgs tw;      // twinning (one-lane form)
  double *ws;       // one-lane form, 48 systems: per-lane matrix workspace (cp::LaneWs), else null
  PhaseArgs ph;     // several phases (one-lane form)
};

template <int NS, typename DT, bool KIN, int G>
__device__ __forceinline__ void qp_tangent_coop_one(const QpArgs &a, long long Q, const coop::Group<G> &g, double *L) {
  typedef coop::Slice<NS, KIN> SL;
  double F[9], Fp_t[9], P[9], T[9], Fe[9], Fp[9];
  qp_F(a.m, a.h, a.u, Q, F);
#pragma unroll
  for (int i = 0; i < 9; ++i) Fp_t[i] = a.Fp_conv[i * a.nq + Q];
  const int l = g.l;
  const bool sys = l < a.ns;
  const double s_a = sys ? a.s_conv[l * a.nq + Q] : 0.0;
  double W_a = 0.0;
  if constexpr (KIN)
    if (sys) W_a = a.W_conv[l * a.nq + Q];
  const double *gt = a.gtab + (long long)a.egrain[Q >> 3] * a.gstride;
  bool plastic = false;
  coop::qp_update<NS, G, DT, KIN>(g, L, a.mat, gt, a.ns, a.max_solver_iter, F, Fp_t, s_a, W_a, 1, P, T, Fe, Fp, (DT *)a.D + Q,
                                  a.nq, a.JxW, plastic, a.nsing);
  if (l < 9) {
    a.Fe_iter[l * a.nq + Q] = coop::pick<9>(Fe, l);
    a.Fp_iter[l * a.nq + Q] = coop::pick<9>(Fp, l);
    a.P[l * a.nq + Q] = a.JxW * coop::pick<9>(P, l);
  }
  if (sys) {
    a.s_iter[l * a.nq + Q] = L[SL::s + l];
    if constexpr (KIN) a.W_iter[l * a.nq + Q] = L[SL::W + l];
    // slipfraction_iter is written only on the plastic path (calculatePlasticity.cc:454-456)
    if (plastic) a.slip_iter[l * a.nq + Q] = a.slip_conv[l * a.nq + Q] + L[SL::x + l];
  }
  coop::gsync();  // the group's next qp reuses the slice
}

// body of k_qp_tangent_coop; lds: 64 / G slices of coop::Slice<NS, KIN>
template <int NS, typename DT, bool KIN>
__device__ __forceinline__ void qp_tangent_coop_pass(const QpArgs &a, double *lds) {
  constexpr int G = coop::group_size<NS>();
  const coop::Group<G> g;
  double *L = lds + ((int)(threadIdx.x & 63) / G) * coop::Slice<NS, KIN>::size;
  qp_queue<G>(a.work, a.nq, [&](long long Q) { qp_tangent_coop_one<NS, DT, KIN, G>(a, Q, g, L); });
}

// ---- post-increment pass
struct PostArgs {
  MeshDims m;
  double h[3], JxW;
  const double *u;
  const unsigned *egrain;
  const double *gtab;
  int gstride;
  const double *mat;
  int ns, max_solver_iter;
  long long nq;
  const double *Fp_conv, *s_conv, *slip_conv, *Fe_conv, *W_conv;
  double *Fe_iter, *Fp_iter, *s_iter, *slip_iter, *T, *rotnew, *W_iter;
  double *partials;
  unsigned *counter;
  double *out;      // 15: E(6) T(6) slip, twin fraction, twinned volume (JxW sums)
  unsigned *queue, *qcount;  // one-lane form: deferred qps
  SingStore sing;            // and their singular systems
  unsigned *work;   // cooperative form: work queue of the qp pass (zeroed before the launch)
  unsigned *nsing;  // pseudo-inverse solves of the pass
  TwinArgs tw;      // twinning (one-lane form)
  double *ws;       // one-lane form, 48 systems: per-lane matrix workspace (cp::LaneWs), else null
  PhaseArgs ph;     // several phases (one-lane form)
};

__device__ __forceinline__ void post_F(const PostArgs &a, long long Q, double F[9]) { qp_F(a.m, a.h, a.u, Q, F); }

// updateAfterIncrement's qp loop (updateAfterIncrement.cc:77-126): calculatePlasticity(.,.,0) at
// the converged displacement, one group of lanes per qp.
template <int NS, bool KIN, int G>
__device__ __forceinline__ void post_qp_coop_one(const PostArgs &a, long long Q, const coop::Group<G> &g, double *L) {
  typedef coop::Slice<NS, KIN> SL;
  double F[9], Fp_t[9], P[9], T[9], Fe[9], Fp[9];
  post_F(a, Q, F);
#pragma unroll
  for (int i = 0; i < 9; ++i) Fp_t[i] = a.Fp_conv[i * a.nq + Q];
  const int l = g.l;
  const bool sys = l < a.ns;
  const double s_a = sys ? a.s_conv[l * a.nq + Q] : 0.0;
  double W_a = 0.0;
  if constexpr (KIN)
    if (sys) W_a = a.W_conv[l * a.nq + Q];
  const double *gt = a.gtab + (long long)a.egrain[Q >> 3] * a.gstride;
  bool plastic = false;
  coop::qp_update<NS, G, double, KIN>(g, L, a.mat, gt, a.ns, a.max_solver_iter, F, Fp_t, s_a, W_a, 0, P, T, Fe, Fp,
                                      (double *)nullptr, 0, 1.0, plastic, a.nsing);
  if (l < 9) {
    a.Fe_iter[l * a.nq + Q] = coop::pick<9>(Fe, l);
    a.Fp_iter[l * a.nq + Q] = coop::pick<9>(Fp, l);
    a.T[l * a.nq + Q] = coop::pick<9>(T, l);
  }
  if (sys) {
    a.s_iter[l * a.nq + Q] = L[SL::s + l];
    if constexpr (KIN) a.W_iter[l * a.nq + Q] = L[SL::W + l];
    if (plastic) a.slip_iter[l * a.nq + Q] = a.slip_conv[l * a.nq + Q] + L[SL::x + l];
  }
  coop::gsync();
}

// body of k_post_qp_coop
template <int NS, bool KIN>
__device__ __forceinline__ void post_qp_coop_pass(const PostArgs &a, double *lds) {
  constexpr int G = coop::group_size<NS>();
  const coop::Group<G> g;
  double *L = lds + ((int)(threadIdx.x & 63) / G) * coop::Slice<NS, KIN>::size;
  qp_queue<G>(a.work, a.nq, [&](long long Q) { post_qp_coop_one<NS, KIN, G>(a, Q, g, L); });
}

// ---- batch of calculatePlasticity calls with host-provided inputs (parity entry point), one group
// of lanes per qp: F, Fpc 9/qp; sc, so, xo n_slip/qp; D 81/qp (flag 1); W (KIN): backstress in
// (W_kh_conv) and out (W_kh_tau), n_slip per qp.  Block b holds qps 64/G * b ...
struct BatchArgs {
  long long n;
  const double *F, *Fpc, *sc;
  const unsigned *grain;
  const double *gtab;
  int gstride;
  const double *mat;
  int ns, msi, flag;
  double *P, *T, *D, *Fe, *Fp, *so, *xo;
  int *na;
  double *Wio;
  int nslip = 1 << 30;  // systems from nslip on are twin systems (one-sided); default: none
};

template <int NS, bool KIN>
__device__ __forceinline__ void qp_batch_body(const BatchArgs &b, double *lds) {
  constexpr int G = coop::group_size<NS>(), GPW = 64 / G;
  typedef coop::Slice<NS, KIN> SL;
  const coop::Group<G> g;
  const int gi = (int)(threadIdx.x & 63) / G;
  double *L = lds + gi * SL::size;
  const long long i = (long long)blockIdx.x * GPW + gi;
  if (i >= b.n) return;  // uniform within the group
  const int l = g.l, ns = b.ns;
  const bool sys = l < ns;
  double f[9], fp[9], p[9], t[9], fe[9], fpo[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) { f[k] = b.F[9 * i + k]; fp[k] = b.Fpc[9 * i + k]; }
  const double s_a = sys ? b.sc[(long long)ns * i + l] : 0.0;
  double W_a = 0.0;
  if constexpr (KIN)
    if (sys) W_a = b.Wio[(long long)ns * i + l];
  const double *gt = b.gtab + (long long)b.grain[i] * b.gstride;
  bool plastic = false;
  const int r = coop::qp_update<NS, G, double, KIN>(g, L, b.mat, gt, ns, b.msi, f, fp, s_a, W_a, b.flag, p, t, fe, fpo,
                                                    b.flag ? b.D + 81 * i : nullptr, 1, 1.0, plastic, nullptr, b.nslip);
  if (l < 9) {
    b.P[9 * i + l] = coop::pick<9>(p, l);
    b.T[9 * i + l] = coop::pick<9>(t, l);
    b.Fe[9 * i + l] = coop::pick<9>(fe, l);
    b.Fp[9 * i + l] = coop::pick<9>(fpo, l);
  }
  if (sys) {
    b.so[(long long)ns * i + l] = L[SL::s + l];
    b.xo[(long long)ns * i + l] = L[SL::x + l];
    if constexpr (KIN) b.Wio[(long long)ns * i + l] = L[SL::W + l];
  }
  if (l == 0) b.na[i] = r;
}

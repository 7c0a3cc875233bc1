// Per-quadrature-point rate-independent crystal-plasticity update and consistent
// tangent (FP64), device side.  Restates calculatePlasticity(cellID, q, flag) of
// src/materialModels/crystalPlasticity/calculatePlasticity.cc:13-959 for one qp
// held by one lane:
//   * per-grain constants (rotated Schmid tensors, stiffness in Voigt form with
//     doubled shear columns, 9x9 TM) come from a table built once per grain
//     (the reference rebuilds them on every call from the grain's orientation,
//     :39-112; the values are the same);
//   * only the n_slip real systems are carried.  The inert twin system the
//     reference appends when twinning is off (s0 = 1e15, zero latent-hardening
//     row/column, userInputParameters.cc:300-311, init.cc:178-185) can never
//     enter the potentially-active set (b = |tau| - 1e15 < 0), so it changes
//     nothing and is dropped;
//   * Max Slip Search Iterations is 1, so the tangent's "previous" terms
//     delFp_delF_prev and dels_delF_prev are zero (:147-148, :476-477) and
//     dels_delF stays zero (:579); Max Solver Iterations may be > 1;
//   * kinematic hardening (C_1, C_2 and the backstress W) is carried by the KIN
//     instantiation of qp_update; without it the backstress is zero.
#pragma once
#include <hip/hip_runtime.h>

namespace cp {

#define CP_DEV __host__ __device__ __forceinline__

// One lane's array in a workspace interleaved over s lanes: element k at p[k * s].  Consecutive
// lanes hold consecutive addresses, so a wavefront's accesses to the same element of its lanes'
// arrays coalesce into one contiguous span.
struct LanePtr {
  double *p;
  long long s;
  CP_DEV double &operator[](long long k) const { return p[k * s]; }
  CP_DEV LanePtr operator+(long long k) const { return LanePtr{p + k * s, s}; }
};

// ---------------------------------------------------------------- 3x3 helpers
CP_DEV void m3_mul(const double *A, const double *B, double *C) {  // C = A B
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
CP_DEV void m3_tmul(const double *A, const double *B, double *C) {  // C = A^T B
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
CP_DEV void m3_mult(const double *A, const double *B, double *C) {  // C = A B^T
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}
CP_DEV double m3_det(const double *A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}
CP_DEV void m3_inv(const double *A, double *B) {
  const double r = 1.0 / m3_det(A);
  B[0] = (A[4] * A[8] - A[5] * A[7]) * r;
  B[1] = (A[2] * A[7] - A[1] * A[8]) * r;
  B[2] = (A[1] * A[5] - A[2] * A[4]) * r;
  B[3] = (A[5] * A[6] - A[3] * A[8]) * r;
  B[4] = (A[0] * A[8] - A[2] * A[6]) * r;
  B[5] = (A[2] * A[3] - A[0] * A[5]) * r;
  B[6] = (A[3] * A[7] - A[4] * A[6]) * r;
  B[7] = (A[1] * A[6] - A[0] * A[7]) * r;
  B[8] = (A[0] * A[4] - A[1] * A[3]) * r;
}
CP_DEV double m3_fro(const double *A) {
  double s = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) s += A[i] * A[i];
  return sqrt(s);
}
CP_DEV double m3_ddot(const double *A, const double *B) {
  double s = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) s += A[i] * B[i];
  return s;
}
CP_DEV void m3_symm(double *A) {
  double t = 0.5 * (A[1] + A[3]); A[1] = A[3] = t;
  t = 0.5 * (A[2] + A[6]); A[2] = A[6] = t;
  t = 0.5 * (A[5] + A[7]); A[5] = A[7] = t;
}
// T = matform(Dmat . vecform(E)), Voigt order (11,22,33,23,13,12)
// (src/materialModels/crystalPlasticity/matrixOperations.cc:133-162)
CP_DEV void dmat_apply(const double *__restrict__ D, const double *E, double *T) {
  const double v[6] = {E[0], E[4], E[8], E[5], E[2], E[1]};
  double t[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) s += D[6 * i + j] * v[j];
    t[i] = s;
  }
  T[0] = t[0]; T[4] = t[1]; T[8] = t[2];
  T[5] = T[7] = t[3]; T[2] = T[6] = t[4]; T[1] = T[3] = t[5];
}
// Taylor series exp with the reference's stop rule ||term||_F <= 1e-10 (matrixOperations.cc:165-186)
CP_DEV void m3_expm(const double *A, double *E) {
  double t[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t2[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) E[i] = t[i];
  double count = 1;
  for (int it = 0; it < 64 && m3_fro(t) > 1.0e-10; ++it) {
    m3_mul(t, A, t2);
    const double r = 1 / count;
#pragma unroll
    for (int i = 0; i < 9; ++i) { t[i] = t2[i] * r; E[i] += t[i]; }
    count += 1.0;
  }
}

// ------------------------------------------------------- small dense solvers
// Relative level below which a pivot (LU) or singular value (SVD fallback) counts as small.
// Near-singular A_PA are common: symmetric multi-slip leaves one singular value at 1e-9..1e-13
// sigma_max (inverted, as the reference does), and ten or more potentially active FCC systems make
// A_PA rank-deficient (latent-hardening rank 4 + elastic coupling rank <= 6: a null space of
// dimension >= 2 at round-off, dropped).
#define CP_PINV_RTOL 1e-10

// LU with partial pivoting on an n x n row-major matrix (ld = NS) with m right-hand
// sides (row-major, ld = m).  Returns false when a pivot is <= CP_PINV_RTOL * max|A|.
template <int NS, typename PA, typename PB>
CP_DEV bool lu_solve(int n, PA A, PB B, int m) {
  double amax = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) amax = fmax(amax, fabs(A[i * NS + j]));
  const double tiny = CP_PINV_RTOL * amax;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double pv = fabs(A[k * NS + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(A[i * NS + k]);
      if (v > pv) { pv = v; p = i; }
    }
    if (!(pv > tiny)) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) { const double t = A[k * NS + j]; A[k * NS + j] = A[p * NS + j]; A[p * NS + j] = t; }
      for (int j = 0; j < m; ++j) { const double t = B[k * m + j]; B[k * m + j] = B[p * m + j]; B[p * m + j] = t; }
    }
    const double inv = 1.0 / A[k * NS + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * NS + k] * inv;
      if (f == 0.0) continue;
      for (int j = k + 1; j < n; ++j) A[i * NS + j] -= f * A[k * NS + j];
      for (int j = 0; j < m; ++j) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    const double inv = 1.0 / A[k * NS + k];
    for (int j = 0; j < m; ++j) {
      double s = B[k * m + j];
      for (int c = k + 1; c < n; ++c) s -= A[k * NS + c] * B[c * m + j];
      B[k * m + j] = s * inv;
    }
  }
  return true;
}

// Pair k of round r of the circle (round-robin) schedule on m = n + (n odd) columns: column m-1
// stays, the others turn; every pair meets once in the m-1 rounds of a sweep, and the m/2 pairs of
// a round share no column, so they may rotate in any order or at once with the same result.
// Returns p < q; q = n is the padding column of odd n (no rotation).
CP_DEV void rr_pair(int m, int r, int k, int &p, int &q) {
  // (r + k) mod (m - 1) and (r - k) mod (m - 1); both sums lie in [0, 2 (m - 1))
  int a = r + k, b = r - k + m - 1;
  if (a >= m - 1) a -= m - 1;
  if (b >= m - 1) b -= m - 1;
  if (k == 0) { a = r; b = m - 1; }
  p = a < b ? a : b;
  q = a < b ? b : a;
}

// x = pinv(A) b by one-sided Jacobi SVD (round-robin sweeps); singular values <= CP_PINV_RTOL *
// sigma_max are dropped unless there is exactly one and its step is determined by the data
// (below).  Used only when LU finds A_PA numerically singular.
// U, V: n x n work matrices (leading dimension NS) of the caller.
template <int NS, typename PA, typename PU>
__host__ __device__ __forceinline__ void svd_pinv_core(int n, PA A, const double *b, double *x, PU U, PU V) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) { U[i * NS + j] = A[i * NS + j]; V[i * NS + j] = (i == j) ? 1.0 : 0.0; }
  const int m = n + (n & 1);
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int r = 0; r < m - 1; ++r)
      for (int k = 0; k < m / 2; ++k) {
        int p, q;
        rr_pair(m, r, k, p, q);
        if (q >= n) continue;
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < n; ++i) {
          const double up = U[i * NS + p], uq = U[i * NS + q];
          al += up * up; be += uq * uq; ga += up * uq;
        }
        if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
        rotated = true;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < n; ++i) {
          const double up = U[i * NS + p], uq = U[i * NS + q];
          U[i * NS + p] = c * up - s * uq; U[i * NS + q] = s * up + c * uq;
          const double vp = V[i * NS + p], vq = V[i * NS + q];
          V[i * NS + p] = c * vp - s * vq; V[i * NS + q] = s * vp + c * vq;
        }
      }
    if (!rotated) break;
  }
  double sig[NS], smax = 0;
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += U[i * NS + j] * U[i * NS + j];
    sig[j] = sqrt(s);
    if (sig[j] > 0)
      for (int i = 0; i < n; ++i) U[i * NS + j] /= sig[j];
    smax = fmax(smax, sig[j]);
  }
  // A single small singular value is inverted, as compute_inverse_svd(0.0) does
  // (inactiveSlipRemoval.cc:26,99), when the step it produces is fixed by the data: the step along
  // v_k is (u_k.b)/sigma_k and the round-off it spreads into the other components is about
  // |u_k.b| eps sigma_max / (sigma_k gap); the pair is kept when that is under 1e-2 |b| / sigma_max.
  // Otherwise (two or more small values, or a round-off one carrying a large u_k.b) the small ones
  // are dropped: the reference's result there depends on round-off.
  double lim = smax * CP_PINV_RTOL;
  int nsmall = 0, ks = -1;
  for (int k = 0; k < n; ++k)
    if (sig[k] <= lim) { nsmall++; ks = k; }
  if (nsmall == 1 && sig[ks] > 0) {
    double s2 = smax, ub = 0, bb = 0;
    for (int k = 0; k < n; ++k)
      if (k != ks) s2 = fmin(s2, sig[k]);
    for (int i = 0; i < n; ++i) { ub += U[i * NS + ks] * b[i]; bb += b[i] * b[i]; }
    const double eps = 2.220446049250313e-16;
    if (fabs(ub) * eps * smax * smax < 1e-2 * sqrt(bb) * sig[ks] * (s2 - sig[ks])) lim = 0.0;
  }
  for (int i = 0; i < n; ++i) x[i] = 0;
  for (int k = 0; k < n; ++k) {
    if (!(sig[k] > lim)) continue;
    double ub = 0;
    for (int i = 0; i < n; ++i) ub += U[i * NS + k] * b[i];
    ub /= sig[k];
    for (int i = 0; i < n; ++i) x[i] += V[i * NS + k] * ub;
  }
}
template <int NS>
__host__ __device__ __attribute__((noinline)) void svd_pinv_solve(int n, const double *A, const double *b, double *x) {
  double U[NS * NS], V[NS * NS];
  svd_pinv_core<NS, const double *, double *>(n, A, b, x, U, V);
}
// The same with U and V in the caller's workspace (2 NS^2 doubles): for the 48-system variant,
// whose matrices do not fit the private segment the runtime grants a lane.
template <int NS>
__host__ __device__ __attribute__((noinline)) void svd_pinv_solve_ws(int n, LanePtr A, const double *b, double *x, LanePtr UV) {
  svd_pinv_core<NS, LanePtr, LanePtr>(n, A, b, x, UV, UV + NS * NS);
}

// What qp_update does when LU finds A_PA numerically singular: the functor receives A_PA
// (n x n, leading dimension NS, intact) and b_PA and either writes x = pinv(A_PA) b_PA and
// returns true, or returns false, which makes qp_update return -1 at once.
template <int NS>
struct SvdInLane {  // the pseudo-inverse in this lane
  LanePtr uv = {nullptr, 1};  // U and V in the lane's workspace (2 NS^2 doubles) when lane_ws<NS>()
  CP_DEV bool operator()(int n, const double *A, const double *b, double *x) const {
    svd_pinv_solve<NS>(n, A, b, x);
    return true;
  }
  CP_DEV bool operator()(int n, LanePtr A, const double *b, double *x) const {
    svd_pinv_solve_ws<NS>(n, A, b, x, uv);
    return true;
  }
};
struct LuOnly {  // give up: the caller handles the qp elsewhere
  template <typename PA>
  CP_DEV bool operator()(int, PA, const double *, double *) const { return false; }
};

// Per-lane workspace of the 48-system variant (doubles): its NS x NS matrices take 4 NS^2 doubles
// (74 KB), more private memory than the runtime grants a lane, so they live in global memory the
// caller shares among its lanes (LanePtr, lane-interleaved): A over the potentially-active set,
// the compacted copy A_PA (then A2 of the tangent), U and V of the pseudo-inverse, and the
// tangent's delgamma and S_PA, at these offsets of each lane's array.
template <int NS> struct LaneWs {
  static constexpr int A = 0, Ap = NS * NS, UV = 2 * NS * NS, rhs = 4 * NS * NS, SV = 4 * NS * NS + 9 * NS,
                       size = 4 * NS * NS + 18 * NS;
};
template <int NS> constexpr bool lane_ws() { return NS > 24; }

// ------------------------------------------------------------ table layouts
// material block (doubles): h0[NS] ss[NS] pw[NS] q[NS*NS] tol
template <int NS> struct MatLayout {
  static constexpr int h0 = 0, ss = NS, pw = 2 * NS, q = 3 * NS, tol = 3 * NS + NS * NS, c1 = tol + 1, c2 = c1 + NS,
                       size = c2 + NS;
};
// grain block (doubles): S[NS][9] Dmat[36] TM[81]
template <int NS> struct GrainLayout {
  static constexpr int S = 0, D = 9 * NS, TM = 9 * NS + 36, R = 9 * NS + 36 + 81, size = 9 * NS + 36 + 81 + 9;
};

// Warm-start inputs of one qp (no counterpart in the reference): feat[0..nf) = (tau_a - W_a)/s_a,
// the trial resolved shear of :142-240 relative to the slip resistance, and feat[nf..nf+6) = the
// increment of the elastic Green-Lagrange strain E(F Fp_t^-1) - E(Fe_conv) rotated to the lattice
// frame (R^T dE R), Voigt order (11,22,33,23,13,12).  Written with stride fstride.
template <int NS>
CP_DEV void qp_features(const double *__restrict__ gt, int nf, const double *F, const double *Fp_t, const double *Fe_c,
                        const double *s, const double *W, float *feat, long long fstride) {
  typedef GrainLayout<NS> GL;
  double Fpinv[9], FE[9], CE[9], E[9], Tst[9], tmp[9];
  m3_inv(Fp_t, Fpinv);
  m3_mul(F, Fpinv, FE);
  m3_tmul(FE, FE, CE);
#pragma unroll
  for (int i = 0; i < 9; ++i) E[i] = 0.5 * (CE[i] - ((i % 4 == 0) ? 1.0 : 0.0));
  dmat_apply(gt + GL::D, E, Tst);
  m3_mul(CE, Tst, tmp);
  for (int a = 0; a < nf; ++a) {
    const double t = m3_ddot(tmp, gt + GL::S + 9 * a) - (W ? W[a] : 0.0);
    feat[a * fstride] = (float)(t / s[a]);
  }
  double dE[9], C0[9];
  m3_tmul(Fe_c, Fe_c, C0);
#pragma unroll
  for (int i = 0; i < 9; ++i) dE[i] = E[i] - 0.5 * (C0[i] - ((i % 4 == 0) ? 1.0 : 0.0));
  const double *R = gt + GL::R;
  m3_tmul(R, dE, tmp);   // R^T dE
  m3_mul(tmp, R, CE);    // R^T dE R
  const int vi[6] = {0, 4, 8, 5, 2, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) feat[(nf + k) * fstride] = (float)CE[vi[k]];
}

// dexp series of calculatePlasticity.cc:645-676 (sign = -1) and :748-779 (sign = +1):
// out = S + sum_{k=1..count3} c_k W_k, W_k = sum_{l=0..k} X^l S X^(k-l), c_k = (+-1)^k / (k+1)!.
// W_k is formed by the recurrence W_k = X W_{k-1} + S X^k (W_0 = S), so only X^k and W_k are kept.
CP_DEV void dexp_series(const double *X, const double *S, int count3, bool neg, double *out) {
  double W[9], Xk[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t1[9], t2[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { out[i] = S[i]; W[i] = S[i]; }
  double fact = 1.0;  // (k+1)!
  for (int k = 1; k <= count3; ++k) {
    fact *= (double)(k + 1);
    m3_mul(Xk, X, t1);
#pragma unroll
    for (int i = 0; i < 9; ++i) Xk[i] = t1[i];
    m3_mul(X, W, t1);
    m3_mul(S, Xk, t2);
    const double c = ((neg && (k & 1)) ? -1.0 : 1.0) / fact;
#pragma unroll
    for (int i = 0; i < 9; ++i) { W[i] = t1[i] + t2[i]; out[i] += c * W[i]; }
  }
}

// nslip: systems a >= nslip are twin systems (twinning), whose resolved shear below the backstress
// is replaced by the backstress (they shear in one sense only, calculatePlasticity.cc:213-216,
// 390-393); the default (no twin systems) leaves every system two-sided.
// Return: number of systems left active after inactive-slip removal (0 = elastic), or -1 when
// the singular-solve functor declined a numerically singular A_PA (nothing was written then).
// s holds s_conv on entry and s_tau on exit; x receives x_beta_old (the slip increments).
// *plastic_path (optional) is set when the potentially-active set was non-empty, the case in
// which the reference writes slipfraction_iter (calculatePlasticity.cc:454-456).
// D (flag = 1) receives dscale * dP/dF with element stride `dstride`:
// D[((3i+k)*9 + 3j+l) * dstride] = dscale * dP_ik / dF_jl, stored as DT.
// KIN: kinematic hardening (calculatePlasticity.cc:294-305, 431-436, with C_1, C_2 of
// multiphaseInit.cc:15-32): W holds the backstress W_kh_conv on entry and W_kh_tau on exit, and
// every sign test on the resolved shear uses tau - W.  Without it the backstress is zero.
// WARM (no counterpart in the reference, which starts from x_beta_old = 0, :311): x0 (ns predicted
// slip increments, stride x0stride; may be null) starts the local loop on the potentially-active
// systems it predicts positive, applied as one pass of the loop body before the first solve.  The
// start stands only if the loop converges on that set without dropping a system; otherwise the
// loop restarts from zero on the trial set.  iters (may be null) receives the local solves of both
// attempts.  Not combined with kinematic hardening.
// ws: this lane's LaneWs<NS> array when lane_ws<NS>() (NS = 48), unused otherwise.
template <int NS, typename DT = double, bool KIN = false, typename SING = SvdInLane<NS>, bool WARM = false>
CP_DEV int qp_update(const double *__restrict__ mat, const double *__restrict__ gt, int ns, int max_solver_iter,
                     const double *F, const double *Fp_t, double *s, double *x, int flag, double *P, double *T,
                     double *Fe_out, double *Fp_out, DT *D, long long dstride, bool *plastic_path = nullptr,
                     double dscale = 1.0, SING sing = SING(), double *W = nullptr, const float *x0 = nullptr,
                     long long x0stride = 1, int *iters = nullptr, int nslip = 1 << 30, LanePtr ws = LanePtr{nullptr, 1}) {
  static_assert(!(WARM && KIN), "the warm start does not cover kinematic hardening");
  constexpr bool WSP = lane_ws<NS>();
  typedef LaneWs<NS> LW;
  // a matrix in the workspace (48 systems) or in this lane's private memory
  auto wsa = [&](double *priv, int off) {
    if constexpr (WSP) return ws + off;
    else return priv;
  };
  typedef MatLayout<NS> ML;
  typedef GrainLayout<NS> GL;
  const double tol1 = mat[ML::tol];
  const double *__restrict__ Dm = gt + GL::D;
  const double *__restrict__ TM = gt + GL::TM;
  const double *__restrict__ Sg = gt + GL::S;

  double Fpinv[9], FE[9], CE[9], E[9], Tst[9], tmp[9], tmp2[9];
  m3_inv(Fp_t, Fpinv);                                   // :142, :160-161
  m3_mul(F, Fpinv, FE);                                  // :143, :164
  m3_tmul(FE, FE, CE);                                   // :168
#pragma unroll
  for (int i = 0; i < 9; ++i) E[i] = 0.5 * (CE[i] - ((i % 4 == 0) ? 1.0 : 0.0));
  dmat_apply(Dm, E, Tst);                                // :184-186
  double Fp[9], Ce[9], TstTau[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { Fp[i] = Fp_t[i]; Ce[i] = CE[i]; TstTau[i] = Tst[i]; }

  double tau_tr[NS], b[NS];
  // backstress at the start of the increment (W_kh_t) and its running value (W_kh_tau)
  double Wt[KIN ? NS : 1];
  auto wtau = [&](int a) -> double { if constexpr (KIN) return W[a]; else return 0.0; };
  auto wt = [&](int a) -> double { if constexpr (KIN) return Wt[a]; else return 0.0; };
  if constexpr (KIN)
    for (int a = 0; a < ns; ++a) Wt[a] = W[a];
  int PA[NS];
  int nPA = 0;
  m3_mul(CE, Tst, tmp);                                  // :203-240
  for (int a = 0; a < ns; ++a) {
    double r = m3_ddot(tmp, Sg + 9 * a);
    if (a >= nslip && (r - wtau(a)) < 0) r = wtau(a);  // twin systems shear one way only (:213-216)
    tau_tr[a] = r;
    b[a] = fabs(r - wtau(a)) - s[a];
    x[a] = 0.0;
    if (b[a] >= tol1) PA[nPA++] = a;
  }
  const bool plastic = nPA > 0;
  if constexpr (WARM) if (iters) *iters = 0;
  if (plastic_path) *plastic_path = plastic;
  double detFE;
  int n_active = 0;

  // ---- plastic correction (:250-468)
  double A_l[WSP ? 1 : NS * NS];
  const auto A = wsa(A_l, LW::A);  // A over the initial potentially-active set, indexed by positions in PA0
  int PA0[NS];
  int n0 = nPA;
  double h[NS];
  double XdelFp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (plastic) {
    for (int i = 0; i < n0; ++i) PA0[i] = PA[i];
    for (int a = 0; a < ns; ++a) h[a] = mat[ML::h0 + a] * pow((1 - s[a] / mat[ML::ss + a]), mat[ML::pw + a]);
    for (int jj = 0; jj < n0; ++jj) {                   // :273-306 restricted to PA x PA
      const int j = PA0[jj];
      double t2[9], t1[9], t3[9];
      m3_mul(CE, Sg + 9 * j, t2);
      m3_symm(t2);
      dmat_apply(Dm, t2, t3);
      m3_mul(CE, t3, t1);
      m3_mul(t2, Tst, t3);
#pragma unroll
      for (int k = 0; k < 9; ++k) t1[k] += 2.0 * t3[k];
      for (int ii = 0; ii < n0; ++ii) {
        const int i = PA0[ii];
        const double d = m3_ddot(Sg + 9 * i, t1);
        const bool opp = ((tau_tr[i] - wtau(i)) * (tau_tr[j] - wtau(j))) < 0.0;
        double v = mat[ML::q + i * NS + j] * h[j] + (opp ? -d : d);
        if constexpr (KIN) {  // :302-305, added to every column of row i
          const double k = mat[ML::c1 + i] - mat[ML::c2 + i] * W[i];
          v += ((tau_tr[i] - W[i]) < 0.0) ? -k : k;
        }
        A[ii * NS + jj] = v;
      }
    }
    // PA as positions into PA0 from here on
    int pos[NS];
    for (int i = 0; i < n0; ++i) pos[i] = i;
    int npos = n0;
    auto linf_b = [&]() {
      double m = 0;
      for (int i = 0; i < npos; ++i) m = fmax(m, fabs(b[PA0[pos[i]]]));
      return m;
    };
    int count1 = 0, n_solves = 0;
    double xb[NS];
    // one solve of inactive_slip_removal on the current set; false: a singular system that the
    // caller's SING policy defers
    auto solve_step = [&]() -> bool {
      // inactive_slip_removal (inactiveSlipRemoval.cc:4-124)
      for (int a = 0; a < ns; ++a) xb[a] = 0.0;
      while (true) {
        double Ap_l[WSP ? 1 : NS * NS], bp[NS];
        const auto Ap = wsa(Ap_l, LW::Ap);
        for (int i = 0; i < npos; ++i) {
          bp[i] = b[PA0[pos[i]]];
          for (int j = 0; j < npos; ++j) Ap[i * NS + j] = A[pos[i] * NS + pos[j]];
        }
        double xs[NS];
        for (int i = 0; i < npos; ++i) xs[i] = bp[i];
        if (npos > 0 && !lu_solve<NS>(npos, Ap, xs, 1)) {
          for (int i = 0; i < npos; ++i)
            for (int j = 0; j < npos; ++j) Ap[i * NS + j] = A[pos[i] * NS + pos[j]];
          if (!sing(npos, Ap, bp, xs)) return false;  // the caller solves this system elsewhere
        }
        for (int a = 0; a < ns; ++a) xb[a] = 0.0;
        for (int i = 0; i < npos; ++i) xb[PA0[pos[i]]] = xs[i];
        int keep = 0;
        bool removed = false;
        for (int i = 0; i < npos; ++i) {
          const int a = PA0[pos[i]];
          if ((x[a] + xb[a]) < 0) removed = true; else pos[keep++] = pos[i];
        }
        if (!removed) break;
        npos = keep;
      }
      for (int a = 0; a < ns; ++a) x[a] += xb[a];        // :119
      return true;
    };
    // one pass of the loop body from xb (increment) and x (total)
    auto update_step = [&]() {
      for (int i = 0; i < 9; ++i) XdelFp[i] = 0.0;      // :338-357
      for (int a = 0; a < ns; ++a) {
        const double g = ((tau_tr[a] - wtau(a)) > 0) ? x[a] : -x[a];
#pragma unroll
        for (int k = 0; k < 9; ++k) XdelFp[k] += g * Sg[9 * a + k];
      }
      double ex[9];
      m3_expm(XdelFp, ex);
      m3_mul(ex, Fp_t, Fp);                              // :359
      m3_inv(Fp, tmp);                                   // :362-373
      m3_mul(F, tmp, FE);
      m3_tmul(FE, FE, Ce);
#pragma unroll
      for (int i = 0; i < 9; ++i) E[i] = 0.5 * (Ce[i] - ((i % 4 == 0) ? 1.0 : 0.0));
      dmat_apply(Dm, E, TstTau);                         // :376-377
      m3_mul(Ce, TstTau, tmp);
      for (int a = 0; a < ns; ++a) {                     // :380-440
        double ta = m3_ddot(tmp, Sg + 9 * a);
        if (a >= nslip && (tau_tr[a] - wtau(a)) < 0) ta = wtau(a);  // :390-393, the backstress before :431
        double h1 = 0;
        for (int c = 0; c < ns; ++c) h1 += mat[ML::q + a * NS + c] * h[c] * xb[c];
        double sa = s[a] + h1;
        if (sa > mat[ML::ss + a]) sa = mat[ML::ss + a];
        s[a] = sa;
        if constexpr (KIN) {  // :431-436
          const double c1 = mat[ML::c1 + a], c2 = mat[ML::c2 + a];
          W[a] += (((tau_tr[a] - W[a]) > 0) ? (c1 - c2 * W[a]) : (-c1 - c2 * W[a])) * xb[a];
        }
        b[a] = fabs(ta - wtau(a)) - sa;
      }
    };
    // the reference's loop from `done` solves already counted; stop_on_drop: stop as soon as
    // inactive-slip removal drops a system.  Returns false when a solve is deferred.
    auto local_loop = [&](int done, bool stop_on_drop) -> bool {
      count1 = done;
      while (linf_b() > tol1) {                          // :322
        count1++;
        if (count1 > max_solver_iter) break;
        const int n_before = npos;
        if (!solve_step()) return false;
        update_step();
        if (stop_on_drop && npos < n_before) break;
      }
      return true;
    };
    auto solves = [&](int done) { return (count1 < max_solver_iter ? count1 : max_solver_iter) - done; };
    bool plain = true;
    if constexpr (WARM) if (x0) {
      // Warm start.  The reference's loop is path dependent: its first solve on the trial set,
      // with inactive-slip removal, fixes the set S1 that later iterations only shrink, and its
      // converged states leave dropped systems above the yield surface, so a converged start on
      // another set is another answer.  The first solve always runs; if the predicted set
      // {a : x0[a] > 0} equals S1 its increment is replaced by x0 and the loop continues, the start
      // standing if the loop converges on S1 without a drop (otherwise it restarts from zero); if
      // the sets differ the loop continues from the first solve as without a prediction.
      bool pred[NS];
      int npred = 0;
      for (int i = 0; i < n0; ++i) {
        pred[i] = (double)x0[PA0[i] * x0stride] > 0.0;
        npred += pred[i];
      }
      if (npred > 0) {
        plain = false;
        if (linf_b() > tol1 && max_solver_iter >= 1) {
          double s_t[NS], b_t[NS];
          for (int a = 0; a < ns; ++a) { s_t[a] = s[a]; b_t[a] = b[a]; }
          if (!solve_step()) return -1;
          n_solves = 1;
          bool same = npos == npred;
          for (int i = 0; i < npos; ++i) same = same && pred[pos[i]];
          if (same) {
            for (int a = 0; a < ns; ++a) xb[a] = 0.0;
            for (int i = 0; i < npos; ++i) { const int a = PA0[pos[i]]; xb[a] = (double)x0[a * x0stride]; }
            for (int a = 0; a < ns; ++a) x[a] = xb[a];
          }
          update_step();
          if (same) {
            const int n_set = npos;
            if (!local_loop(0, true)) return -1;
            const int c = solves(0);
            n_solves += c;
            bool ok = npos == n_set && c < max_solver_iter;
            for (int i = 0; i < npos; ++i)
              if (!(fabs(b[PA0[pos[i]]]) <= tol1)) ok = false;
            if (!ok) {  // restart from zero on the trial set
              for (int i = 0; i < n0; ++i) pos[i] = i;
              npos = n0;
              for (int a = 0; a < ns; ++a) { s[a] = s_t[a]; b[a] = b_t[a]; x[a] = 0.0; }
#pragma unroll
              for (int i = 0; i < 9; ++i) { Fp[i] = Fp_t[i]; Ce[i] = CE[i]; TstTau[i] = Tst[i]; }
              if (!local_loop(0, false)) return -1;
              n_solves += solves(0);
            }
          } else {
            if (!local_loop(1, false)) return -1;
            n_solves += solves(1);
          }
        }
      }
    }
    if (plain) {
      if (!local_loop(0, false)) return -1;
      n_solves = solves(0);
    }
    if constexpr (WARM)
      if (iters) *iters = n_solves;
    n_active = npos;
    // final active set as system indices
    for (int i = 0; i < npos; ++i) PA[i] = PA0[pos[i]];
    nPA = npos;
    m3_inv(Fp, Fpinv);                                   // :458-460
    m3_mul(F, Fpinv, FE);
  }

  // ---- stresses (:188-196 elastic, :458-468 plastic)
  detFE = m3_det(FE);
  m3_mul(FE, TstTau, tmp);
  {
    const double r = 1.0 / detFE;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] *= r;
  }
  m3_mult(tmp, FE, T);
  double Finv[9];
  m3_inv(F, Finv);
  m3_mult(T, Finv, tmp2);
  const double pf = plastic ? m3_det(F) : detFE;
#pragma unroll
  for (int i = 0; i < 9; ++i) P[i] = pf * tmp2[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) { Fe_out[i] = FE[i]; Fp_out[i] = Fp[i]; }
  if (!flag) return n_active;

  // ---- tangent: in-loop block (:473-828) with delFp_delF_prev = 0 and dels_delF = 0
  // delFp_delF = sum_j SV_j (x) delgamma_j (rank nPA): SV[j] (9) and delgamma (rhs, 9 per j) are
  // kept and each column kl of it is formed where the final tangent needs it
  double rhs_l[WSP ? 1 : NS * 9], SV_l[WSP ? 1 : NS * 9];
  const auto rhs = wsa(rhs_l, LW::rhs);
  const auto SV = wsa(SV_l, LW::SV);
  const int nsv = (plastic && nPA > 0) ? nPA : 0;
  if (nsv > 0) {
    // count3 (:607-619): number of series terms until ||X^k / k!||_F <= max(tol/1e4, 1e-12)
    int count3 = 0;
    {
      double t2[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t1[9];
      double tol2 = 1.0;
      const double lim = fmax(tol1 / 1e4, 1e-12);
      while (tol2 > lim && count3 < 64) {
        count3++;
        m3_mul(XdelFp, t2, t1);
        const double r = 1.0 / count3;
#pragma unroll
        for (int i = 0; i < 9; ++i) t1[i] *= r;
        tol2 = m3_fro(t1);
#pragma unroll
        for (int i = 0; i < 9; ++i) t2[i] = t1[i];
      }
    }
    // delb (:581-604): delb(k, kl) = +- S_k : deltau_delF(:, kl), streamed per column kl
    for (int kl = 0; kl < 9; ++kl) {
      const int k = kl / 3, l = kl % 3;
      double dE[9], dT[9], dtau[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dE[3 * i + j] = 0.5 * (Fpinv[3 * l + i] * FE[3 * k + j] + Fpinv[3 * l + j] * FE[3 * k + i]);
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        double sacc = 0;
#pragma unroll
        for (int c = 0; c < 9; ++c) sacc += TM[9 * r + c] * dE[c];
        dT[r] = sacc;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double sacc = 0;
#pragma unroll
          for (int a = 0; a < 3; ++a) sacc += 2 * dE[3 * i + a] * TstTau[3 * a + j] + Ce[3 * i + a] * dT[3 * a + j];
          dtau[3 * i + j] = sacc;
        }
      for (int kk = 0; kk < nPA; ++kk) {
        const int a = PA[kk];
        const double v = m3_ddot(dtau, Sg + 9 * a);
        rhs[kk * 9 + kl] = ((tau_tr[a] - wt(a)) < 0) ? -v : v;
      }
    }
    // A2 (:622-709)
    double A2_l[WSP ? 1 : NS * NS];
    const auto A2 = wsa(A2_l, LW::Ap);  // the compacted A_PA is no longer needed
    double eneg[9], Xneg[9], t6[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Xneg[i] = -XdelFp[i];
    m3_expm(Xneg, eneg);
    m3_mul(CE, eneg, t6);
    for (int jj = 0; jj < nPA; ++jj) {
      const int j = PA[jj];
      double diff[9], t2[9], t3[9], tm[9];
      dexp_series(XdelFp, Sg + 9 * j, count3, true, diff);
      m3_tmul(diff, t6, t2);
      m3_symm(t2);
      dmat_apply(Dm, t2, t3);
      m3_mul(Ce, t3, tm);
      m3_mul(t2, TstTau, t3);
#pragma unroll
      for (int k = 0; k < 9; ++k) tm[k] += 2.0 * t3[k];
      const bool nj = (tau_tr[j] - wt(j)) < 0.0;
      for (int ii = 0; ii < nPA; ++ii) {
        const int i = PA[ii];
        const double d = m3_ddot(Sg + 9 * i, tm);
        const bool ni = (tau_tr[i] - wt(i)) < 0.0;
        double v = mat[ML::q + i * NS + j] * h[j] + ((ni ^ nj) ? -d : d);
        if constexpr (KIN) {  // :700-703: sign from W_kh_t, value from W_kh_tau
          const double k = mat[ML::c1 + i] - mat[ML::c2 + i] * W[i];
          v += ni ? -k : k;
        }
        A2[ii * NS + jj] = v;
      }
    }
    // delgamma = A2^-1 (delb - dels_PA) with dels = 0 (:712-722)
    if (!lu_solve<NS>(nPA, A2, rhs, 9)) {
      for (int i = 0; i < nPA * 9; ++i) rhs[i] = 0.0;  // singular A2: no plastic tangent contribution
    }
    // S_PA and delFp_delF = S_PA delgamma (:734-821)
    for (int jj = 0; jj < nPA; ++jj) {
      const int j = PA[jj];
      double diff[9], t1[9];
      dexp_series(XdelFp, Sg + 9 * j, count3, false, diff);
      m3_mul(diff, Fp_t, t1);
      const double sg = ((tau_tr[j] - wt(j)) < 0) ? -1.0 : 1.0;
#pragma unroll
      for (int r = 0; r < 9; ++r) SV[jj * 9 + r] = sg * t1[r];
    }
  }

  // ---- final tangent (:835-940), streamed per column kl
  double q1[9], q2[9], q3[9], q5[9];
  m3_mul(Finv, FE, q2);                                  // temp = temp2 = F^-1 Fe
  m3_mult(TstTau, q2, q1);                               // T* (F^-1 Fe)^T
  m3_mul(FE, TstTau, q3);
  m3_mult(q3, F, q5);
  for (int kl = 0; kl < 9; ++kl) {
    const int k = kl / 3, l = kl % 3;
    double dFe[9], W[9], dTs[9], dFp[9];
    // column kl of delFp_delF, summed over the active systems in order
#pragma unroll
    for (int r = 0; r < 9; ++r) dFp[r] = 0.0;
    for (int jj = 0; jj < nsv; ++jj) {
      const double g = rhs[jj * 9 + kl];
#pragma unroll
      for (int r = 0; r < 9; ++r) dFp[r] += SV[jj * 9 + r] * g;
    }
    // dFe(ij,kl) = d_ik Fpinv(l,j) - sum_ab FE(i,a) dFp(ab,kl) Fpinv(b,j)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sacc = 0;
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) sacc += dFp[3 * a + bb] * Fpinv[3 * bb + j];
        W[3 * a + j] = sacc;  // (dFp . Fpinv)(a, j)
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sacc = (i == k) ? Fpinv[3 * l + j] : 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) sacc -= FE[3 * i + a] * W[3 * a + j];
        dFe[3 * i + j] = sacc;
      }
    // dTs(ij,kl) = sum_ab TM(ij,ab) (dFe^T FE)(a,b)
    double G[9];
    m3_tmul(dFe, FE, G);
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      double sacc = 0;
#pragma unroll
      for (int c = 0; c < 9; ++c) sacc += TM[9 * r + c] * G[c];
      dTs[r] = sacc;
    }
    // PK(ij,kl) = (dFe q1)(i,j) + (FE dTs q2^T)(i,j) - (q3 dFe Finv^T)(i,j) - sum_a q5(i,a) Finv(j,k) Finv(l,a)
    double t1[9], t2[9], t3[9], t4[9];
    m3_mul(dFe, q1, t1);
    m3_mul(FE, dTs, t4);
    m3_mult(t4, q2, t2);
    m3_mul(q3, dFe, t4);
    m3_mult(t4, Finv, t3);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double qa = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) qa += q5[3 * i + a] * Finv[3 * l + a];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        D[((3 * i + j) * 9 + kl) * dstride] = (DT)(dscale * (t1[3 * i + j] + t2[3 * i + j] - t3[3 * i + j] - qa * Finv[3 * j + k]));
    }
  }
  return n_active;
}

}  // namespace cp

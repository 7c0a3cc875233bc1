// CPU FP64 restatement of the PRISMS-Plasticity crystal-plasticity Newton step.
//
// TEST INFRASTRUCTURE ONLY (see cp_oracle.h): loaded by tests/, smoke() and the
// cpu_baseline leg of bench.py; never by the product.
//
// Reference files cited below are relative to the reference repository root:
//   CP  = src/materialModels/crystalPlasticity/
//   BVP = src/ellipticBVP/
#include "cp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>
#include <functional>
#include <map>
#ifdef _OPENMP
#include <omp.h>
#endif

// Hooks of the flop-counting build of this file (cp_oracle_flops.cpp); empty in the oracle itself.
#ifndef OR_FLOP_QP_BEGIN
#define OR_FLOP_QP_BEGIN(flag)
#define OR_FLOP_QP_END(flag)
#define OR_FLOP_STEP(inc, it, max_newton)
#define OR_FLOP_SINGULAR()
#endif

namespace {

// Test-only knob (or_tangent_perturbation): scales one block of the consistent tangent by
// (1 + g_pert_rel) so that a test can show that its checks see an error of that size there.
// Block ids: 1 delb, 2 the elastic part of A2, 3 the hardening part of A2, 4 the k >= 1 terms of
// the d exp series, 5 S_PA, 6 dT*/dF in the final tangent, 7 the plastic part of dFe/dF,
// 8 the q5 term of the final tangent; 9 (rel ignored) replaces the kinematic-hardening term of A2 by
// the exact linearisation of b_i = |tau_i - W_i| - s_i with dW_i = (sgn_i C_1 - C_2 W_i) dgamma_i,
// i.e. C_1 - sgn_i C_2 W_i on the diagonal, where the reference adds sgn_i (C_1 - C_2 W_i) to every
// column of the row (:700-703).  0 (the default) leaves the tangent as the reference forms it.
int g_pert_block = 0;
double g_pert_rel = 0.0;
inline double pert(int block) { return g_pert_block == block ? 1.0 + g_pert_rel : 1.0; }

// wall-clock split of the latest or_simulate (ns): assembly (qp passes, K_e, scatter), Krylov
// solves, post-increment passes; read by or_last_timings
long long g_phase_ns[3];
long long now_ns() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (long long)t.tv_sec * 1000000000LL + t.tv_nsec;
}

// ---------------------------------------------------------------- 3x3 algebra
struct M3 {
  double a[3][3];
  double *operator[](int i) { return a[i]; }
  const double *operator[](int i) const { return a[i]; }
};

M3 zero3() { M3 r; std::memset(r.a, 0, sizeof(r.a)); return r; }
M3 eye3() { M3 r = zero3(); r[0][0] = r[1][1] = r[2][2] = 1.0; return r; }
M3 load3(const double *p) { M3 r; for (int i = 0; i < 9; ++i) r.a[i / 3][i % 3] = p[i]; return r; }
void store3(const M3 &m, double *p) { for (int i = 0; i < 9; ++i) p[i] = m.a[i / 3][i % 3]; }

// deal.II FullMatrix::mmult (C = A B), Tmmult (C = A^T B), mTmult (C = A B^T)
M3 mul(const M3 &A, const M3 &B) {
  M3 C = zero3();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
  return C;
}
M3 tmul(const M3 &A, const M3 &B) {
  M3 C = zero3();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[k][i] * B[k][j];
      C[i][j] = s;
    }
  return C;
}
M3 mult(const M3 &A, const M3 &B) {
  M3 C = zero3();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i][k] * B[j][k];
      C[i][j] = s;
    }
  return C;
}
M3 scale(const M3 &A, double s) { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r[i][j] = s * A[i][j]; return r; }
void addto(M3 &A, double s, const M3 &B) { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) A[i][j] += s * B[i][j]; }
double det3(const M3 &A) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
M3 inv3(const M3 &A) {
  const double d = det3(A), r = 1.0 / d;
  M3 B;
  B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * r;
  B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * r;
  B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * r;
  B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * r;
  B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * r;
  B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * r;
  B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * r;
  B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * r;
  B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * r;
  return B;
}
double fro3(const M3 &A) { double s = 0; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s += A[i][j] * A[i][j]; return std::sqrt(s); }
M3 symm3(const M3 &A) {  // FullMatrix::symmetrize
  M3 r = A;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < i; ++j) { const double t = (A[i][j] + A[j][i]) / 2; r[i][j] = r[j][i] = t; }
  return r;
}

// vecform / matform, CP matrixOperations.cc:133-162 (Voigt order 11,22,33,23,13,12)
void vecform(const M3 &A, double v[6]) { v[0] = A[0][0]; v[1] = A[1][1]; v[2] = A[2][2]; v[3] = A[1][2]; v[4] = A[0][2]; v[5] = A[0][1]; }
M3 matform(const double v[6]) {
  M3 A; A[0][0] = v[0]; A[1][1] = v[1]; A[2][2] = v[2];
  A[1][2] = A[2][1] = v[3]; A[0][2] = A[2][0] = v[4]; A[0][1] = A[1][0] = v[5];
  return A;
}
M3 dmat_apply(const double D[6][6], const M3 &E) {  // matform(D . vecform(E))
  double v[6], t[6];
  vecform(E, v);
  for (int i = 0; i < 6; ++i) { double s = 0; for (int j = 0; j < 6; ++j) s += D[i][j] * v[j]; t[i] = s; }
  return matform(t);
}

// matrixExponential, CP matrixOperations.cc:165-186 (Taylor series, stop when ||term||_F <= 1e-10)
M3 expm3(const M3 &A) {
  M3 E = eye3(), t = eye3();
  double count = 1;
  while (fro3(t) > 1.0e-10) {
    M3 t2 = mul(t, A);
    t2 = scale(t2, 1 / count);
    t = t2;
    addto(E, 1.0, t);
    count += 1.0;
  }
  return E;
}

// odfpoint, CP rotationOperations.cc:4-51 (Rodrigues -> rotation)
M3 odfpoint(const double r[3]) {
  double rr = 0;
  for (int i = 0; i < 3; ++i) rr += r[i] * r[i];
  const double t1 = 1.0 - rr, t2 = 1.0 + rr;
  M3 R = eye3();
  for (int i = 0; i < 3; ++i) R[i][i] *= t1;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) R[i][j] += 2.0 * r[i] * r[j];
  R[0][1] -= 2.0 * r[2]; R[0][2] += 2.0 * r[1]; R[1][2] -= 2.0 * r[0];
  R[1][0] += 2.0 * r[2]; R[2][0] -= 2.0 * r[1]; R[2][1] += 2.0 * r[0];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) R[i][j] *= 1.0 / t2;
  return R;
}

// elasticmoduli, CP matrixOperations.cc:189-234: Ar = N C N^T with the 6x6 Bond matrix N(R)
void elasticmoduli(const M3 &R, const double C[6][6], double Ar[6][6]) {
  double N[6][6];
  const int pi[3][2] = {{1, 2}, {2, 0}, {0, 1}};  // (23),(31),(12)
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) N[i][j] = R[i][j] * R[i][j];
    for (int j = 0; j < 3; ++j) N[i][3 + j] = 2.0 * R[i][pi[j][0]] * R[i][pi[j][1]];
  }
  for (int i = 0; i < 3; ++i) {
    const int a = pi[i][0], b = pi[i][1];
    for (int j = 0; j < 3; ++j) N[3 + i][j] = R[a][j] * R[b][j];
    for (int j = 0; j < 3; ++j) {
      const int c = pi[j][0], d = pi[j][1];
      N[3 + i][3 + j] = R[a][c] * R[b][d] + R[a][d] * R[b][c];
    }
  }
  double T[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { double s = 0; for (int k = 0; k < 6; ++k) s += N[i][k] * C[k][j]; T[i][j] = s; }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { double s = 0; for (int k = 0; k < 6; ++k) s += T[i][k] * N[j][k]; Ar[i][j] = s; }
}

// ------------------------------------------------------ small dense solvers
// Pair k of round r of the circle (round-robin) schedule on m = n + (n odd) columns: column m-1
// stays, the others turn; every pair meets once in the m-1 rounds of a sweep, and the m/2 pairs of
// a round share no column.  Returns p < q; q = n is the padding column of odd n (no rotation).
void rr_pair(int m, int r, int k, int &p, int &q) {
  // (r + k) mod (m - 1) and (r - k) mod (m - 1); both sums lie in [0, 2 (m - 1))
  int a = r + k, b = r - k + m - 1;
  if (a >= m - 1) a -= m - 1;
  if (b >= m - 1) b -= m - 1;
  if (k == 0) { a = r; b = m - 1; }
  p = std::min(a, b);
  q = std::max(a, b);
}

// one-sided (Hestenes) Jacobi SVD of an n x n row-major matrix: A = U diag(sig) V^T, sweeps in
// round-robin order
void jacobi_svd(int n, const double *A, std::vector<double> &U, std::vector<double> &sig, std::vector<double> &V) {
  U.assign(A, A + n * n);
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  const int m = n + (n & 1);
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int r = 0; r < m - 1; ++r)
      for (int k = 0; k < m / 2; ++k) {
        int p, q;
        rr_pair(m, r, k, p, q);
        if (q >= n) continue;
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < n; ++i) { const double up = U[i * n + p], uq = U[i * n + q]; al += up * up; be += uq * uq; ga += up * uq; }
        if (ga == 0.0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
        rotated = true;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < n; ++i) {
          const double up = U[i * n + p], uq = U[i * n + q];
          U[i * n + p] = c * up - s * uq; U[i * n + q] = s * up + c * uq;
          const double vp = V[i * n + p], vq = V[i * n + q];
          V[i * n + p] = c * vp - s * vq; V[i * n + q] = s * vp + c * vq;
        }
      }
    if (!rotated) break;
  }
  sig.assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += U[i * n + j] * U[i * n + j];
    sig[j] = std::sqrt(s);
    if (sig[j] > 0) for (int i = 0; i < n; ++i) U[i * n + j] /= sig[j];
  }
}

// x = pinv(A) b over the nonzero singular values (rel = 0: compute_inverse_svd(0.0)).  With rel > 0,
// singular values <= rel*sigma_max are "small".  A single small singular value is inverted, as the
// reference does, when the result is determined by the data: the step along v_k is
// (u_k.b)/sigma_k, and the round-off it carries into the other components is about
// |u_k.b| eps sigma_max / (sigma_k gap), gap = distance to the next singular value; the pair is
// kept when that is under 1e-2 |b| / sigma_max (symmetric multi-slip: sigma_k ~ 1e-9..1e-16,
// margin >= 100x).  Otherwise (two or more small singular values, or a round-off one carrying a
// large u_k.b) the small ones are dropped: the reference's result there depends on round-off.
void svd_pinv_solve(int n, const double *A, const double *b, double *x, double rel) {
  std::vector<double> U, s, V;
  jacobi_svd(n, A, U, s, V);
  double smax = 0;
  for (int i = 0; i < n; ++i) smax = std::max(smax, s[i]);
  double lim = smax * rel;
  int nsmall = 0, ks = -1;
  for (int i = 0; i < n; ++i) if (s[i] <= lim) { nsmall++; ks = i; }
  if (nsmall == 1 && s[ks] > 0) {
    double s2 = smax, ub = 0, bb = 0;
    for (int i = 0; i < n; ++i) if (i != ks) s2 = std::min(s2, s[i]);
    for (int i = 0; i < n; ++i) { ub += U[i * n + ks] * b[i]; bb += b[i] * b[i]; }
    const double eps = 2.220446049250313e-16;
    if (std::fabs(ub) * eps * smax * smax < 1e-2 * std::sqrt(bb) * s[ks] * (s2 - s[ks])) lim = 0.0;
  }
  for (int i = 0; i < n; ++i) x[i] = 0;
  for (int k = 0; k < n; ++k) {
    if (!(s[k] > lim)) continue;
    double ub = 0;
    for (int i = 0; i < n; ++i) ub += U[i * n + k] * b[i];
    ub /= s[k];
    for (int i = 0; i < n; ++i) x[i] += V[i * n + k] * ub;
  }
}

// Relative level below which a pivot (LU) or singular value (SVD fallback) counts as small.
// Symmetric multi-slip leaves one singular value at 1e-9..1e-13 sigma_max (inverted, as the
// reference does); rank-deficient active sets show a null space of dimension >= 2 at round-off.
double kPinvRtol = 1e-10;

// LU with partial pivoting, multiple right-hand sides (row-major B, n x m, overwritten by X).
// Returns false if a pivot is <= kPinvRtol * max|A|.
bool lu_solve(int n, std::vector<double> A, double *B, int m) {
  double amax = 0;
  for (double v : A) amax = std::max(amax, std::fabs(v));
  const double tiny = kPinvRtol * amax;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double pv = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i) if (std::fabs(A[i * n + k]) > pv) { pv = std::fabs(A[i * n + k]); p = i; }
    if (!(pv > tiny)) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      for (int j = 0; j < m; ++j) std::swap(B[k * m + j], B[p * m + j]);
    }
    const double inv = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] * inv;
      if (f == 0.0) continue;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      for (int j = 0; j < m; ++j) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    const double inv = 1.0 / A[k * n + k];
    for (int j = 0; j < m; ++j) {
      double s = B[k * m + j];
      for (int c = k + 1; c < n; ++c) s -= A[k * n + c] * B[c * m + j];
      B[k * m + j] = s * inv;
    }
  }
  return true;
}

long long g_pinv_calls = 0, g_pinv_singular = 0;  // diagnostics (single-threaded runs only)
int g_sample_n = 0;                                 // first matrix that took the fallback
double g_sample_A[OR_MAXT * OR_MAXT], g_sample_b[OR_MAXT];
int g_diag = 0;                                     // histogram of sigma_min / sigma_max (decades)
long long g_hist[24];
std::vector<double> g_records;                      // n, s_min/s_max, s_2nd/s_max, |u_min.b|/|b|

void pinv_solve(int n, const double *A, const double *b, double *x, int mode) {
  if (n == 0) return;
  g_pinv_calls++;
  if (g_diag) {
    std::vector<double> U, sg, V;
    jacobi_svd(n, A, U, sg, V);
    int kmin = 0, k2 = -1;
    double hi = 0;
    for (int k = 0; k < n; ++k) { hi = std::max(hi, sg[k]); if (sg[k] < sg[kmin]) kmin = k; }
    for (int k = 0; k < n; ++k) if (k != kmin && (k2 < 0 || sg[k] < sg[k2])) k2 = k;
    const double r = hi > 0 ? sg[kmin] / hi : 0.0;
    int k = r > 0 ? (int)std::floor(-std::log10(r)) : 23;
    g_hist[std::min(std::max(k, 0), 23)]++;
    if (r < 1e-8 && g_records.size() < 4 * 200000) {
      double ub = 0, bb = 0;
      for (int i = 0; i < n; ++i) { ub += U[i * n + kmin] * b[i]; bb += b[i] * b[i]; }
      g_records.push_back(n);
      g_records.push_back(r);
      g_records.push_back(k2 >= 0 && hi > 0 ? sg[k2] / hi : 1.0);
      g_records.push_back(bb > 0 ? std::fabs(ub) / std::sqrt(bb) : 0.0);
    }
  }
  if (mode == 1) {
    std::vector<double> Av(A, A + n * n);
    for (int i = 0; i < n; ++i) x[i] = b[i];
    if (lu_solve(n, Av, x, 1)) return;
    OR_FLOP_SINGULAR();
    if (g_pinv_singular++ == 0) {
      g_sample_n = n;
      for (int i = 0; i < n * n; ++i) g_sample_A[i] = A[i];
      for (int i = 0; i < n; ++i) g_sample_b[i] = b[i];
    }
    svd_pinv_solve(n, A, b, x, kPinvRtol);
    return;
  }
  svd_pinv_solve(n, A, b, x, 0.0);
}

// inactive_slip_removal, CP inactiveSlipRemoval.cc:4-124
void inactive_slip_removal(int nT, std::vector<double> &x_old, std::vector<double> &x_beta, std::vector<int> &PA,
                           const std::vector<double> &b, const std::vector<double> &A, int mode) {
  auto solve_on = [&](std::vector<double> &x1) {
    const int n = (int)PA.size();
    std::vector<double> Ap(n * n), bp(n), xp(n);
    for (int i = 0; i < n; ++i) {
      bp[i] = b[PA[i]];
      for (int j = 0; j < n; ++j) Ap[i * n + j] = A[PA[i] * nT + PA[j]];
    }
    pinv_solve(n, Ap.data(), bp.data(), xp.data(), mode);
    x1.assign(nT, 0.0);
    for (int i = 0; i < n; ++i) x1[PA[i]] = xp[i];
  };
  std::vector<double> x1;
  solve_on(x1);
  while (true) {  // :44-115
    std::vector<int> row;
    for (int i = 0; i < nT; ++i) if ((x_old[i] + x1[i]) < 0) row.push_back(i);
    if (row.empty()) break;
    std::vector<int> PA_new;
    for (int p : PA) if (std::find(row.begin(), row.end(), p) == row.end()) PA_new.push_back(p);
    PA = PA_new;
    solve_on(x1);
  }
  x_beta = x1;                                   // :118
  for (int i = 0; i < nT; ++i) x_old[i] += x_beta[i];  // :119
}

// ------------------------------------------------------------ material state
struct Grain {
  M3 R;
  double Dmat[6][6];  // rotated stiffness with shear columns doubled (calculatePlasticity.cc:55-68)
  double TM[9][9];    // 9x9 re-indexing of the unscaled rotated stiffness (:70-77)
  std::vector<M3> S;  // Schmid tensors in the sample frame (:83-104)
};

void grain_setup(const or_material *mat, const double rot[3], Grain &g) {
  g.R = odfpoint(rot);
  double D2[6][6];
  elasticmoduli(g.R, mat->C, D2);
  for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) g.Dmat[i][j] = D2[i][j];
  for (int i = 0; i < 6; ++i) for (int j = 3; j < 6; ++j) g.Dmat[i][j] = 2 * g.Dmat[i][j];
  const int v2[9] = {0, 5, 4, 5, 1, 3, 4, 3, 2};
  for (int i = 0; i < 9; ++i) for (int j = 0; j < 9; ++j) g.TM[i][j] = D2[v2[i]][v2[j]];
  g.S.resize(mat->nT);
  for (int a = 0; a < mat->nT; ++a) {
    M3 mn;
    for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) mn[j][k] = mat->m[a][j] * mat->n[a][k];
    M3 t2 = mul(g.R, mn);
    g.S[a] = mult(t2, g.R);
  }
}

// M^p by repeated right multiplication starting from I (calculatePlasticity.cc:657-665)
M3 mpow(const M3 &X, int p) { M3 r = eye3(); for (int m = 0; m < p; ++m) r = mul(r, X); return r; }

// d exp series term of calculatePlasticity.cc:645-676 (sign=-1) and :748-779 (sign=+1)
M3 dexp_series(const M3 &X, const M3 &Sj, int count3, double sign) {
  M3 diff = Sj;
  for (int k = 1; k <= count3; ++k) {
    M3 t4 = zero3();
    for (int l = 0; l <= k; ++l) {
      M3 t5 = mul(mpow(X, l), Sj);
      M3 t6 = mul(t5, mpow(X, k - l));
      addto(t4, 1.0, t6);
    }
    const double c = pert(4) * (sign < 0 ? std::pow(-1.0, k) : 1.0) / std::tgamma((double)(k + 2));
    t4 = scale(t4, c);
    addto(diff, 1.0, t4);
  }
  return diff;
}

// calculatePlasticity(cellID, q, flag), CP calculatePlasticity.cc:13-959 (twinning branch :962-1019 off)
int qp_update_body(const or_material *mat, const Grain &g, const double *Fin, const double *Fp_conv, const double *s_conv,
                   const double *W_conv, const double *slip_conv, int flag, double *Pout, double *Tout, double *dPdF,
                   double *Fe_out, double *Fp_out, double *s_out, double *W_out, double *slip_iter, const double *x0,
                   int *iters, const double *tw_conv, double *tw_iter) {
  const int nT = mat->nT, ns = mat->n_slip, ntw = mat->n_twin;
  const double tol1 = mat->tol;
  if (iters) *iters = 0;
  const M3 F = load3(Fin);
  const M3 FP_t = load3(Fp_conv);
  std::vector<double> s_tau(s_conv, s_conv + nT), W_tau(W_conv, W_conv + nT), W_t(W_conv, W_conv + nT);

  M3 FP_tau = FP_t;                        // :140
  M3 Fpn_inv = inv3(FP_t);                 // :142
  M3 FE_tau = mul(F, Fpn_inv);             // :143
  std::vector<double> delFp(81, 0.0), dels(nT * 9, 0.0);  // :147-148
  M3 FP_t2, CE_tr, Ce_tau = zero3(), Tstar_tr, Tstar_tau, T_tau, P_tau, del_FP = zero3();
  std::vector<double> tau_tr(nT), tau(nT), b(nT), x_old(nT, 0.0), x_beta(nT, 0.0), h(nT), hab(nT * nT), A(nT * nT);
  std::vector<int> PA;
  double det_FE_tau = 1.0;
  int n_active = 0;

  unsigned iter1 = 1;
  while (true) {                                          // :153
    if (iter1 > (unsigned)mat->max_slip_search) break;    // :155-158
    FP_t2 = FP_tau;                                       // :160
    Fpn_inv = inv3(FP_t2);
    M3 FE_tr = mul(F, Fpn_inv);
    CE_tr = tmul(FE_tr, FE_tr);                           // :168
    M3 Ee;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Ee[i][j] = 0.5 * (CE_tr[i][j] - (i == j ? 1.0 : 0.0));
    Tstar_tr = dmat_apply(g.Dmat, Ee);                    // :184-185
    Tstar_tau = Tstar_tr;
    det_FE_tau = det3(FE_tau);                            // :188-196
    M3 temp = mul(FE_tau, Tstar_tr);
    temp = scale(temp, 1.0 / det_FE_tau);
    T_tau = mult(temp, FE_tau);
    M3 Finv = inv3(F);
    P_tau = scale(mult(T_tau, Finv), det_FE_tau);

    temp = mul(CE_tr, Tstar_tr);                          // :203
    PA.clear();
    for (int i = 0; i < nT; ++i) {                        // :207-240
      double r = 0;
      for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) r += temp[j][k] * g.S[i][j][k];
      tau_tr[i] = r;
      if (i > ns - 1) if ((tau_tr[i] - W_tau[i]) < 0) tau_tr[i] = W_tau[i];
      b[i] = std::fabs(tau_tr[i] - W_tau[i]) - s_tau[i];
      if (b[i] >= tol1) PA.push_back(i);
    }
    tau = tau_tr;
    if (PA.empty()) break;                                // :246-247

    for (int i = 0; i < ns; ++i) h[i] = mat->h0[i] * std::pow((1 - s_tau[i] / mat->ss[i]), mat->pw[i]);  // :255-262
    for (int i = 0; i < ntw; ++i) h[ns + i] = mat->h0[ns + i] * std::pow((1 - s_tau[ns + i] / mat->ss[ns + i]), mat->pw[ns + i]);
    for (int i = 0; i < nT; ++i)
      for (int j = 0; j < nT; ++j) { hab[i * nT + j] = mat->q[i][j] * h[j]; A[i * nT + j] = hab[i * nT + j]; }

    std::vector<M3> Mj(nT);                               // :273-291 (the j-dependent factor)
    for (int j = 0; j < nT; ++j) {
      M3 t2 = symm3(mul(CE_tr, g.S[j]));
      M3 t3 = dmat_apply(g.Dmat, t2);
      M3 t1 = mul(CE_tr, t3);
      t3 = mul(t2, Tstar_tr);
      addto(t1, 2.0, t3);
      Mj[j] = t1;
    }
    for (int i = 0; i < nT; ++i)
      for (int j = 0; j < nT; ++j) {
        double &a = A[i * nT + j];
        const bool opp = ((tau_tr[i] - W_tau[i]) * (tau_tr[j] - W_tau[j])) < 0.0;
        for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) {
          if (opp) a -= g.S[i][k][l] * Mj[j][k][l]; else a += g.S[i][k][l] * Mj[j][k][l];
        }
        if ((tau_tr[i] - W_tau[i]) < 0.0) a -= mat->C1[i] - mat->C2[i] * W_tau[i]; else a += mat->C1[i] - mat->C2[i] * W_tau[i];
      }

    std::fill(x_old.begin(), x_old.end(), 0.0);           // :311
    auto linf_bPA = [&]() { double m = 0; for (int p : PA) m = std::max(m, std::fabs(b[p])); return m; };
    // one update of the loop body from x_beta (increment) and x_old (total) (:338-440)
    auto update = [&]() {
      del_FP = zero3();                                   // :338-357
      for (int i = 0; i < nT; ++i)
        for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) {
          if ((tau_tr[i] - W_tau[i]) > 0) del_FP[j][k] = del_FP[j][k] + x_old[i] * g.S[i][j][k];
          else del_FP[j][k] = del_FP[j][k] - x_old[i] * g.S[i][j][k];
        }
      FP_tau = mul(expm3(del_FP), FP_t2);                 // :359
      FE_tau = mul(F, inv3(FP_tau));                      // :362-363
      Ce_tau = tmul(FE_tau, FE_tau);                      // :365
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Ee[i][j] = 0.5 * (Ce_tau[i][j] - (i == j ? 1.0 : 0.0));
      Tstar_tau = dmat_apply(g.Dmat, Ee);                 // :376-377
      temp = mul(Ce_tau, Tstar_tau);
      for (int i = 0; i < nT; ++i) {                      // :384-396
        double r = 0;
        for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) r += temp[j][k] * g.S[i][j][k];
        tau[i] = r;
        if (i > ns - 1) if ((tau_tr[i] - W_tau[i]) < 0) tau[i] = W_tau[i];
      }
      for (int i = 0; i < nT; ++i) {                      // :401-409
        double h1 = 0;
        for (int j = 0; j < nT; ++j) h1 = h1 + hab[i * nT + j] * x_beta[j];
        s_tau[i] = s_tau[i] + h1;
      }
      for (int i = 0; i < ns; ++i) if (s_tau[i] > mat->ss[i]) s_tau[i] = mat->ss[i];  // :411-429
      for (int i = 0; i < ntw; ++i) if (s_tau[ns + i] > mat->ss[ns + i]) s_tau[ns + i] = mat->ss[ns + i];
      for (int i = 0; i < nT; ++i) {                      // :431-436
        if ((tau_tr[i] - W_tau[i]) > 0) W_tau[i] = W_tau[i] + (mat->C1[i] - mat->C2[i] * W_tau[i]) * x_beta[i];
        else W_tau[i] = W_tau[i] + (-mat->C1[i] - mat->C2[i] * W_tau[i]) * x_beta[i];
      }
      for (int i = 0; i < nT; ++i) b[i] = std::fabs(tau[i] - W_tau[i]) - s_tau[i];  // :438-440
    };
    // the local loop (:322-441), `done` solves already counted against Max Solver Iterations;
    // `stop_on_drop`: stop as soon as inactive-slip removal drops a system.  Returns the solves.
    auto local_loop = [&](bool, bool stop_on_drop, unsigned done = 0) {
      unsigned count1 = done;
      while (linf_bPA() > tol1) {                         // :322
        count1++;
        if (count1 > (unsigned)mat->max_solver_iter) break;
        std::fill(x_beta.begin(), x_beta.end(), 0.0);
        const size_t n_before = PA.size();
        inactive_slip_removal(nT, x_old, x_beta, PA, b, A, mat->pinv_mode);  // :336
        if (stop_on_drop && PA.size() < n_before) { update(); break; }
        update();
      }
      return (int)(std::min(count1, (unsigned)mat->max_solver_iter) - done);
    };
    // Warm start (no counterpart in the reference, which always starts from x_beta_old = 0, :311).
    // The reference's loop is path dependent: its first solve on the trial set, with inactive-slip
    // removal, fixes the set S1 the later iterations work in (they only drop systems), and its
    // converged states routinely leave dropped systems above the yield surface, so a converged
    // start on another set is another answer.  The first solve therefore always runs.  If the
    // predicted set {a in PA : x0[a] > 0} equals S1, its increment is replaced by x0 (applied as one
    // pass of the loop body) and the loop continues on S1; the start stands if the loop converges
    // without dropping a system, otherwise the loop restarts from zero.  If the sets differ the
    // loop simply continues from the first solve, as without a prediction.  iters counts every
    // solve, the restarted ones included.
    int n_solves = 0;
    std::vector<int> PAw;
    if (x0)
      for (int p : PA) if (p < ns && x0[p] > 0) PAw.push_back(p);
    if (!PAw.empty()) {
      const std::vector<int> PA_trial = PA;
      const std::vector<double> s_start = s_tau, W_start = W_tau, b_start = b;
      const M3 FE_start = FE_tau;
      bool warm_used = false;
      if (linf_bPA() > tol1 && mat->max_solver_iter >= 1) {
        std::fill(x_beta.begin(), x_beta.end(), 0.0);
        inactive_slip_removal(nT, x_old, x_beta, PA, b, A, mat->pinv_mode);  // the first solve (:336)
        n_solves = 1;
        std::vector<int> S1 = PA, P2 = PAw;
        std::sort(S1.begin(), S1.end());
        if (S1 == P2) {
          warm_used = true;
          std::fill(x_beta.begin(), x_beta.end(), 0.0);
          for (int p : PAw) x_beta[p] = x0[p];
          x_old = x_beta;
        }
        update();
        if (warm_used) {
          const size_t n_set = PA.size();
          const int c = local_loop(false, true);
          n_solves += c;
          bool ok = PA.size() == n_set && c < mat->max_solver_iter;
          for (int p : PA) if (!(std::fabs(b[p]) <= tol1)) ok = false;
          if (!ok) {
            PA = PA_trial;
            s_tau = s_start; W_tau = W_start; b = b_start;
            FP_tau = FP_t2; FE_tau = FE_start; Ce_tau = zero3(); Tstar_tau = Tstar_tr; tau = tau_tr;
            std::fill(x_old.begin(), x_old.end(), 0.0);
            std::fill(x_beta.begin(), x_beta.end(), 0.0);
            n_solves += local_loop(false, false);
          }
        } else {
          n_solves += local_loop(false, false, 1);
        }
      } else {
        n_solves = local_loop(false, false);
      }
    } else {
      n_solves = local_loop(false, false);
    }

    if (iters) *iters = n_solves;
    if (mat->twinning && tw_iter)                         // :450-452
      for (int i = 0; i < ntw; ++i) tw_iter[i] = tw_conv[i] + x_old[ns + i] / mat->twin_shear;
    for (int i = 0; i < ns; ++i) slip_iter[i] = slip_conv[i] + x_old[i];  // :454-456

    Fpn_inv = inv3(FP_tau);                               // :458-468
    FE_tau = mul(F, Fpn_inv);
    det_FE_tau = det3(FE_tau);
    temp = scale(mul(FE_tau, Tstar_tau), 1.0 / det_FE_tau);
    T_tau = mult(temp, FE_tau);
    const double det_F = det3(F);
    P_tau = scale(mult(T_tau, inv3(F)), det_F);
    n_active = (int)PA.size();

    if (flag == 1) {                                      // :473-828
      const int nPA = (int)PA.size();
      std::vector<double> delFp_prev = delFp, dels_prev = dels;
      M3 t1 = mul(F, Fpn_inv);
      double dFe[81] = {0}, dE[81] = {0}, dT[81] = {0}, dtau[81] = {0};
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) {
        double &v = dFe[(3 * i + j) * 9 + 3 * k + l];
        for (int a = 0; a < 3; ++a) for (int bb = 0; bb < 3; ++bb) v = v - pert(7) * t1[i][a] * delFp[(3 * a + bb) * 9 + 3 * k + l] * Fpn_inv[bb][j];
        if (i == k) v = v + Fpn_inv[l][j];
      }
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l)
        for (int a = 0; a < 3; ++a)
          dE[(3 * i + j) * 9 + 3 * k + l] += 0.5 * (dFe[(3 * a + i) * 9 + 3 * k + l] * FE_tau[a][j] + dFe[(3 * a + j) * 9 + 3 * k + l] * FE_tau[a][i]);
      for (int r = 0; r < 9; ++r) for (int c = 0; c < 9; ++c) {  // delT_delF = TM . delEtrial_delF (:531)
        double s = 0;
        for (int k = 0; k < 9; ++k) s += g.TM[r][k] * dE[k * 9 + c];
        dT[r * 9 + c] = s;
      }
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l)
        for (int a = 0; a < 3; ++a)                       // :533-545
          dtau[(3 * i + j) * 9 + 3 * k + l] = dtau[(3 * i + j) * 9 + 3 * k + l] + 2 * dE[(3 * i + a) * 9 + 3 * k + l] * Tstar_tau[a][j] +
                                              Ce_tau[i][a] * dT[(3 * a + j) * 9 + 3 * k + l];
      std::vector<double> dh(nT, 0.0);                     // :553-560
      for (int i = 0; i < ns; ++i) dh[i] = mat->h0[i] * std::pow((1 - s_tau[i] / mat->ss[i]), (mat->pw[i] - 1)) * (-1.0 / mat->ss[i]);
      for (int i = 0; i < ntw; ++i) dh[ns + i] = mat->h0[ns + i] * std::pow((1 - s_tau[ns + i] / mat->ss[ns + i]), (mat->pw[ns + i] - 1)) * (-1.0 / mat->ss[ns + i]);
      {                                                    // :562-579 dels = (I - term_ds)^-1 dels_prev
        std::vector<double> M(nT * nT);
        for (int k = 0; k < nT; ++k) for (int l = 0; l < nT; ++l) M[k * nT + l] = (k == l ? 1.0 : 0.0) - x_old[l] * mat->q[k][l] * dh[l];
        dels = dels_prev;
        lu_solve(nT, M, dels.data(), 9);
      }
      std::vector<double> delb(nPA * 9);                   // :581-604
      for (int k = 0; k < nPA; ++k) {
        const M3 &S = g.S[PA[k]];
        for (int l = 0; l < 9; ++l) {
          double s = 0;
          for (int m = 0; m < 9; ++m) s += dtau[m * 9 + l] * S[m / 3][m % 3];
          delb[k * 9 + l] = pert(1) * s;
        }
        if ((tau_tr[PA[k]] - W_t[PA[k]]) < 0) for (int l = 0; l < 9; ++l) delb[k * 9 + l] = -1.0 * delb[k * 9 + l];
      }
      double tol2 = 1.0;                                   // :607-619
      int count3 = 0;
      {
        M3 s1 = zero3(), s2 = eye3();
        while (tol2 > std::max<double>(tol1 / 1e4, 1e-12)) {
          count3 = count3 + 1;
          s1 = mul(del_FP, s2);
          s1 = scale(s1, (1.0 / count3));
          tol2 = fro3(s1);
          s2 = s1;
        }
      }
      std::vector<double> A2(nPA * nPA);                   // :622-709
      for (int i = 0; i < nPA; ++i) for (int j = 0; j < nPA; ++j) A2[i * nPA + j] = pert(3) * hab[PA[i] * nT + PA[j]];
      const M3 eneg = expm3(scale(del_FP, -1.0));
      for (int j = 0; j < nPA; ++j) {
        const M3 diff = dexp_series(del_FP, g.S[PA[j]], count3, -1.0);
        for (int i = 0; i < nPA; ++i) {
          M3 t6 = mul(CE_tr, eneg);
          M3 t2 = symm3(tmul(diff, t6));
          M3 t3 = dmat_apply(g.Dmat, t2);
          M3 tm = mul(Ce_tau, t3);
          t3 = mul(t2, Tstar_tau);
          addto(tm, 2.0, t3);
          double &a = A2[i * nPA + j];
          const bool ni = (tau_tr[PA[i]] - W_t[PA[i]]) < 0.0, nj = (tau_tr[PA[j]] - W_t[PA[j]]) < 0.0;
          for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) {
            const double e = pert(2) * g.S[PA[i]][k][l] * tm[k][l];
            if (ni ^ nj) a -= e; else a += e;
          }
          if (g_pert_block == 9) {
            if (i == j) a += mat->C1[PA[i]] - (ni ? -1.0 : 1.0) * mat->C2[PA[i]] * W_tau[PA[i]];
          } else {
            if (ni) a -= mat->C1[PA[i]] - mat->C2[PA[i]] * W_tau[PA[i]]; else a += mat->C1[PA[i]] - mat->C2[PA[i]] * W_tau[PA[i]];
          }
        }
      }
      std::vector<double> dg(nPA * 9);                     // :712-722 delgamma = A2^-1 (delb - dels_PA)
      for (int i = 0; i < nPA; ++i) for (int j = 0; j < 9; ++j) dg[i * 9 + j] = delb[i * 9 + j] - dels[PA[i] * 9 + j];
      // a numerically singular A2 contributes no plastic tangent term (same rule as the HIP kernel)
      if (nPA > 0 && !lu_solve(nPA, A2, dg.data(), 9)) std::fill(dg.begin(), dg.end(), 0.0);
      std::vector<double> dg2(nT * 9, 0.0);                // :724-729
      for (int i = 0; i < nPA; ++i) for (int j = 0; j < 9; ++j) dg2[PA[i] * 9 + j] = dg[i * 9 + j];
      std::vector<double> SPA(9 * nPA);                    // :734-796
      for (int j = 0; j < nPA; ++j) {
        const M3 diff = dexp_series(del_FP, g.S[PA[j]], count3, +1.0);
        const M3 t = mul(diff, FP_t2);
        const bool nj = (tau_tr[PA[j]] - W_t[PA[j]]) < 0;
        for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) SPA[(3 * k + l) * nPA + j] = pert(5) * (nj ? -t[k][l] : t[k][l]);
      }
      const M3 epos = expm3(del_FP);                       // :799-821
      std::vector<double> dFp2(81, 0.0);
      for (int r = 0; r < 9; ++r) for (int c = 0; c < 9; ++c) {
        double s = 0;
        for (int k = 0; k < nPA; ++k) s += SPA[r * nPA + k] * dg[k * 9 + c];
        dFp2[r * 9 + c] = s;
      }
      std::fill(delFp.begin(), delFp.end(), 0.0);
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l)
        for (int a = 0; a < 3; ++a)
          delFp[(3 * i + j) * 9 + 3 * k + l] = delFp[(3 * i + j) * 9 + 3 * k + l] + epos[i][a] * delFp_prev[(3 * a + j) * 9 + 3 * k + l];
      for (int r = 0; r < 81; ++r) delFp[r] += dFp2[r];
      // :823-827 computes dels_delF_prev = dels_delF + A_ds*delgamma_delF2, which the
      // next block overwrites with dels_delF at :477 before reading it; dels keeps its value.
    }
    iter1 = iter1 + 1;                                    // :830
  }

  if (flag == 1) {                                        // :835-940 final tangent
    M3 t1 = mul(F, Fpn_inv);
    double dFe[81] = {0}, dTs[81] = {0};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) {
      double &v = dFe[(3 * i + j) * 9 + 3 * k + l];
      for (int a = 0; a < 3; ++a) for (int bb = 0; bb < 3; ++bb) v = v - pert(7) * t1[i][a] * delFp[(3 * a + bb) * 9 + 3 * k + l] * Fpn_inv[bb][j];
      if (i == k) v = v + Fpn_inv[l][j];
    }
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l)
      for (int a = 0; a < 3; ++a) for (int bb = 0; bb < 3; ++bb) for (int c = 0; c < 3; ++c)
        dTs[(3 * i + j) * 9 + 3 * k + l] = dTs[(3 * i + j) * 9 + 3 * k + l] + pert(6) * g.TM[3 * i + j][3 * a + bb] * dFe[(3 * c + a) * 9 + 3 * k + l] * FE_tau[c][bb];
    const M3 t4 = inv3(F);                                // :885-897
    const M3 tt = mul(t4, FE_tau);
    const M3 q1 = mult(Tstar_tau, tt);
    const M3 q2 = mul(t4, FE_tau);
    const M3 q3 = mul(FE_tau, Tstar_tau);
    const M3 q5 = mult(q3, F);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) {
      double v = 0;                                       // :899-913
      for (int a = 0; a < 3; ++a) {
        for (int bb = 0; bb < 3; ++bb)
          v = v + (i == a ? 1.0 : 0.0) * dFe[(3 * a + bb) * 9 + 3 * k + l] * q1[bb][j] + FE_tau[i][a] * dTs[(3 * a + bb) * 9 + 3 * k + l] * q2[j][bb] -
              q3[i][a] * dFe[(3 * a + bb) * 9 + 3 * k + l] * t4[j][bb];
        v = v - pert(8) * q5[i][a] * t4[j][k] * t4[l][a];
      }
      dPdF[(3 * i + j) * 9 + 3 * k + l] = v;              // :916-939 with L = I
    }
  }

  store3(P_tau, Pout);                                    // :942-959
  store3(T_tau, Tout);
  store3(FE_tau, Fe_out);
  store3(FP_tau, Fp_out);
  for (int i = 0; i < nT; ++i) { s_out[i] = s_tau[i]; W_out[i] = W_tau[i]; }
  return n_active;
}

int qp_update(const or_material *mat, const Grain &g, const double *Fin, const double *Fp_conv, const double *s_conv,
              const double *W_conv, const double *slip_conv, int flag, double *Pout, double *Tout, double *dPdF,
              double *Fe_out, double *Fp_out, double *s_out, double *W_out, double *slip_iter,
              const double *x0 = nullptr, int *iters = nullptr, const double *tw_conv = nullptr,
              double *tw_iter = nullptr) {
  OR_FLOP_QP_BEGIN(flag);
  const int rc = qp_update_body(mat, g, Fin, Fp_conv, s_conv, W_conv, slip_conv, flag, Pout, Tout, dPdF, Fe_out, Fp_out,
                                s_out, W_out, slip_iter, x0, iters, tw_conv, tw_iter);
  OR_FLOP_QP_END(flag);
  return rc;
}

// ------------------------------------------------------------ twinning
// Quaternion helpers of updateAfterIncrement.cc:780-821: q = cos(phi/2) (1, r) from a Rodrigues
// vector; the product p = q2 q1 (R(p) = R(q2) R(q1)) with the sign chosen so that p0 >= 0; and
// r = q_vec / q0 with |q0| raised to 1e-4 when smaller.
void rod2quat(const double r[3], double q[4]) {
  const double c = std::cos(std::atan(std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])));
  q[0] = c; q[1] = c * r[0]; q[2] = c * r[1]; q[3] = c * r[2];
}
void quatproduct(const double q2[4], const double q1[4], double p[4]) {
  const double a = q2[0], b = q1[0];
  const double dot1 = q1[1] * q2[1] + q1[2] * q2[2] + q1[3] * q2[3];
  p[0] = a * b - dot1;
  p[1] = a * q1[1] + b * q2[1] + q2[2] * q1[3] - q1[2] * q2[3];
  p[2] = a * q1[2] + b * q2[2] - q2[1] * q1[3] + q1[1] * q2[3];
  p[3] = a * q1[3] + b * q2[3] + q2[1] * q1[2] - q1[1] * q2[2];
  if (p[0] < 0) for (int i = 0; i < 4; ++i) p[i] = -p[i];
}
void quat2rod(double q[4], double r[3]) {
  if (std::fabs(q[0]) < 1e-4) q[0] = 1e-4;
  const double inv = 1 / q[0];
  for (int i = 0; i < 3; ++i) r[i] = q[i + 1] * inv;
}
// the orientation after twinning on system k: rot rotated by 180 degrees about the twin normal
// (calculatePlasticity.cc:995-1008)
void twin_rotation(const or_material *mat, int k, const double rot[3], double out[3]) {
  double q2[4], q1[4] = {0.0, mat->n[mat->n_slip + k][0], mat->n[mat->n_slip + k][1], mat->n[mat->n_slip + k][2]}, p[4];
  rod2quat(rot, q2);
  quatproduct(q2, q1, p);
  quat2rod(p, out);
}
// calculatePlasticity.cc:962-1019 for one qp (see or_twin_event)
int twin_event(const or_material *mat, double F_r, double F_e, const double *tw_iter, int twin_conv, const double rot_conv[3],
               double rot_iter[3], int *twin_iter, double *s_iter, bool multiphase = false) {
  if (!mat->twinning) return -1;
  // :970-980; with several phases F_T is the phase's threshold fraction alone
  double F_T = (!multiphase && F_r > 0) ? mat->twin_threshold + (mat->twin_saturation * F_e / F_r) : mat->twin_threshold;
  if (F_T > 1.0) F_T = 1.0;                                                                             // :983-985
  int pos = 0;                                                                                          // :986-990
  for (int i = 1; i < mat->n_twin; ++i) if (tw_iter[i] > tw_iter[pos]) pos = i;
  if (twin_conv == 1 || !(F_r > 0) || !(tw_iter[pos] > F_T)) return -1;                                // :991-993
  twin_rotation(mat, pos, rot_conv, rot_iter);
  *twin_iter = 1;
  for (int i = 0; i < mat->n_twin; ++i) s_iter[mat->n_slip + i] = 100000;                              // :1013-1015
  return pos;
}

// ----------------------------------------------------------- finite elements
// Q1 hexahedron, QGauss(2), axis-aligned uniform cell of size h.  Local dof d
// = 3*v + comp with lexicographic vertex v = vx + 2 vy + 4 vz (FESystem<FE_Q(1)^3>).
struct Shape {
  double grad[8][8][3];  // [q][v][dir]
  double JxW;
};
Shape make_shape(const double h[3]) {
  Shape s;
  const double g[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  for (int q = 0; q < 8; ++q) {
    const double xi[3] = {g[q & 1], g[(q >> 1) & 1], g[(q >> 2) & 1]};
    for (int v = 0; v < 8; ++v) {
      const int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
      double N[3], dN[3];
      for (int d = 0; d < 3; ++d) { N[d] = b[d] ? xi[d] : 1.0 - xi[d]; dN[d] = (b[d] ? 1.0 : -1.0) / h[d]; }
      s.grad[q][v][0] = dN[0] * N[1] * N[2];
      s.grad[q][v][1] = N[0] * dN[1] * N[2];
      s.grad[q][v][2] = N[0] * N[1] * dN[2];
    }
  }
  s.JxW = h[0] * h[1] * h[2] / 8.0;
  return s;
}

// F at a quadrature point (getElementalValues.cc:129-138)
M3 def_grad(const Shape &sh, int q, const double *ue) {
  M3 F = zero3();
  for (int d = 0; d < 24; ++d) {
    const int i = d % 3, v = d / 3;
    for (int j = 0; j < 3; ++j) F[i][j] += ue[d] * sh.grad[q][v][j];
  }
  for (int i = 0; i < 3; ++i) F[i][i] += 1;
  return F;
}

struct QPState {  // SoA-free per-qp history (oracle only)
  std::vector<double> Fp_conv, Fe_conv, s_conv, W_conv, slip_conv, rotnew_conv;
  std::vector<double> Fp_iter, Fe_iter, s_iter, W_iter, slip_iter;
  std::vector<double> T;  // Cauchy stress of the last flag-0 pass
  // twinning: orientation of the constitutive update (rot_conv/iter), output orientation before
  // reorient() (rotnew_iter), twinned flags, twin fractions (n_twin per qp)
  std::vector<double> rot_conv, rot_iter, rotnew_iter, tw_conv, tw_iter;
  std::vector<int> twin_conv, twin_iter;
  std::vector<int> tgrain;  // index into the twinned grains (-1: the element's grain)
};

struct Mesh {
  int nx, ny, nz, nnx, nny, nnz;
  long long nnode, ndof, nelem;
  double h[3];
  int node(int i, int j, int k) const { return i + nnx * (j + nny * k); }
  void elem_dofs(long long e, int *dofs) const {
    const int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((long long)nx * ny));
    for (int v = 0; v < 8; ++v) {
      const int n = node(ex + (v & 1), ey + ((v >> 1) & 1), ez + ((v >> 2) & 1));
      for (int c = 0; c < 3; ++c) dofs[3 * v + c] = 3 * n + c;
    }
  }
};
Mesh make_mesh(const or_problem *pb) {
  Mesh m;
  m.nx = pb->nx; m.ny = pb->ny; m.nz = pb->nz;
  m.nnx = m.nx + 1; m.nny = m.ny + 1; m.nnz = m.nz + 1;
  m.nnode = (long long)m.nnx * m.nny * m.nnz;
  m.ndof = 3 * m.nnode;
  m.nelem = (long long)m.nx * m.ny * m.nz;
  m.h[0] = pb->span[0] / m.nx; m.h[1] = pb->span[1] / m.ny; m.h[2] = pb->span[2] / m.nz;
  return m;
}

struct CSR {
  std::vector<long long> rp;
  std::vector<int> ci;
  std::vector<double> val;
  long long find(long long r, int c) const {
    const int *b = ci.data() + rp[r], *e = ci.data() + rp[r + 1];
    const int *p = std::lower_bound(b, e, c);
    return p - ci.data();
  }
};

// DoFTools::make_sparsity_pattern over the 27-node stencil (BVP init.cc:183-199)
CSR make_csr(const Mesh &m) {
  CSR A;
  A.rp.assign(m.ndof + 1, 0);
  for (int k = 0; k < m.nnz; ++k) for (int j = 0; j < m.nny; ++j) for (int i = 0; i < m.nnx; ++i) {
    const int cnt = (std::min(i + 1, m.nnx - 1) - std::max(i - 1, 0) + 1) * (std::min(j + 1, m.nny - 1) - std::max(j - 1, 0) + 1) *
                    (std::min(k + 1, m.nnz - 1) - std::max(k - 1, 0) + 1);
    const long long n = m.node(i, j, k);
    for (int c = 0; c < 3; ++c) A.rp[3 * n + c + 1] = 3 * cnt;
  }
  for (long long r = 0; r < m.ndof; ++r) A.rp[r + 1] += A.rp[r];
  A.ci.resize(A.rp[m.ndof]);
  for (int k = 0; k < m.nnz; ++k) for (int j = 0; j < m.nny; ++j) for (int i = 0; i < m.nnx; ++i) {
    const long long n = m.node(i, j, k);
    std::vector<int> cols;
    for (int dk = -1; dk <= 1; ++dk) for (int dj = -1; dj <= 1; ++dj) for (int di = -1; di <= 1; ++di) {
      const int a = i + di, b = j + dj, c = k + dk;
      if (a < 0 || b < 0 || c < 0 || a >= m.nnx || b >= m.nny || c >= m.nnz) continue;
      const int nb = m.node(a, b, c);
      for (int cc = 0; cc < 3; ++cc) cols.push_back(3 * nb + cc);
    }
    std::sort(cols.begin(), cols.end());
    for (int c = 0; c < 3; ++c) std::copy(cols.begin(), cols.end(), A.ci.begin() + A.rp[3 * n + c]);
  }
  A.val.assign(A.ci.size(), 0.0);
  return A;
}

// Table segment of increment inc (boundaryConditions.cc:162-170): timeCounter starts at 1 and
// moves on while the increment's end time delT*(inc+1) lies past the table time it points at; the
// reference re-tests it at every boundary node, which is a while loop over one increment.
int tabular_segment(const or_problem *pb, int inc) {
  int tc = 1;
  for (int k = 0; k <= inc; ++k) {
    const double t = pb->delT * (k + 1);
    while (tc < pb->n_table - 1 && t > pb->table_time[tc]) ++tc;
  }
  return tc;
}

// setBoundaryValues + applyDirichletBCs (BVP boundaryConditions.cc:6-114, 452-491): constrained
// flag and the per-increment value deluConstraint (BVP init.cc:327-348) of every dof.
void dirichlet(const Mesh &m, const or_problem *pb, int inc, std::vector<char> &con, std::vector<double> &delu) {
  con.assign(m.ndof, 0);
  delu.assign(m.ndof, 0.0);
  double dU[6][3];
  for (int f = 0; f < 6; ++f) for (int d = 0; d < 3; ++d) dU[f][d] = pb->face_total_u[f][d] / pb->total_increments;
  if (pb->tabular) {  // boundaryConditions.cc:153-267: slope of the current table segment times delT
    const int tc = tabular_segment(pb, inc);
    for (int f = 0; f < 6; ++f) for (int d = 0; d < 3; ++d)
      dU[f][d] = (pb->table_u[f][d][tc] - pb->table_u[f][d][tc - 1]) / (pb->table_time[tc] - pb->table_time[tc - 1]) * pb->delT;
  }
  // velocity-gradient BC (updateBeforeIncrement.cc:10-36, boundaryConditions.cc:269-275): F_(k+1) =
  // (I - L dt)^-1 F_k from F_0 = I, and every boundary dof moves by (F_(inc+1) - F_inc) X at
  // iteration 0
  M3 dF = zero3();
  if (pb->velgrad) {
    M3 IL = eye3();
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) IL[a][b] -= pb->vel_grad[a][b] * pb->delT;
    const M3 Minv = inv3(IL);
    M3 Fk = eye3();
    for (int k = 0; k <= inc; ++k) {
      const M3 Fn = mul(Minv, Fk);
      for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) dF[a][b] = Fn[a][b] - Fk[a][b];
      Fk = Fn;
    }
  }
  double cyc_sign = 1.0;
  if (pb->cyclic) {
    const int cf = pb->cyclic_face - 1, cd = pb->cyclic_dof - 1;
    dU[cf][cd] = dU[cf][cd] * pb->total_increments * pb->delT / pb->quarter_cycle_time;  // BVP init.cc:343-345
    const double cycleTime = 4 * pb->quarter_cycle_time, t = std::fmod(inc * pb->delT, cycleTime);
    if (t < pb->quarter_cycle_time) cyc_sign = 1.0; else if (t < 3 * pb->quarter_cycle_time) cyc_sign = -1.0; else cyc_sign = 1.0;
  }
  for (int k = 0; k < m.nnz; ++k) for (int j = 0; j < m.nny; ++j) for (int i = 0; i < m.nnx; ++i) {
    const bool onb[6] = {i == 0, i == m.nx, j == 0, j == m.ny, k == 0, k == m.nz};
    if (!(onb[0] || onb[1] || onb[2] || onb[3] || onb[4] || onb[5])) continue;  // boundary faces only
    const long long n = m.node(i, j, k);
    for (int d = 0; d < 3; ++d) {
      bool flag = false;
      double val = 0.0;
      if (pb->cyclic && d == pb->cyclic_dof - 1 && onb[pb->cyclic_face - 1]) {  // boundaryConditions.cc:16-77
        flag = true; val = cyc_sign * dU[pb->cyclic_face - 1][d];
      } else {
        for (int f = 0; f < 6; ++f)
          if (pb->face_dof_constrained[f][d] && onb[f]) { flag = true; val = dU[f][d]; break; }
      }
      if (!flag && pb->velgrad) {  // the last rule of setBoundaryValues: every boundary dof
        const double X[3] = {i * m.h[0], j * m.h[1], k * m.h[2]};
        flag = true;
        for (int c = 0; c < 3; ++c) val += dF[d][c] * X[c];
      }
      if (flag) { con[3 * n + d] = 1; delu[3 * n + d] = val; }
    }
  }
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

void spmv(const CSR &A, const std::vector<double> &x, std::vector<double> &y) {
  const long long n = (long long)A.rp.size() - 1;
#pragma omp parallel for schedule(static)
  for (long long r = 0; r < n; ++r) {
    double s = 0;
    for (long long p = A.rp[r]; p < A.rp[r + 1]; ++p) s += A.val[p] * x[A.ci[p]];
    y[r] = s;
  }
}
void spmv_t(const CSR &A, const std::vector<double> &x, std::vector<double> &y) {
  const long long n = (long long)A.rp.size() - 1;
  std::fill(y.begin(), y.end(), 0.0);
  for (long long r = 0; r < n; ++r)
    for (long long p = A.rp[r]; p < A.rp[r + 1]; ++p) y[A.ci[p]] += A.val[p] * x[r];
}

// PETSc KSPBICG with PCJACOBI, preconditioned-norm convergence test of deal.II's
// SolverControl (residual <= tol), zero initial guess (BVP solveLinearSystem.cc:13-16).
int bicg(const CSR &A, const std::vector<double> &b, std::vector<double> &x, double tol, int maxit) {
  const size_t n = b.size();
  std::vector<double> dinv(n);
  for (size_t r = 0; r < n; ++r) {
    const long long p = A.find(r, (int)r);
    const double d = A.val[p];
    dinv[r] = (d == 0.0) ? 1.0 : 1.0 / d;
  }
  x.assign(n, 0.0);
  std::vector<double> Rr = b, Rl = b, Zr(n), Zl(n), Pr(n), Pl(n);
  for (size_t i = 0; i < n; ++i) { Zr[i] = Rr[i] * dinv[i]; Zl[i] = Rl[i] * dinv[i]; }
  double dp = std::sqrt(dot(Zr, Zr));
  if (dp <= tol) return 0;
  double betaold = 1.0;
  int i = 0;
  do {
    const double beta = dot(Zr, Rl);
    if (i == 0) {
      if (beta == 0.0) return 0;
      Pr = Zr; Pl = Zl;
    } else {
      const double bb = beta / betaold;
      for (size_t k = 0; k < n; ++k) { Pr[k] = Zr[k] + bb * Pr[k]; Pl[k] = Zl[k] + bb * Pl[k]; }
    }
    betaold = beta;
    spmv(A, Pr, Zr);
    spmv_t(A, Pl, Zl);
    const double dpi = dot(Zr, Pl);
    const double a = beta / dpi;
    for (size_t k = 0; k < n; ++k) { x[k] += a * Pr[k]; Rr[k] -= a * Zr[k]; Rl[k] -= a * Zl[k]; }
    for (size_t k = 0; k < n; ++k) { Zr[k] = Rr[k] * dinv[k]; Zl[k] = Rl[k] * dinv[k]; }
    dp = std::sqrt(dot(Zr, Zr));
    ++i;
    if (dp <= tol) return i;
  } while (i < maxit);
  return i;
}

// Jacobi-PCG (the alternative the product offers); stopping test on the preconditioned norm as above.
int pcg(const CSR &A, const std::vector<double> &b, std::vector<double> &x, double tol, int maxit) {
  const size_t n = b.size();
  std::vector<double> dinv(n);
  for (size_t r = 0; r < n; ++r) { const double d = A.val[A.find(r, (int)r)]; dinv[r] = (d == 0.0) ? 1.0 : 1.0 / d; }
  x.assign(n, 0.0);
  std::vector<double> r = b, z(n), p(n), q(n);
  for (size_t i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
  double dp = std::sqrt(dot(z, z));
  if (dp <= tol) return 0;
  double rz = dot(r, z);
  p = z;
  int it = 0;
  while (it < maxit) {
    spmv(A, p, q);
    const double a = rz / dot(p, q);
    for (size_t k = 0; k < n; ++k) { x[k] += a * p[k]; r[k] -= a * q[k]; }
    for (size_t k = 0; k < n; ++k) z[k] = r[k] * dinv[k];
    ++it;
    dp = std::sqrt(dot(z, z));
    if (dp <= tol) break;
    const double rz2 = dot(r, z), bt = rz2 / rz;
    rz = rz2;
    for (size_t k = 0; k < n; ++k) p[k] = z[k] + bt * p[k];
  }
  return it;
}

// symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (the reference calls
// LAPACK dsyevx through compute_eigenvalues_symmetric, reorient.cc:35,48).  The sweeps end when
// the off-diagonal is zero or a sweep leaves the eigenvalues and V unchanged (later rotations are
// smaller still, down to the identity once th * th overflows).
void sym_eig3(const M3 &Cin, double lam[3], M3 &V) {
  M3 C = Cin;
  V = eye3();
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = C[0][1] * C[0][1] + C[0][2] * C[0][2] + C[1][2] * C[1][2];
    if (off == 0.0) break;
    const M3 V0 = V;
    const double d0[3] = {C[0][0], C[1][1], C[2][2]};
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (C[p][q] == 0.0) continue;
        const double th = (C[q][q] - C[p][p]) / (2.0 * C[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // C <- J^T C J
          const double ckp = C[k][p], ckq = C[k][q];
          C[k][p] = c * ckp - s * ckq; C[k][q] = s * ckp + c * ckq;
        }
        for (int k = 0; k < 3; ++k) {
          const double cpk = C[p][k], cqk = C[q][k];
          C[p][k] = c * cpk - s * cqk; C[q][k] = s * cpk + c * cqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq; V[k][q] = s * vkp + c * vkq;
        }
      }
    bool same = d0[0] == C[0][0] && d0[1] == C[1][1] && d0[2] == C[2][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) same = same && V0[i][j] == V[i][j];
    if (same) break;
  }
  for (int i = 0; i < 3; ++i) lam[i] = C[i][i];
}

// R of the polar decomposition Fe = R U (reorient.cc:33-57)
M3 polar_R(const M3 &Fe) {
  const M3 C = tmul(Fe, Fe);
  double lam[3];
  M3 V;
  sym_eig3(C, lam, V);
  M3 L = zero3();
  for (int k = 0; k < 3; ++k) L[k][k] = std::sqrt(lam[k]);
  M3 U = mult(mul(V, L), V);
  return mul(Fe, inv3(U));
}

// reorient, CP reorient.cc:18-96 (one qp)
void reorient_qp(const M3 &Fe_old, const M3 &Fe_new, double r[3]) {
  const M3 R_old = polar_R(Fe_old), R_new = polar_R(Fe_new);
  M3 Om = R_new;
  addto(Om, -1.0, R_old);
  Om = mult(Om, R_new);
  const double w[3] = {-0.5 * (Om[1][2] - Om[2][1]), 0.5 * (Om[0][2] - Om[2][0]), -0.5 * (Om[0][1] - Om[1][0])};
  const double dt = w[0] * r[0] + w[1] * r[1] + w[2] * r[2];
  const double cr[3] = {w[1] * r[2] - w[2] * r[1], w[2] * r[0] - w[0] * r[2], w[0] * r[1] - w[1] * r[0]};
  double rn[3];
  for (int i = 0; i < 3; ++i) rn[i] = r[i] + 0.5 * (w[i] + dt * r[i] + cr[i]);
  const double nrm = std::sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]), mx = 10000;
  if (nrm > mx) for (int i = 0; i < 3; ++i) rn[i] = rn[i] * mx / nrm;
  for (int i = 0; i < 3; ++i) r[i] = rn[i];
}

}  // namespace

// ---------------------------------------------------------------- periodic BCs
// The affine constraints of Use Periodic BCs (BVP periodicBCs.cc), closed as deal.II's
// AffineConstraints::close() leaves them: every constrained dof d reads u_d = sum_k w_k u_{m_k} + s g_d
// over unconstrained masters m_k, s = the load scale of the Newton iteration (1 for
// setPeriodicityConstraintsInc0, -1 for Inc0Neg, 0 for IncNot0: the inhomogeneities are linear in
// the per-increment vertex values).
struct PLine {
  std::vector<std::pair<long long, double>> e;
  double g = 0.0;
};
struct PeriodicCons {
  std::vector<char> con;
  std::vector<PLine> line;  // closed lines, indexed by dof
  std::vector<double> tab_scale;  // tabular periodic BCs: load scale of increment 0..n-1
};

// A double converted to `unsigned int` as x86-64 code does it (truncation to a 64-bit integer, low
// 32 bits kept): the reference accumulates the derived vertex value in an unsigned int
// (periodicBCs.cc:56, 515-517), so values in (-1, 1) become 0.
static double as_unsigned(double x) { return (double)(uint32_t)(int64_t)(long long)x; }

// Node of vertex v (rows 3v..3v+2: V000 V100 V010 V001 V110 V101 V011 V111)
static int pbc_vertex_node(const Mesh &m, int v) {
  static const int c[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  return m.node(c[v][0] * m.nx, c[v][1] * m.ny, c[v][2] * m.nz);
}
// Node t (0..n-2, sorted along the edge, periodicBCs.cc:775-800) of edge e (EX000 EX001 EX010 EX011
// EY000 EY001 EY100 EY101 EZ000 EZ010 EZ100 EZ110: EXab at y = a, z = b; EYab at x = a, z = b;
// EZab at x = a, y = b)
static int pbc_edge_node(const Mesh &m, int e, int t) {
  const int dir = e / 4, a = (e % 4) >> 1, b = e & 1;
  if (dir == 0) return m.node(t + 1, a * m.ny, b * m.nz);
  if (dir == 1) return m.node(a * m.nx, t + 1, b * m.nz);
  return m.node(a * m.nx, b * m.ny, t + 1);
}

// setPeriodicityConstraintsInit (periodicBCs.cc:23-864, the parts that matter on one process) and
// setFace/Edge/NodeConstraints (:866-1204) at unit load scale, then closed.  Returns false if
// the input is outside what this restatement covers (unequal edge counts, a row out of range, a
// constraint cycle).
static bool build_periodic(const Mesh &m, const or_problem *pb, PeriodicCons &pc) {
  const or_periodic *P = pb->periodic;
  if (m.nx != m.ny || m.ny != m.nz || m.nx < 2) return false;  // one edge-node count for all edges (:49)
  if (P->n_vertex_lines > OR_MAX_PBC_LINES || P->n_edge_lines > OR_MAX_PBC_LINES || P->n_face_lines > OR_MAX_PBC_LINES) return false;
  const int n = m.nx;
  // per-increment vertex values (:108-117)
  double val[24];
  int known[24];
  const double incs = P->tabular ? (double)(unsigned)(P->tabular_time / pb->delT) : (double)pb->total_increments;
  for (int i = 0; i < 24; ++i) { val[i] = P->vertex_total[i] / incs; known[i] = P->vertex_known[i]; }
  // lines with a single unknown vertex dof fix it (:501-527); a line with all dofs known is done
  const int nV = P->n_vertex_lines;
  std::vector<int> vknown(nV, 0);
  for (int j = 0; j < nV; ++j) {
    const int cnt = P->vertex_dofs[j][0];
    if (cnt < 1 || cnt > 4) return false;
    for (int k = 0; k < cnt; ++k) if (P->vertex_dofs[j][k + 1] < 0 || P->vertex_dofs[j][k + 1] >= 24) return false;
  }
  for (int pass = 0; pass < nV; ++pass)
    for (int j = 0; j < nV; ++j) {
      const int cnt = P->vertex_dofs[j][0];
      int check1 = 0;
      for (int k = 0; k < cnt; ++k) check1 += known[P->vertex_dofs[j][k + 1]] == 1;
      if (check1 == cnt) vknown[j] = 1;
      if (check1 == cnt - 1) {
        double check2 = 0.0;
        int kf = 0;
        for (int k = 0; k < cnt; ++k) {
          const int d = P->vertex_dofs[j][k + 1];
          if (known[d] == 1) check2 = as_unsigned(check2 + val[d] * P->vertex_coef[j][k + 1]);
          else kf = k;
        }
        const int d = P->vertex_dofs[j][kf + 1];
        known[d] = 1;
        val[d] = check2 * (-1.0 / P->vertex_coef[j][kf + 1]);
        vknown[j] = 1;
      }
    }
  const long long N = m.ndof;
  std::vector<char> con(N, 0);
  std::vector<PLine> raw(N);
  auto vdof = [&](int r) -> long long { return 3LL * pbc_vertex_node(m, r / 3) + r % 3; };
  auto add_line = [&](long long d) { con[d] = 1; };  // AffineConstraints::add_line keeps an existing line
  auto add_entry = [&](long long d, long long c, double w) {
    for (auto &pr : raw[d].e) if (pr.first == c) return;  // an existing entry stays
    raw[d].e.push_back({c, w});
  };
  // the vertex terms of a face or edge line: known vertex dofs give the inhomogeneity, unknown ones
  // entries (periodicBCs.cc:945-1003, 1133-1203)
  auto vertex_terms = [&](long long d, int v1, int v2, double c0, double c2, double c3) {
    const bool f1 = known[v1] == 1, f2 = known[v2] == 1;
    if (f1 && f2) raw[d].g = (-1 / c0) * c2 * val[v1] + (-1 / c0) * c3 * val[v2];
    else if (f1) { raw[d].g = (-1 / c0) * c2 * val[v1]; add_entry(d, vdof(v2), (-1 / c0) * c3); }
    else if (f2) { raw[d].g = (-1 / c0) * c3 * val[v2]; add_entry(d, vdof(v1), (-1 / c0) * c2); }
    else { add_entry(d, vdof(v2), (-1 / c0) * c3); add_entry(d, vdof(v1), (-1 / c0) * c2); }
  };
  // faces: make_periodicity_constraints ties each dof of a positive face (cells away from the box
  // edges, :120-150: the face's interior nodes) to its partner on the negative face (:1047-1075), then
  // the vertex terms (:1077-1203)
  for (int i = 0; i < P->n_face_lines; ++i) {
    const int r = P->face_dofs[i][0];
    if (r < 0 || r >= 18 || ((r / 3) & 1) == 0) return false;  // the line starts with FXP, FYP or FZP
    for (int k = 2; k < 4; ++k) if (P->face_dofs[i][k] < 0 || P->face_dofs[i][k] >= 24) return false;
  }
  auto face_nodes = [&](int dir, int side, std::vector<int> &out) {
    out.clear();
    for (int b = 1; b < n; ++b)
      for (int a = 1; a < n; ++a) {
        const int x = side * n;
        out.push_back(dir == 0 ? m.node(x, a, b) : dir == 1 ? m.node(a, x, b) : m.node(a, b, x));
      }
  };
  std::vector<int> fp, fn;
  for (int i = 0; i < P->n_face_lines; ++i) {
    const int r = P->face_dofs[i][0], dir = (r / 3) / 2, c = r % 3;
    face_nodes(dir, 1, fp);
    face_nodes(dir, 0, fn);
    for (size_t k = 0; k < fp.size(); ++k) {
      const long long d = 3LL * fp[k] + c;
      if (con[d]) continue;
      add_line(d);
      add_entry(d, 3LL * fn[k] + c, 1.0);
    }
  }
  for (int i = 0; i < P->n_face_lines; ++i) {
    const int r = P->face_dofs[i][0], dir = (r / 3) / 2, c = r % 3;
    face_nodes(dir, 1, fp);
    for (int node : fp) {
      const long long d = 3LL * node + c;
      if (con[d]) vertex_terms(d, P->face_dofs[i][2], P->face_dofs[i][3], P->face_coef[i][0], P->face_coef[i][2], P->face_coef[i][3]);
    }
  }
  // edges (:922-1004): edge row E1 node t = -(c1/c0) E2 node t + vertex terms
  for (int j = 0; j < P->n_edge_lines; ++j) {
    const int *r = P->edge_dofs[j];
    const double *c = P->edge_coef[j];
    if (r[0] < 0 || r[0] >= 36 || r[1] < 0 || r[1] >= 36 || r[2] < 0 || r[2] >= 24 || r[3] < 0 || r[3] >= 24) return false;
    for (int t = 0; t < n - 1; ++t) {
      const long long d1 = 3LL * pbc_edge_node(m, r[0] / 3, t) + r[0] % 3, d2 = 3LL * pbc_edge_node(m, r[1] / 3, t) + r[1] % 3;
      add_line(d1);
      add_entry(d1, d2, (-1 / c[0]) * c[1]);
      vertex_terms(d1, r[2], r[3], c[0], c[2], c[3]);
    }
  }
  // vertices (:866-920): a line with unknown dofs constrains its first dof; known dofs take their value
  for (int j = 0; j < nV; ++j) {
    if (vknown[j]) continue;
    const int cnt = P->vertex_dofs[j][0];
    const long long d1 = vdof(P->vertex_dofs[j][1]);
    if (known[P->vertex_dofs[j][1]] == 1) return false;  // the known value would overwrite the line's inhomogeneity
    const double c1 = P->vertex_coef[j][1];
    add_line(d1);
    double inh = 0;
    for (int k = 1; k < cnt; ++k) {
      const int rv = P->vertex_dofs[j][k + 1];
      if (known[rv] == 1) inh += P->vertex_coef[j][k + 1] * val[rv];
      else add_entry(d1, vdof(rv), (-1 / c1) * P->vertex_coef[j][k + 1]);
    }
    raw[d1].g = (-1 / c1) * inh;
  }
  for (int i = 0; i < 24; ++i)
    if (known[i] == 1) { const long long d = vdof(i); add_line(d); raw[d].g = val[i]; }
  // AffineConstraints::close(): entries that are themselves constrained are replaced by their lines
  pc.con = con;
  pc.line.assign(N, PLine());
  std::vector<char> state(N, 0);  // 0 open, 1 in progress, 2 closed
  std::function<bool(long long)> close = [&](long long d) -> bool {
    if (state[d] == 2) return true;
    if (state[d] == 1) return false;  // cycle
    state[d] = 1;
    PLine out;
    out.g = raw[d].g;
    std::map<long long, double> acc;
    for (auto &pr : raw[d].e) {
      if (con[pr.first]) {
        if (!close(pr.first)) return false;
        const PLine &l = pc.line[pr.first];
        out.g += pr.second * l.g;
        for (auto &q : l.e) acc[q.first] += pr.second * q.second;
      } else acc[pr.first] += pr.second;
    }
    for (auto &q : acc) out.e.push_back(q);
    pc.line[d] = out;
    state[d] = 2;
    return true;
  };
  for (long long d = 0; d < N; ++d) if (con[d] && !close(d)) return false;
  // tabular periodic BCs (setPeriodicityConstraints, :1276-1310): the coefficient of the table
  // segment of each increment's first Newton iteration; the counter moves at most once per increment
  if (P->tabular) {
    pc.tab_scale.assign(pb->total_increments, 0.0);
    int tc = 1;
    for (int inc = 0; inc < pb->total_increments; ++inc) {
      const double t = pb->delT * (inc + 1) - pb->delT / 1000;
      if (inc == 0) tc = 1;
      if (tc < P->n_table && t > P->table_time[tc]) ++tc;
      pc.tab_scale[inc] = (tc - 1 < P->n_table) ? (double)P->table_coef[tc - 1] : 0.0;
    }
  }
  return true;
}

// Load scale of the constraints at Newton iteration `it` of increment `inc`
static double periodic_scale(const PeriodicCons &pc, const or_problem *pb, int inc, int it) {
  if (it != 0) return 0.0;
  if (pb->periodic->tabular) return pc.tab_scale[inc];
  return 1.0;
}

// One Newton system under the closed periodic constraints: distribute_local_to_global of every
// element (condensed matrix C^T K C and right-hand side C^T (R_e - K_e s g) on the unconstrained
// rows; a constrained row keeps |K_ii| on its diagonal, or the mean |diag| of the element), the
// Krylov solve (solveLinearSystem.cc:13-31) and constraints.distribute (:32).  Returns ||rhs||.
static double periodic_newton_solve(const Mesh &m, const PeriodicCons &pc, double sc, const std::vector<double> &Kbuf,
                                    const std::vector<double> &Rbuf, const or_problem *pb, std::vector<double> &du, int *iters) {
  const long long N = m.ndof;
  std::vector<std::map<long long, double>> J(N);
  std::vector<double> b(N, 0.0);
  std::vector<std::pair<long long, double>> ex[24];
  for (long long e = 0; e < m.nelem; ++e) {
    int dofs[24];
    m.elem_dofs(e, dofs);
    const double *Ke = &Kbuf[576 * e], *Re = &Rbuf[24 * e];
    for (int i = 0; i < 24; ++i) {
      ex[i].clear();
      if (pc.con[dofs[i]]) ex[i] = pc.line[dofs[i]].e;
      else ex[i].push_back({dofs[i], 1.0});
    }
    double avg = 0;
    for (int i = 0; i < 24; ++i) avg += std::fabs(Ke[i * 24 + i]);
    avg /= 24.0;
    for (int i = 0; i < 24; ++i) {
      double ri = Re[i];
      for (int j = 0; j < 24; ++j)
        if (pc.con[dofs[j]] && sc != 0.0 && pc.line[dofs[j]].g != 0.0) ri -= Ke[i * 24 + j] * sc * pc.line[dofs[j]].g;
      for (auto &a : ex[i]) {
        b[a.first] += a.second * ri;
        for (int j = 0; j < 24; ++j)
          for (auto &c : ex[j]) J[a.first][c.first] += a.second * Ke[i * 24 + j] * c.second;
      }
      if (pc.con[dofs[i]]) {
        const double nd = std::fabs(Ke[i * 24 + i]) != 0 ? std::fabs(Ke[i * 24 + i]) : avg;
        J[dofs[i]][dofs[i]] += nd;
      }
    }
  }
  CSR A;
  A.rp.assign(N + 1, 0);
  for (long long r = 0; r < N; ++r) {
    J[r][r] += 0.0;  // every row holds its diagonal
    A.rp[r + 1] = A.rp[r] + (long long)J[r].size();
  }
  A.ci.resize(A.rp[N]);
  A.val.resize(A.rp[N]);
  for (long long r = 0; r < N; ++r) {
    long long k = A.rp[r];
    for (auto &pr : J[r]) { A.ci[k] = (int)pr.first; A.val[k] = pr.second; ++k; }
  }
  const double rn = std::sqrt(dot(b, b));
  std::vector<double> dx;
  const int its = (pb->solver == 1) ? pcg(A, b, dx, pb->lin_tol * rn, pb->lin_maxit) : bicg(A, b, dx, pb->lin_tol * rn, pb->lin_maxit);
  if (iters) *iters = its;
  du.assign(N, 0.0);
  for (long long d = 0; d < N; ++d) {
    if (!pc.con[d]) { du[d] = dx[d]; continue; }
    double v = sc * pc.line[d].g;
    for (auto &pr : pc.line[d].e) v += pr.second * dx[pr.first];
    du[d] = v;
  }
  return rn;
}

long long or_periodic_lines(const or_problem *pb, int *constrained, long long *rowptr, long long *master, double *weight,
                            double *g, long long cap) {
  if (!pb->periodic) return -1;
  const Mesh m = make_mesh(pb);
  PeriodicCons pc;
  if (!build_periodic(m, pb, pc)) return -1;
  long long k = 0;
  for (long long d = 0; d < m.ndof; ++d) {
    if (constrained) constrained[d] = pc.con[d];
    if (rowptr) rowptr[d] = k;
    if (g) g[d] = pc.con[d] ? pc.line[d].g : 0.0;
    if (!pc.con[d]) continue;
    for (auto &pr : pc.line[d].e) {
      if (k < cap) { master[k] = pr.first; weight[k] = pr.second; }
      ++k;
    }
  }
  if (rowptr) rowptr[m.ndof] = k;
  return k;
}

// ======================================================================= C API
extern "C" {

void or_tangent_perturbation(int block, double rel) {
  g_pert_block = block;
  g_pert_rel = rel;
}

int or_qp_update(const or_material *mat, const double rot[3], const double F[9], const double Fp_conv[9],
                 const double *s_conv, const double *W_conv, const double *slip_conv, int flag, double P[9], double T[9],
                 double dPdF[81], double Fe_iter[9], double Fp_iter[9], double *s_iter, double *W_iter, double *slip_iter) {
  Grain g;
  grain_setup(mat, rot, g);
  return qp_update(mat, g, F, Fp_conv, s_conv, W_conv, slip_conv, flag, P, T, dPdF, Fe_iter, Fp_iter, s_iter, W_iter, slip_iter);
}

int or_twin_event(const or_material *mat, double F_r, double F_e, const double *tw_iter, int twin_conv,
                  const double rot_conv[3], double rot_iter[3], int *twin_iter, double *s_iter) {
  return twin_event(mat, F_r, F_e, tw_iter, twin_conv, rot_conv, rot_iter, twin_iter, s_iter);
}

int or_qp_update_twin(const or_material *mat, const double rot[3], const double F[9], const double Fp_conv[9],
                      const double *s_conv, const double *W_conv, const double *slip_conv, const double *tw_conv, int flag,
                      double P[9], double T[9], double dPdF[81], double Fe_iter[9], double Fp_iter[9], double *s_iter,
                      double *W_iter, double *slip_iter, double *tw_iter) {
  Grain g;
  grain_setup(mat, rot, g);
  return qp_update(mat, g, F, Fp_conv, s_conv, W_conv, slip_conv, flag, P, T, dPdF, Fe_iter, Fp_iter, s_iter, W_iter, slip_iter,
                   nullptr, nullptr, tw_conv, tw_iter);
}

int or_qp_update_warm(const or_material *mat, const double rot[3], const double F[9], const double Fp_conv[9],
                      const double *s_conv, const double *W_conv, const double *slip_conv, int flag, const double *x0,
                      double P[9], double T[9], double dPdF[81], double Fe_iter[9], double Fp_iter[9], double *s_iter,
                      double *W_iter, double *slip_iter, int *iters) {
  Grain g;
  grain_setup(mat, rot, g);
  return qp_update(mat, g, F, Fp_conv, s_conv, W_conv, slip_conv, flag, P, T, dPdF, Fe_iter, Fp_iter, s_iter, W_iter,
                   slip_iter, x0, iters);
}

void or_pinv_stats(long long *calls, long long *singular) {
  *calls = g_pinv_calls; *singular = g_pinv_singular;
  g_pinv_calls = g_pinv_singular = 0;
}

void or_set_pinv_rtol(double rtol) { kPinvRtol = rtol; }
void or_last_timings(long long ns[3]) { for (int i = 0; i < 3; ++i) ns[i] = g_phase_ns[i]; }

int or_pinv_records(double *out, int max_records) {
  const int n = std::min((int)(g_records.size() / 4), max_records);
  for (int i = 0; i < 4 * n; ++i) out[i] = g_records[i];
  g_records.clear();
  return n;
}

void or_pinv_histogram(int enable, long long hist[24]) {
  if (hist) for (int k = 0; k < 24; ++k) { hist[k] = g_hist[k]; g_hist[k] = 0; }
  g_diag = enable;
}

int or_pinv_sample(double *A, double *b) {
  for (int i = 0; i < g_sample_n * g_sample_n; ++i) A[i] = g_sample_A[i];
  for (int i = 0; i < g_sample_n; ++i) b[i] = g_sample_b[i];
  return g_sample_n;
}

void or_odfpoint(const double r[3], double R[9]) { store3(odfpoint(r), R); }
void or_matrix_exponential(const double A[9], double E[9]) { store3(expm3(load3(A)), E); }
void or_polar_rotation(const double Fe[9], double R[9]) { store3(polar_R(load3(Fe)), R); }
int or_pinv_solve(int n, const double *A, const double *b, double *x, int mode) { pinv_solve(n, A, b, x, mode); return 0; }

// getElementalValues for one element at the converged state (F from ue), K_e and R_e
// (getElementalValues.cc:127-163).
int or_element_matrix(const or_material *mat, const double rot[3], const double h[3], const double ue[24],
                      const double *Fp_conv, const double *s_conv, double Ke[576], double Re[24]) {
  Grain g;
  grain_setup(mat, rot, g);
  const Shape sh = make_shape(h);
  const int nT = mat->nT;
  std::vector<double> W(nT, 0.0), s_it(nT), W_it(nT), slip(mat->n_slip, 0.0), slip_it(mat->n_slip, 0.0);
  std::memset(Ke, 0, 576 * sizeof(double));
  std::memset(Re, 0, 24 * sizeof(double));
  for (int q = 0; q < 8; ++q) {
    const M3 F = def_grad(sh, q, ue);
    double Fv[9], P[9], T[9], D[81], Fe[9], Fp[9];
    store3(F, Fv);
    qp_update(mat, g, Fv, Fp_conv + 9 * q, s_conv + nT * q, W.data(), slip.data(), 1, P, T, D, Fe, Fp, s_it.data(), W_it.data(), slip_it.data());
    for (int d = 0; d < 24; ++d) {
      const int i = d % 3, v = d / 3;
      for (int j = 0; j < 3; ++j) Re[d] -= sh.grad[q][v][j] * P[3 * i + j] * sh.JxW;
    }
    for (int d1 = 0; d1 < 24; ++d1) {
      const int i = d1 % 3, v1 = d1 / 3;
      for (int d2 = 0; d2 < 24; ++d2) {
        const int j = d2 % 3, v2 = d2 / 3;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) Ke[d1 * 24 + d2] += sh.grad[q][v1][k] * D[(3 * i + k) * 9 + 3 * j + l] * sh.grad[q][v2][l] * sh.JxW;
      }
    }
  }
  return 0;
}

unsigned or_material_id(const unsigned *vox, const int nv[3], const double span[3], const double p[3]) {
  // keys are voxel centres x*stencil + stencil/2 (crystalOrientationsIO.cc:77-105); the search
  // below mirrors std::map::lower_bound and the tie rule "dist_to_upper >= dist_to_lower -> lower"
  int idx[3];
  for (int d = 0; d < 3; ++d) {
    const double st = span[d] / (nv[d]);
    std::vector<double> keys(nv[d]);
    for (int x = 0; x < nv[d]; ++x) keys[x] = x * st + st / 2;
    const int up = (int)(std::lower_bound(keys.begin(), keys.end(), p[d]) - keys.begin());
    int sel;
    if (up == nv[d]) sel = up - 1;
    else if (up == 0) sel = 0;
    else {
      const double dl = std::abs(keys[up - 1] - p[d]), du = std::abs(keys[up] - p[d]);
      sel = (du >= dl) ? up - 1 : up;
    }
    idx[d] = sel;
  }
  return vox[((long long)idx[0] * nv[1] + idx[1]) * nv[2] + idx[2]];
}

void or_element_grains(const unsigned *vox, const int nv[3], const or_problem *pb, unsigned *elem_grain) {
  const Mesh m = make_mesh(pb);
  for (long long e = 0; e < m.nelem; ++e) {
    const int ex = (int)(e % m.nx), ey = (int)((e / m.nx) % m.ny), ez = (int)(e / ((long long)m.nx * m.ny));
    // deal.II cell->center(): mean of the 8 vertices
    const double xs[2][3] = {{ex * m.h[0], ey * m.h[1], ez * m.h[2]}, {(ex + 1) * m.h[0], (ey + 1) * m.h[1], (ez + 1) * m.h[2]}};
    double c[3];
    for (int d = 0; d < 3; ++d) c[d] = (xs[0][d] + xs[1][d]) / 2;
    elem_grain[e] = or_material_id(vox, nv, pb->span, c);
  }
}

long long or_sparsity(const or_problem *pb, long long *row_ptr, int *col_idx) {
  const Mesh m = make_mesh(pb);
  CSR A = make_csr(m);
  if (row_ptr) std::copy(A.rp.begin(), A.rp.end(), row_ptr);
  if (col_idx) std::copy(A.ci.begin(), A.ci.end(), col_idx);
  return A.rp[m.ndof];
}

static double *g_dump_K = nullptr, *g_dump_R = nullptr, *g_dump_du = nullptr;
static int g_dump_inc = 0, g_dump_it = 0;  // the Newton system that is dumped

int or_krylov(int method, long long n, const long long *rp, const int *ci, const double *val, const double *b,
              double *x, double tol, int maxit) {
  CSR A;
  A.rp.assign(rp, rp + n + 1);
  A.ci.assign(ci, ci + rp[n]);
  A.val.assign(val, val + rp[n]);
  std::vector<double> bv(b, b + n), xv;
  const int it = (method == 1) ? pcg(A, bv, xv, tol, maxit) : bicg(A, bv, xv, tol, maxit);
  std::copy(xv.begin(), xv.end(), x);
  return it;
}

int or_newton_system(const or_material *mat, const or_problem *pb, const unsigned *elem_grain, const double *grain_rod,
                     int n_grain_rows, int inc, int it, double *K_vals, double *R, double *du) {
  if (inc < 0 || it < 0 || it >= pb->max_newton) return 1;
  if (pb->periodic) return 2;  // the condensed periodic matrix has another pattern
  g_dump_K = K_vals; g_dump_R = R; g_dump_du = du;
  g_dump_inc = inc; g_dump_it = it;
  or_problem p1 = *pb;
  p1.n_increments = inc + 1;
  std::vector<double> rows(15 * (size_t)(inc + 1));
  const int rc = or_simulate(mat, &p1, elem_grain, grain_rod, n_grain_rows, rows.data(), nullptr, nullptr, nullptr, nullptr);
  g_dump_K = g_dump_R = g_dump_du = nullptr;
  g_dump_inc = g_dump_it = 0;
  return rc;
}

int or_first_newton(const or_material *mat, const or_problem *pb, const unsigned *elem_grain, const double *grain_rod,
                    int n_grain_rows, double *K_vals, double *R, double *du) {
  if (pb->periodic) return 2;  // the condensed periodic matrix has another pattern
  g_dump_K = K_vals; g_dump_R = R; g_dump_du = du;
  g_dump_inc = g_dump_it = 0;
  or_problem p1 = *pb;
  p1.n_increments = 1;
  p1.max_newton = 1;
  double rows[15];
  const int rc = or_simulate(mat, &p1, elem_grain, grain_rod, n_grain_rows, rows, nullptr, nullptr, nullptr, nullptr);
  g_dump_K = g_dump_R = g_dump_du = nullptr;
  return rc;
}

// or_simulate over nph phases: grain_phase (1-based, per grain row; null = one phase) selects the
// material of each grain (multiphaseInit.cc:68-361).  Per-qp state arrays have the largest phase's
// size; each qp uses its phase's leading part.
int simulate(const or_material *mats, int nph, const int *grain_phase, const or_problem *pb, const unsigned *elem_grain,
             const double *grain_rod, int n_grain_rows, double *rows, int *lin_iters, double *res_norms, double *state_out,
             double *u_out) {
#ifdef _OPENMP
  if (pb->nthreads > 0) omp_set_num_threads(pb->nthreads);
#endif
  const Mesh m = make_mesh(pb);
  const Shape sh = make_shape(m.h);
  const bool multiphase = grain_phase != nullptr;
  int nT = 0, ns = 0, ntw = 0;
  bool twinning = false;
  for (int k = 0; k < nph; ++k) {
    nT = std::max(nT, mats[k].nT);
    ns = std::max(ns, mats[k].n_slip);
    if (mats[k].twinning) { twinning = true; ntw = std::max(ntw, mats[k].n_twin); }
  }
  const long long nq = m.nelem * 8;
  g_phase_ns[0] = g_phase_ns[1] = g_phase_ns[2] = 0;
  for (long long e = 0; e < m.nelem; ++e) {
    if ((int)elem_grain[e] >= n_grain_rows) return -1;
    if (multiphase && (grain_phase[elem_grain[e]] < 1 || grain_phase[elem_grain[e]] > nph)) return -1;
  }
  auto gmat = [&](int g) -> const or_material * { return mats + (multiphase ? grain_phase[g] - 1 : 0); };
  auto emat = [&](long long e) -> const or_material * { return gmat((int)elem_grain[e]); };
  // per-grain constants, identical to what calculatePlasticity rebuilds on every call
  std::vector<Grain> grains(n_grain_rows);
  std::vector<char> used(n_grain_rows, 0);
  for (long long e = 0; e < m.nelem; ++e) used[elem_grain[e]] = 1;
  for (int g = 0; g < n_grain_rows; ++g) if (used[g]) grain_setup(gmat(g), grain_rod + 3 * g, grains[g]);

  QPState st;  // crystalPlasticity::init (CP init.cc:302-330)
  st.Fp_conv.assign(9 * nq, 0.0); st.Fe_conv.assign(9 * nq, 0.0);
  for (long long q = 0; q < nq; ++q) for (int i = 0; i < 3; ++i) { st.Fp_conv[9 * q + 4 * i] = 1.0; st.Fe_conv[9 * q + 4 * i] = 1.0; }
  st.Fp_iter = st.Fp_conv; st.Fe_iter = st.Fe_conv;
  st.s_conv.assign(nT * nq, 0.0);
  for (long long q = 0; q < nq; ++q) {
    const or_material *mq = emat(q / 8);
    for (int a = 0; a < mq->nT; ++a) st.s_conv[nT * q + a] = mq->s0[a];
  }
  st.s_iter = st.s_conv;
  st.W_conv.assign(nT * nq, 0.0); st.W_iter = st.W_conv;
  st.slip_conv.assign(ns * nq, 0.0); st.slip_iter = st.slip_conv;
  st.T.assign(9 * nq, 0.0);
  st.rotnew_conv.resize(3 * nq);
  for (long long e = 0; e < m.nelem; ++e)
    for (int q = 0; q < 8; ++q) for (int i = 0; i < 3; ++i) st.rotnew_conv[3 * (8 * e + q) + i] = grain_rod[3 * elem_grain[e] + i];
  double F_r = 0.0, F_e = 0.0;  // volume-averaged twin fraction / reoriented volume of the last increment
  // the averages sum the first numTwinSystems1 twin and numSlipSystems1 slip fractions of every qp,
  // phase 1's counts whatever the qp's phase (updateAfterIncrement.cc:222-236); with several phases
  // the sums are then discarded (:722-725)
  const int nTw1 = mats[0].twinning ? mats[0].n_twin : 1, nSl1 = mats[0].n_slip;
  std::vector<Grain> twin_grains;
  if (twinning) {
    st.rot_conv = st.rotnew_conv; st.rot_iter = st.rotnew_conv; st.rotnew_iter = st.rotnew_conv;
    st.tw_conv.assign((size_t)ntw * nq, 0.0); st.tw_iter = st.tw_conv;
    st.twin_conv.assign(nq, 0); st.twin_iter = st.twin_conv;
    st.tgrain.assign(nq, -1);
  }
  // the orientation tables of qp Q: its grain's, or its own once it has twinned
  auto qgrain = [&](long long e, long long Q) -> const Grain & {
    return (twinning && st.tgrain[Q] >= 0) ? twin_grains[st.tgrain[Q]] : grains[elem_grain[e]];
  };
  // reorientation by twinning after a qp update (calculatePlasticity.cc:962-1019)
  auto twin_step = [&](long long Q) {
    if (!twinning) return;
    if (twin_event(emat(Q / 8), F_r, F_e, &st.tw_iter[(size_t)ntw * Q], st.twin_conv[Q], &st.rot_conv[3 * Q], &st.rot_iter[3 * Q],
                   &st.twin_iter[Q], &st.s_iter[(size_t)nT * Q], multiphase) >= 0)
      for (int i = 0; i < 3; ++i) st.rotnew_iter[3 * Q + i] = st.rot_iter[3 * Q + i];
  };

  const bool periodic = pb->periodic != nullptr;
  PeriodicCons pcons;
  if (periodic && !build_periodic(m, pb, pcons)) return -3;
  CSR A = make_csr(m);
  std::vector<double> x(m.ndof, 0.0), R(m.ndof), du;
  std::vector<double> Kbuf(576 * m.nelem), Rbuf(24 * m.nelem);
  std::vector<char> con;
  std::vector<double> delu;
  double vol_total = 0;

  for (int inc = 0; inc < pb->n_increments; ++inc) {
    if (periodic) { con.assign(m.ndof, 0); delu.assign(m.ndof, 0.0); }
    else dirichlet(m, pb, inc, con, delu);
    for (int it = 0; it < pb->max_newton; ++it) {
      OR_FLOP_STEP(inc, it, pb->max_newton);
      const long long t_asm = now_ns();
      // assemble (BVP assemble.cc:6-104) with getElementalValues (CP getElementalValues.cc:5-168)
#pragma omp parallel for schedule(dynamic, 64)
      for (long long e = 0; e < m.nelem; ++e) {
        int dofs[24];
        m.elem_dofs(e, dofs);
        double ue[24];
        for (int d = 0; d < 24; ++d) ue[d] = x[dofs[d]];
        double *Ke = &Kbuf[576 * e], *Re = &Rbuf[24 * e];
        std::memset(Ke, 0, 576 * sizeof(double));
        std::memset(Re, 0, 24 * sizeof(double));
        for (int q = 0; q < 8; ++q) {
          const long long Q = 8 * e + q;
          const Grain &g = qgrain(e, Q);
          const M3 F = def_grad(sh, q, ue);
          double Fv[9], P[9], T[9], D[81];
          store3(F, Fv);
          qp_update(emat(e), g, Fv, &st.Fp_conv[9 * Q], &st.s_conv[nT * Q], &st.W_conv[nT * Q], &st.slip_conv[ns * Q], 1, P, T, D,
                    &st.Fe_iter[9 * Q], &st.Fp_iter[9 * Q], &st.s_iter[nT * Q], &st.W_iter[nT * Q], &st.slip_iter[ns * Q], nullptr,
                    nullptr, twinning ? &st.tw_conv[(size_t)ntw * Q] : nullptr, twinning ? &st.tw_iter[(size_t)ntw * Q] : nullptr);
          twin_step(Q);
          for (int d = 0; d < 24; ++d) {
            const int i = d % 3, v = d / 3;
            for (int j = 0; j < 3; ++j) Re[d] -= sh.grad[q][v][j] * P[3 * i + j] * sh.JxW;
          }
          for (int d1 = 0; d1 < 24; ++d1) {
            const int i = d1 % 3, v1 = d1 / 3;
            for (int d2 = 0; d2 < 24; ++d2) {
              const int j = d2 % 3, v2 = d2 / 3;
              double s = Ke[d1 * 24 + d2];
              for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) s += sh.grad[q][v1][k] * D[(3 * i + k) * 9 + 3 * j + l] * sh.grad[q][v2][l] * sh.JxW;
              Ke[d1 * 24 + d2] = s;
            }
          }
        }
      }
      if (periodic) {
        const double sc = periodic_scale(pcons, pb, inc, it);
        const double rn = periodic_newton_solve(m, pcons, sc, Kbuf, Rbuf, pb, du,
                                                (lin_iters ? &lin_iters[inc * pb->max_newton + it] : nullptr));
        if (res_norms) res_norms[inc * pb->max_newton + it] = rn;
        for (long long r = 0; r < m.ndof; ++r) x[r] += du[r];
        g_phase_ns[1] += now_ns() - t_asm;
        continue;
      }
      // AffineConstraints::distribute_local_to_global with inhomogeneous Dirichlet lines:
      // free rows get K_ff and R - K_fc g; constrained rows only a diagonal |K_ii| (or the mean |diag|).
      std::fill(A.val.begin(), A.val.end(), 0.0);
      std::fill(R.begin(), R.end(), 0.0);
      for (long long e = 0; e < m.nelem; ++e) {
        int dofs[24];
        m.elem_dofs(e, dofs);
        const double *Ke = &Kbuf[576 * e], *Re = &Rbuf[24 * e];
        bool anycon = false;
        for (int d = 0; d < 24; ++d) anycon |= con[dofs[d]] != 0;
        for (int i = 0; i < 24; ++i) {
          const int gi = dofs[i];
          if (con[gi]) continue;
          double val = Re[i];
          for (int j = 0; j < 24; ++j)
            if (con[dofs[j]] && it == 0 && delu[dofs[j]] != 0.0) val -= Ke[i * 24 + j] * delu[dofs[j]];
          if (val != 0.0) R[gi] += val;
          for (int j = 0; j < 24; ++j) {
            if (con[dofs[j]]) continue;
            A.val[A.find(gi, dofs[j])] += Ke[i * 24 + j];
          }
        }
        if (anycon) {
          double avg = 0;
          for (int i = 0; i < 24; ++i) avg += std::fabs(Ke[i * 24 + i]);
          avg /= 24.0;
          for (int i = 0; i < 24; ++i) {
            const int gi = dofs[i];
            if (!con[gi]) continue;
            const double nd = std::fabs(Ke[i * 24 + i]) != 0 ? std::fabs(Ke[i * 24 + i]) : avg;
            A.val[A.find(gi, gi)] += nd;
          }
        }
      }
      const bool dump = inc == g_dump_inc && it == g_dump_it;
      if (g_dump_K && dump) std::copy(A.val.begin(), A.val.end(), g_dump_K);
      if (g_dump_R && dump) std::copy(R.begin(), R.end(), g_dump_R);
      const double rn = std::sqrt(dot(R, R));
      const long long t_solve = now_ns();
      g_phase_ns[0] += t_solve - t_asm;
      if (res_norms) res_norms[inc * pb->max_newton + it] = rn;
      const int its = (pb->solver == 1) ? pcg(A, R, du, pb->lin_tol * rn, pb->lin_maxit) : bicg(A, R, du, pb->lin_tol * rn, pb->lin_maxit);
      if (lin_iters) lin_iters[inc * pb->max_newton + it] = its;
      for (long long r = 0; r < m.ndof; ++r) if (con[r]) du[r] = (it == 0) ? delu[r] : 0.0;  // constraints.distribute
      if (g_dump_du && dump) std::copy(du.begin(), du.end(), g_dump_du);
      for (long long r = 0; r < m.ndof; ++r) x[r] += du[r];
      g_phase_ns[1] += now_ns() - t_solve;
    }
    const long long t_post = now_ns();
    // updateAfterIncrement (CP updateAfterIncrement.cc:6-305, 710-771)
    double sT[3][3] = {{0}}, sE[3][3] = {{0}}, Fs = 0, Fr_sum = 0, Fe_sum = 0;
    vol_total = 0;
    for (long long e = 0; e < m.nelem; ++e) {
      int dofs[24];
      m.elem_dofs(e, dofs);
      double ue[24];
      for (int d = 0; d < 24; ++d) ue[d] = x[dofs[d]];
      for (int q = 0; q < 8; ++q) {
        const long long Q = 8 * e + q;
        const Grain &g = qgrain(e, Q);
        const M3 F = def_grad(sh, q, ue);
        double Fv[9], P[9], D[81];
        store3(F, Fv);
        qp_update(emat(e), g, Fv, &st.Fp_conv[9 * Q], &st.s_conv[nT * Q], &st.W_conv[nT * Q], &st.slip_conv[ns * Q], 0, P, &st.T[9 * Q], D,
                  &st.Fe_iter[9 * Q], &st.Fp_iter[9 * Q], &st.s_iter[nT * Q], &st.W_iter[nT * Q], &st.slip_iter[ns * Q], nullptr,
                  nullptr, twinning ? &st.tw_conv[(size_t)ntw * Q] : nullptr, twinning ? &st.tw_iter[(size_t)ntw * Q] : nullptr);
        twin_step(Q);
        if (twinning) {  // updateAfterIncrement.cc:176, 222-228: the first numTwinSystems1 fractions (phase 1's count, 1 without its twinning)
          for (int i = 0; i < std::min(nTw1, ntw); ++i) Fr_sum += st.tw_iter[(size_t)ntw * Q + i] * sh.JxW;
          Fe_sum += st.twin_iter[Q] * sh.JxW;
        }
        const M3 C = tmul(F, F);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const double E = 0.5 * (C[i][j] - (i == j ? 1.0 : 0.0));
            sT[i][j] += st.T[9 * Q + 3 * i + j] * sh.JxW;
            sE[i][j] += E * sh.JxW;
          }
        vol_total += sh.JxW;
        for (int a = 0; a < nSl1; ++a) Fs += st.slip_iter[ns * Q + a] * sh.JxW;  // :234-236: phase 1's slip count
      }
    }
    // reorient (CP reorient.cc) then the _conv <- _iter commit (:287-305); with twinning the output
    // orientation first takes the twinned one (rotnew_conv = rotnew_iter, :287)
    if (twinning) st.rotnew_conv = st.rotnew_iter;
    for (long long Q = 0; Q < nq; ++Q) reorient_qp(load3(&st.Fe_conv[9 * Q]), load3(&st.Fe_iter[9 * Q]), &st.rotnew_conv[3 * Q]);
    st.Fe_conv = st.Fe_iter; st.Fp_conv = st.Fp_iter; st.s_conv = st.s_iter; st.W_conv = st.W_iter; st.slip_conv = st.slip_iter;
    if (twinning) {
      st.rotnew_iter = st.rotnew_conv;                    // :293
      st.tw_conv = st.tw_iter;                            // :302-305
      for (long long Q = 0; Q < nq; ++Q)
        if (st.twin_iter[Q] && !st.twin_conv[Q]) {        // newly twinned: its own orientation tables
          st.tgrain[Q] = (int)twin_grains.size();
          twin_grains.emplace_back();
          grain_setup(emat(Q / 8), &st.rot_iter[3 * Q], twin_grains.back());
        }
      st.rot_conv = st.rot_iter;
      st.twin_conv = st.twin_iter;
      F_r = Fr_sum / vol_total;                           // :719-720
      F_e = Fe_sum / vol_total;
      // with Enable Multiphase both are set to zero instead (:722-725), so calculatePlasticity's
      // F_r > 0 test (:991-993) never lets a qp reorient by twinning
      if (multiphase) F_r = F_e = 0.0;
    }
    double *row = rows + 15 * inc;
    const int vi[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    for (int c = 0; c < 6; ++c) { row[c] = sE[vi[c][0]][vi[c][1]] / vol_total; row[6 + c] = sT[vi[c][0]][vi[c][1]] / vol_total; }
    g_phase_ns[2] += now_ns() - t_post;
    row[12] = twinning ? F_r : 0.0;  // TwinRealVF (without twinning the inert system never activates)
    row[13] = twinning ? F_e : 0.0;  // TwinMade
    row[14] = multiphase ? 0.0 : Fs / vol_total;  // F_s = 0 with Enable Multiphase (:722-725)
  }
  if (state_out) {
    const int W = 9 + 9 + 9 + nT + 3 + ns + (twinning ? 4 + ntw : 0);
    for (long long Q = 0; Q < nq; ++Q) {
      double *o = state_out + W * Q;
      std::copy(&st.T[9 * Q], &st.T[9 * Q] + 9, o);
      std::copy(&st.Fp_conv[9 * Q], &st.Fp_conv[9 * Q] + 9, o + 9);
      std::copy(&st.Fe_conv[9 * Q], &st.Fe_conv[9 * Q] + 9, o + 18);
      std::copy(&st.s_conv[nT * Q], &st.s_conv[nT * Q] + nT, o + 27);
      std::copy(&st.rotnew_conv[3 * Q], &st.rotnew_conv[3 * Q] + 3, o + 27 + nT);
      std::copy(&st.slip_conv[ns * Q], &st.slip_conv[ns * Q] + ns, o + 30 + nT);
      if (twinning) {
        std::copy(&st.rot_conv[3 * Q], &st.rot_conv[3 * Q] + 3, o + 30 + nT + ns);
        o[33 + nT + ns] = st.twin_conv[Q];
        std::copy(&st.tw_conv[(size_t)ntw * Q], &st.tw_conv[(size_t)ntw * Q] + ntw, o + 34 + nT + ns);
      }
    }
  }
  if (u_out) std::copy(x.begin(), x.end(), u_out);
  return 0;
}

int or_simulate(const or_material *mat, const or_problem *pb, const unsigned *elem_grain, const double *grain_rod,
                int n_grain_rows, double *rows, int *lin_iters, double *res_norms, double *state_out, double *u_out) {
  return simulate(mat, 1, nullptr, pb, elem_grain, grain_rod, n_grain_rows, rows, lin_iters, res_norms, state_out, u_out);
}

int or_simulate_phases(const or_material *mats, int n_phases, const int *grain_phase, const or_problem *pb,
                       const unsigned *elem_grain, const double *grain_rod, int n_grain_rows, double *rows, int *lin_iters,
                       double *res_norms, double *state_out, double *u_out) {
  if (n_phases < 1 || n_phases > 4 || !grain_phase) return -1;
  return simulate(mats, n_phases, grain_phase, pb, elem_grain, grain_rod, n_grain_rows, rows, lin_iters, res_norms, state_out,
                  u_out);
}

}  // extern "C"

// Wave-cooperative quadrature-point update (FP64), device side: calculatePlasticity(cellID, q,
// flag) of src/materialModels/crystalPlasticity/calculatePlasticity.cc:13-959 for one qp held by
// a group of G lanes of one wavefront (G = 16 for up to 16 slip systems, 32 for up to 32, 64 for
// up to 64).
//
// Work split inside a group:
//   * lane a owns slip system a: its trial resolved shear, slip resistance, hardening, slip
//     increment, column a of the active-set matrix A (:273-306) and of the tangent's A2
//     (:622-709), and row a of the pivoted LU that solves either;
//   * lane kl < 9 owns column kl of the tangent's right-hand sides delb (:581-604) and of
//     dP/dF (:835-940);
//   * the 3x3 kinematics (trial state, the exponential update of Fp, stresses) are computed by
//     every lane of the group on identical data.
// Per-system vectors and the NS x NS matrices live in one LDS slice per group; registers hold
// only statically indexed 3x3 arrays, so a qp keeps no private (scratch) memory.  Control flow is
// uniform within a group: every branch depends on a ballot mask, a group reduction, or values
// that all lanes of the group compute from the same data.  The arithmetic of each entry is the
// one-lane routine's (cp_qp.h qp_update): the same operations in the same order, so both agree
// to rounding; the singular branch (the pseudo-inverse of inactiveSlipRemoval.cc:26,99) runs as a
// group-cooperative one-sided Jacobi SVD in the one-lane routine's round-robin order.
#pragma once
#include <hip/hip_runtime.h>

#include "cp_qp.h"

namespace coop {

template <int NS>
constexpr int group_size() { return NS <= 16 ? 16 : (NS <= 32 ? 32 : 64); }

// Orders a group's LDS writes before its next reads.  The lanes of a group are in one wavefront,
// which executes its LDS instructions in order; the fences keep the compiler from moving memory
// accesses across this point.
__device__ __forceinline__ void gsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G>
struct Group {
  int l;     // lane within the group
  int base;  // first lane of the group within the wavefront
  __device__ Group() {
    const int lane = (int)(threadIdx.x & 63);
    l = lane & (G - 1);
    base = lane & ~(G - 1);
  }
  // bit i set when lane i of the group has p
  __device__ unsigned long long ballot(bool p) const {
    const unsigned long long v = __ballot(p);
    if constexpr (G == 64) return v;
    else return (v >> base) & ((1ull << G) - 1ull);
  }
  __device__ double max(double v) const {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, G));
    return v;
  }
  // largest v over the group and the lowest lane index holding it (v not NaN)
  __device__ void argmax(double v, int i, double &vm, int &im) const {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(v, o, G);
      const int i2 = __shfl_xor(i, o, G);
      if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
    }
    vm = v;
    im = i;
  }
};

// LDS slice of one group (doubles)
template <int NS, bool KIN>
struct Slice {
  static constexpr int NP = (NS + 1) / 2;
  // per-system vectors: trial shear, b, slip resistance, hardening h_beta, total and current slip
  // increments, the LU right-hand side / solution, singular values, compacted b (SVD), and the
  // running and start-of-increment backstress (kinematic hardening only)
  static constexpr int tau = 0, b = NS, s = 2 * NS, h = 3 * NS, x = 4 * NS, xb = 5 * NS, y = 6 * NS, sig = 7 * NS,
                       bc = 8 * NS, W = 9 * NS, Wt = 10 * NS, nvec = KIN ? 11 * NS : 9 * NS;
  // M0: A, then A2 (factored in place), then S_PA (NS x 9); M1: the LU copy of A_PA / U of the SVD,
  // then delb / delgamma (NS x 9); M2: V of the SVD
  static constexpr int M0 = nvec, M1 = M0 + NS * NS, M2 = M1 + NS * NS, cs = M2 + NS * NS, fl = cs + 2 * NP,
                       size = fl + NP;
};

__device__ __forceinline__ int nth_bit(unsigned long long m, int i) {
  for (int k = 0; k < i; ++k) m &= m - 1;
  return __ffsll((long long)m) - 1;
}

// v[i] for a lane-dependent i < N without dynamic register indexing
template <int N>
__device__ __forceinline__ double pick(const double *v, int i) {
  double r = v[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = (i == k) ? v[k] : r;
  return r;
}

// LU with partial pivoting restricted to the systems in `mask` (the rows and columns of the
// compacted matrix, in ascending system order), row i held by lane i; M (ld NS) and Y (ld m, m <= G
// right-hand sides) in LDS are overwritten with the factor and the solution.  Each entry sees
// cp::lu_solve's operations on the compacted matrix, in the same order.  Returns false (for the
// whole group) when a pivot is <= CP_PINV_RTOL * max|M|.
template <int NS, int G>
__device__ bool glu_solve(const Group<G> &g, unsigned long long mask, double *M, double *Y, int m) {
  const int i = g.l;
  const bool row = i < NS && ((mask >> i) & 1);
  double rmax = 0.0;
  if (row)
    for (unsigned long long r = mask; r; r &= r - 1) rmax = fmax(rmax, fabs(M[i * NS + (__ffsll((long long)r) - 1)]));
  const double tiny = CP_PINV_RTOL * g.max(rmax);
  unsigned long long rem = mask;
  while (rem) {
    const int c = __ffsll((long long)rem) - 1;  // pivot column: the next system of the set
    double v = (i < NS && ((rem >> i) & 1)) ? fabs(M[i * NS + c]) : -1.0;
    if (v != v) v = -1.0;
    double pv;
    int p;
    g.argmax(v, i, pv, p);
    const double dcc = fabs(M[c * NS + c]);
    if (dcc != dcc || !(pv > tiny)) return false;
    rem &= rem - 1;
    if (p != c) {  // lane j exchanges column j of rows c and p, and right-hand side j
      if (i < NS) {
        const double t = M[c * NS + i];
        M[c * NS + i] = M[p * NS + i];
        M[p * NS + i] = t;
      }
      if (i < m) {
        const double t = Y[c * m + i];
        Y[c * m + i] = Y[p * m + i];
        Y[p * m + i] = t;
      }
      gsync();
    }
    const double inv = 1.0 / M[c * NS + c];
    if (i < NS && ((rem >> i) & 1)) {
      const double f = M[i * NS + c] * inv;
      if (f != 0.0) {
        for (unsigned long long r = rem; r; r &= r - 1) {
          const int j = __ffsll((long long)r) - 1;
          M[i * NS + j] -= f * M[c * NS + j];
        }
        for (int j = 0; j < m; ++j) Y[i * m + j] -= f * Y[c * m + j];
      }
    }
    gsync();
  }
  // back substitution, lane j for right-hand side j
  if (i < m) {
    for (unsigned long long t = mask; t;) {
      const int c = 63 - __clzll((long long)t);
      t &= ~(1ull << c);
      const double inv = 1.0 / M[c * NS + c];
      double s = Y[c * m + i];
      for (unsigned long long r = mask & ~((2ull << c) - 1ull); r; r &= r - 1) {
        const int c2 = __ffsll((long long)r) - 1;
        s -= M[c * NS + c2] * Y[c2 * m + i];
      }
      Y[c * m + i] = s * inv;
    }
  }
  gsync();
  return true;
}

// xout[a] = (pinv(A_PA) b_PA)_a for the systems a in `mask` (A: ld NS, original indices; b: per
// system): cp::svd_pinv_solve on the compacted system, lane i holding row i of U and V (compact
// order).  Each round-robin round takes two steps: lane k forms the rotation of pair k (sums in
// the one-lane order), then every lane rotates its row.  The pairs of a round share no column, so
// the result is the one-lane routine's, bit for bit.
template <int NS, int G, bool KIN>
__device__ void gsvd_pinv(const Group<G> &g, unsigned long long mask, const double *A, const double *b, double *L, double *xout) {
  typedef Slice<NS, KIN> SL;
  double *U = L + SL::M1, *V = L + SL::M2, *bc = L + SL::bc, *sig = L + SL::sig, *cs = L + SL::cs, *fl = L + SL::fl;
  const int n = __popcll((long long)mask);
  const int i = g.l;
  const bool row = i < n;
  int pi = 0;
  if (row) {
    pi = nth_bit(mask, i);
    int j = 0;
    for (unsigned long long r = mask; r; r &= r - 1, ++j) {
      U[j * NS + i] = A[pi * NS + (__ffsll((long long)r) - 1)];  // column j, row i at j * NS + i
      V[j * NS + i] = (i == j) ? 1.0 : 0.0;
    }
    bc[i] = b[pi];
  }
  gsync();
  const int m = n + (n & 1), np = m / 2;
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int r = 0; r < m - 1; ++r) {
      if (i < np) {
        int p, q;
        cp::rr_pair(m, r, i, p, q);
        double f = 0.0, c = 1.0, sn = 0.0;
        if (q < n) {
          double al = 0, be = 0, ga = 0;
          for (int k = 0; k < n; ++k) {
            const double up = U[p * NS + k], uq = U[q * NS + k];
            al += up * up; be += uq * uq; ga += up * uq;
          }
          if (!(ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be))) {
            f = 1.0;
            const double zeta = (be - al) / (2.0 * ga);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
          }
        }
        cs[2 * i] = c;
        cs[2 * i + 1] = sn;
        fl[i] = f;
      }
      gsync();
      for (int k = 0; k < np; ++k) {
        if (fl[k] == 0.0) continue;
        rotated = true;
        int p, q;
        cp::rr_pair(m, r, k, p, q);
        if (row) {
          const double c = cs[2 * k], sn = cs[2 * k + 1];
          const double up = U[p * NS + i], uq = U[q * NS + i];
          U[p * NS + i] = c * up - sn * uq;
          U[q * NS + i] = sn * up + c * uq;
          const double vp = V[p * NS + i], vq = V[q * NS + i];
          V[p * NS + i] = c * vp - sn * vq;
          V[q * NS + i] = sn * vp + c * vq;
        }
      }
      gsync();
    }
    if (!rotated) break;
  }
  if (row) {  // lane j: singular value j
    double ss = 0.0;
    for (int k = 0; k < n; ++k) ss += U[i * NS + k] * U[i * NS + k];
    sig[i] = sqrt(ss);
  }
  gsync();
  double smax = 0.0;
  for (int j = 0; j < n; ++j) smax = fmax(smax, sig[j]);
  if (row)
    for (int j = 0; j < n; ++j)
      if (sig[j] > 0) U[j * NS + i] /= sig[j];
  gsync();
  // a single small singular value is inverted when the data fix its step (cp::svd_pinv_solve)
  double lim = smax * CP_PINV_RTOL;
  int nsmall = 0, ks = -1;
  for (int k = 0; k < n; ++k)
    if (sig[k] <= lim) { nsmall++; ks = k; }
  if (nsmall == 1 && sig[ks] > 0) {
    double s2 = smax, ub = 0, bb = 0;
    for (int k = 0; k < n; ++k)
      if (k != ks) s2 = fmin(s2, sig[k]);
    for (int r = 0; r < n; ++r) { ub += U[ks * NS + r] * bc[r]; bb += bc[r] * bc[r]; }
    const double eps = 2.220446049250313e-16;
    if (fabs(ub) * eps * smax * smax < 1e-2 * sqrt(bb) * sig[ks] * (s2 - sig[ks])) lim = 0.0;
  }
  double xi = 0.0;
  for (int k = 0; k < n; ++k) {
    if (!(sig[k] > lim)) continue;
    double ub = 0;
    for (int r = 0; r < n; ++r) ub += U[k * NS + r] * bc[r];
    ub /= sig[k];
    if (row) xi += V[k * NS + i] * ub;
  }
  if (row) xout[pi] = xi;
  gsync();
}

// The update of one qp by the group.  F, Fp_t: the deformation gradient and Fp_conv (every lane);
// s_a, W_a: slip resistance and backstress of this lane's system (lanes a < ns).  On return every
// lane holds P, T, Fe, Fp; the slice holds s_tau (Slice::s), x_beta_old (Slice::x) and W_kh_tau
// (Slice::W) per system; with flag = 1, lane kl < 9 has written column kl of dscale * dP/dF,
// D[((3i+j)*9 + kl) * dstride].  Returns the number of systems left active (0 = elastic);
// `plastic` is set when the potentially-active set was non-empty (the reference then writes
// slipfraction_iter, :454-456).  nsing (optional) counts the pseudo-inverse solves.
// KIN: kinematic hardening (:294-305, 431-436, 700-703).
template <int NS, int G, typename DT, bool KIN>
__device__ __forceinline__ int qp_update(const Group<G> &g, double *__restrict__ L, const double *__restrict__ mat,
                                         const double *__restrict__ gt, int ns, int max_solver_iter, const double *F,
                                         const double *Fp_t, double s_a, double W_a, int flag, double *P, double *T,
                                         double *Fe_out, double *Fp_out, DT *D, long long dstride, double dscale, bool &plastic,
                                         unsigned *nsing, int nslip = 1 << 30) {
  using namespace cp;
  typedef MatLayout<NS> ML;
  typedef GrainLayout<NS> GL;
  typedef Slice<NS, KIN> SL;
  const int l = g.l;
  const bool sys = l < ns;
  const double tol1 = mat[ML::tol];
  const double *__restrict__ Dm = gt + GL::D;
  const double *__restrict__ TM = gt + GL::TM;
  const double *__restrict__ Sg = gt + GL::S;
  double *tau = L + SL::tau, *bv = L + SL::b, *sv = L + SL::s, *hv = L + SL::h, *xv = L + SL::x, *xbv = L + SL::xb,
         *yv = L + SL::y, *Wv = L + SL::W, *Wtv = L + SL::Wt, *M0 = L + SL::M0, *M1 = L + SL::M1, *M2 = L + SL::M2;
  auto wtau = [&](int a) -> double { if constexpr (KIN) return Wv[a]; else return 0.0; };
  auto wt = [&](int a) -> double { if constexpr (KIN) return Wtv[a]; else return 0.0; };

  // ---- trial state (:142-240)
  double Fpinv[9], FE[9], CE[9], E[9], Tst[9], tmp[9];
  m3_inv(Fp_t, Fpinv);
  m3_mul(F, Fpinv, FE);
  m3_tmul(FE, FE, CE);
#pragma unroll
  for (int i = 0; i < 9; ++i) E[i] = 0.5 * (CE[i] - ((i % 4 == 0) ? 1.0 : 0.0));
  dmat_apply(Dm, E, Tst);
  double Fp[9], Ce[9], TstTau[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { Fp[i] = Fp_t[i]; Ce[i] = CE[i]; TstTau[i] = Tst[i]; }
  m3_mul(CE, Tst, tmp);
  bool act = false;
  if (sys) {
    double r = m3_ddot(tmp, Sg + 9 * l);
    const double w = KIN ? W_a : 0.0;
    if (l >= nslip && (r - w) < 0) r = w;  // twin systems shear one way only (:213-216)
    const double bb = fabs(r - w) - s_a;
    tau[l] = r;
    bv[l] = bb;
    sv[l] = s_a;
    xv[l] = 0.0;
    xbv[l] = 0.0;
    if constexpr (KIN) { Wv[l] = W_a; Wtv[l] = W_a; }
    act = bb >= tol1;
  }
  const unsigned long long pa0 = g.ballot(act);
  plastic = pa0 != 0ull;
  gsync();
  double XdelFp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long mask = 0ull;

  // ---- plastic correction (:250-468)
  if (plastic) {
    if (sys) hv[l] = mat[ML::h0 + l] * pow((1 - s_a / mat[ML::ss + l]), mat[ML::pw + l]);
    gsync();
    if ((pa0 >> l) & 1) {  // column l of A over the potentially-active set (:273-306)
      const int j = l;
      double t2[9], t1[9], t3[9];
      m3_mul(CE, Sg + 9 * j, t2);
      m3_symm(t2);
      dmat_apply(Dm, t2, t3);
      m3_mul(CE, t3, t1);
      m3_mul(t2, Tst, t3);
#pragma unroll
      for (int k = 0; k < 9; ++k) t1[k] += 2.0 * t3[k];
      const double tj = tau[j] - wtau(j), hj = hv[j];
      for (unsigned long long r = pa0; r; r &= r - 1) {
        const int i = __ffsll((long long)r) - 1;
        const double d = m3_ddot(Sg + 9 * i, t1);
        const double ti = tau[i] - wtau(i);
        const bool opp = (ti * tj) < 0.0;
        double v = mat[ML::q + i * NS + j] * hj + (opp ? -d : d);
        if constexpr (KIN) {  // :302-305, added to every column of row i
          const double k = mat[ML::c1 + i] - mat[ML::c2 + i] * Wv[i];
          v += (ti < 0.0) ? -k : k;
        }
        M0[i * NS + j] = v;
      }
    }
    gsync();
    mask = pa0;
    auto linf_b = [&]() {
      double v = ((mask >> l) & 1) ? fabs(bv[l]) : 0.0;
      if (v != v) v = 0.0;
      return g.max(v);
    };
    int count1 = 0;
    while (linf_b() > tol1) {  // :322
      count1++;
      if (count1 > max_solver_iter) break;
      // inactive_slip_removal (inactiveSlipRemoval.cc:4-124) on the current set
      while (true) {
        if ((mask >> l) & 1) {
          for (unsigned long long r = mask; r; r &= r - 1) {
            const int j = __ffsll((long long)r) - 1;
            M1[l * NS + j] = M0[l * NS + j];
          }
          yv[l] = bv[l];
        }
        gsync();
        if (!glu_solve<NS, G>(g, mask, M1, yv, 1)) {
          if (nsing && l == 0) atomicAdd(nsing, 1u);
          gsvd_pinv<NS, G, KIN>(g, mask, M0, bv, L, yv);
        }
        double xbl = 0.0;
        if (sys) {
          xbl = ((mask >> l) & 1) ? yv[l] : 0.0;
          xbv[l] = xbl;
        }
        const unsigned long long rm = g.ballot(((mask >> l) & 1) && (xv[l] + xbl) < 0);
        gsync();
        if (!rm) break;
        mask &= ~rm;
      }
      if (sys) xv[l] += xbv[l];  // :119
      gsync();
      // one pass of the loop body (:338-440)
#pragma unroll
      for (int i = 0; i < 9; ++i) XdelFp[i] = 0.0;
      for (int a = 0; a < ns; ++a) {
        const double ga = ((tau[a] - wtau(a)) > 0) ? xv[a] : -xv[a];
#pragma unroll
        for (int k = 0; k < 9; ++k) XdelFp[k] += ga * Sg[9 * a + k];
      }
      double ex[9];
      m3_expm(XdelFp, ex);
      m3_mul(ex, Fp_t, Fp);  // :359
      m3_inv(Fp, tmp);       // :362-373
      m3_mul(F, tmp, FE);
      m3_tmul(FE, FE, Ce);
#pragma unroll
      for (int i = 0; i < 9; ++i) E[i] = 0.5 * (Ce[i] - ((i % 4 == 0) ? 1.0 : 0.0));
      dmat_apply(Dm, E, TstTau);  // :376-377
      m3_mul(Ce, TstTau, tmp);
      gsync();  // every lane has read the backstress before it changes
      if (sys) {  // :380-440
        double ta = m3_ddot(tmp, Sg + 9 * l);
        if (l >= nslip && (tau[l] - wtau(l)) < 0) ta = wtau(l);  // :390-393, the backstress before :431
        double h1 = 0;
        for (int c = 0; c < ns; ++c) h1 += mat[ML::q + l * NS + c] * hv[c] * xbv[c];
        double sa = sv[l] + h1;
        if (sa > mat[ML::ss + l]) sa = mat[ML::ss + l];
        sv[l] = sa;
        double wl = 0.0;
        if constexpr (KIN) {  // :431-436
          const double c1 = mat[ML::c1 + l], c2 = mat[ML::c2 + l];
          wl = Wv[l];
          wl += (((tau[l] - wl) > 0) ? (c1 - c2 * wl) : (-c1 - c2 * wl)) * xbv[l];
          Wv[l] = wl;
        }
        bv[l] = fabs(ta - wl) - sa;
      }
      gsync();
    }
    m3_inv(Fp, Fpinv);  // :458-460
    m3_mul(F, Fpinv, FE);
  }
  const int n_active = plastic ? __popcll((long long)mask) : 0;

  // ---- stresses (:188-196 elastic, :458-468 plastic)
  const double detFE = m3_det(FE);
  m3_mul(FE, TstTau, tmp);
  {
    const double r = 1.0 / detFE;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] *= r;
  }
  m3_mult(tmp, FE, T);
  double Finv[9], tmp2[9];
  m3_inv(F, Finv);
  m3_mult(T, Finv, tmp2);
  const double pf = plastic ? m3_det(F) : detFE;
#pragma unroll
  for (int i = 0; i < 9; ++i) P[i] = pf * tmp2[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) { Fe_out[i] = FE[i]; Fp_out[i] = Fp[i]; }
  if (!flag) return n_active;

  // ---- tangent: in-loop block (:473-828) with delFp_delF_prev = 0 and dels_delF = 0
  const int kl = l < 9 ? l : 0, kk = kl / 3, ll = kl % 3;
  // rows / columns of the 3x3 factors that column kl needs
  double FpinvL[3], FEK[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double r0[3] = {Fpinv[i], Fpinv[3 + i], Fpinv[6 + i]};
    FpinvL[i] = pick<3>(r0, ll);
    const double r1[3] = {FE[i], FE[3 + i], FE[6 + i]};
    FEK[i] = pick<3>(r1, kk);
  }
  const bool tang = n_active > 0;
  if (tang) {
    // count3 (:607-619): series terms until ||X^k / k!||_F <= max(tol/1e4, 1e-12)
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
    // delb (:581-604), column kl: delb(a, kl) = +- S_a : deltau_delF(:, kl)
    if (l < 9) {
      double dE[9], dT[9], dtau[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dE[3 * i + j] = 0.5 * (FpinvL[i] * FEK[j] + FpinvL[j] * FEK[i]);
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
      for (unsigned long long r = mask; r; r &= r - 1) {
        const int a = __ffsll((long long)r) - 1;
        const double v = m3_ddot(dtau, Sg + 9 * a);
        M1[a * 9 + kl] = ((tau[a] - wt(a)) < 0) ? -v : v;
      }
    }
    // A2 (:622-709), column l
    if ((mask >> l) & 1) {
      double eneg[9], Xneg[9], t6[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) Xneg[i] = -XdelFp[i];
      m3_expm(Xneg, eneg);
      m3_mul(CE, eneg, t6);
      const int j = l;
      double diff[9], t2[9], t3[9], tm[9];
      dexp_series(XdelFp, Sg + 9 * j, count3, true, diff);
      m3_tmul(diff, t6, t2);
      m3_symm(t2);
      dmat_apply(Dm, t2, t3);
      m3_mul(Ce, t3, tm);
      m3_mul(t2, TstTau, t3);
#pragma unroll
      for (int k = 0; k < 9; ++k) tm[k] += 2.0 * t3[k];
      const bool nj = (tau[j] - wt(j)) < 0.0;
      const double hj = hv[j];
      for (unsigned long long r = mask; r; r &= r - 1) {
        const int i = __ffsll((long long)r) - 1;
        const double d = m3_ddot(Sg + 9 * i, tm);
        const bool ni = (tau[i] - wt(i)) < 0.0;
        double v = mat[ML::q + i * NS + j] * hj + ((ni ^ nj) ? -d : d);
        if constexpr (KIN) {  // :700-703: sign from W_kh_t, value from W_kh_tau
          const double k = mat[ML::c1 + i] - mat[ML::c2 + i] * Wv[i];
          v += ni ? -k : k;
        }
        M0[i * NS + j] = v;
      }
    }
    gsync();
    // delgamma = A2^-1 (delb - dels_PA) with dels = 0 (:712-722)
    if (!glu_solve<NS, G>(g, mask, M0, M1, 9)) {  // singular A2: no plastic tangent contribution
      if ((mask >> l) & 1)
#pragma unroll
        for (int k = 0; k < 9; ++k) M1[l * 9 + k] = 0.0;
      gsync();
    }
    // S_PA (:734-821): row j = sign_j * dexp(X, S_j) Fp_t, into M0 (the factor of A2 is no longer needed)
    if ((mask >> l) & 1) {
      const int j = l;
      double diff[9], t1[9];
      dexp_series(XdelFp, Sg + 9 * j, count3, false, diff);
      m3_mul(diff, Fp_t, t1);
      const double sg = ((tau[j] - wt(j)) < 0) ? -1.0 : 1.0;
#pragma unroll
      for (int r = 0; r < 9; ++r) M0[j * 9 + r] = sg * t1[r];
    }
    gsync();
  }

  // ---- final tangent (:835-940), column kl in lane kl
  if (l < 9) {
    double q1[9], q2[9], q3[9], q5[9];
    m3_mul(Finv, FE, q2);  // temp = temp2 = F^-1 Fe
    m3_mult(TstTau, q2, q1);
    m3_mul(FE, TstTau, q3);
    m3_mult(q3, F, q5);
    double FinvL[3], FinvK[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double r0[3] = {Finv[i], Finv[3 + i], Finv[6 + i]};
      FinvL[i] = pick<3>(r0, ll);
      const double r1[3] = {Finv[3 * i], Finv[3 * i + 1], Finv[3 * i + 2]};
      FinvK[i] = pick<3>(r1, kk);  // Finv(i, k)
    }
    double dFe[9], Wm[9], dTs[9], dFp[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) dFp[r] = 0.0;
    if (tang)
      for (unsigned long long r = mask; r; r &= r - 1) {
        const int j = __ffsll((long long)r) - 1;
        const double gk = M1[j * 9 + kl];
#pragma unroll
        for (int q = 0; q < 9; ++q) dFp[q] += M0[j * 9 + q] * gk;
      }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sacc = 0;
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) sacc += dFp[3 * a + bb] * Fpinv[3 * bb + j];
        Wm[3 * a + j] = sacc;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sacc = (i == kk) ? FpinvL[j] : 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) sacc -= FE[3 * i + a] * Wm[3 * a + j];
        dFe[3 * i + j] = sacc;
      }
    double Gm[9];
    m3_tmul(dFe, FE, Gm);
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      double sacc = 0;
#pragma unroll
      for (int c = 0; c < 9; ++c) sacc += TM[9 * r + c] * Gm[c];
      dTs[r] = sacc;
    }
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
      for (int a = 0; a < 3; ++a) qa += q5[3 * i + a] * FinvL[a];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        D[((3 * i + j) * 9 + kl) * dstride] = (DT)(dscale * (t1[3 * i + j] + t2[3 * i + j] - t3[3 * i + j] - qa * FinvK[j]));
    }
  }
  return n_active;
}

}  // namespace coop

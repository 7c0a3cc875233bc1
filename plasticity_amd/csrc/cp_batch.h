// Batched RVEs: many independent crystal-plasticity boundary-value problems of uneven size (each an
// n_x x n_y x n_z Q1 box with its own grains and material) advanced together, every Newton step of
// every RVE in the same launches.  The reference runs one RVE per job: its cell loop
// (src/ellipticBVP/assemble.cc:64-90), Newton loop (solveNonLinearSystem.cc:14-75) and
// updateAfterIncrement (updateAfterIncrement.cc:6-305) are the per-RVE work done here for all RVEs at
// once.  Included by cp_api.hip (uses its mesh helpers, reductions and qp routines).
//
// Layout: RVEs are concatenated (qps, elements, nodes, dofs); a table of BRve gives each RVE's mesh,
// offsets and block ranges; every element carries its RVE id.  Quadrature points of all RVEs and
// material classes (12, 24 or 48 slip systems) form one work queue, drained by one persistent
// launch of one-wave blocks: a wavefront claims 64 consecutive qps at a time, so qps with many
// active systems or local iterations do not hold up a fixed share of the grid.  Each RVE's
// Jacobian is its own node-block matrix; the Krylov solver (Jacobi-BiCG, the reference's, or
// Jacobi-PCG) runs all RVEs in the same launches, every block belonging to one RVE, with per-RVE
// scalars, stopping tests and iteration counts.  A Newton iteration is replayed from HIP graphs (assembly, chunks of Krylov iterations,
// update), captured once per batch; an autotuner picks the qp pass's grid and the Krylov chunk.
#pragma once

namespace {

constexpr int BATCH_MAX_CLASSES = 4;
constexpr int BATCH_MAX_RVES = 65535;  // RVE ids are 16-bit
constexpr int BATCH_PBLK = 256;        // qps per block of the post-increment reduction

struct BRve {
  MeshDims m;
  double h[3], JxW, vol;
  long long node_off, elem_off, kof;  // first node, first element, first Jacobian value (243 per node)
  int cls;                            // material class
  int kblk0, nkblk;                   // Krylov blocks (VB nodes each)
  int pblk0, npblk;                   // post-reduction blocks (BATCH_PBLK qps each)
};

struct BClass {
  const double *mat;   // material block (cp::MatLayout<NS>)
  const double *gtab;  // grain tables of the class (rows: grain rows of all its RVEs)
  int gstride, ns, NS, msi;
};

// Numerically singular local systems of one material class: the queue of qps deferred by the
// first pass and the store of their systems (the single-RVE passes' pipeline, SingAny: queued in
// pass 1, solved by the cooperative k_sing_svd, replayed; later singular solves in the lane).
struct BSing {
  SingStore st;
  unsigned *queue, *count;
};

struct BArgs {
  const BRve *rve;
  const BClass *cls;
  const BSing *sing;           // [class]
  const unsigned short *erve;  // [nelem] RVE of each element
  const unsigned *egrain;      // [nelem] row of the element's grain in its class table
  long long nq;
  const double *u;
  const double *Fp_conv, *s_conv, *slip_conv;
  double *Fe_iter, *Fp_iter, *s_iter, *slip_iter, *T;
  void *D;     // [81][nq] JxW dP/dF (DT)
  double *P;   // [9][nq] JxW P
  double *ws;  // 48-system workspace: cp::LaneWs<48> per lane, wavefront blocks interleaved by lane
  long long ws_lane;
  unsigned *work;  // work-queue counter (zeroed before the launch)
};

// calculatePlasticity(cellID, q, flag) at qp Q of the batch (flag 1: assembly stage 1, writes the
// iter history, JxW dP/dF and JxW P; flag 0: the post-increment pass, writes the history and T).
// False when the singular-system functor deferred the qp (nothing written).
template <int NS, typename DT>
__device__ __attribute__((noinline)) bool bq_one(const BArgs &a, long long Q, int flag, SingAny<NS> &sing) {
  const BRve R = a.rve[a.erve[Q >> 3]];
  const BClass C = a.cls[R.cls];
  double F[9];
  qp_F(R.m, R.h, a.u + 3 * R.node_off, Q - 8 * R.elem_off, F);
  double Fp_t[9], s[NS], x[NS], P[9], T[9], Fe[9], Fp[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fp_t[i] = a.Fp_conv[i * a.nq + Q];
  for (int i = 0; i < C.ns; ++i) s[i] = a.s_conv[i * a.nq + Q];
  const double *gt = C.gtab + (long long)a.egrain[Q >> 3] * C.gstride;
  bool plastic = false;
  const int na = cp::qp_update<NS, DT, false>(C.mat, gt, C.ns, C.msi, F, Fp_t, s, x, flag, P, T, Fe, Fp, flag ? (DT *)a.D + Q : nullptr,
                                              a.nq, &plastic, R.JxW, SingRef<NS>{&sing}, nullptr, nullptr, 1, nullptr, 1 << 30, sing.ws);
  if (na < 0) return false;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    a.Fe_iter[i * a.nq + Q] = Fe[i];
    a.Fp_iter[i * a.nq + Q] = Fp[i];
    if (flag) a.P[i * a.nq + Q] = R.JxW * P[i];
    else a.T[i * a.nq + Q] = T[i];
  }
  for (int i = 0; i < C.ns; ++i) a.s_iter[i * a.nq + Q] = s[i];
  if (plastic)  // slipfraction_iter is written on the plastic path only (calculatePlasticity.cc:454-456)
    for (int i = 0; i < C.ns; ++i) a.slip_iter[i * a.nq + Q] = a.slip_conv[i * a.nq + Q] + x[i];
  return true;
}
// U and V of the in-lane pseudo-inverse live in the lane's workspace for every class
template <int NS>
__device__ __forceinline__ SingAny<NS> bsing(const BSing &S, int mode, unsigned Q, unsigned e, int have, cp::LanePtr lw) {
  SingAny<NS> f{mode, S.st, S.count, S.queue, Q, e, have, 0, lw};
  return f;
}

template <int NS, typename DT>
__device__ __forceinline__ void bq_claim(const BArgs &a, long long Q, int k, int flag, cp::LanePtr lw) {
  SingAny<NS> f = bsing<NS>(a.sing[k], SING_CLAIM, (unsigned)Q, 0u, 0, lw);
  bq_one<NS, DT>(a, Q, flag, f);  // a singular system defers the qp to the passes below
}
// The persistent qp pass: one-wave blocks drain the queue of all qps, 64 at a time per wavefront;
// qps that meet a numerically singular A_PA are queued per class.
template <typename DT>
__global__ void __launch_bounds__(64) QP_WAVES kb_qp_pass(BArgs a, int flag) {
  __shared__ unsigned sbase;
  const cp::LanePtr lw = wave_block_ptr(a.ws, a.ws_lane);
  for (;;) {
    if (threadIdx.x == 0) sbase = atomicAdd(a.work, 64u);
    __syncthreads();
    const long long base = sbase;
    __syncthreads();
    if (base >= a.nq) break;
    const long long Q = base + threadIdx.x;
    if (Q < a.nq) {
      const int k = a.rve[a.erve[Q >> 3]].cls;
      const int NS = a.cls[k].NS;
      if (NS == 48) bq_claim<48, DT>(a, Q, k, flag, lw);
      else if (NS == 24) bq_claim<24, DT>(a, Q, k, flag, lw);
      else bq_claim<12, DT>(a, Q, k, flag, lw);
    }
  }
}
// class k's queued qps with the cooperative solutions of their first singular system
template <int NS, typename DT>
__global__ void __launch_bounds__(64) QP_WAVES kb_qp_replay(BArgs a, int k, int flag) {
  const BSing S = a.sing[k];
  const cp::LanePtr lw = wave_block_ptr(a.ws, a.ws_lane);
  const unsigned n = min(*S.count, S.st.cap);
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    if (S.st.done[e]) continue;
    SingAny<NS> f = bsing<NS>(S, SING_REPLAY, 0u, e, S.st.have[e], lw);
    if (bq_one<NS, DT>(a, S.queue[e], flag, f)) S.st.done[e] = 1;
  }
}
// the rest of class k's queue (entries beyond the store, or deferred again): pseudo-inverse in the lane
template <int NS, typename DT>
__global__ void __launch_bounds__(64) QP_WAVES kb_qp_deferred(BArgs a, int k, int flag) {
  const BSing S = a.sing[k];
  const cp::LanePtr lw = wave_block_ptr(a.ws, a.ws_lane);
  const unsigned n = *S.count;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (i >= S.st.cap || !S.st.done[i]) {
      SingAny<NS> f = bsing<NS>(S, SING_LANE, 0u, i, 0, lw);
      bq_one<NS, DT>(a, S.queue[i], flag, f);
    }
}

// ---- assembly stage 2, Dirichlet lines, diagonal, residual norms

struct BElemArgs {
  const BRve *rve;
  const unsigned short *col_rve;  // block -> RVE, colours concatenated
  const int *col_blk0;            // [8][nrve + 1] first block of each RVE within each colour
  int nrve;
  long long nq;
  const void *D;
  const double *P;
  const unsigned char *con;
  const double *delu;
  int iter0;
  void *K;
  double *R;
};
// one colour of the element loop over every RVE; every block holds ASM_EL elements of one RVE
template <typename KT>
__global__ void __launch_bounds__(64) kb_element(BElemArgs b, int col, int cbase) {
  __shared__ KT sD[64 * 81];
  __shared__ double sP[64 * 9];
  __shared__ double sdiag[ASM_EL * 24];
  const int r = b.col_rve[cbase + blockIdx.x];
  const BRve R = b.rve[r];
  ElemArgs a;
  a.m = R.m;
  for (int d = 0; d < 3; ++d) a.h[d] = R.h[d];
  a.cx = col & 1; a.cy = (col >> 1) & 1; a.cz = (col >> 2) & 1;
  a.ncx = (R.m.nx - a.cx + 1) / 2; a.ncy = (R.m.ny - a.cy + 1) / 2; a.ncz = (R.m.nz - a.cz + 1) / 2;
  a.nq = b.nq;
  a.D = b.D; a.P = b.P;
  a.con = b.con + 3 * R.node_off; a.delu = b.delu + 3 * R.node_off;
  a.iter0 = b.iter0;
  a.K = (void *)((KT *)b.K + R.kof);
  a.R = b.R + 3 * R.node_off;
  const long long blk = (long long)blockIdx.x - b.col_blk0[col * (b.nrve + 1) + r];
  element_body<KT>(a, blk * ASM_EL, 8 * R.elem_off, sD, sP, sdiag);
}

// Dirichlet lines of every RVE (boundaryConditions.cc:6-114, 452-491 with the simple face rules):
// node rank n of the batch, RVE from the node table
__global__ void kb_dirichlet(const BRve *rve, const unsigned short *nrve, long long nnode, BCArgs bc, unsigned char *con,
                             double *delu) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const BRve R = rve[nrve[n]];
  int i, j, k;
  node_ijk(R.m, n - R.node_off, i, j, k);
  const bool onb[6] = {i == 0, i == R.m.nx, j == 0, j == R.m.ny, k == 0, k == R.m.nz};
  for (int d = 0; d < 3; ++d) {
    bool flag = false;
    double val = 0.0;
    for (int f = 0; f < 6; ++f)
      if (bc.fdc[f][d] && onb[f]) { flag = true; val = bc.dU[f][d]; break; }
    con[3 * n + d] = flag ? 1 : 0;
    delu[3 * n + d] = flag ? val : 0.0;
  }
}

template <typename KT>
__global__ void kb_dinv(const BRve *rve, const unsigned short *nrve, long long nnode, const KT *__restrict__ K, double *dinv) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const BRve R = rve[nrve[n]];
  const long long nl = n - R.node_off;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double d = (double)K[R.kof + (13 * R.m.nnode + nl) * 9 + 4 * c];
    dinv[3 * n + c] = (d == 0.0) ? 1.0 : 1.0 / d;
  }
}

__global__ void k_zero_words(size_t n, unsigned *p) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// ---- Krylov solvers over every RVE, per RVE the single-RVE solvers' recurrences (Jacobi-PCG; and
// Jacobi-BiCG, whose left sequence rides along in the same kernels): block b of the Krylov kernels
// covers VB nodes of RVE krve[b]; each kernel sums its RVE's block partials in the same order in
// every block, so all blocks of an RVE take the same decision.
struct BKState {
  double tol, bnorm, dp, rzh[2];
  int it, done, breakdown, maxit, pad;
};
struct BKArgs {
  const BRve *rve;
  const unsigned short *krve;  // Krylov block -> RVE
  BKState *st;
  const void *K;
  const double *B, *dinv;
  double *X, *R, *Z, *P, *Q;     // CG; BiCG: R, Z, P the right sequence (Q unused)
  double *Rl, *Zl, *Pl;          // BiCG: the left sequence (null for CG)
  double *ipart, *opart, *upart;  // 3, 1 and 2 partials per block
};

// Sum of NV partials over blocks [b0, b0 + nb), identical in every block of the RVE.
template <int NV>
__device__ __forceinline__ void rve_partials(const double *part, int b0, int nb, double (&v)[NV], double *lds) {
  __shared__ double tot[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += VB)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += part[(size_t)(b0 + b) * NV + k];
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = tot[k];
}

// x = 0, r = b, z = D^-1 b; partials of b.b, z.z, z.b
__global__ void __launch_bounds__(VB) kb_cg_init(BKArgs a) {
  __shared__ double lds[48];
  const BRve R = a.rve[a.krve[blockIdx.x]];
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (nl < R.m.nnode) {
    const long long g0 = 3 * (R.node_off + nl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double b = a.B[g0 + c], z = b * a.dinv[g0 + c];
      a.X[g0 + c] = 0.0; a.R[g0 + c] = b; a.Z[g0 + c] = z;
      if (a.Rl) { a.Rl[g0 + c] = b; a.Zl[g0 + c] = z; }
      acc[0] += b * b; acc[1] += z * z; acc[2] += z * b;
    }
  }
  block_partials<3>(acc, a.ipart, lds);
}
// per RVE (one block each): tolerance lin_tol ||b|| and the state of iteration 0
__global__ void __launch_bounds__(VB) kb_cg_init2(BKArgs a, double lin_tol, int maxit) {
  __shared__ double lds[48];
  const BRve R = a.rve[blockIdx.x];
  double v[3];
  rve_partials<3>(a.ipart, R.kblk0, R.nkblk, v, lds);
  if (threadIdx.x == 0) {
    BKState &s = a.st[blockIdx.x];
    s.bnorm = sqrt(v[0]);
    s.tol = lin_tol * s.bnorm;
    s.dp = sqrt(v[1]);
    s.rzh[0] = v[2];
    s.rzh[1] = 1.0;
    s.it = 0;
    s.maxit = maxit;
    s.breakdown = 0;
    s.done = (s.dp <= s.tol) ? 1 : 0;
    if (!s.done && v[2] == 0.0) { s.done = 1; s.breakdown = 1; }
    if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) { s.done = 1; s.breakdown = KRYLOV_NONFINITE; }
  }
}

// iteration it (parity par = it & 1): it = 0: p = z; it >= 1: stopping test on ||z|| and
// b = (r.z)_it / (r.z)_(it-1) from the update kernel's partials, p = z + b p (BiCG: both sequences,
// r.z meaning z.rl)
__global__ void __launch_bounds__(VB) kb_cg_p(BKArgs a, int first, int par) {
  __shared__ double lds[32];
  const int r = a.krve[blockIdx.x];
  BKState *st = a.st + r;
  if (__syncthreads_or(st->done != 0)) return;
  const BRve R = a.rve[r];
  double b = 0.0;
  if (!first) {
    double v[2];
    rve_partials<2>(a.upart, R.kblk0, R.nkblk, v, lds);
    const double dp = sqrt(v[0]);
    const bool breakdown = (v[1] == 0.0), nonfinite = !isfinite(dp) || !isfinite(v[1]);
    const bool done = (dp <= st->tol) || st->it >= st->maxit || breakdown || nonfinite;
    if (blockIdx.x == R.kblk0 && threadIdx.x == 0) {
      st->dp = dp;
      if (done) {
        st->done = 1;
        st->breakdown = nonfinite ? KRYLOV_NONFINITE : ((breakdown && !(dp <= st->tol)) ? 1 : 0);
      } else {
        st->rzh[par] = v[1];
      }
    }
    if (done) return;
    b = v[1] / st->rzh[par ^ 1];
  }
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  if (nl < R.m.nnode) {
    const long long g0 = 3 * (R.node_off + nl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a.P[g0 + c] = first ? a.Z[g0 + c] : a.Z[g0 + c] + b * a.P[g0 + c];
      if (a.Pl) a.Pl[g0 + c] = first ? a.Zl[g0 + c] : a.Zl[g0 + c] + b * a.Pl[g0 + c];
    }
  }
}

// q = K p (the RVE's node-block matrix); partials of p.q
template <typename VT>
__global__ void __launch_bounds__(VB) kb_cg_spmv(BKArgs a) {
  __shared__ double lds[16];
  const int r = a.krve[blockIdx.x];
  if (__syncthreads_or(a.st[r].done != 0)) return;
  const BRve R = a.rve[r];
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  double acc[1] = {0.0};
  if (nl < R.m.nnode) {
    int i, j, k;
    node_ijk32(R.m, nl, i, j, k);
    const VT *K = (const VT *)a.K + R.kof;
    const double *P = a.P + 3 * R.node_off;
    const long long nn = R.m.nnode;
    double y[3] = {0, 0, 0};
    constexpr int UNR = sizeof(VT) == 4 ? 9 : 3;
#pragma unroll UNR
    for (int s = 0; s < 27; ++s) {
      bool ok;
      const long long mn = nb_masked(R.m, nl, i, j, k, s, ok);
      const double t0 = P[3 * mn], t1 = P[3 * mn + 1], t2 = P[3 * mn + 2];
      const double w = ok ? 1.0 : 0.0;
      block_mv(K + ((long long)s * nn + nl) * 9, w * t0, w * t1, w * t2, y);
    }
    double *Q = a.Q + 3 * R.node_off;
#pragma unroll
    for (int c = 0; c < 3; ++c) Q[3 * nl + c] = y[c];
    acc[0] = y[0] * P[3 * nl] + y[1] * P[3 * nl + 1] + y[2] * P[3 * nl + 2];
  }
  block_partials<1>(acc, a.opart, lds);
}

// a = (r.z)_it / (p.q); x += a p; r -= a q; z = D^-1 r; partials of z.z and r.z; it += 1
__global__ void __launch_bounds__(VB) kb_cg_update(BKArgs a, int par) {
  __shared__ double lds[32];
  const int r = a.krve[blockIdx.x];
  BKState *st = a.st + r;
  if (__syncthreads_or(st->done != 0)) return;
  const BRve R = a.rve[r];
  double pq[1];
  rve_partials<1>(a.opart, R.kblk0, R.nkblk, pq, lds);
  const double al = st->rzh[par] / pq[0];
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (nl < R.m.nnode) {
    const long long g0 = 3 * (R.node_off + nl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long g = g0 + c;
      const double rn = a.R[g] - al * a.Q[g], z = rn * a.dinv[g];
      a.X[g] += al * a.P[g];
      a.R[g] = rn;
      a.Z[g] = z;
      acc[0] += z * z;
      acc[1] += rn * z;
    }
  }
  __syncthreads();
  block_partials<2>(acc, a.upart, lds);
  if (blockIdx.x == R.kblk0 && threadIdx.x == 0) st->it += 1;
}

// BiCG (the reference's KSPBICG with PCJACOBI, solveLinearSystem.cc:13-16, per RVE): Zr = K Pr,
// Zl = K^T Pl on the RVE's node-block matrix; partials of Zr.Pl
template <typename VT>
__global__ void __launch_bounds__(VB) kb_bicg_spmv(BKArgs a) {
  __shared__ double lds[16];
  const int r = a.krve[blockIdx.x];
  if (__syncthreads_or(a.st[r].done != 0)) return;
  const BRve R = a.rve[r];
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  double acc[1] = {0.0};
  if (nl < R.m.nnode) {
    int i, j, k;
    node_ijk32(R.m, nl, i, j, k);
    const VT *K = (const VT *)a.K + R.kof;
    const double *Pr = a.P + 3 * R.node_off, *Pl = a.Pl + 3 * R.node_off;
    const long long nn = R.m.nnode;
    double y[3] = {0, 0, 0}, yt[3] = {0, 0, 0};
#pragma unroll 3
    for (int s = 0; s < 27; ++s) {
      bool ok;
      const long long mn = nb_masked(R.m, nl, i, j, k, s, ok);
      const double w = ok ? 1.0 : 0.0;
      double pr[3], pl[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) { pr[c] = w * Pr[3 * mn + c]; pl[c] = w * Pl[3 * mn + c]; }
      block_mv(K + ((long long)s * nn + nl) * 9, pr[0], pr[1], pr[2], y);
      block_tmv(K + ((long long)(26 - s) * nn + mn) * 9, pl[0], pl[1], pl[2], yt);
    }
    double *Zr = a.Z + 3 * R.node_off, *Zl = a.Zl + 3 * R.node_off;
#pragma unroll
    for (int c = 0; c < 3; ++c) { Zr[3 * nl + c] = y[c]; Zl[3 * nl + c] = yt[c]; }
    acc[0] = y[0] * Pl[3 * nl] + y[1] * Pl[3 * nl + 1] + y[2] * Pl[3 * nl + 2];
  }
  block_partials<1>(acc, a.opart, lds);
}

// a = (r.z)_it / (Zr.Pl); x += a Pr; r -= a Zr; rl -= a Zl; z = D^-1 r; zl = D^-1 rl; partials of
// z.z and z.rl; it += 1
__global__ void __launch_bounds__(VB) kb_bicg_update(BKArgs a, int par) {
  __shared__ double lds[32];
  const int r = a.krve[blockIdx.x];
  BKState *st = a.st + r;
  if (__syncthreads_or(st->done != 0)) return;
  const BRve R = a.rve[r];
  double dpi[1];
  rve_partials<1>(a.opart, R.kblk0, R.nkblk, dpi, lds);
  const double al = st->rzh[par] / dpi[0];
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (nl < R.m.nnode) {
    const long long g0 = 3 * (R.node_off + nl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long g = g0 + c;
      const double d = a.dinv[g];
      a.X[g] += al * a.P[g];
      const double rrn = a.R[g] - al * a.Z[g], rln = a.Rl[g] - al * a.Zl[g];
      a.R[g] = rrn; a.Rl[g] = rln;
      const double zrn = rrn * d, zln = rln * d;
      a.Z[g] = zrn; a.Zl[g] = zln;
      acc[0] += zrn * zrn;
      acc[1] += zrn * rln;
    }
  }
  __syncthreads();
  block_partials<2>(acc, a.upart, lds);
  if (blockIdx.x == R.kblk0 && threadIdx.x == 0) st->it += 1;
}

// block partials of |R|^2 (Krylov block layout), for the per-RVE residual norms
__global__ void __launch_bounds__(VB) kb_norm_parts(BKArgs a, double *part) {
  __shared__ double lds[16];
  const BRve R = a.rve[a.krve[blockIdx.x]];
  const long long nl = (long long)(blockIdx.x - R.kblk0) * VB + threadIdx.x;
  double acc[1] = {0.0};
  if (nl < R.m.nnode) {
    const long long g0 = 3 * (R.node_off + nl);
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[0] += a.B[g0 + c] * a.B[g0 + c];
  }
  block_partials<1>(acc, part, lds);
}

// ---- post-increment reduction (updateAfterIncrement.cc:107-126, 218-245) and reorient()
// (reorient.cc:18-96) per qp; block b covers BATCH_PBLK qps of RVE prve[b]; partials of the 13
// JxW-weighted sums (E 6, T 6, slip) per block
struct BPostArgs {
  const BRve *rve;
  const BClass *cls;
  const unsigned short *prve;
  long long nq;
  const double *u, *T, *slip_iter, *Fe_conv, *Fe_iter;
  double *rotnew, *part;
};
__global__ void __launch_bounds__(BATCH_PBLK) kb_post_reduce(BPostArgs a) {
  __shared__ double lds[13 * 16];
  const BRve R = a.rve[a.prve[blockIdx.x]];
  const long long ql = (long long)(blockIdx.x - R.pblk0) * BATCH_PBLK + threadIdx.x;
  double acc[13];
#pragma unroll
  for (int i = 0; i < 13; ++i) acc[i] = 0.0;
  if (ql < 8 * R.m.nelem) {
    const long long Q = 8 * R.elem_off + ql;
    double F[9], C[9], T[9];
    qp_F(R.m, R.h, a.u + 3 * R.node_off, ql, F);
    cp::m3_tmul(F, F, C);
#pragma unroll
    for (int i = 0; i < 9; ++i) T[i] = a.T[i * a.nq + Q];
    const int vi[6] = {0, 4, 8, 5, 2, 1};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      acc[c] = 0.5 * (C[vi[c]] - ((vi[c] % 4 == 0) ? 1.0 : 0.0)) * R.JxW;
      acc[6 + c] = T[vi[c]] * R.JxW;
    }
    const int ns = a.cls[R.cls].ns;
    double slip_sum = 0;
    for (int i = 0; i < ns; ++i) slip_sum += a.slip_iter[i * a.nq + Q] * R.JxW;
    acc[12] = slip_sum;
    double Fo[9], Fn[9], Ro[9], Rn[9], Om[9], Om2[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) { Fo[i] = a.Fe_conv[i * a.nq + Q]; Fn[i] = a.Fe_iter[i * a.nq + Q]; }
    polar_R(Fo, Ro);
    polar_R(Fn, Rn);
#pragma unroll
    for (int i = 0; i < 9; ++i) Om[i] = Rn[i] - Ro[i];
    cp::m3_mult(Om, Rn, Om2);
    const double w[3] = {-0.5 * (Om2[5] - Om2[7]), 0.5 * (Om2[2] - Om2[6]), -0.5 * (Om2[1] - Om2[3])};
    double rr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) rr[i] = a.rotnew[i * a.nq + Q];
    const double dt = w[0] * rr[0] + w[1] * rr[1] + w[2] * rr[2];
    const double cr[3] = {w[1] * rr[2] - w[2] * rr[1], w[2] * rr[0] - w[0] * rr[2], w[0] * rr[1] - w[1] * rr[0]};
    double rn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) rn[i] = rr[i] + 0.5 * (w[i] + dt * rr[i] + cr[i]);
    const double nrm = sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]);
    if (nrm > 10000.0)
      for (int i = 0; i < 3; ++i) rn[i] = rn[i] * 10000.0 / nrm;
#pragma unroll
    for (int i = 0; i < 3; ++i) a.rotnew[i * a.nq + Q] = rn[i];
  }
  block_partials<13>(acc, a.part, lds);
}

}  // namespace

// ---------------------------------------------------------------- host side
struct pa_batch {
  std::vector<pa_material> mats;
  pa_problem pb;
  int nrve = 0, ncls = 0, nsmax = 0;
  int precision = PA_PREC_FP64;
  int method = PA_SOLVER_BICG;  // PA_SOLVER_BICG (the reference's) or PA_SOLVER_CG
  std::vector<BRve> hrve;
  std::vector<int> rve_mat;
  long long nq = 0, nelem = 0, nnode = 0, ndof = 0;
  int nkblk = 0, npblk = 0;
  std::vector<int> col_blk0, col_base;  // [8][nrve + 1]; [9]
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t kev[2] = {nullptr, nullptr};
  // grains (host copies until the first Newton step builds the tables)
  std::vector<std::vector<unsigned>> egrain;
  std::vector<std::vector<double>> grod;
  bool tables = false;
  // device
  BRve *d_rve = nullptr;
  BClass *d_cls = nullptr;
  std::vector<BClass> hcls;
  unsigned short *d_erve = nullptr, *d_nrve = nullptr, *d_krve = nullptr, *d_prve = nullptr, *d_colrve = nullptr;
  int *d_colblk0 = nullptr;
  unsigned *d_egrain = nullptr, *d_work = nullptr;
  double *d_mat[BATCH_MAX_CLASSES] = {nullptr}, *d_gtab[BATCH_MAX_CLASSES] = {nullptr};
  double *Fp_conv = nullptr, *Fp_iter = nullptr, *Fe_conv = nullptr, *Fe_iter = nullptr, *s_conv = nullptr, *s_iter = nullptr;
  double *slip_conv = nullptr, *slip_iter = nullptr, *rotnew = nullptr, *T = nullptr;
  double *u = nullptr, *R = nullptr, *dinv = nullptr, *delu = nullptr, *P = nullptr;
  unsigned char *con = nullptr;
  void *K = nullptr, *D = nullptr;
  double *kv[8] = {nullptr};  // X R Z P Q Rl Zl Pl
  BKState *st = nullptr, *hst = nullptr;  // device state; pinned host copies (2 x nrve)
  double *ipart = nullptr, *opart = nullptr, *upart = nullptr, *npart = nullptr, *ppart = nullptr;
  double *ws = nullptr;
  long long ws_lane = 0;
  std::vector<BSing> hsing;  // per class: queue, count and store of the deferred singular systems
  BSing *d_sing = nullptr;
  unsigned *d_counts = nullptr;  // [class]: queue lengths; [4 + class]: replay misses
  // launch choices (pa_batch_tune) and graphs
  int qp_grid = 512, chunk = 16;
  bool use_graphs = true;
  hipGraphExec_t g_asm[2][2][2] = {};  // [BiCG][mixed][iter0]
  hipGraphExec_t g_chunk[2][2] = {};   // [BiCG][mixed]
  double last_ms[4] = {0, 0, 0, 0};  // assembly, Krylov, post-increment, qp pass (eager runs)
  std::vector<double> tune;          // last autotuner report
  std::vector<void *> allocs;
};

namespace {

constexpr int BATCH_MAX_GRID = 1024;  // one-wave blocks of the persistent qp pass (tuner's largest)

template <typename T>
int balloc(pa_batch *b, T **p, size_t count) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, count * sizeof(T) + 16));
  b->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}
template <typename T>
int bupload(pa_batch *b, T **p, const std::vector<T> &h) {
  int rc = balloc(b, p, std::max<size_t>(h.size(), 1));
  if (rc) return rc;
  if (!h.empty()) HIPCHK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}
bool bmixed(const pa_batch *b) { return b->precision == PA_PREC_MIXED; }
size_t bkbytes(const pa_batch *b) { return (bmixed(b) ? sizeof(float) : sizeof(double)) * 243 * (size_t)b->nnode; }

bool bbicg(const pa_batch *b) { return b->method == PA_SOLVER_BICG; }
void batch_drop_graphs(pa_batch *b) {
  for (int m = 0; m < 2; ++m)
    for (int f = 0; f < 2; ++f) {
      for (int i = 0; i < 2; ++i)
        if (b->g_asm[m][f][i]) { (void)hipGraphExecDestroy(b->g_asm[m][f][i]); b->g_asm[m][f][i] = nullptr; }
      if (b->g_chunk[m][f]) { (void)hipGraphExecDestroy(b->g_chunk[m][f]); b->g_chunk[m][f] = nullptr; }
    }
}

BArgs batch_qp_args(pa_batch *b) {
  BArgs a;
  a.rve = b->d_rve; a.cls = b->d_cls; a.sing = b->d_sing; a.erve = b->d_erve; a.egrain = b->d_egrain;
  a.nq = b->nq; a.u = b->u;
  a.Fp_conv = b->Fp_conv; a.s_conv = b->s_conv; a.slip_conv = b->slip_conv;
  a.Fe_iter = b->Fe_iter; a.Fp_iter = b->Fp_iter; a.s_iter = b->s_iter; a.slip_iter = b->slip_iter; a.T = b->T;
  a.D = b->D; a.P = b->P;
  a.ws = b->ws; a.ws_lane = b->ws_lane;
  a.work = b->d_work;
  return a;
}
BKArgs batch_k_args(pa_batch *b) {
  BKArgs a;
  a.rve = b->d_rve; a.krve = b->d_krve; a.st = b->st; a.K = b->K;
  a.B = b->R; a.dinv = b->dinv;
  a.X = b->kv[0]; a.R = b->kv[1]; a.Z = b->kv[2]; a.P = b->kv[3]; a.Q = b->kv[4];
  const bool bi = bbicg(b);
  a.Rl = bi ? b->kv[5] : nullptr; a.Zl = bi ? b->kv[6] : nullptr; a.Pl = bi ? b->kv[7] : nullptr;
  a.ipart = b->ipart; a.opart = b->opart; a.upart = b->upart;
  return a;
}

template <int NS, typename DT>
void batch_sing_passes(pa_batch *b, const BArgs &a, int k, int flag, int grid) {
  hipLaunchKernelGGL(k_sing_svd<NS>, dim3(SVD_GRID), dim3(sing_per_block<NS>() * sing_group<NS>()), 0, b->stream, b->hsing[k].st,
                     (const unsigned *)b->hsing[k].count);
  hipLaunchKernelGGL((kb_qp_replay<NS, DT>), dim3(grid), dim3(64), 0, b->stream, a, k, flag);
  hipLaunchKernelGGL((kb_qp_deferred<NS, DT>), dim3(grid), dim3(64), 0, b->stream, a, k, flag);
}
template <typename DT>
void batch_qp_passes(pa_batch *b, const BArgs &a, int flag, int grid) {
  hipLaunchKernelGGL(kb_qp_pass<DT>, dim3(grid), dim3(64), 0, b->stream, a, flag);
  for (int k = 0; k < b->ncls; ++k) {
    const int NS = b->hcls[k].NS;
    if (NS == 48) batch_sing_passes<48, DT>(b, a, k, flag, grid);
    else if (NS == 24) batch_sing_passes<24, DT>(b, a, k, flag, grid);
    else batch_sing_passes<12, DT>(b, a, k, flag, grid);
  }
}
// the qp passes over every qp of the batch (flag 1: assembly stage 1; flag 0: post-increment pass):
// the persistent pass, then per class the cooperative pseudo-inverse round, the replay and the
// in-lane rest of the qps it deferred
int batch_qp_pass(pa_batch *b, int flag, int grid) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, b->stream, b->d_work);
  for (int k = 0; k < b->ncls; ++k) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, b->stream, b->d_counts + k);
  const BArgs a = batch_qp_args(b);
  if (flag && bmixed(b)) batch_qp_passes<float>(b, a, flag, grid);
  else batch_qp_passes<double>(b, a, flag, grid);
  HIPCHK(hipGetLastError());
  return 0;
}

// operator and update of one Krylov iteration (parity par)
void batch_op_update(pa_batch *b, const BKArgs &k, int par) {
  if (bbicg(b)) {
    if (bmixed(b)) hipLaunchKernelGGL(kb_bicg_spmv<float>, dim3(b->nkblk), dim3(VB), 0, b->stream, k);
    else hipLaunchKernelGGL(kb_bicg_spmv<double>, dim3(b->nkblk), dim3(VB), 0, b->stream, k);
    hipLaunchKernelGGL(kb_bicg_update, dim3(b->nkblk), dim3(VB), 0, b->stream, k, par);
  } else {
    if (bmixed(b)) hipLaunchKernelGGL(kb_cg_spmv<float>, dim3(b->nkblk), dim3(VB), 0, b->stream, k);
    else hipLaunchKernelGGL(kb_cg_spmv<double>, dim3(b->nkblk), dim3(VB), 0, b->stream, k);
    hipLaunchKernelGGL(kb_cg_update, dim3(b->nkblk), dim3(VB), 0, b->stream, k, par);
  }
}
// Krylov initialisation and iteration 0 on the assembled system
void batch_krylov_start(pa_batch *b, const BKArgs &k) {
  hipLaunchKernelGGL(kb_cg_init, dim3(b->nkblk), dim3(VB), 0, b->stream, k);
  hipLaunchKernelGGL(kb_cg_init2, dim3(b->nrve), dim3(VB), 0, b->stream, k, b->pb.lin_tol, b->pb.lin_maxit);
  hipLaunchKernelGGL(kb_cg_p, dim3(b->nkblk), dim3(VB), 0, b->stream, k, 1, 0);
  batch_op_update(b, k, 0);
}

// assembly of every RVE and Krylov iteration 0: no memset (fill kernels only), so the sequence can
// be captured into a graph
int batch_assemble_seq(pa_batch *b, int iter0) {
  const size_t kw = bkbytes(b) / 4;
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>(65535, (kw + 255) / 256)), dim3(256), 0, b->stream, kw, (unsigned *)b->K);
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>(65535, (2 * b->ndof + 255) / 256)), dim3(256), 0, b->stream,
                     (size_t)(2 * b->ndof), (unsigned *)b->R);
  if (int rc = batch_qp_pass(b, 1, b->qp_grid)) return rc;
  BElemArgs e;
  e.rve = b->d_rve; e.col_rve = b->d_colrve; e.col_blk0 = b->d_colblk0; e.nrve = b->nrve;
  e.nq = b->nq; e.D = b->D; e.P = b->P; e.con = b->con; e.delu = b->delu; e.iter0 = iter0; e.K = b->K; e.R = b->R;
  for (int col = 0; col < 8; ++col) {
    const int nb = b->col_base[col + 1] - b->col_base[col];
    if (nb == 0) continue;
    if (bmixed(b)) hipLaunchKernelGGL(kb_element<float>, dim3(nb), dim3(64), 0, b->stream, e, col, b->col_base[col]);
    else hipLaunchKernelGGL(kb_element<double>, dim3(nb), dim3(64), 0, b->stream, e, col, b->col_base[col]);
  }
  const unsigned gn = nblk(b->nnode, 256);
  if (bmixed(b)) hipLaunchKernelGGL(kb_dinv<float>, dim3(gn), dim3(256), 0, b->stream, b->d_rve, b->d_nrve, b->nnode, (const float *)b->K, b->dinv);
  else hipLaunchKernelGGL(kb_dinv<double>, dim3(gn), dim3(256), 0, b->stream, b->d_rve, b->d_nrve, b->nnode, (const double *)b->K, b->dinv);
  const BKArgs k = batch_k_args(b);
  hipLaunchKernelGGL(kb_norm_parts, dim3(b->nkblk), dim3(VB), 0, b->stream, k, b->npart);
  batch_krylov_start(b, k);
  HIPCHK(hipGetLastError());
  return 0;
}
// Krylov iterations 1 + 16j .. chunk + 16j (chunk even, so iteration 1 + k has parity (1 + k) & 1)
int batch_chunk_seq(pa_batch *b) {
  const BKArgs k = batch_k_args(b);
  for (int i = 0; i < b->chunk; ++i) {
    const int par = (1 + i) & 1;
    hipLaunchKernelGGL(kb_cg_p, dim3(b->nkblk), dim3(VB), 0, b->stream, k, 0, par);
    batch_op_update(b, k, par);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename F>
int batch_capture(pa_batch *b, hipGraphExec_t *out, F &&seq) {
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = seq();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : fail(PA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return fail(PA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
  *out = ex;
  return 0;
}
int batch_assemble(pa_batch *b, int iter0) {
  if (!b->use_graphs) return batch_assemble_seq(b, iter0);
  hipGraphExec_t &g = b->g_asm[bbicg(b) ? 1 : 0][bmixed(b) ? 1 : 0][iter0 ? 1 : 0];
  if (!g)
    if (int rc = batch_capture(b, &g, [&] { return batch_assemble_seq(b, iter0); })) return rc;
  HIPCHK(hipGraphLaunch(g, b->stream));
  return 0;
}
int batch_chunk(pa_batch *b) {
  if (!b->use_graphs) return batch_chunk_seq(b);
  hipGraphExec_t &g = b->g_chunk[bbicg(b) ? 1 : 0][bmixed(b) ? 1 : 0];
  if (!g)
    if (int rc = batch_capture(b, &g, [&] { return batch_chunk_seq(b); })) return rc;
  HIPCHK(hipGraphLaunch(g, b->stream));
  return 0;
}

// The Krylov loop after iteration 0: chunks until every RVE's solve has ended; the host reads the
// states one chunk behind, so the stream never idles on the round trip.
int batch_krylov(pa_batch *b, int *lin_iters) {
  for (int j = 0;; ++j) {
    if (int rc = batch_chunk(b)) return rc;
    BKState *h = b->hst + (size_t)(j & 1) * b->nrve;
    HIPCHK(hipMemcpyAsync(h, b->st, sizeof(BKState) * b->nrve, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipEventRecord(b->kev[j & 1], b->stream));
    if (j == 0) continue;
    HIPCHK(hipEventSynchronize(b->kev[(j - 1) & 1]));
    const BKState *hp = b->hst + (size_t)((j - 1) & 1) * b->nrve;
    bool all = true;
    for (int r = 0; r < b->nrve && all; ++r) all = hp[r].done != 0;
    if (!all) continue;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int r = 0; r < b->nrve; ++r) {
      if (lin_iters) lin_iters[r] = hp[r].it;
      if (hp[r].breakdown == KRYLOV_NONFINITE)
        return fail(PA_ERR_NONFINITE, "batch: Krylov solve of RVE " + std::to_string(r) + " stopped on a NaN/Inf norm");
    }
    return 0;
  }
}

// per-class grain tables, element grain rows, the initial state (crystalPlasticity::init,
// init.cc:302-330: Fp = Fe = I, s = s0, slip = 0, rotnew = the grain's orientation)
int batch_build_tables(pa_batch *b) {
  std::vector<std::vector<double>> tab(b->ncls);
  std::vector<int> rows(b->ncls, 0);
  std::vector<unsigned> eg(b->nelem);
  std::vector<double> rn(3 * (size_t)b->nq);
  for (int r = 0; r < b->nrve; ++r) {
    const BRve &R = b->hrve[r];
    const int k = R.cls;
    const pa_material &mat = b->mats[b->rve_mat[r]];
    const int ns = mat.n_slip, NS = ns_template(ns);
    std::vector<double> mv(3 * ns), nv(3 * ns);
    for (int a = 0; a < ns; ++a) {
      double mn = 0, nn = 0;
      for (int j = 0; j < 3; ++j) { mn += mat.m[a][j] * mat.m[a][j]; nn += mat.n[a][j] * mat.n[a][j]; }
      mn = std::sqrt(mn); nn = std::sqrt(nn);
      for (int j = 0; j < 3; ++j) { mv[3 * a + j] = mat.m[a][j] / mn; nv[3 * a + j] = mat.n[a][j] / nn; }
    }
    const int ng = (int)(b->grod[r].size() / 3);
    const int gstride = b->hcls[k].gstride;
    const size_t t0 = tab[k].size();
    tab[k].resize(t0 + (size_t)ng * gstride, 0.0);
    for (int g = 0; g < ng; ++g)
      grain_table(mat, (const double(*)[3])mv.data(), (const double(*)[3])nv.data(), &b->grod[r][3 * g], ns, NS,
                  &tab[k][t0 + (size_t)g * gstride]);
    for (long long e = 0; e < R.m.nelem; ++e) {
      const unsigned g = b->egrain[r][e];
      eg[R.elem_off + e] = (unsigned)rows[k] + g;
      for (int q = 0; q < 8; ++q)
        for (int i = 0; i < 3; ++i) rn[i * b->nq + 8 * (R.elem_off + e) + q] = b->grod[r][3 * g + i];
    }
    rows[k] += ng;
  }
  for (int k = 0; k < b->ncls; ++k) {
    if (int rc = bupload(b, &b->d_gtab[k], tab[k])) return rc;
    b->hcls[k].gtab = b->d_gtab[k];
  }
  if (int rc = bupload(b, &b->d_cls, b->hcls)) return rc;
  HIPCHK(hipMemcpy(b->d_egrain, eg.data(), eg.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->rotnew, rn.data(), rn.size() * sizeof(double), hipMemcpyHostToDevice));
  // state
  const long long nq = b->nq;
  std::vector<double> h(9 * nq, 0.0);
  for (int i = 0; i < 3; ++i) std::fill(h.begin() + (4 * i) * nq, h.begin() + (4 * i + 1) * nq, 1.0);
  for (double *p : {b->Fp_conv, b->Fp_iter, b->Fe_conv, b->Fe_iter}) HIPCHK(hipMemcpy(p, h.data(), 9 * nq * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> s((size_t)b->nsmax * nq, 0.0);
  for (int r = 0; r < b->nrve; ++r) {
    const pa_material &mat = b->mats[b->rve_mat[r]];
    const long long q0 = 8 * b->hrve[r].elem_off, q1 = q0 + 8 * b->hrve[r].m.nelem;
    for (int a = 0; a < mat.n_slip; ++a) std::fill(s.begin() + a * nq + q0, s.begin() + a * nq + q1, mat.s0[a]);
  }
  HIPCHK(hipMemcpy(b->s_conv, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->s_iter, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
  const size_t bs = (size_t)b->nsmax * nq * sizeof(double);
  HIPCHK(hipMemset(b->slip_conv, 0, bs));
  HIPCHK(hipMemset(b->slip_iter, 0, bs));
  HIPCHK(hipMemset(b->u, 0, sizeof(double) * b->ndof));
  // Dirichlet lines: simple face rules, increment deluConstraint = totalU / totalIncrements (init.cc:341)
  BCArgs bc;
  std::memset(&bc, 0, sizeof(bc));
  for (int f = 0; f < 6; ++f)
    for (int d = 0; d < 3; ++d) {
      bc.fdc[f][d] = b->pb.face_dof_constrained[f][d];
      bc.dU[f][d] = b->pb.face_total_u[f][d] / b->pb.total_increments;
    }
  hipLaunchKernelGGL(kb_dirichlet, dim3(nblk(b->nnode, 256)), dim3(256), 0, b->stream, (const BRve *)b->d_rve,
                     (const unsigned short *)b->d_nrve, b->nnode, bc, b->con, b->delu);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  b->tables = true;
  return 0;
}
int batch_ready(pa_batch *b) {
  if (b->tables) return 0;
  for (int r = 0; r < b->nrve; ++r)
    if (b->grod[r].empty()) return fail(PA_ERR_STATE, "pa_batch_set_grains has not been called for RVE " + std::to_string(r));
  return batch_build_tables(b);
}

}  // namespace

extern "C" {

void pa_batch_destroy(pa_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graphs(b);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->hst) (void)hipHostFree(b->hst);
  for (auto &e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : b->kev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int pa_batch_create(const pa_material *mats, int n_mats, const pa_problem *pb, int n_rve, const int *rve_mat, const int *rve_nel,
                    pa_batch **out) {
  if (!mats || !pb || !rve_mat || !rve_nel || !out) return fail(PA_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_mats < 1 || n_mats > BATCH_MAX_CLASSES) return fail(PA_ERR_ARG, "a batch takes 1 to 4 materials");
  if (n_rve < 1 || n_rve > BATCH_MAX_RVES) return fail(PA_ERR_ARG, "a batch takes 1 to 65535 RVEs");
  for (int k = 0; k < n_mats; ++k) {
    const pa_material &m = mats[k];
    if (m.n_slip < 1 || m.n_slip > PA_MAX_SLIP) return fail(PA_ERR_ARG, "n_slip out of range");
    if (m.twinning || m.kinematic_hardening) return fail(PA_ERR_UNSUPPORTED, "batch: twinning and kinematic hardening are not carried");
    if (m.max_slip_search != 1) return fail(PA_ERR_UNSUPPORTED, "Max Slip Search Iterations must be 1");
    if (m.max_solver_iter < 1) return fail(PA_ERR_ARG, "Max Solver Iterations must be >= 1");
  }
  if (pb->tabular || pb->cyclic || pb->velgrad) return fail(PA_ERR_UNSUPPORTED, "batch: simple face boundary conditions only");
  if (pb->total_increments < 1 || pb->max_newton < 1 || pb->lin_maxit < 1) return fail(PA_ERR_ARG, "bad increment/solver counts");
  for (int r = 0; r < n_rve; ++r) {
    if (rve_mat[r] < 0 || rve_mat[r] >= n_mats) return fail(PA_ERR_ARG, "RVE material index out of range");
    for (int d = 0; d < 3; ++d)
      if (rve_nel[3 * r + d] < 1 || rve_nel[3 * r + d] > 1024) return fail(PA_ERR_ARG, "RVE mesh sizes must be 1..1024");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PA_ERR_HIP, "no HIP device available");
  pa_batch *b = new pa_batch();
  b->mats.assign(mats, mats + n_mats);
  b->pb = *pb;
  b->nrve = n_rve;
  b->rve_mat.assign(rve_mat, rve_mat + n_rve);
  b->ncls = n_mats;
  b->hcls.resize(n_mats);
  int rc = 0;
  if (hipStreamCreate(&b->stream) != hipSuccess) { delete b; return fail(PA_ERR_HIP, "hipStreamCreate failed"); }
  for (auto &e : b->ev) hipEventCreate(&e);
  for (auto &e : b->kev) hipEventCreateWithFlags(&e, hipEventDisableTiming);
  bool has48 = false;
  for (int k = 0; k < n_mats; ++k) {  // material blocks (cp::MatLayout<NS>)
    const pa_material &m = mats[k];
    const int ns = m.n_slip, NS = ns_template(ns);
    has48 = has48 || NS == 48;
    b->nsmax = std::max(b->nsmax, ns);
    const int tol = 3 * NS + NS * NS, c1 = tol + 1, c2 = c1 + NS;
    std::vector<double> mb(c2 + NS, 0.0);
    for (int a = 0; a < ns; ++a) {
      mb[a] = m.h0[a]; mb[NS + a] = m.ss[a]; mb[2 * NS + a] = m.pw[a];
      for (int c = 0; c < ns; ++c) mb[3 * NS + a * NS + c] = m.q[a][c];
    }
    mb[tol] = m.stress_tol;
    if ((rc = bupload(b, &b->d_mat[k], mb))) { pa_batch_destroy(b); return rc; }
    b->hcls[k].mat = b->d_mat[k];
    b->hcls[k].gstride = 9 * NS + 36 + 81 + 9;
    b->hcls[k].ns = ns;
    b->hcls[k].NS = NS;
    b->hcls[k].msi = m.max_solver_iter;
  }
  // RVE table: meshes, offsets, block ranges
  b->hrve.resize(n_rve);
  std::vector<unsigned short> erve, nrve, krve, prve;
  b->col_blk0.assign(8 * (n_rve + 1), 0);
  b->col_base.assign(9, 0);
  for (int r = 0; r < n_rve; ++r) {
    BRve &R = b->hrve[r];
    MeshDims &m = R.m;
    m.nx = rve_nel[3 * r]; m.ny = rve_nel[3 * r + 1]; m.nz = rve_nel[3 * r + 2];
    m.nnx = m.nx + 1; m.nny = m.ny + 1; m.nnz = m.nz + 1;
    m.nnode = (long long)m.nnx * m.nny * m.nnz;
    m.nelem = (long long)m.nx * m.ny * m.nz;
    for (int d = 0; d < 3; ++d) R.h[d] = pb->span[d] / (d == 0 ? m.nx : d == 1 ? m.ny : m.nz);
    R.JxW = R.h[0] * R.h[1] * R.h[2] / 8.0;
    R.vol = 0.0;  // the sum of JxW over every qp in order (updateAfterIncrement.cc:221)
    for (long long q = 0; q < 8 * m.nelem; ++q) R.vol += R.JxW;
    R.node_off = b->nnode; R.elem_off = b->nelem; R.kof = 243 * b->nnode;
    R.cls = rve_mat[r];
    R.kblk0 = b->nkblk; R.nkblk = (int)nblk(m.nnode, VB);
    R.pblk0 = b->npblk; R.npblk = (int)nblk(8 * m.nelem, BATCH_PBLK);
    b->nnode += m.nnode; b->nelem += m.nelem;
    b->nkblk += R.nkblk; b->npblk += R.npblk;
    erve.insert(erve.end(), (size_t)m.nelem, (unsigned short)r);
    nrve.insert(nrve.end(), (size_t)m.nnode, (unsigned short)r);
    krve.insert(krve.end(), (size_t)R.nkblk, (unsigned short)r);
    prve.insert(prve.end(), (size_t)R.npblk, (unsigned short)r);
  }
  b->nq = 8 * b->nelem;
  b->ndof = 3 * b->nnode;
  std::vector<unsigned short> colrve;
  for (int col = 0; col < 8; ++col) {
    const int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
    int acc = 0;
    for (int r = 0; r < n_rve; ++r) {
      const MeshDims &m = b->hrve[r].m;
      const long long ncol = (long long)((m.nx - cx + 1) / 2) * ((m.ny - cy + 1) / 2) * ((m.nz - cz + 1) / 2);
      const int nb = (int)nblk(ncol, ASM_EL);
      b->col_blk0[col * (n_rve + 1) + r] = acc;
      colrve.insert(colrve.end(), (size_t)nb, (unsigned short)r);
      acc += nb;
    }
    b->col_blk0[col * (n_rve + 1) + n_rve] = acc;
    b->col_base[col + 1] = b->col_base[col] + acc;
  }
  const long long nq = b->nq, N = b->ndof;
  const int nsm = b->nsmax;
  b->ws_lane = has48 ? cp::LaneWs<48>::size : 0;  // only the 48-system variant keeps matrices there
  if ((rc = bupload(b, &b->d_rve, b->hrve)) || (rc = bupload(b, &b->d_erve, erve)) || (rc = bupload(b, &b->d_nrve, nrve)) ||
      (rc = bupload(b, &b->d_krve, krve)) || (rc = bupload(b, &b->d_prve, prve)) || (rc = bupload(b, &b->d_colrve, colrve)) ||
      (rc = bupload(b, &b->d_colblk0, b->col_blk0)) || (rc = balloc(b, &b->d_egrain, b->nelem)) || (rc = balloc(b, &b->d_work, 4)) ||
      (rc = balloc(b, &b->Fp_conv, 9 * nq)) || (rc = balloc(b, &b->Fp_iter, 9 * nq)) || (rc = balloc(b, &b->Fe_conv, 9 * nq)) ||
      (rc = balloc(b, &b->Fe_iter, 9 * nq)) || (rc = balloc(b, &b->s_conv, nsm * nq)) || (rc = balloc(b, &b->s_iter, nsm * nq)) ||
      (rc = balloc(b, &b->slip_conv, nsm * nq)) || (rc = balloc(b, &b->slip_iter, nsm * nq)) || (rc = balloc(b, &b->rotnew, 3 * nq)) ||
      (rc = balloc(b, &b->T, 9 * nq)) || (rc = balloc(b, &b->u, N)) || (rc = balloc(b, &b->R, N)) || (rc = balloc(b, &b->dinv, N)) ||
      (rc = balloc(b, &b->delu, N)) || (rc = balloc(b, &b->con, N)) || (rc = balloc(b, &b->P, 9 * nq)) ||
      (rc = balloc(b, (double **)&b->K, 243 * b->nnode)) || (rc = balloc(b, (double **)&b->D, 81 * nq)) || (rc = balloc(b, &b->st, n_rve)) ||
      (rc = balloc(b, &b->ipart, 3 * (size_t)b->nkblk)) || (rc = balloc(b, &b->opart, (size_t)b->nkblk)) ||
      (rc = balloc(b, &b->upart, 2 * (size_t)b->nkblk)) || (rc = balloc(b, &b->npart, (size_t)b->nkblk)) ||
      (rc = balloc(b, &b->ppart, 13 * (size_t)b->npblk)) || (rc = balloc(b, &b->ws, std::max<size_t>(64, (size_t)BATCH_MAX_GRID * 64 * b->ws_lane))) ||
      (rc = balloc(b, &b->d_counts, 8))) {
    pa_batch_destroy(b);
    return rc;
  }
  for (int i = 0; i < 8; ++i)
    if ((rc = balloc(b, &b->kv[i], N))) { pa_batch_destroy(b); return rc; }
  // per class: deferred-qp queue and singular-system store, as many entries as fit 4 GiB (a qp the
  // store cannot hold solves its singular systems in its own lane, with U and V in global memory for
  // 48 systems: a 48-system entry takes 19 KB, 223 k entries)
  if (hipMemset(b->d_counts, 0, 8 * sizeof(unsigned)) != hipSuccess) { pa_batch_destroy(b); return fail(PA_ERR_HIP, "memset"); }
  b->hsing.resize(n_mats);
  for (int k = 0; k < n_mats; ++k) {
    long long nqk = 0;
    for (int r = 0; r < n_rve; ++r)
      if (rve_mat[r] == k) nqk += 8 * b->hrve[r].m.nelem;
    const size_t NS = (size_t)b->hcls[k].NS;
    const size_t per = sizeof(double) * (NS * NS + NS + SING_ROUNDS * NS) + SING_ROUNDS * 8 + 3 * sizeof(int);
    const size_t cap = std::max<size_t>(1, std::min<size_t>((size_t)nqk, ((size_t)4 << 30) / per));
    BSing &S = b->hsing[k];
    S.count = b->d_counts + k;
    S.st.cap = (unsigned)cap;
    S.st.miss = b->d_counts + 4 + k;
    if ((rc = balloc(b, &S.queue, std::max<long long>(1, nqk))) || (rc = balloc(b, &S.st.A, cap * NS * NS)) ||
        (rc = balloc(b, &S.st.b, cap * NS)) || (rc = balloc(b, &S.st.n, cap)) || (rc = balloc(b, &S.st.X, cap * SING_ROUNDS * NS)) ||
        (rc = balloc(b, &S.st.key, cap * SING_ROUNDS)) || (rc = balloc(b, &S.st.have, cap)) || (rc = balloc(b, &S.st.done, cap))) {
      pa_batch_destroy(b);
      return rc;
    }
  }
  if ((rc = bupload(b, &b->d_sing, b->hsing))) { pa_batch_destroy(b); return rc; }
  if (hipHostMalloc((void **)&b->hst, 2 * sizeof(BKState) * n_rve, hipHostMallocDefault) != hipSuccess) {
    pa_batch_destroy(b);
    return fail(PA_ERR_HIP, "hipHostMalloc failed");
  }
  b->egrain.resize(n_rve);
  b->grod.resize(n_rve);
  *out = b;
  return 0;
}

int pa_batch_set_grains(pa_batch *b, int rve, const unsigned *elem_grain, const double *grain_rod, int n_grain_rows) {
  if (!b || !elem_grain || !grain_rod) return fail(PA_ERR_ARG, "null argument");
  if (rve < 0 || rve >= b->nrve) return fail(PA_ERR_ARG, "RVE index out of range");
  if (b->tables) return fail(PA_ERR_STATE, "grains are fixed once the batch has taken a Newton step");
  if (n_grain_rows < 1) return fail(PA_ERR_ARG, "no orientation rows");
  const long long ne = b->hrve[rve].m.nelem;
  for (long long e = 0; e < ne; ++e)
    if ((int)elem_grain[e] >= n_grain_rows) return fail(PA_ERR_ARG, "grain id without an orientation row");
  b->egrain[rve].assign(elem_grain, elem_grain + ne);
  b->grod[rve].assign(grain_rod, grain_rod + 3 * (size_t)n_grain_rows);
  return 0;
}

int pa_batch_set_solver(pa_batch *b, int method, int precision) {
  if (!b) return fail(PA_ERR_ARG, "null batch");
  if (method != PA_SOLVER_BICG && method != PA_SOLVER_CG) return fail(PA_ERR_ARG, "batch: Jacobi-BiCG or Jacobi-PCG");
  if (precision != PA_PREC_FP64 && precision != PA_PREC_MIXED) return fail(PA_ERR_ARG, "unknown precision");
  b->method = method;
  b->precision = precision;
  return 0;
}

int pa_batch_set_graphs(pa_batch *b, int on) {
  if (!b) return fail(PA_ERR_ARG, "null batch");
  b->use_graphs = on != 0;
  return 0;
}

int pa_batch_newton_iteration(pa_batch *b, int inc, int it, double *res_norms, int *lin_iters) {
  if (!b) return fail(PA_ERR_ARG, "null batch");
  (void)inc;
  if (int rc = batch_ready(b)) return rc;
  HIPCHK(hipEventRecord(b->ev[0], b->stream));
  if (int rc = batch_assemble(b, it == 0)) return rc;
  HIPCHK(hipEventRecord(b->ev[1], b->stream));
  std::vector<double> np(b->nkblk);
  HIPCHK(hipMemcpyAsync(np.data(), b->npart, np.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  if (int rc = batch_krylov(b, lin_iters)) return rc;
  hipLaunchKernelGGL(k_apply_du, dim3(nblk(b->ndof, 256)), dim3(256), 0, b->stream, b->ndof, b->u, (const double *)b->kv[0],
                     (const unsigned char *)b->con, (const double *)b->delu, (int)(it == 0));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->ev[2], b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, b->ev[0], b->ev[1]) == hipSuccess) b->last_ms[0] = ms;
  if (hipEventElapsedTime(&ms, b->ev[1], b->ev[2]) == hipSuccess) b->last_ms[1] = ms;
  for (int r = 0; r < b->nrve; ++r) {
    double s = 0;
    for (int k = 0; k < b->hrve[r].nkblk; ++k) s += np[b->hrve[r].kblk0 + k];
    if (res_norms) res_norms[r] = std::sqrt(s);
    if (!std::isfinite(s)) return fail(PA_ERR_NONFINITE, "batch: assembled residual of RVE " + std::to_string(r) + " is not finite");
  }
  return 0;
}

int pa_batch_commit_increment(pa_batch *b, double *rows) {
  if (!b) return fail(PA_ERR_ARG, "null batch");
  if (int rc = batch_ready(b)) return rc;
  HIPCHK(hipEventRecord(b->ev[3], b->stream));
  if (int rc = batch_qp_pass(b, 0, b->qp_grid)) return rc;
  BPostArgs a;
  a.rve = b->d_rve; a.cls = b->d_cls; a.prve = b->d_prve; a.nq = b->nq; a.u = b->u; a.T = b->T; a.slip_iter = b->slip_iter;
  a.Fe_conv = b->Fe_conv; a.Fe_iter = b->Fe_iter; a.rotnew = b->rotnew; a.part = b->ppart;
  hipLaunchKernelGGL(kb_post_reduce, dim3(b->npblk), dim3(BATCH_PBLK), 0, b->stream, a);
  HIPCHK(hipGetLastError());
  const size_t b9 = 9 * b->nq * sizeof(double), bs = (size_t)b->nsmax * b->nq * sizeof(double);
  HIPCHK(hipMemcpyAsync(b->Fe_conv, b->Fe_iter, b9, hipMemcpyDeviceToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->Fp_conv, b->Fp_iter, b9, hipMemcpyDeviceToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->s_conv, b->s_iter, bs, hipMemcpyDeviceToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->slip_conv, b->slip_iter, bs, hipMemcpyDeviceToDevice, b->stream));
  std::vector<double> pp(13 * (size_t)b->npblk);
  HIPCHK(hipMemcpyAsync(pp.data(), b->ppart, pp.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipEventRecord(b->ev[4], b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, b->ev[3], b->ev[4]) == hipSuccess) b->last_ms[2] = ms;
  bool finite = true;
  for (int r = 0; r < b->nrve; ++r) {
    const BRve &R = b->hrve[r];
    double s[13] = {0};
    for (int k = 0; k < R.npblk; ++k)
      for (int i = 0; i < 13; ++i) s[i] += pp[13 * (size_t)(R.pblk0 + k) + i];
    if (rows) {
      double *row = rows + 15 * (size_t)r;
      for (int i = 0; i < 12; ++i) row[i] = s[i] / R.vol;
      row[12] = row[13] = 0.0;
      row[14] = s[12] / R.vol;
    }
    for (int i = 0; i < 13; ++i) finite = finite && std::isfinite(s[i]);
  }
  if (!finite) return fail(PA_ERR_NONFINITE, "batch: post-increment averages are not finite");
  return 0;
}

// Autotuner: the persistent qp pass at 256, 512, 768 and 1024 one-wave blocks and the Krylov chunk
// at 8, 16 and 32 iterations, each timed on the current state (the qp pass of the next assembly and
// a complete solve of the current Jacobian; both leave the state as they found it), the fastest
// kept.  report (may be null): grid, chunk, then the qp pass in ms per grid candidate (4) and the
// solve in ms per chunk candidate (3).
int pa_batch_tune(pa_batch *b, double *report) {
  if (!b) return fail(PA_ERR_ARG, "null batch");
  if (int rc = batch_ready(b)) return rc;
  const int grids[4] = {256, 512, 768, 1024}, chunks[3] = {8, 16, 32};
  double tq[4], tk[3];
  const bool graphs = b->use_graphs;
  b->use_graphs = false;
  int best = 0;
  for (int i = 0; i < 4; ++i) {
    float best_ms = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      HIPCHK(hipEventRecord(b->ev[0], b->stream));
      if (int rc = batch_qp_pass(b, 1, grids[i])) { b->use_graphs = graphs; return rc; }
      HIPCHK(hipEventRecord(b->ev[1], b->stream));
      HIPCHK(hipEventSynchronize(b->ev[1]));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
      best_ms = std::min(best_ms, ms);
    }
    tq[i] = best_ms;
    if (tq[i] < tq[best]) best = i;
  }
  b->qp_grid = grids[best];
  // a Jacobian and right-hand side to solve: the assembly of a first Newton iteration (with the
  // Dirichlet increment) at the current state; u is not touched
  if (int rc = batch_assemble_seq(b, 1)) { b->use_graphs = graphs; return rc; }
  int bk = 0;
  for (int i = 0; i < 3; ++i) {
    b->chunk = chunks[i];
    batch_krylov_start(b, batch_k_args(b));
    HIPCHK(hipEventRecord(b->ev[0], b->stream));
    if (int rc = batch_krylov(b, nullptr)) { b->use_graphs = graphs; return rc; }
    HIPCHK(hipEventRecord(b->ev[1], b->stream));
    HIPCHK(hipEventSynchronize(b->ev[1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    tk[i] = ms;
    if (tk[i] < tk[bk]) bk = i;
  }
  b->chunk = chunks[bk];
  b->use_graphs = graphs;
  batch_drop_graphs(b);  // captured with the old choices
  b->tune.assign({(double)b->qp_grid, (double)b->chunk, tq[0], tq[1], tq[2], tq[3], tk[0], tk[1], tk[2]});
  if (report)
    for (size_t i = 0; i < b->tune.size(); ++i) report[i] = b->tune[i];
  return 0;
}

int pa_batch_qp_pass_ms(pa_batch *b, int flag, double *ms_out) {
  if (!b || !ms_out) return fail(PA_ERR_ARG, "null argument");
  if (int rc = batch_ready(b)) return rc;
  HIPCHK(hipEventRecord(b->ev[0], b->stream));
  if (int rc = batch_qp_pass(b, flag ? 1 : 0, b->qp_grid)) return rc;
  HIPCHK(hipEventRecord(b->ev[1], b->stream));
  HIPCHK(hipEventSynchronize(b->ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  *ms_out = ms;
  b->last_ms[3] = ms;
  return 0;
}

int pa_batch_last_timings(pa_batch *b, double ms[4]) {
  if (!b || !ms) return fail(PA_ERR_ARG, "null argument");
  for (int i = 0; i < 4; ++i) ms[i] = b->last_ms[i];
  return 0;
}

int pa_batch_get_state(pa_batch *b, int rve, int which, double *host, long long n) {
  if (!b || !host) return fail(PA_ERR_ARG, "null argument");
  if (rve < 0 || rve >= b->nrve) return fail(PA_ERR_ARG, "RVE index out of range");
  const BRve &R = b->hrve[rve];
  const long long nq = b->nq, q0 = 8 * R.elem_off, nqr = 8 * R.m.nelem;
  const int ns = b->mats[b->rve_mat[rve]].n_slip;
  const double *src = nullptr;
  int comps = 0;
  switch (which) {
    case 0: src = b->T; comps = 9; break;
    case 1: src = b->Fp_conv; comps = 9; break;
    case 2: src = b->Fe_conv; comps = 9; break;
    case 3: src = b->s_conv; comps = ns; break;
    case 4: src = b->rotnew; comps = 3; break;
    case 5: src = b->slip_conv; comps = ns; break;
    case 6: case 7: {
      const long long cnt = 3 * R.m.nnode;
      if (n < cnt) return fail(PA_ERR_ARG, "host buffer too small");
      HIPCHK(hipStreamSynchronize(b->stream));
      HIPCHK(hipMemcpy(host, (which == 6 ? b->u : b->R) + 3 * R.node_off, cnt * sizeof(double), hipMemcpyDeviceToHost));
      return 0;
    }
    default: return fail(PA_ERR_ARG, "unknown state id");
  }
  if (n < comps * nqr) return fail(PA_ERR_ARG, "host buffer too small");
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<double> tmp((size_t)nqr);
  for (int k = 0; k < comps; ++k) {
    HIPCHK(hipMemcpy(tmp.data(), src + k * nq + q0, nqr * sizeof(double), hipMemcpyDeviceToHost));
    for (long long q = 0; q < nqr; ++q) host[q * comps + k] = tmp[q];
  }
  return 0;
}

int pa_batch_sizes(pa_batch *b, long long out[4]) {
  if (!b || !out) return fail(PA_ERR_ARG, "null argument");
  out[0] = b->nq; out[1] = b->nelem; out[2] = b->nnode; out[3] = b->ndof;
  return 0;
}

}  // extern "C"

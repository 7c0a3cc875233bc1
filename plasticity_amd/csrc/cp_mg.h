// Geometric multigrid preconditioner for the node-block Jacobian of the structured Q1 mesh.
//
// Included by cp_api.hip after the Krylov helpers (MeshDims, KState, solve_ended, nb_masked,
// node_ijk32, block_mv, block_partials, sum_partials).  Not a reference algorithm: the reference
// solves with Jacobi-BiCG (solveLinearSystem.cc:13-16); this preconditioner changes how the
// Krylov solve reaches the same stopping test ||D^-1 r|| <= tol ||b||, not what it solves.
//
// Hierarchy: level l+1 halves every axis of level l; the coarsest level has at most
// MG_COARSE_NODES nodes.  Prolongation P is trilinear interpolation per displacement component;
// on level 0 its rows of constrained dofs are zero, so corrections never touch a Dirichlet dof
// and the Galerkin operators A_(l+1) = P^T A_l P see only free rows and columns.  Smoothing is
// Chebyshev of degree 2 on D^-1 A over [g/30, g], g = the Gershgorin bound max_i sum_j |A_ij / A_ii|
// (an upper bound of the spectrum, no eigenvalue iteration needed), applied before and after the
// coarse correction, so the V-cycle is the same linear operator on every call.  The coarsest
// level is solved exactly by a dense inverse (Gauss-Jordan with partial pivoting in LDS).
//
// Layouts: every level operator is [27 slots][nodes][9] row-major 3x3 blocks, like the Jacobian
// (slot s couples node n to n + (s%3-1, s/3%3-1, s/9-1)); vectors are [node][3].
#pragma once

constexpr int MG_MAX_LEVELS = 12;
constexpr int MG_COARSE_NODES = 27;             // 81 dofs (2^3 elements): the dense inverse fits 64 KiB of LDS
constexpr int MG_COARSE_DOFS = 3 * MG_COARSE_NODES;
constexpr double MG_CHEB_RATIO = 30.0;          // smoothed band [g / 30, g]

// Chebyshev(2) coefficients for the band [g / MG_CHEB_RATIO, g]: the first step is d0 = D^-1 r0 / theta,
// the second d1 = c1 d0 + c2 D^-1 r1 (r the unpreconditioned residual).
struct Cheb2 {
  double inv_theta, c1, c2;
};
__device__ __forceinline__ Cheb2 cheb2(double g) {
  const double hi = g, lo = g / MG_CHEB_RATIO;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
  const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
  Cheb2 c;
  c.inv_theta = 1.0 / theta;
  c.c1 = rho1 * rho0;
  c.c2 = 2.0 * rho1 / delta;
  return c;
}

__device__ __forceinline__ bool in_mesh(const MeshDims &m, int i, int j, int k) {
  return ((unsigned)i < (unsigned)m.nnx) & ((unsigned)j < (unsigned)m.nny) & ((unsigned)k < (unsigned)m.nnz);
}
__device__ __forceinline__ long long node_at(const MeshDims &m, int i, int j, int k) {
  return i + (long long)m.nnx * (j + (long long)m.nny * k);
}

// ---- setup

// Galerkin product A_c = P^T A_f P, one lane per (coarse node I, coarse slot S).  Fine node
// i = 2I + a (a in {-1,0,1}^3, weight prod(a_d ? 1/2 : 1)) couples to fine node j = 2J + b with
// J = I + S; only pairs with j - i inside the 27-point stencil contribute.  con (level 0 only)
// masks constrained rows and columns.  Sums are taken in FP64 in a fixed order.
template <typename VT>
__global__ void __launch_bounds__(256) k_mg_galerkin(MeshDims mf, const VT *__restrict__ Af, const unsigned char *__restrict__ con,
                                                     MeshDims mc, VT *__restrict__ Ac) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 27 * mc.nnode) return;
  const int S = (int)(t / mc.nnode);
  const long long I = t - (long long)S * mc.nnode;
  int ix, iy, iz;
  node_ijk32(mc, I, ix, iy, iz);
  const int sx = S % 3 - 1, sy = (S / 3) % 3 - 1, sz = S / 9 - 1;
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  if (in_mesh(mc, ix + sx, iy + sy, iz + sz)) {
    for (int az = -1; az <= 1; ++az)
      for (int ay = -1; ay <= 1; ++ay)
        for (int ax = -1; ax <= 1; ++ax) {
          const int fx = 2 * ix + ax, fy = 2 * iy + ay, fz = 2 * iz + az;
          if (!in_mesh(mf, fx, fy, fz)) continue;
          const long long fi = node_at(mf, fx, fy, fz);
          const double wa = (ax ? 0.5 : 1.0) * (ay ? 0.5 : 1.0) * (az ? 0.5 : 1.0);
          double mr[3] = {1.0, 1.0, 1.0};
          if (con)
            for (int r = 0; r < 3; ++r) mr[r] = con[3 * fi + r] ? 0.0 : 1.0;
          // b_d ranges so that (2 S_d + b_d) - a_d lies in {-1, 0, 1}
          for (int bz = max(-1, az - 2 * sz - 1); bz <= min(1, az - 2 * sz + 1); ++bz)
            for (int by = max(-1, ay - 2 * sy - 1); by <= min(1, ay - 2 * sy + 1); ++by)
              for (int bx = max(-1, ax - 2 * sx - 1); bx <= min(1, ax - 2 * sx + 1); ++bx) {
                const int jx = 2 * (ix + sx) + bx, jy = 2 * (iy + sy) + by, jz = 2 * (iz + sz) + bz;
                if (!in_mesh(mf, jx, jy, jz)) continue;
                const int ox = jx - fx, oy = jy - fy, oz = jz - fz;
                const int slot = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
                const double w = wa * (bx ? 0.5 : 1.0) * (by ? 0.5 : 1.0) * (bz ? 0.5 : 1.0);
                const VT *B = Af + ((long long)slot * mf.nnode + fi) * 9;
                double mcol[3] = {1.0, 1.0, 1.0};
                if (con) {
                  const long long fj = node_at(mf, jx, jy, jz);
                  for (int cc = 0; cc < 3; ++cc) mcol[cc] = con[3 * fj + cc] ? 0.0 : 1.0;
                }
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                  for (int cc = 0; cc < 3; ++cc) acc[3 * r + cc] += w * mr[r] * mcol[cc] * (double)B[3 * r + cc];
              }
        }
  }
  VT *out = Ac + ((long long)S * mc.nnode + I) * 9;
#pragma unroll
  for (int q = 0; q < 9; ++q) out[q] = (VT)acc[q];
}

// 1/diag (levels >= 1; level 0 keeps the solver's own) and block maxima of the Gershgorin ratio
// sum_j |A_ij| / |A_ii| of each row (rows with a zero diagonal are left out and get dinv 0).
template <typename VT>
__global__ void __launch_bounds__(256) k_mg_diag(MeshDims m, const VT *__restrict__ A, double *__restrict__ dinv, double *part) {
  __shared__ double red[4];
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double g = 0.0;
  if (n < m.nnode) {
    double rs[3] = {0, 0, 0};
    for (int s = 0; s < 27; ++s) {
      const VT *B = A + ((long long)s * m.nnode + n) * 9;
#pragma unroll
      for (int q = 0; q < 9; ++q) rs[q / 3] += fabs((double)B[q]);
    }
    const VT *D = A + (13 * m.nnode + n) * 9;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double d = (double)D[4 * r];
      if (dinv) dinv[3 * n + r] = (d == 0.0) ? 0.0 : 1.0 / d;
      if (d != 0.0) g = fmax(g, rs[r] / fabs(d));
    }
  }
  // block maximum
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) g = fmax(g, __shfl_down(g, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, red[w]);
    part[blockIdx.x] = v;
  }
}

// lam = max of the block maxima (one block)
__global__ void __launch_bounds__(256) k_mg_lam(const double *part, unsigned n, double *lam) {
  __shared__ double red[4];
  double g = 0.0;
  for (unsigned b = threadIdx.x; b < n; b += blockDim.x) g = fmax(g, part[b]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) g = fmax(g, __shfl_down(g, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, red[w]);
    *lam = v;
  }
}

// Dense inverse of the coarsest operator (nd = 3 * nodes <= MG_COARSE_DOFS), one block:
// in-place Gauss-Jordan with partial pivoting; row interchanges are undone as column
// interchanges of the inverse in reverse order.  A zero pivot is replaced by 1.
template <typename VT>
__global__ void __launch_bounds__(256) k_mg_factor(MeshDims m, const VT *__restrict__ A, double *__restrict__ Minv) {
  __shared__ double M[MG_COARSE_DOFS * MG_COARSE_DOFS];
  __shared__ double fcol[MG_COARSE_DOFS];
  __shared__ int perm[MG_COARSE_DOFS];
  const int nn = (int)m.nnode, nd = 3 * nn;
  for (int e = threadIdx.x; e < nd * nd; e += blockDim.x) M[e] = 0.0;
  __syncthreads();
  // dense row (I, r), column (J, c) = block of slot(J - I) at node I
  for (int e = threadIdx.x; e < 27 * nn; e += blockDim.x) {
    const int s = e / nn, I = e - s * nn;
    int i, j, k;
    node_ijk32(m, I, i, j, k);
    bool ok;
    const long long J = nb_masked(m, I, i, j, k, s, ok);
    if (!ok) continue;
    const VT *B = A + ((long long)s * nn + I) * 9;
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) M[(3 * I + r) * nd + 3 * (int)J + cc] = (double)B[3 * r + cc];
  }
  __syncthreads();
  for (int k = 0; k < nd; ++k) {
    if (threadIdx.x == 0) {
      int p = k;
      double best = fabs(M[k * nd + k]);
      for (int i = k + 1; i < nd; ++i)
        if (fabs(M[i * nd + k]) > best) { best = fabs(M[i * nd + k]); p = i; }
      perm[k] = p;
    }
    __syncthreads();
    const int p = perm[k];
    if (p != k)
      for (int j = threadIdx.x; j < nd; j += blockDim.x) {
        const double t = M[k * nd + j];
        M[k * nd + j] = M[p * nd + j];
        M[p * nd + j] = t;
      }
    __syncthreads();
    double piv = M[k * nd + k];
    if (piv == 0.0) piv = 1.0;
    __syncthreads();
    for (int j = threadIdx.x; j < nd; j += blockDim.x) M[k * nd + j] = (j == k ? 1.0 : M[k * nd + j]) / piv;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) fcol[i] = (i == k) ? 0.0 : M[i * nd + k];
    __syncthreads();
    for (int e = threadIdx.x; e < nd * nd; e += blockDim.x) {
      const int i = e / nd, j = e - i * nd;
      if (i == k) continue;
      M[e] = (j == k ? 0.0 : M[e]) - fcol[i] * M[k * nd + j];
    }
    __syncthreads();
  }
  for (int k = nd - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k)
      for (int i = threadIdx.x; i < nd; i += blockDim.x) {
        const double t = M[i * nd + k];
        M[i * nd + k] = M[i * nd + p];
        M[i * nd + p] = t;
      }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nd * nd; e += blockDim.x) Minv[e] = M[e];
}

// ---- V-cycle
//
// Every kernel returns at once after the solve has ended.  Only the Krylov kernels set that
// flag, never while a V-cycle kernel runs, so a plain load of it decides for the whole grid.
__device__ __forceinline__ bool solve_done(const KState *st) { return *(volatile const int *)&st->done != 0; }

// Node-level bodies (lane = node n of level m), shared by the per-level kernels and the
// one-block kernel of the coarse levels; UNR slots of loads in flight (nine for FP32 operators
// in the per-level kernels, three where registers are scarcer).
template <typename VT>
constexpr int mg_unroll() { return sizeof(VT) == 4 ? 9 : 3; }

// y = A v at node n, v given by a functor of the neighbour node (masked neighbours give zero)
template <int UNR, typename VT, typename V>
__device__ __forceinline__ void mg_apply(const MeshDims &m, const VT *__restrict__ A, long long n, V v, double y[3]) {
  int i, j, k;
  node_ijk32(m, n, i, j, k);
  y[0] = y[1] = y[2] = 0.0;
  const long long nn = m.nnode;
#pragma unroll UNR
  for (int s = 0; s < 27; ++s) {
    bool ok;
    const long long mn = nb_masked(m, n, i, j, k, s, ok);
    double t[3];
    v(mn, t);
    const double w = ok ? 1.0 : 0.0;
    block_mv(A + ((long long)s * nn + n) * 9, w * t[0], w * t[1], w * t[2], y);
  }
}

// Pre-smoothing from x = 0: d0 = D^-1 b / theta, res = b - A d0, d1 = c1 d0 + c2 D^-1 res,
// x = d0 + d1.  Writes x, d = d1 and res.
template <int UNR, typename VT>
__device__ __forceinline__ void mg_pre_node(const MeshDims &m, const VT *__restrict__ A, const double *__restrict__ dinv,
                                            const Cheb2 &ch, const double *__restrict__ b, double *__restrict__ x,
                                            double *__restrict__ d, double *__restrict__ res, long long n) {
  double y[3];
  mg_apply<UNR>(m, A, n, [&](long long q, double t[3]) {
    for (int c = 0; c < 3; ++c) t[c] = ch.inv_theta * dinv[3 * q + c] * b[3 * q + c];
  }, y);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double bi = b[3 * n + r], di = dinv[3 * n + r];
    const double d0 = di * bi * ch.inv_theta, rs = bi - y[r];
    const double d1 = ch.c1 * d0 + ch.c2 * di * rs;
    x[3 * n + r] = d0 + d1;
    d[3 * n + r] = d1;
    res[3 * n + r] = rs;
  }
}
// res <- res - A d (the residual b - A x after pre-smoothing)
template <int UNR, typename VT>
__device__ __forceinline__ void mg_resid_node(const MeshDims &m, const VT *__restrict__ A, const double *__restrict__ d,
                                              double *__restrict__ res, long long n) {
  double y[3];
  mg_apply<UNR>(m, A, n, [&](long long q, double t[3]) {
    for (int c = 0; c < 3; ++c) t[c] = d[3 * q + c];
  }, y);
#pragma unroll
  for (int r = 0; r < 3; ++r) res[3 * n + r] -= y[r];
}
// Post-smoothing, step 1: res = b - A x, d = D^-1 res / theta (x is updated in step 2, since
// neighbours still read it here)
template <int UNR, typename VT>
__device__ __forceinline__ void mg_post1_node(const MeshDims &m, const VT *__restrict__ A, const double *__restrict__ dinv,
                                              const Cheb2 &ch, const double *__restrict__ b, const double *__restrict__ x,
                                              double *__restrict__ d, double *__restrict__ res, long long n) {
  double y[3];
  mg_apply<UNR>(m, A, n, [&](long long q, double t[3]) {
    for (int c = 0; c < 3; ++c) t[c] = x[3 * q + c];
  }, y);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double rs = b[3 * n + r] - y[r];
    res[3 * n + r] = rs;
    d[3 * n + r] = dinv[3 * n + r] * rs * ch.inv_theta;
  }
}
// Post-smoothing, step 2: res <- res - A d, x += d + (c1 d + c2 D^-1 res); returns b.x_new at n
template <int UNR, typename VT>
__device__ __forceinline__ double mg_post2_node(const MeshDims &m, const VT *__restrict__ A, const double *__restrict__ dinv,
                                                const Cheb2 &ch, const double *__restrict__ b, double *__restrict__ x,
                                                const double *__restrict__ d, const double *__restrict__ res, long long n) {
  double y[3];
  mg_apply<UNR>(m, A, n, [&](long long q, double t[3]) {
    for (int c = 0; c < 3; ++c) t[c] = d[3 * q + c];
  }, y);
  double bx = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double d0 = d[3 * n + r], rs = res[3 * n + r] - y[r];
    const double xn = x[3 * n + r] + d0 + ch.c1 * d0 + ch.c2 * dinv[3 * n + r] * rs;
    x[3 * n + r] = xn;
    bx += b[3 * n + r] * xn;
  }
  return bx;
}
// b_c = P^T r_f at coarse node (ix, iy, iz): the 27 fine nodes around 2I, con masking level 0
__device__ __forceinline__ void restrict_node(const MeshDims &mf, const double *__restrict__ rf, const unsigned char *__restrict__ con,
                                              int ix, int iy, int iz, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  for (int az = -1; az <= 1; ++az)
    for (int ay = -1; ay <= 1; ++ay)
      for (int ax = -1; ax <= 1; ++ax) {
        const int fx = 2 * ix + ax, fy = 2 * iy + ay, fz = 2 * iz + az;
        if (!in_mesh(mf, fx, fy, fz)) continue;
        const long long fi = node_at(mf, fx, fy, fz);
        const double w = (ax ? 0.5 : 1.0) * (ay ? 0.5 : 1.0) * (az ? 0.5 : 1.0);
#pragma unroll
        for (int r = 0; r < 3; ++r) out[r] += (con && con[3 * fi + r]) ? 0.0 : w * rf[3 * fi + r];
      }
}
// x_f += P x_c at fine node n (at most 8 coarse parents), constrained dofs untouched
__device__ __forceinline__ void prolong_node(const MeshDims &mf, double *__restrict__ xf, const unsigned char *__restrict__ con,
                                             const MeshDims &mc, const double *__restrict__ xc, long long n) {
  int i, j, k;
  node_ijk32(mf, n, i, j, k);
  const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
  const int i1 = (i + 1) >> 1, j1 = (j + 1) >> 1, k1 = (k + 1) >> 1;  // equal to i0.. when the index is even
  const double wi = (i & 1) ? 0.5 : 1.0, wj = (j & 1) ? 0.5 : 1.0, wk = (k & 1) ? 0.5 : 1.0;
  double o[3] = {0, 0, 0};
  for (int c = 0; c < 8; ++c) {
    const int cx = (c & 1) ? i1 : i0, cy = (c & 2) ? j1 : j0, cz = (c & 4) ? k1 : k0;
    // an even index has a single parent: count it once (the second copy gets weight 0)
    const double w = (((c & 1) && !(i & 1)) || ((c & 2) && !(j & 1)) || ((c & 4) && !(k & 1))) ? 0.0 : wi * wj * wk;
    const long long J = node_at(mc, cx, cy, cz);
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] += w * xc[3 * J + r];
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
    if (!(con && con[3 * n + r])) xf[3 * n + r] += o[r];
}

// ---- per-level kernels (lane = node; blocks of at most VB lanes, exactly VB for the block partials of
// level 0)

template <typename VT>
__global__ void __launch_bounds__(VB) k_mg_pre(MeshDims m, const VT *__restrict__ A, const double *__restrict__ dinv,
                                              const double *__restrict__ lam, const double *__restrict__ b, double *__restrict__ x,
                                              double *__restrict__ d, double *__restrict__ res, const KState *st) {
  if (solve_done(st)) return;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n < m.nnode) mg_pre_node<mg_unroll<VT>()>(m, A, dinv, cheb2(*lam), b, x, d, res, n);
}
template <typename VT>
__global__ void __launch_bounds__(VB) k_mg_resid(MeshDims m, const VT *__restrict__ A, const double *__restrict__ d,
                                                double *__restrict__ res, const KState *st) {
  if (solve_done(st)) return;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n < m.nnode) mg_resid_node<mg_unroll<VT>()>(m, A, d, res, n);
}
__global__ void __launch_bounds__(VB) k_mg_restrict(MeshDims mf, const double *__restrict__ rf, const unsigned char *__restrict__ con,
                                                   MeshDims mc, double *__restrict__ bc, const KState *st) {
  if (solve_done(st)) return;
  const long long I = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (I >= mc.nnode) return;
  int ix, iy, iz;
  node_ijk32(mc, I, ix, iy, iz);
  double o[3];
  restrict_node(mf, rf, con, ix, iy, iz, o);
#pragma unroll
  for (int r = 0; r < 3; ++r) bc[3 * I + r] = o[r];
}
__global__ void __launch_bounds__(VB) k_mg_prolong(MeshDims mf, double *__restrict__ xf, const unsigned char *__restrict__ con,
                                                  MeshDims mc, const double *__restrict__ xc, const KState *st) {
  if (solve_done(st)) return;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n < mf.nnode) prolong_node(mf, xf, con, mc, xc, n);
}
template <typename VT>
__global__ void __launch_bounds__(VB) k_mg_post1(MeshDims m, const VT *__restrict__ A, const double *__restrict__ dinv,
                                                const double *__restrict__ lam, const double *__restrict__ b,
                                                const double *__restrict__ x, double *__restrict__ d, double *__restrict__ res,
                                                const KState *st) {
  if (solve_done(st)) return;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n < m.nnode) mg_post1_node<mg_unroll<VT>()>(m, A, dinv, cheb2(*lam), b, x, d, res, n);
}
// on level 0 (ROOT) the block partials of b.x (= r.z of the Krylov solve) follow
template <typename VT, bool ROOT>
__global__ void __launch_bounds__(VB) k_mg_post2(MeshDims m, const VT *__restrict__ A, const double *__restrict__ dinv,
                                                const double *__restrict__ lam, const double *__restrict__ b, double *__restrict__ x,
                                                const double *__restrict__ d, const double *__restrict__ res, const KState *st,
                                                double *partials) {
  __shared__ double lds[16];
  if (solve_done(st)) return;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (n < m.nnode) acc[0] = mg_post2_node<mg_unroll<VT>()>(m, A, dinv, cheb2(*lam), b, x, d, res, n);
  if (ROOT) block_partials<1>(acc, partials, lds);
}

// ---- the coarse levels in one block
//
// Levels lt..L-1 (each at most MG_TAIL_NODES nodes) are smoothed and the coarsest level L solved
// by one block of MG_TAIL_THREADS lanes, phases separated by block barriers: restriction of the
// residual of level lt-1, the downward leg, the dense solve, the upward leg.  The result is x of
// level lt; its prolongation to level lt-1 is a per-level kernel again.
constexpr int MG_TAIL_NODES = 256;
constexpr int MG_TAIL_THREADS = 256;
struct MgTail {
  int lt, L;
  MeshDims mf;                    // level lt-1
  const double *rf;               // its residual
  const unsigned char *con;       // its constraint mask (level 0 only)
  MeshDims m[MG_MAX_LEVELS];      // levels lt..L at [l]
  const void *A[MG_MAX_LEVELS];
  const double *dinv[MG_MAX_LEVELS], *lam;
  double *b[MG_MAX_LEVELS], *x[MG_MAX_LEVELS], *r[MG_MAX_LEVELS], *d[MG_MAX_LEVELS];
  const double *minv;
};
template <typename VT>
__global__ void __launch_bounds__(MG_TAIL_THREADS) k_mg_tail(MgTail t, const KState *st) {
  constexpr int TU = sizeof(VT) == 4 ? 3 : 2;
  __shared__ double bc[MG_COARSE_DOFS];
  if (solve_done(st)) return;
  const int L = t.L;
  // b_lt = P^T r_(lt-1)
  for (long long I = threadIdx.x; I < t.m[t.lt].nnode; I += blockDim.x) {
    int ix, iy, iz;
    node_ijk32(t.m[t.lt], I, ix, iy, iz);
    double o[3];
    restrict_node(t.mf, t.rf, t.con, ix, iy, iz, o);
    double *dst = (t.lt == L) ? nullptr : t.b[t.lt];
    for (int c = 0; c < 3; ++c) {
      if (dst) dst[3 * I + c] = o[c];
      else bc[3 * I + c] = o[c];
    }
  }
  __syncthreads();
  for (int l = t.lt; l < L; ++l) {
    const MeshDims &m = t.m[l];
    const VT *A = (const VT *)t.A[l];
    const Cheb2 ch = cheb2(t.lam[l]);
    for (long long n = threadIdx.x; n < m.nnode; n += blockDim.x) mg_pre_node<TU>(m, A, t.dinv[l], ch, t.b[l], t.x[l], t.d[l], t.r[l], n);
    __syncthreads();
    for (long long n = threadIdx.x; n < m.nnode; n += blockDim.x) mg_resid_node<TU>(m, A, t.d[l], t.r[l], n);
    __syncthreads();
    const MeshDims &mc = t.m[l + 1];
    for (long long I = threadIdx.x; I < mc.nnode; I += blockDim.x) {
      int ix, iy, iz;
      node_ijk32(mc, I, ix, iy, iz);
      double o[3];
      restrict_node(m, t.r[l], nullptr, ix, iy, iz, o);
      for (int c = 0; c < 3; ++c) {
        if (l + 1 == L) bc[3 * I + c] = o[c];
        else t.b[l + 1][3 * I + c] = o[c];
      }
    }
    __syncthreads();
  }
  // coarsest: x_L = Minv b_L
  {
    const int nd = 3 * (int)t.m[L].nnode;
    for (int row = threadIdx.x; row < nd; row += blockDim.x) {
      double s = 0.0;
      for (int col = 0; col < nd; ++col) s += t.minv[row * nd + col] * bc[col];
      t.x[L][row] = s;
    }
  }
  __syncthreads();
  for (int l = L - 1; l >= t.lt; --l) {
    const MeshDims &m = t.m[l];
    const VT *A = (const VT *)t.A[l];
    const Cheb2 ch = cheb2(t.lam[l]);
    for (long long n = threadIdx.x; n < m.nnode; n += blockDim.x) prolong_node(m, t.x[l], nullptr, t.m[l + 1], t.x[l + 1], n);
    __syncthreads();
    for (long long n = threadIdx.x; n < m.nnode; n += blockDim.x) mg_post1_node<TU>(m, A, t.dinv[l], ch, t.b[l], t.x[l], t.d[l], t.r[l], n);
    __syncthreads();
    for (long long n = threadIdx.x; n < m.nnode; n += blockDim.x) mg_post2_node<TU>(m, A, t.dinv[l], ch, t.b[l], t.x[l], t.d[l], t.r[l], n);
    __syncthreads();
  }
}

// ---- Krylov kernels of the multigrid-preconditioned CG (the operator kernel is k_cg_spmv)

// Iteration it: the stopping test from the update's partials of ||D^-1 r||^2 (it >= 1), r.z of the
// last V-cycle, b = (r.z)_it / (r.z)_(it-1); p <- z (it = 0) or z + b p.
__global__ void __launch_bounds__(VB) k_mgcg_p(long long N, const double *__restrict__ Z, double *__restrict__ P, KState *st,
                                               const double *upart, unsigned n_upart, const double *zpart, unsigned n_zpart, int it) {
  __shared__ double lds[32];
  if (solve_ended(st)) return;
  double rz[1];
  sum_partials<1>(zpart, n_zpart, rz, lds);
  double b = 0.0;
  if (it == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rzh[0] = rz[0];
  } else {
    double dd[1];
    sum_partials<1>(upart, n_upart, dd, lds);
    const double dp = sqrt(dd[0]);
    const bool breakdown = (rz[0] == 0.0);
    const bool nonfinite = !isfinite(dp) || !isfinite(rz[0]);
    const bool done = (dp <= st->tol) || it >= st->maxit || breakdown || nonfinite;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->dp = dp;
      if (done) {
        st->done = 1;
        st->it = it;
        st->breakdown = nonfinite ? KRYLOV_NONFINITE : ((breakdown && !(dp <= st->tol)) ? 1 : 0);
      } else {
        st->rzh[it & 1] = rz[0];
      }
    }
    if (done) return;
    b = rz[0] / st->rzh[(it - 1) & 1];
  }
  const long long stride = (long long)gridDim.x * VB;
  for (long long r0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; r0 < N; r0 += VPT * stride) {
    double z[VPT], p[VPT];
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const long long r = r0 + j * stride;
      const bool ok = r < N;
      z[j] = ok ? Z[r] : 0.0;
      p[j] = (ok && it > 0) ? P[r] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const long long r = r0 + j * stride;
      if (r < N) P[r] = (it == 0) ? z[j] : z[j] + b * p[j];
    }
  }
}

// a = (r.z)_it / (p.q); x += a p; r -= a q; block partials of ||D^-1 r||^2
__global__ void __launch_bounds__(VB) k_mgcg_update(long long N, double *__restrict__ X, const double *__restrict__ P,
                                                    double *__restrict__ R, const double *__restrict__ Q,
                                                    const double *__restrict__ dinv, const KState *st, const double *opart,
                                                    unsigned n_opart, double *upart, int it) {
  __shared__ double lds[32];
  if (solve_ended(st)) return;
  double pq[1];
  sum_partials<1>(opart, n_opart, pq, lds);
  const double a = st->rzh[it & 1] / pq[0];
  const long long stride = (long long)gridDim.x * VB;
  double acc[1] = {0.0};
  for (long long r0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; r0 < N; r0 += VPT * stride) {
    double x[VPT], p[VPT], rr[VPT], q[VPT], d[VPT];
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const long long r = r0 + j * stride;
      const bool ok = r < N;
      x[j] = ok ? X[r] : 0.0; p[j] = ok ? P[r] : 0.0; rr[j] = ok ? R[r] : 0.0; q[j] = ok ? Q[r] : 0.0; d[j] = ok ? dinv[r] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const long long r = r0 + j * stride;
      if (r < N) {
        const double rn = rr[j] - a * q[j];
        X[r] = x[j] + a * p[j];
        R[r] = rn;
        acc[0] += (rn * d[j]) * (rn * d[j]);
      }
    }
  }
  __syncthreads();
  block_partials<1>(acc, upart, lds);
}

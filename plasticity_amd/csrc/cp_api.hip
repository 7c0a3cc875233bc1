// plasticity_amd: C ABI, mesh/DoF integer work, assembly, post-increment and Krylov kernels
// for gfx950.  Declarations and reference citations: include/plasticity_amd.h.
//
// Data layout in HBM (one context):
//   per-qp state, structure of arrays, component-major:  X[c * nq + Q]  (Q = 8 e + q)
//     Fp_conv/Fp_iter/Fe_conv/Fe_iter 9, s_conv/s_iter ns, slip_conv/slip_iter ns, rotnew 3, T 9
//   Jacobian in 27-point node-block form: K[(slot * nnode + n) * 9 + 3a + b] is the 3x3 block
//     coupling node n (rows 3n+a) to its neighbour n + off(slot) (cols 3m+b),
//     slot = (dz+1)*9 + (dy+1)*3 + (dx+1).  Entries outside the mesh stay zero.  Each block is
//     72 (FP64) or 36 (FP32) contiguous bytes, so one lane loads it with wide loads and the 64
//     lanes of a wave, holding consecutive nodes, read one contiguous span per slot.
//   vectors: dof = 3 n + component, FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/plasticity_amd.h"
#include "cp_qp.h"
#include "cp_coop.h"
#include "cp_tables.h"
#include "cp_mlp.h"

namespace {

thread_local std::string g_err = "";

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// Inverse of a 3x3 matrix by cofactors; false when the determinant is zero.
bool host_inv3(const double A[3][3], double R[3][3]) {
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1], c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2],
               c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  if (det == 0.0) return false;
  const double id = 1.0 / det;
  R[0][0] = c00 * id;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  R[1][0] = c01 * id;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  R[2][0] = c02 * id;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  return true;
}

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(PA_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int NSMAX = 48;
constexpr int RED_BLOCK = 256;

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block sum of up to NV values (blockDim.x multiple of 64, <= 1024); result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* NV * 16 */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) lds[k * 16 + w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0;
      for (int i = 0; i < nw; ++i) s += lds[k * 16 + i];
      v[k] = s;
    }
  }
}

// Deterministic grid reduction: each block writes its NV partial sums; the last block to
// arrive sums the partials, every thread a fixed strided subset, then a block sum.  Returns
// true in thread 0 of the last block, with v holding the totals.
template <int NV>
__device__ bool grid_sum(double (&v)[NV], double *partials, unsigned *counter, double *lds) {
  __shared__ bool last;
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[blockIdx.x * NV + k] = v[k];
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
      s += __hip_atomic_load(&partials[b * NV + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v[k] = s;
  }
  __syncthreads();  // lds reuse
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) *counter = 0;
  return threadIdx.x == 0;
}

// ------------------------------------------------------------ mesh helpers
#include "cp_passes.h"

__device__ __forceinline__ void node_ijk(const MeshDims &m, long long n, int &i, int &j, int &k) {
  i = (int)(n % m.nnx);
  j = (int)((n / m.nnx) % m.nny);
  k = (int)(n / ((long long)m.nnx * m.nny));
}

// ---------------------------------------------------- grain ids (integer work)
// Nearest voxel centre to the cell centre per axis, ties to the lower voxel
// (crystalOrientationsIO.cc:121-193).  Arithmetic is spelled with _rn intrinsics so no
// fused multiply-add changes a key or a centre.
__device__ int nearest_voxel(double p, int nv, double span) {
  const double st = __ddiv_rn(span, (double)nv);
  const double half = __ddiv_rn(st, 2.0);
  int lo = 0, hi = nv;  // lower_bound: first key >= p
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const double key = __dadd_rn(__dmul_rn((double)mid, st), half);
    if (key < p) lo = mid + 1; else hi = mid;
  }
  if (lo == nv) return nv - 1;
  if (lo == 0) return 0;
  const double kl = __dadd_rn(__dmul_rn((double)(lo - 1), st), half), ku = __dadd_rn(__dmul_rn((double)lo, st), half);
  const double dl = fabs(__dsub_rn(kl, p)), du = fabs(__dsub_rn(ku, p));
  return (du >= dl) ? lo - 1 : lo;
}

__global__ void k_element_grains(MeshDims m, double sx, double sy, double sz, const unsigned *__restrict__ vox, int nv0,
                                 int nv1, int nv2, unsigned *__restrict__ eg) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= m.nelem) return;
  const int ex = (int)(e % m.nx), ey = (int)((e / m.nx) % m.ny), ez = (int)(e / ((long long)m.nx * m.ny));
  const double h[3] = {__ddiv_rn(sx, (double)m.nx), __ddiv_rn(sy, (double)m.ny), __ddiv_rn(sz, (double)m.nz)};
  const int ie[3] = {ex, ey, ez};
  const double span[3] = {sx, sy, sz};
  const int nv[3] = {nv0, nv1, nv2};
  int idx[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double c = __ddiv_rn(__dadd_rn(__dmul_rn((double)ie[d], h[d]), __dmul_rn((double)(ie[d] + 1), h[d])), 2.0);
    idx[d] = nearest_voxel(c, nv[d], span[d]);
  }
  eg[e] = vox[((long long)idx[0] * nv1 + idx[1]) * nv2 + idx[2]];
}

// ------------------------------------------------------ sparsity (integer work)
// Full 27-point pattern, columns ascending (DoFTools::make_sparsity_pattern, init.cc:183-199).
// Entries before node (i,j,k) follow in closed form from per-axis neighbour counts
// c(i) = 2 at the ends, 3 inside (1 for a single node) and prefix P(i) = sum_{i'<i} c(i').
__device__ __forceinline__ long long cnt1(int i, int nn) { return (long long)(min(i + 1, nn - 1) - max(i - 1, 0) + 1); }
__device__ __forceinline__ long long pre1(int i, int nn) {
  long long s = 0;
  if (i > 0) s = cnt1(0, nn) + (long long)(i - 1) * 3;
  if (i == nn && nn >= 2) s = s - 3 + cnt1(nn - 1, nn);
  return s;
}
__global__ void k_csr_rowptr(MeshDims m, long long *__restrict__ rp) {
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long ndof = 3 * m.nnode;
  if (r > ndof) return;
  if (r == ndof) {
    rp[r] = 9 * pre1(m.nnx, m.nnx) * pre1(m.nny, m.nny) * pre1(m.nnz, m.nnz);
    return;
  }
  const long long n = r / 3;
  const int c = (int)(r % 3);
  int i, j, k;
  node_ijk(m, n, i, j, k);
  const long long Tx = pre1(m.nnx, m.nnx), Ty = pre1(m.nny, m.nny);
  const long long before = pre1(k, m.nnz) * Ty * Tx + cnt1(k, m.nnz) * pre1(j, m.nny) * Tx + cnt1(k, m.nnz) * cnt1(j, m.nny) * pre1(i, m.nnx);
  const long long len = 3 * cnt1(i, m.nnx) * cnt1(j, m.nny) * cnt1(k, m.nnz);
  rp[r] = 9 * before + c * len;
}
// one thread per (row, position): column of each entry, and the SB source of its value
template <typename KT>
__global__ void k_csr_cols(MeshDims m, const long long *__restrict__ rp, int *__restrict__ ci, const KT *__restrict__ K,
                           double *__restrict__ val) {
  const long long r = blockIdx.x * (long long)blockDim.y + threadIdx.y;
  if (r >= 3 * m.nnode) return;
  const long long n = r / 3;
  const int a = (int)(r % 3);
  int i, j, k;
  node_ijk(m, n, i, j, k);
  const int p = threadIdx.x;  // 0..80
  const int xl = (i > 0) ? -1 : 0, xh = (i < m.nnx - 1) ? 1 : 0;
  const int yl = (j > 0) ? -1 : 0, yh = (j < m.nny - 1) ? 1 : 0;
  const int zl = (k > 0) ? -1 : 0, zh = (k < m.nnz - 1) ? 1 : 0;
  const int ax = xh - xl + 1, ay = yh - yl + 1, az = zh - zl + 1;
  if (p >= 3 * ax * ay * az) return;
  const int b = p % 3, nb = p / 3;
  const int dx = xl + nb % ax, dy = yl + (nb / ax) % ay, dz = zl + nb / (ax * ay);
  const long long mnode = n + dx + (long long)dy * m.nnx + (long long)dz * m.nnx * m.nny;
  const long long pos = rp[r] + p;
  if (ci) ci[pos] = (int)(3 * mnode + b);
  if (val) {
    const int slot = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
    val[pos] = (double)K[((long long)slot * m.nnode + n) * 9 + 3 * a + b];
  }
}

// ------------------------------------------------------------ Dirichlet lines
// setBoundaryValues + applyDirichletBCs (boundaryConditions.cc:6-114, 269-275, 452-491): the first
// matching face rule wins; cyclic rule first when enabled; the velocity-gradient rule, when
// enabled, takes every boundary dof no face rule took, with value dF X at the node's position X.
struct BCArgs {
  int fdc[6][3];
  double dU[6][3];
  int cyclic, cface, cdof;
  double csign;
  int velgrad;
  double dF[3][3], h[3];
};
__global__ void k_dirichlet(MeshDims m, BCArgs bc, unsigned char *__restrict__ con, double *__restrict__ delu) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= m.nnode) return;
  int i, j, k;
  node_ijk(m, n, i, j, k);
  const bool onb[6] = {i == 0, i == m.nx, j == 0, j == m.ny, k == 0, k == m.nz};
  const bool bnd = onb[0] || onb[1] || onb[2] || onb[3] || onb[4] || onb[5];
  for (int d = 0; d < 3; ++d) {
    bool flag = false;
    double val = 0.0;
    if (bnd) {
      if (bc.cyclic && d == bc.cdof - 1 && onb[bc.cface - 1]) {
        flag = true;
        val = bc.csign * bc.dU[bc.cface - 1][d];
      } else {
        for (int f = 0; f < 6; ++f)
          if (bc.fdc[f][d] && onb[f]) { flag = true; val = bc.dU[f][d]; break; }
      }
      if (!flag && bc.velgrad) {
        flag = true;
        val = bc.dF[d][0] * (i * bc.h[0]) + bc.dF[d][1] * (j * bc.h[1]) + bc.dF[d][2] * (k * bc.h[2]);
      }
    }
    con[3 * n + d] = flag ? 1 : 0;
    delu[3 * n + d] = flag ? val : 0.0;
  }
}

// ------------------------------------------------------------ singular local systems
// The pseudo-inverse branch of inactive-slip removal (inactiveSlipRemoval.cc:26,99), for the
// qps whose A_PA LU finds numerically singular.  Pass 1 runs every qp on the LU path; such a qp
// claims entry e of the deferred queue and, below the store's capacity, leaves its system
// (A_PA, b_PA, n) in the store.  Then SING_ROUNDS rounds of: k_sing_svd solves every pending
// system with a group of lanes per system (lane = row, U and V in LDS) and appends x to the
// entry's solutions; a replay kernel reruns each pending qp, feeding the stored solutions to its
// singular solves in order; in the last round a further singular solve runs in the lane itself
// (few qps meet a second one, and a round costs the latency of a whole qp), otherwise the qp
// leaves its next system.  Entries above capacity are finished with the pseudo-inverse in their
// own lane.  Every path computes the same x (same Jacobi sweep order and rules).
// FNV-1a over the bit patterns of n, A_PA and b_PA
template <int NS, typename PA>
__device__ __forceinline__ unsigned long long sing_key(int n, PA A, const double *b) {
  unsigned long long h = 1469598103934665603ull ^ (unsigned long long)n;
  auto mix = [&](double v) { h = (h ^ (unsigned long long)__double_as_longlong(v)) * 1099511628211ull; };
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) mix(A[i * NS + j]);
    mix(b[i]);
  }
  return h;
}
template <int NS, typename PA>
__device__ __forceinline__ void sing_write(const SingStore &st, unsigned e, int round, int n, PA A, const double *b) {
  double *dA = st.A + (size_t)e * NS * NS;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) dA[i * NS + j] = A[i * NS + j];
    st.b[(size_t)e * NS + i] = b[i];
  }
  st.n[e] = n;
  if (round < SING_ROUNDS) st.key[(size_t)e * SING_ROUNDS + round] = sing_key<NS>(n, A, b);
}
// What a qp does at a numerically singular A_PA, by pass: queue it (pass 1), replay the stored
// solutions (rounds), or solve in its own lane (the last pass).  One functor type for all passes
// means one compiled qp_update, so a rerun meets bit-identical systems in the same order; a
// stored solution is still used only for the system it was computed from (checked by hash).
enum { SING_CLAIM = 0, SING_REPLAY = 1, SING_LANE = 2 };
template <int NS>
struct SingAny {
  int mode;
  SingStore st;
  unsigned *count, *queue;
  unsigned Q, e;
  int have, k;
  cp::LanePtr ws = {nullptr, 1};  // this lane's cp::LaneWs array (48-system variant), else unused
  __device__ void svd(int n, const double *A, const double *b, double *x) const { cp::svd_pinv_solve<NS>(n, A, b, x); }
  __device__ void svd(int n, cp::LanePtr A, const double *b, double *x) const {
    cp::svd_pinv_solve_ws<NS>(n, A, b, x, ws + cp::LaneWs<NS>::UV);
  }
  template <typename PA>
  __device__ bool operator()(int n, PA A, const double *b, double *x) {
    if (mode == SING_LANE) {
      svd(n, A, b, x);
      return true;
    }
    if (mode == SING_CLAIM) {
      const unsigned ee = atomicAdd(count, 1u);
      queue[ee] = Q;
      if (ee < st.cap) {
        sing_write<NS>(st, ee, 0, n, A, b);
        st.have[ee] = 0;
        st.done[ee] = 0;
      }
      return false;
    }
    if (k < have) {
      if (sing_key<NS>(n, A, b) == st.key[(size_t)e * SING_ROUNDS + k]) {
        const double *src = st.X + ((size_t)e * SING_ROUNDS + k) * NS;
        for (int i = 0; i < n; ++i) x[i] = src[i];
      } else {
        atomicAdd(st.miss, 1u);
        svd(n, A, b, x);
      }
      ++k;
      return true;
    }
    if (have >= SING_ROUNDS) {  // last round: solve here
      svd(n, A, b, x);
      ++k;
      return true;
    }
    sing_write<NS>(st, e, k, n, A, b);
    return false;
  }
};

template <int NS>
struct SingRef {
  SingAny<NS> *f;
  template <typename PA>
  __device__ bool operator()(int n, PA A, const double *b, double *x) const { return (*f)(n, A, b, x); }
};

template <int NS>
constexpr int sing_group() { return NS <= 16 ? 16 : (NS <= 32 ? 32 : 64); }
// systems per block of k_sing_svd: U and V take 2 NS^2 doubles of LDS each
template <int NS>
constexpr int sing_per_block() { return NS <= 24 ? 4 : 1; }

// Orders this lane's LDS writes before the group's next reads: the lanes of a group are in one
// wavefront, which executes its LDS instructions in order; the fences keep the compiler from
// moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x = pinv(A_PA) b_PA for every pending entry: svd_pinv_solve (cp_qp.h) with a group of lanes per
// system.  Each round-robin round takes two steps: lane k forms the rotation of pair k (its sums
// in the in-lane order, from LDS), then lane i applies the round's rotations to row i of U and V.
// The pairs of a round share no column, so the result is the one the in-lane routine gives.
// Control flow is uniform within a group; a group sits inside one wavefront.
template <int NS>
__global__ void __launch_bounds__(sing_per_block<NS>() * sing_group<NS>()) k_sing_svd(SingStore st, const unsigned *count) {
  constexpr int G = sing_group<NS>(), SPB = sing_per_block<NS>(), NP = (NS + 1) / 2;
  static_assert(NP <= G, "a group needs a lane per pair");
  __shared__ double lds[SPB][2 * NS * NS + NS + 2 * NP];
  __shared__ int flag[SPB][NP];
  const int grp = threadIdx.x / G, i = threadIdx.x % G;
  double *U = lds[grp], *V = lds[grp] + NS * NS, *bv = lds[grp] + 2 * NS * NS;  // column j, row r at j * NS + r
  double *cs = lds[grp] + 2 * NS * NS + NS;
  int *fl = flag[grp];
  const unsigned ne = min(*count, st.cap);
  for (unsigned e = blockIdx.x * SPB + grp; e < ne; e += gridDim.x * SPB) {
    const int have = st.have[e];
    if (st.done[e] || have >= SING_ROUNDS) continue;
    const int n = st.n[e];
    const bool row = i < n;
    const double *A = st.A + (size_t)e * NS * NS;
    if (row) {
      for (int j = 0; j < n; ++j) {
        U[j * NS + i] = A[i * NS + j];
        V[j * NS + i] = (i == j) ? 1.0 : 0.0;
      }
      bv[i] = st.b[(size_t)e * NS + i];
    }
    wave_lds_sync();
    const int m = n + (n & 1), np = m / 2;
    for (int sweep = 0; sweep < 80; ++sweep) {
      bool rotated = false;
      for (int r = 0; r < m - 1; ++r) {
        if (i < np) {
          int p, q;
          cp::rr_pair(m, r, i, p, q);
          int f = 0;
          double c = 1.0, sn = 0.0;
          if (q < n) {
            double al = 0, be = 0, ga = 0;
            for (int k = 0; k < n; ++k) {
              const double up = U[p * NS + k], uq = U[q * NS + k];
              al += up * up; be += uq * uq; ga += up * uq;
            }
            if (!(ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be))) {
              f = 1;
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
        wave_lds_sync();
        for (int k = 0; k < np; ++k) {
          if (!fl[k]) continue;
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
        wave_lds_sync();
      }
      if (!rotated) break;
    }
    double sig[NS], smax = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      sig[j] = 0.0;
      if (j < n) {
        double ss = 0.0;
        for (int k = 0; k < n; ++k) ss += U[j * NS + k] * U[j * NS + k];
        sig[j] = sqrt(ss);
        smax = fmax(smax, sig[j]);
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < NS; ++j)
      if (j < n && row && sig[j] > 0) U[j * NS + i] /= sig[j];
    wave_lds_sync();
    // a single small singular value is kept when the data fix its step (svd_pinv_solve)
    double lim = smax * CP_PINV_RTOL;
    int nsmall = 0, ks = -1;
#pragma unroll
    for (int k = 0; k < NS; ++k)
      if (k < n && sig[k] <= lim) { nsmall++; ks = k; }
    if (nsmall == 1) {
      double sks = 0.0, s2 = smax;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        if (k == ks) sks = sig[k];
        else if (k < n) s2 = fmin(s2, sig[k]);
      }
      if (sks > 0) {
        double ub = 0, bb = 0;
        for (int r = 0; r < n; ++r) { ub += U[ks * NS + r] * bv[r]; bb += bv[r] * bv[r]; }
        const double eps = 2.220446049250313e-16;
        if (fabs(ub) * eps * smax * smax < 1e-2 * sqrt(bb) * sks * (s2 - sks)) lim = 0.0;
      }
    }
    double xi = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      if (k >= n || !(sig[k] > lim)) continue;
      double ub = 0;
      for (int r = 0; r < n; ++r) ub += U[k * NS + r] * bv[r];
      ub /= sig[k];
      if (row) xi += V[k * NS + i] * ub;
    }
    if (row) st.X[((size_t)e * SING_ROUNDS + have) * NS + i] = xi;
    if (i == 0) st.have[e] = have + 1;
    wave_lds_sync();  // the group's next system overwrites U, V, b
  }
}

// svd_pinv_solve in one lane per system, for pa_pinv_solve (A: [count][NS*NS], b and x: [count][NS])
template <int NS>
__global__ void __launch_bounds__(64) k_pinv_lane(unsigned count, int n, const double *A, const double *b, double *x) {
  const unsigned e = blockIdx.x * 64 + threadIdx.x;
  if (e >= count) return;
  double xl[NS];
  cp::svd_pinv_solve<NS>(n, A + (size_t)e * NS * NS, b + (size_t)e * NS, xl);
  for (int i = 0; i < n; ++i) x[(size_t)e * NS + i] = xl[i];
}

// ------------------------------------------------------------ assembly
// ---- assembly, stage 1: one lane per qp.  F from u, calculatePlasticity(.,.,1)
// (getElementalValues.cc:129-141); writes the iter history and JxW * dP/dF (as DT) and JxW * P.

// Lanes whose local problem meets a numerically singular A_PA (the pseudo-inverse branch of
// inactive-slip removal) are deferred: pass 1 runs every qp on the LU path and appends the
// singular ones to a device queue; pass 2 runs only the queued qps, so the slow branch never
// holds up a wavefront of regular qps.  Results do not depend on which lane computes a qp.
template <int NS, typename DT, bool KIN>
__device__ __attribute__((noinline)) bool qp_tangent_one(const QpArgs &a, long long Q, SingAny<NS> &sing) {
  const long long e = Q >> 3;
  const int q = (int)(Q & 7);
  const int ex = (int)(e % a.m.nx), ey = (int)((e / a.m.nx) % a.m.ny), ez = (int)(e / ((long long)a.m.nx * a.m.ny));
  double F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const long long nd = (ex + (v & 1)) + (long long)a.m.nnx * ((ey + ((v >> 1) & 1)) + (long long)a.m.nny * (ez + ((v >> 2) & 1)));
    double g[3];
    shape_grad(q, v, a.h, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double uc = a.u[3 * nd + c];
#pragma unroll
      for (int j = 0; j < 3; ++j) F[3 * c + j] += uc * g[j];
    }
  }
  F[0] += 1; F[4] += 1; F[8] += 1;
  double Fp_t[9], s[NS], x[NS], P[9], T[9], Fe[9], Fp[9], W[KIN ? NS : 1];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fp_t[i] = a.Fp_conv[i * a.nq + Q];
  const double *mat;
  int ns;
  qp_phase(a.ph, a.mat, a.ns, a.egrain[e], mat, ns);
  for (int i = 0; i < ns; ++i) s[i] = a.s_conv[i * a.nq + Q];
  if constexpr (KIN)
    for (int i = 0; i < ns; ++i) W[i] = a.W_conv[i * a.nq + Q];
  const double *gt = a.gtab + table_row(a.tw, a.egrain, Q) * a.gstride;
  TwinArgs tw = a.tw;
  if (a.ph.gphase && tw.on) phase_twin(a.ph, a.egrain[e], tw);
  const int nslip = tw.on ? tw.nslip : ns;
  bool plastic = false;
  const int na = cp::qp_update<NS, DT, KIN>(mat, gt, ns, a.max_solver_iter, F, Fp_t, s, x, 1, P, T, Fe, Fp, (DT *)a.D + Q,
                                            a.nq, &plastic, a.JxW, SingRef<NS>{&sing}, W, nullptr, 1, nullptr, nslip, sing.ws);
  if (na < 0) return false;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    a.Fe_iter[i * a.nq + Q] = Fe[i];
    a.Fp_iter[i * a.nq + Q] = Fp[i];
    a.P[i * a.nq + Q] = a.JxW * P[i];
  }
  for (int i = 0; i < ns; ++i) a.s_iter[i * a.nq + Q] = s[i];
  if constexpr (KIN)
    for (int i = 0; i < ns; ++i) a.W_iter[i * a.nq + Q] = W[i];
  // slipfraction_iter is written only on the plastic path (calculatePlasticity.cc:454-456)
  if (plastic)
    for (int i = 0; i < nslip; ++i) a.slip_iter[i * a.nq + Q] = a.slip_conv[i * a.nq + Q] + x[i];
  if (tw.on && tw.ntw > 0) twin_epilogue(tw, Q, a.nq, plastic, x, a.s_iter);
  return true;
}

// The qp passes keep their active-set and tangent matrices in per-lane private memory and are
// bound by its latency: two waves per SIMD (256 VGPRs, more spilled) run the first pass 7 % faster
// than one wave with 512 registers; three or four waves are no faster than one.  The cap is set
// on every kernel that calls the per-qp function, which is compiled once for all of them.
#define QP_WAVES __attribute__((amdgpu_waves_per_eu(2)))

// this lane's array in the 48-system variant's workspace (unused for the others, whose matrices
// stay in private memory): each wavefront owns a block of 64 arrays interleaved lane by lane, so a
// wavefront's accesses to one element coalesce and its data stays within a few pages; the grids of
// those passes are capped at WS_BLOCKS one-wave blocks
constexpr unsigned WS_BLOCKS = 512;
__device__ __forceinline__ cp::LanePtr wave_block_ptr(double *ws, long long per_lane) {
  const unsigned lane = blockIdx.x * blockDim.x + threadIdx.x;
  return cp::LanePtr{ws + (size_t)(lane >> 6) * 64 * per_lane + (lane & 63), 64};
}
template <int NS>
__device__ __forceinline__ cp::LanePtr lane_ws(double *ws) {
  if constexpr (cp::lane_ws<NS>()) return wave_block_ptr(ws, cp::LaneWs<NS>::size);
  else return cp::LanePtr{nullptr, 1};
}

// pass 1: lane = qp (grid-stride), LU path only; singular local problems are queued
template <int NS, typename DT, bool KIN>
__global__ void __launch_bounds__(256) QP_WAVES k_qp_tangent(QpArgs a) {
  for (long long Q = blockIdx.x * (long long)blockDim.x + threadIdx.x; Q < a.nq; Q += (long long)gridDim.x * blockDim.x) {
    SingAny<NS> f{SING_CLAIM, a.sing, a.qcount, a.queue, (unsigned)Q, 0u, 0, 0, lane_ws<NS>(a.ws)};
    qp_tangent_one<NS, DT, KIN>(a, Q, f);  // singular: queued
  }
}
// rounds: the pending entries with their stored solutions
template <int NS, typename DT, bool KIN>
__global__ void __launch_bounds__(64) QP_WAVES k_qp_tangent_replay(QpArgs a) {
  const unsigned n = min(*a.qcount, a.sing.cap);
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    if (a.sing.done[e]) continue;
    SingAny<NS> f{SING_REPLAY, a.sing, a.qcount, a.queue, 0u, e, a.sing.have[e], 0, lane_ws<NS>(a.ws)};
    if (qp_tangent_one<NS, DT, KIN>(a, a.queue[e], f)) a.sing.done[e] = 1;
  }
}
// pass 2: the queued qps, grid-stride, pseudo-inverse allowed
template <int NS, typename DT, bool KIN>
__global__ void __launch_bounds__(64) QP_WAVES k_qp_tangent_deferred(QpArgs a) {
  const unsigned n = *a.qcount;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (i >= a.sing.cap || !a.sing.done[i]) {
      SingAny<NS> f{SING_LANE, a.sing, a.qcount, a.queue, 0u, i, 0, 0, lane_ws<NS>(a.ws)};
      qp_tangent_one<NS, DT, KIN>(a, a.queue[i], f);
    }
}

__global__ void k_zero_u32(unsigned *p) { *p = 0u; }

// ---- assembly, stage 2: one colour of the element loop of assemble() (assemble.cc:64-90).
// Lane = (element, vertex v1) forms 3 rows of K_e and R_e (getElementalValues.cc:144-162) from
// the staged JxW dP/dF (KT arithmetic) and JxW P (FP64), and adds them into the node-block
// Jacobian with deal.II's condensation of Dirichlet lines
// (AffineConstraints::distribute_local_to_global): free-free entries added, free rows get
// -K_ij g_j for constrained columns at iteration 0, constrained rows get |K_ii| (or the mean
// |diag| of K_e when K_ii = 0) on the diagonal and nothing in R.  Elements of one colour share
// no node, so the read-modify-write needs no atomics and the sum order is fixed.
struct ElemArgs {
  MeshDims m;
  double h[3];
  int cx, cy, cz, ncx, ncy, ncz;
  long long nq;
  const void *D;
  const double *P;
  const unsigned char *con;
  const double *delu;
  int iter0;
  void *K;
  double *R;
};

constexpr int ASM_EL = 8;  // elements per 64-lane workgroup

// One block's ASM_EL elements of a colour: element ci0 + le of the colour (lexicographic over the
// colour's elements), qp 8 e + q at qbase + 8 e + q of the staged arrays; nodes, K, R, con and delu
// as in the mesh of `a` (a batch passes its RVE's mesh and pointers).  lds: the block's staging.
template <typename KT>
__device__ __forceinline__ void element_body(const ElemArgs &a, long long ci0, long long qbase, KT *sD, double *sP, double *sdiag) {
  const int t = threadIdx.x, le = t >> 3, sub = t & 7;
  const long long ncol = (long long)a.ncx * a.ncy * a.ncz;
  const long long ci = ci0 + le;
  const bool active = ci < ncol;
  long long e = 0;
  int ex = 0, ey = 0, ez = 0;
  if (active) {
    ex = a.cx + 2 * (int)(ci % a.ncx);
    ey = a.cy + 2 * (int)((ci / a.ncx) % a.ncy);
    ez = a.cz + 2 * (int)(ci / ((long long)a.ncx * a.ncy));
    e = ex + (long long)a.m.nx * (ey + (long long)a.m.ny * ez);
  }
  // stage the 8 elements' 64 qps: lane t loads qp (le, sub)
  const KT *Dg = (const KT *)a.D;
  if (active) {
    const long long Q = qbase + 8 * e + sub;
#pragma unroll 9
    for (int c = 0; c < 81; ++c) sD[c * 64 + t] = Dg[c * a.nq + Q];
#pragma unroll
    for (int c = 0; c < 9; ++c) sP[t * 9 + c] = a.P[c * a.nq + Q];
  }
  __syncthreads();
  long long nodes[8];
#pragma unroll
  for (int v = 0; v < 8; ++v)
    nodes[v] = (ex + (v & 1)) + (long long)a.m.nnx * ((ey + ((v >> 1) & 1)) + (long long)a.m.nny * (ez + ((v >> 2) & 1)));
  const int v1 = sub;
  KT Kr[3][24];
  double Rr[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Rr[i] = 0;
#pragma unroll
    for (int c = 0; c < 24; ++c) Kr[i][c] = 0;
  }
  if (active) {
    for (int q = 0; q < 8; ++q) {
      const KT *Dq = sD + (le * 8 + q);
      const double *Pq = sP + (le * 8 + q) * 9;
      double g1d[3];
      shape_grad(q, v1, a.h, g1d);
      KT g1[3] = {(KT)g1d[0], (KT)g1d[1], (KT)g1d[2]};
      KT g2[8][3];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        double g[3];
        shape_grad(q, v, a.h, g);
        g2[v][0] = (KT)g[0]; g2[v][1] = (KT)g[1]; g2[v][2] = (KT)g[2];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Rr[i] -= g1d[0] * Pq[3 * i] + g1d[1] * Pq[3 * i + 1] + g1d[2] * Pq[3 * i + 2];
        KT w[9];
#pragma unroll
        for (int jl = 0; jl < 9; ++jl)
          w[jl] = g1[0] * Dq[((3 * i + 0) * 9 + jl) * 64] + g1[1] * Dq[((3 * i + 1) * 9 + jl) * 64] + g1[2] * Dq[((3 * i + 2) * 9 + jl) * 64];
#pragma unroll
        for (int v2 = 0; v2 < 8; ++v2)
#pragma unroll
          for (int j = 0; j < 3; ++j) Kr[i][3 * v2 + j] += w[3 * j] * g2[v2][0] + w[3 * j + 1] * g2[v2][1] + w[3 * j + 2] * g2[v2][2];
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) sdiag[le * 24 + 3 * v1 + i] = (double)Kr[i][3 * v1 + i];
  }
  __syncthreads();
  if (!active) return;
  double avg = 0;
  for (int d = 0; d < 24; ++d) avg += fabs(sdiag[le * 24 + d]);
  avg /= 24.0;
  const long long n1 = nodes[v1];
  const long long nn = a.m.nnode;
  KT *K = (KT *)a.K;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const long long gi = 3 * n1 + i;
    if (a.con[gi]) {
      const double kd = fabs((double)Kr[i][3 * v1 + i]);
      K[(13 * nn + n1) * 9 + 4 * i] += (KT)(kd != 0 ? kd : avg);
      continue;
    }
    double rv = Rr[i];
#pragma unroll
    for (int v2 = 0; v2 < 8; ++v2) {
      const int slot = (((v2 >> 2) & 1) - ((v1 >> 2) & 1) + 1) * 9 + (((v2 >> 1) & 1) - ((v1 >> 1) & 1) + 1) * 3 + ((v2 & 1) - (v1 & 1) + 1);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const long long gj = 3 * nodes[v2] + j;
        if (a.con[gj]) {
          if (a.iter0) {
            const double g = a.delu[gj];
            if (g != 0.0) rv -= (double)Kr[i][3 * v2 + j] * g;
          }
        } else {
          K[((long long)slot * nn + n1) * 9 + 3 * i + j] += Kr[i][3 * v2 + j];
        }
      }
    }
    if (rv != 0.0) a.R[gi] += rv;
  }
}
template <typename KT>
__global__ void __launch_bounds__(64) k_element(ElemArgs a) {
  __shared__ KT sD[64 * 81];
  __shared__ double sP[64 * 9];
  __shared__ double sdiag[ASM_EL * 24];
  element_body<KT>(a, (long long)blockIdx.x * ASM_EL, 0, sD, sP, sdiag);
}

// ------------------------------------------------ post-increment qp pass

// symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations.  The sweeps end when the
// off-diagonal is zero or a sweep leaves the eigenvalues (diagonal of C) and V unchanged: the
// rotations only get smaller from there (an off-diagonal below ~1e-154 of the diagonal even gives
// the identity rotation, since th * th overflows), so further sweeps return the same result.
__device__ void sym_eig3(const double *Cin, double lam[3], double V[9]) {
  double C[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { C[i] = Cin[i]; V[i] = (i % 4 == 0) ? 1.0 : 0.0; }
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = C[1] * C[1] + C[2] * C[2] + C[5] * C[5];
    if (off == 0.0) break;
    double prev[12];
#pragma unroll
    for (int i = 0; i < 9; ++i) prev[i] = V[i];
    prev[9] = C[0]; prev[10] = C[4]; prev[11] = C[8];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = p + 1; q < 3; ++q) {
        if (C[3 * p + q] == 0.0) continue;
        const double th = (C[3 * q + q] - C[3 * p + p]) / (2.0 * C[3 * p + q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double ckp = C[3 * k + p], ckq = C[3 * k + q];
          C[3 * k + p] = c * ckp - s * ckq; C[3 * k + q] = s * ckp + c * ckq;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double cpk = C[3 * p + k], cqk = C[3 * q + k];
          C[3 * p + k] = c * cpk - s * cqk; C[3 * q + k] = s * cpk + c * cqk;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[3 * k + p], vkq = V[3 * k + q];
          V[3 * k + p] = c * vkp - s * vkq; V[3 * k + q] = s * vkp + c * vkq;
        }
      }
    bool same = prev[9] == C[0] && prev[10] == C[4] && prev[11] == C[8];
#pragma unroll
    for (int i = 0; i < 9; ++i) same = same && prev[i] == V[i];
    if (same) break;
  }
  lam[0] = C[0]; lam[1] = C[4]; lam[2] = C[8];
}
// rotation of the polar decomposition Fe = R U (reorient.cc:33-57)
__device__ void polar_R(const double *Fe, double *R) {
  double C[9], lam[3], V[9], VL[9], U[9], Ui[9];
  cp::m3_tmul(Fe, Fe, C);
  sym_eig3(C, lam, V);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) VL[3 * i + j] = V[3 * i + j] * sqrt(lam[j]);
  cp::m3_mult(VL, V, U);
  cp::m3_inv(U, Ui);
  cp::m3_mul(Fe, Ui, R);
}

// updateAfterIncrement's qp loop (updateAfterIncrement.cc:77-126): calculatePlasticity(.,.,0) at
// the converged displacement; singular local problems deferred as in assembly.
template <int NS, bool KIN>
__device__ __attribute__((noinline)) bool post_qp_one(const PostArgs &a, long long Q, SingAny<NS> &sing) {
  double F[9];
  post_F(a, Q, F);
  double Fp_t[9], s[NS], x[NS], P[9], T[9], Fe[9], Fp[9], W[KIN ? NS : 1];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fp_t[i] = a.Fp_conv[i * a.nq + Q];
  const double *mat;
  int ns;
  qp_phase(a.ph, a.mat, a.ns, a.egrain[Q >> 3], mat, ns);
  for (int i = 0; i < ns; ++i) s[i] = a.s_conv[i * a.nq + Q];
  if constexpr (KIN)
    for (int i = 0; i < ns; ++i) W[i] = a.W_conv[i * a.nq + Q];
  const double *gt = a.gtab + table_row(a.tw, a.egrain, Q) * a.gstride;
  TwinArgs tw = a.tw;
  if (a.ph.gphase && tw.on) phase_twin(a.ph, a.egrain[Q >> 3], tw);
  const int nslip = tw.on ? tw.nslip : ns;
  bool plastic = false;
  const int na = cp::qp_update<NS, double, KIN>(mat, gt, ns, a.max_solver_iter, F, Fp_t, s, x, 0, P, T, Fe, Fp, (double *)nullptr,
                                                0, &plastic, 1.0, SingRef<NS>{&sing}, W, nullptr, 1, nullptr, nslip, sing.ws);
  if (na < 0) return false;
#pragma unroll
  for (int i = 0; i < 9; ++i) { a.Fe_iter[i * a.nq + Q] = Fe[i]; a.Fp_iter[i * a.nq + Q] = Fp[i]; a.T[i * a.nq + Q] = T[i]; }
  for (int i = 0; i < ns; ++i) a.s_iter[i * a.nq + Q] = s[i];
  if constexpr (KIN)
    for (int i = 0; i < ns; ++i) a.W_iter[i * a.nq + Q] = W[i];
  if (plastic)
    for (int i = 0; i < nslip; ++i) a.slip_iter[i * a.nq + Q] = a.slip_conv[i * a.nq + Q] + x[i];
  if (tw.on && tw.ntw > 0) twin_epilogue(tw, Q, a.nq, plastic, x, a.s_iter);
  return true;
}
// lane = qp, grid-stride over one-wave blocks (as the first assembly pass)
template <int NS, bool KIN>
__global__ void __launch_bounds__(64) QP_WAVES k_post_qp(PostArgs a) {
  for (long long Q = blockIdx.x * (long long)blockDim.x + threadIdx.x; Q < a.nq; Q += (long long)gridDim.x * blockDim.x) {
    SingAny<NS> f{SING_CLAIM, a.sing, a.qcount, a.queue, (unsigned)Q, 0u, 0, 0, lane_ws<NS>(a.ws)};
    post_qp_one<NS, KIN>(a, Q, f);
  }
}
template <int NS, bool KIN>
__global__ void __launch_bounds__(64) QP_WAVES k_post_qp_replay(PostArgs a) {
  const unsigned n = min(*a.qcount, a.sing.cap);
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    if (a.sing.done[e]) continue;
    SingAny<NS> f{SING_REPLAY, a.sing, a.qcount, a.queue, 0u, e, a.sing.have[e], 0, lane_ws<NS>(a.ws)};
    if (post_qp_one<NS, KIN>(a, a.queue[e], f)) a.sing.done[e] = 1;
  }
}
template <int NS, bool KIN>
__global__ void __launch_bounds__(64) QP_WAVES k_post_qp_deferred(PostArgs a) {
  const unsigned n = *a.qcount;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (i >= a.sing.cap || !a.sing.done[i]) {
      SingAny<NS> f{SING_LANE, a.sing, a.qcount, a.queue, 0u, i, 0, 0, lane_ws<NS>(a.ws)};
      post_qp_one<NS, KIN>(a, a.queue[i], f);
    }
}

// Green-Lagrange strain, JxW-weighted sums (updateAfterIncrement.cc:107-126, 218-245) and
// reorient() (reorient.cc:18-96) with Fe_old = Fe_conv, Fe_new = Fe_iter; deterministic sums.
__global__ void __launch_bounds__(256) k_post_reduce(PostArgs a) {
  __shared__ double lds[15 * 16];
  const long long Q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double acc[15];
#pragma unroll
  for (int i = 0; i < 15; ++i) acc[i] = 0.0;
  if (Q < a.nq) {
    double F[9], C[9], T[9];
    post_F(a, Q, F);
    cp::m3_tmul(F, F, C);
#pragma unroll
    for (int i = 0; i < 9; ++i) T[i] = a.T[i * a.nq + Q];
    const int vi[6] = {0, 4, 8, 5, 2, 1};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      acc[c] = 0.5 * (C[vi[c]] - ((vi[c] % 4 == 0) ? 1.0 : 0.0)) * a.JxW;
      acc[6 + c] = T[vi[c]] * a.JxW;
    }
    // SlipTotal, TwinRealVF and TwinMade numerators (updateAfterIncrement.cc:176, 222-236); the
    // host discards them with several phases, as the reference does (:722-725)
    double slip_sum = 0;
    const int nslip = a.tw.on ? a.tw.nslip : a.ns;
    for (int i = 0; i < nslip; ++i) slip_sum += a.slip_iter[i * a.nq + Q] * a.JxW;
    acc[12] = slip_sum;
    if (a.tw.on) {
      double fr = 0;
      for (int k = 0; k < a.tw.ntw; ++k) fr += a.tw.tw_iter[k * a.nq + Q] * a.JxW;
      acc[13] = fr;
      acc[14] = (a.tw.var_iter[Q] != 0 ? 1.0 : 0.0) * a.JxW;
    }
    double Fo[9], Fn[9], Ro[9], Rn[9], Om[9], Om2[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) { Fo[i] = a.Fe_conv[i * a.nq + Q]; Fn[i] = a.Fe_iter[i * a.nq + Q]; }
    polar_R(Fo, Ro);
    polar_R(Fn, Rn);
#pragma unroll
    for (int i = 0; i < 9; ++i) Om[i] = Rn[i] - Ro[i];
    cp::m3_mult(Om, Rn, Om2);
    const double w[3] = {-0.5 * (Om2[5] - Om2[7]), 0.5 * (Om2[2] - Om2[6]), -0.5 * (Om2[1] - Om2[3])};
    double r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = a.rotnew[i * a.nq + Q];
    if (a.tw.on && a.tw.var_iter[Q] != a.tw.var_conv[Q]) {  // reoriented by twinning in this increment (:1009-1010, :287)
      const long long row = (long long)a.egrain[Q >> 3] * a.tw.nvar + a.tw.var_iter[Q];
#pragma unroll
      for (int i = 0; i < 3; ++i) r[i] = a.tw.vrod[3 * row + i];
    }
    const double dt = w[0] * r[0] + w[1] * r[1] + w[2] * r[2];
    const double cr[3] = {w[1] * r[2] - w[2] * r[1], w[2] * r[0] - w[0] * r[2], w[0] * r[1] - w[1] * r[0]};
    double rn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) rn[i] = r[i] + 0.5 * (w[i] + dt * r[i] + cr[i]);
    const double nrm = sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]);
    if (nrm > 10000.0)
      for (int i = 0; i < 3; ++i) rn[i] = rn[i] * 10000.0 / nrm;
#pragma unroll
    for (int i = 0; i < 3; ++i) a.rotnew[i * a.nq + Q] = rn[i];
  }
  if (grid_sum<15>(acc, a.partials, a.counter, lds)) {
#pragma unroll
    for (int i = 0; i < 15; ++i) a.out[i] = acc[i];
  }
}

// ---- cooperative qp kernels (bodies in cp_passes.h): one-wave blocks, one LDS slice per group of
// lanes.  The batch entry point (pa_qp_update) always runs in this form; the Newton-step passes run
// in it when the context's qp form is PA_QP_COOP.
template <int NS, bool KIN>
__global__ void __launch_bounds__(64) k_qp_batch(BatchArgs b) {
  __shared__ double lds[64 / coop::group_size<NS>() * coop::Slice<NS, KIN>::size];
  qp_batch_body<NS, KIN>(b, lds);
}
template <int NS, typename DT, bool KIN>
__global__ void __launch_bounds__(64) k_qp_tangent_coop(QpArgs a) {
  __shared__ double lds[64 / coop::group_size<NS>() * coop::Slice<NS, KIN>::size];
  qp_tangent_coop_pass<NS, DT, KIN>(a, lds);
}
template <int NS, bool KIN>
__global__ void __launch_bounds__(64) k_post_qp_coop(PostArgs a) {
  __shared__ double lds[64 / coop::group_size<NS>() * coop::Slice<NS, KIN>::size];
  post_qp_coop_pass<NS, KIN>(a, lds);
}

template <int NS>
void launch_qp_batch(hipStream_t st, bool kin, const BatchArgs &b) {
  const unsigned grid = (unsigned)((b.n + 64 / coop::group_size<NS>() - 1) / (64 / coop::group_size<NS>()));
  if (kin) hipLaunchKernelGGL((k_qp_batch<NS, true>), dim3(grid), dim3(64), 0, st, b);
  else hipLaunchKernelGGL((k_qp_batch<NS, false>), dim3(grid), dim3(64), 0, st, b);
}

// x = pinv(A) b for dense systems, one group of lanes per system: the cooperative pseudo-inverse
// (coop::gsvd_pinv) of the cooperative qp kernels, for pa_pinv_solve (A: [count][NS*NS], b and x:
// [count][NS]).
template <int NS>
__global__ void __launch_bounds__(64) k_pinv_group(unsigned count, int n, const double *A, const double *b, double *x) {
  constexpr int G = coop::group_size<NS>(), GPW = 64 / G;
  typedef coop::Slice<NS, false> SL;
  __shared__ double lds[GPW * SL::size];
  const coop::Group<G> g;
  const int gi = threadIdx.x / G;
  double *L = lds + gi * SL::size;
  const unsigned e = blockIdx.x * GPW + gi;
  if (e >= count) return;  // uniform within the group
  const unsigned long long mask = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  double *M0 = L + SL::M0, *bv = L + SL::b, *yv = L + SL::y;
  if (g.l < n) {
    for (int j = 0; j < n; ++j) M0[g.l * NS + j] = A[(size_t)e * NS * NS + g.l * NS + j];
    bv[g.l] = b[(size_t)e * NS + g.l];
  }
  coop::gsync();
  coop::gsvd_pinv<NS, G, false>(g, mask, M0, bv, L, yv);
  if (g.l < n) x[(size_t)e * NS + g.l] = yv[g.l];
}

// Per-grain sums of JxW and JxW * Cauchy stress (updateAfterIncrement.cc:611-668): one block per
// grain id g = blockIdx.x + 1, its 256 lanes visiting the elements in index order (stride 256), then
// a fixed-order tree reduction, so the sums are deterministic.  out[(g-1)*7 + c]: c = 0 sum of JxW,
// 1..6 sums of JxW * (Txx, Tyy, Tzz, Txy, Txz, Tyz).
__global__ void __launch_bounds__(256) k_grain_sums(long long nelem, long long nq, const unsigned *__restrict__ egrain,
                                                    const double *__restrict__ T, double jxw, double *__restrict__ out) {
  __shared__ double red[7][256];
  const unsigned g = blockIdx.x + 1;
  const int t = threadIdx.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int comp[6] = {0, 4, 8, 1, 2, 5};
  for (long long e = t; e < nelem; e += 256) {
    if (egrain[e] != g) continue;
    for (int q = 0; q < 8; ++q) {
      const long long Q = 8 * e + q;
      acc[0] += jxw;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[1 + c] += T[comp[c] * nq + Q] * jxw;
    }
  }
#pragma unroll
  for (int c = 0; c < 7; ++c) red[c][t] = acc[c];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int c = 0; c < 7; ++c) red[c][t] += red[c][t + w];
    __syncthreads();
  }
  if (t < 7) out[(long long)blockIdx.x * 7 + t] = red[t][0];
}

// warm-start inputs of a batch of qps (FCC: 12 ratios + 6 strain components per qp, row-major)
template <int NS>
__global__ void __launch_bounds__(64) k_qp_features(long long n, const double *__restrict__ F, const double *__restrict__ Fpc,
                                                    const double *__restrict__ Fec, const double *__restrict__ sc,
                                                    const double *__restrict__ Wc, const unsigned *__restrict__ grain,
                                                    const double *gtab, int gstride, int ns, float *feat) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[9], fp[9], fe[9], s[NS], w[NS];
  for (int k = 0; k < 9; ++k) { f[k] = F[9 * i + k]; fp[k] = Fpc[9 * i + k]; fe[k] = Fec[9 * i + k]; }
  for (int k = 0; k < ns; ++k) { s[k] = sc[(long long)ns * i + k]; w[k] = Wc ? Wc[(long long)ns * i + k] : 0.0; }
  cp::qp_features<NS>(gtab + (long long)grain[i] * gstride, mlp::NOUT, f, fp, fe, s, w, feat + (long long)mlp::NIN * i, 1);
}

// batch of calculatePlasticity calls whose local loop starts from x0 (12 per qp, may be null = the
// reference's start from zero); iters receives the local solves of each qp
template <int NS>
__global__ void __launch_bounds__(64) k_qp_batch_warm(long long n, const double *__restrict__ F, const double *__restrict__ Fpc,
                                                      const double *__restrict__ sc, const unsigned *__restrict__ grain,
                                                      const double *gtab, int gstride, const double *mat, int ns, int msi,
                                                      int flag, double *P, double *T, double *D, double *Fe, double *Fp,
                                                      double *so, double *xo, int *na, const float *x0, int *iters) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[9], fp[9], s[NS], x[NS], p[9], t[9], fe[9], fpo[9];
  for (int k = 0; k < 9; ++k) { f[k] = F[9 * i + k]; fp[k] = Fpc[9 * i + k]; }
  for (int k = 0; k < ns; ++k) s[k] = sc[(long long)ns * i + k];
  const double *gt = gtab + (long long)grain[i] * gstride;
  int it = 0;
  const int r = cp::qp_update<NS, double, false, cp::SvdInLane<NS>, true>(
      mat, gt, ns, msi, f, fp, s, x, flag, p, t, fe, fpo, flag ? D + 81 * i : nullptr, 1, nullptr, 1.0, cp::SvdInLane<NS>(),
      nullptr, x0 ? x0 + (long long)mlp::NOUT * i : nullptr, 1, &it);
  for (int k = 0; k < 9; ++k) { P[9 * i + k] = p[k]; T[9 * i + k] = t[k]; Fe[9 * i + k] = fe[k]; Fp[9 * i + k] = fpo[k]; }
  for (int k = 0; k < ns; ++k) { so[(long long)ns * i + k] = s[k]; xo[(long long)ns * i + k] = x[k]; }
  na[i] = r;
  iters[i] = it;
}

// ------------------------------------------------------------ Krylov (FP64)
// Solver state on the device.  Reductions never use atomics or fences: a kernel writes one
// partial sum per block, and the next kernel on the stream (the kernel boundary orders the
// writes) adds them in every block in the same fixed order, so every block holds the same
// bitwise total.  rzh[i & 1] holds r.z of iteration i (PETSc's beta), written by block 0 of the
// kernel that first forms it and read one kernel later.
// breakdown: 0 none, 1 r.z = 0 before convergence, KRYLOV_NONFINITE a norm or r.z that is NaN or
// infinite (the solve stops at once, as PETSc's KSP_DIVERGED_NANORINF)
constexpr int KRYLOV_NONFINITE = 2;
struct KState {
  double beta, dp, tol, bnorm;
  double rzh[2];
  int it, done, maxit, breakdown;
};

constexpr int VB = 256;   // threads per block of the Krylov kernels
constexpr int VPT = 4;    // vector entries per thread of the update kernels

// Block-uniform test of the "solve ended" flag.  Block 0 of a kernel that detects convergence
// sets it while other blocks of the same kernel may be starting, so each block takes one
// decision for all its threads.
__device__ __forceinline__ bool solve_ended(const KState *st) { return __syncthreads_or(st->done != 0); }

__device__ __forceinline__ long long nb_node(const MeshDims &m, long long n, int i, int j, int k, int s, bool &ok) {
  const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
  ok = (i + dx >= 0) && (i + dx < m.nnx) && (j + dy >= 0) && (j + dy < m.nny) && (k + dz >= 0) && (k + dz < m.nnz);
  return n + dx + (long long)dy * m.nnx + (long long)dz * m.nnx * m.nny;
}

// Neighbour of node n = (i, j, k) in slot s (dx, dy, dz = s % 3 - 1, s / 3 % 3 - 1, s / 9 - 1),
// without branches: a neighbour outside the mesh is reported as node n with ok = false, and its
// vector entries are masked to zero by the caller.
__device__ __forceinline__ long long nb_masked(const MeshDims &m, long long n, int i, int j, int k, int s, bool &ok) {
  const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
  // one unsigned compare per axis covers both bounds; bitwise & keeps it free of branches
  ok = ((unsigned)(i + dx) < (unsigned)m.nnx) & ((unsigned)(j + dy) < (unsigned)m.nny) & ((unsigned)(k + dz) < (unsigned)m.nnz);
  return ok ? n + dx + (long long)dy * m.nnx + (long long)dz * m.nnx * m.nny : n;
}
// (i, j, k) of node n in 32-bit arithmetic (node counts stay below 2^31)
__device__ __forceinline__ void node_ijk32(const MeshDims &m, long long n, int &i, int &j, int &k) {
  const unsigned u = (unsigned)n, nx = (unsigned)m.nnx, ny = (unsigned)m.nny;
  const unsigned r = u / nx;
  i = (int)(u - r * nx);
  k = (int)(r / ny);
  j = (int)(r - (unsigned)k * ny);
}

// Block sums written to partials[blockIdx.x * NV + k] (thread 0).
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partials, double *lds) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[blockIdx.x * NV + k] = v[k];
  }
}
// Totals of nparts blocks' NV partials, identical in every block, returned to every thread.
// Four independent loads per thread per round keep the latency of the (mostly remote-L2)
// reads to about one round trip.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *partials, unsigned nparts, double (&v)[NV], double *lds) {
  __shared__ double tot[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (unsigned b0 = threadIdx.x; b0 < nparts; b0 += 4 * VB) {
    double t[4][NV];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned b = b0 + j * VB;
#pragma unroll
      for (int k = 0; k < NV; ++k) t[j][k] = (b < nparts) ? partials[(size_t)b * NV + k] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < NV; ++k) v[k] += t[j][k];
  }
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = tot[k];
}

// Convergence test and direction coefficient at the start of Krylov iteration it >= 1, from the
// update kernel's partials (z.z, r.z) of iteration it-1: PETSc's test on the preconditioned norm
// ||z|| <= tol (KSPConvergedDefault), and b = (r.z)_it / (r.z)_{it-1}.  Returns false when the
// solve has ended (block 0 then records it, the norm and the breakdown flag).
__device__ __forceinline__ bool krylov_next(KState *st, const double *upart, unsigned n_upart, int it, double *lds, double &b) {
  double v[2];
  sum_partials<2>(upart, n_upart, v, lds);
  const double dp = sqrt(v[0]);
  const bool breakdown = (v[1] == 0.0);
  const bool nonfinite = !isfinite(dp) || !isfinite(v[1]);
  const bool done = (dp <= st->tol) || it >= st->maxit || breakdown || nonfinite;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->dp = dp;
    if (done) {
      st->done = 1;
      st->it = it;
      st->breakdown = nonfinite ? KRYLOV_NONFINITE : ((breakdown && !(dp <= st->tol)) ? 1 : 0);
    } else {
      st->rzh[it & 1] = v[1];
    }
  }
  if (done) return false;
  b = v[1] / st->rzh[(it - 1) & 1];
  return true;
}

// 3x3 block product of one slot: y += B x (B row-major, 9 contiguous values)
template <typename VT>
__device__ __forceinline__ void block_mv(const VT *__restrict__ B, double x0, double x1, double x2, double y[3]) {
  VT b[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) b[c] = B[c];
#pragma unroll
  for (int r = 0; r < 3; ++r) y[r] += (double)b[3 * r] * x0 + (double)b[3 * r + 1] * x1 + (double)b[3 * r + 2] * x2;
}
// y += B^T x
template <typename VT>
__device__ __forceinline__ void block_tmv(const VT *__restrict__ B, double x0, double x1, double x2, double y[3]) {
  VT b[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) b[c] = B[c];
#pragma unroll
  for (int r = 0; r < 3; ++r) y[r] += (double)b[r] * x0 + (double)b[3 + r] * x1 + (double)b[6 + r] * x2;
}

// ---- BiCG (PETSc KSPBICG with PCJACOBI, solveLinearSystem.cc:13-16): three kernels per iteration

// Iteration it: convergence test / b from the previous update's partials, then
// p <- z (it = 0) or p <- z + b p, for both sequences.
__global__ void __launch_bounds__(VB) k_bicg_p(long long N, const double *__restrict__ Zr, const double *__restrict__ Zl,
                                               double *__restrict__ Pr, double *__restrict__ Pl, KState *st,
                                               const double *upart, unsigned n_upart, int it) {
  __shared__ double lds[32];
  if (solve_ended(st)) return;
  double b = 0.0;
  if (it > 0 && !krylov_next(st, upart, n_upart, it, lds, b)) return;
  const long long stride = (long long)gridDim.x * VB;
  for (long long r0 = blockIdx.x * (long long)VB + threadIdx.x; r0 < N; r0 += VPT * stride) {
    double zr[VPT], zl[VPT], pr[VPT], pl[VPT];
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const long long r = r0 + j * stride;
      const bool ok = r < N, old = ok && it > 0;
      zr[j] = ok ? Zr[r] : 0.0; zl[j] = ok ? Zl[r] : 0.0;
      pr[j] = old ? Pr[r] : 0.0; pl[j] = old ? Pl[r] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const long long r = r0 + j * stride;
      if (r < N) {
        Pr[r] = (it == 0) ? zr[j] : zr[j] + b * pr[j];
        Pl[r] = (it == 0) ? zl[j] : zl[j] + b * pl[j];
      }
    }
  }
}

// Zr = K Pr, Zl = K^T Pl on the node-block matrix (lane = node); block partials of Zr.Pl
template <typename VT>
__global__ void __launch_bounds__(VB) k_bicg_spmv(MeshDims m, const VT *__restrict__ K, const double *__restrict__ Pr,
                                                  const double *__restrict__ Pl, double *__restrict__ Zr, double *__restrict__ Zl,
                                                  const KState *st, double *partials) {
  __shared__ double lds[16];
  if (solve_ended(st)) return;
  const long long n = blockIdx.x * (long long)VB + threadIdx.x;
  double acc[1] = {0.0};
  if (n < m.nnode) {
    int i, j, k;
    node_ijk32(m, n, i, j, k);
    double y[3] = {0, 0, 0}, yt[3] = {0, 0, 0};
    const long long nn = m.nnode;
    // a masked neighbour reads the K^T block of node n itself, which multiplies zero
#pragma unroll 3
    for (int s = 0; s < 27; ++s) {
      bool ok;
      const long long mn = nb_masked(m, n, i, j, k, s, ok);
      const double w = ok ? 1.0 : 0.0;
      double pr[3], pl[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) { pr[c] = w * Pr[3 * mn + c]; pl[c] = w * Pl[3 * mn + c]; }
      block_mv(K + ((long long)s * nn + n) * 9, pr[0], pr[1], pr[2], y);
      block_tmv(K + ((long long)(26 - s) * nn + mn) * 9, pl[0], pl[1], pl[2], yt);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) { Zr[3 * n + r] = y[r]; Zl[3 * n + r] = yt[r]; }
    acc[0] = y[0] * Pl[3 * n] + y[1] * Pl[3 * n + 1] + y[2] * Pl[3 * n + 2];
  }
  block_partials<1>(acc, partials, lds);
}

// a = (r.z)_it / (Zr.Pl); x += a p; r -= a z; rl -= a zl; z = D^-1 r; zl = D^-1 rl;
// block partials of z.z and z.rl
__global__ void __launch_bounds__(VB) k_bicg_update(long long N, double *__restrict__ X, const double *__restrict__ Pr,
                                                    double *__restrict__ Rr, double *__restrict__ Rl, double *__restrict__ Zr,
                                                    double *__restrict__ Zl, const double *__restrict__ dinv, const KState *st,
                                                    const double *opart, unsigned n_opart, double *upart, int it) {
  __shared__ double lds[32];
  if (solve_ended(st)) return;
  double dpi[1];
  sum_partials<1>(opart, n_opart, dpi, lds);
  const double a = st->rzh[it & 1] / dpi[0];
  const long long stride = (long long)gridDim.x * VB;
  double acc[2] = {0.0, 0.0};
  for (long long r0 = blockIdx.x * (long long)VB + threadIdx.x; r0 < N; r0 += 2 * stride) {
    // two entries per thread, all loads issued before use
    double d[2], x[2], p[2], rr[2], rl[2], zr[2], zl[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const long long r = r0 + j * stride;
      const bool ok = r < N;
      d[j] = ok ? dinv[r] : 0.0; x[j] = ok ? X[r] : 0.0; p[j] = ok ? Pr[r] : 0.0;
      rr[j] = ok ? Rr[r] : 0.0; rl[j] = ok ? Rl[r] : 0.0; zr[j] = ok ? Zr[r] : 0.0; zl[j] = ok ? Zl[r] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const long long r = r0 + j * stride;
      if (r < N) {
        X[r] = x[j] + a * p[j];
        const double rrn = rr[j] - a * zr[j], rln = rl[j] - a * zl[j];
        Rr[r] = rrn; Rl[r] = rln;
        const double zrn = rrn * d[j], zln = rln * d[j];
        Zr[r] = zrn; Zl[r] = zln;
        acc[0] += zrn * zrn;
        acc[1] += zrn * rln;
      }
    }
  }
  __syncthreads();
  block_partials<2>(acc, upart, lds);
}

// r = rl = b; z = zl = D^-1 b; x = 0; tol = lin_tol * ||b||, ||z||, r.z (once per solve; the
// last block to finish reduces the partials)
__global__ void __launch_bounds__(VB) k_krylov_init(long long N, const double *__restrict__ B, double *__restrict__ X,
                                                    double *__restrict__ Rr, double *__restrict__ Rl, double *__restrict__ Zr,
                                                    double *__restrict__ Zl, const double *__restrict__ dinv, KState *st,
                                                    double lin_tol, int maxit, double *partials, unsigned *counter) {
  __shared__ double lds[48];
  const long long stride = (long long)gridDim.x * VB;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long r = blockIdx.x * (long long)VB + threadIdx.x; r < N; r += stride) {
    const double b = B[r], z = b * dinv[r];
    X[r] = 0.0; Rr[r] = b; Zr[r] = z;
    if (Rl) { Rl[r] = b; Zl[r] = z; }
    acc[0] += b * b; acc[1] += z * z; acc[2] += z * b;
  }
  if (grid_sum<3>(acc, partials, counter, lds)) {
    st->bnorm = sqrt(acc[0]);
    st->tol = lin_tol * st->bnorm;
    st->dp = sqrt(acc[1]);
    st->beta = acc[2];
    st->rzh[0] = acc[2];
    st->rzh[1] = 1.0;
    st->it = 0;
    st->maxit = maxit;
    st->breakdown = 0;
    st->done = (st->dp <= st->tol) ? 1 : 0;
    if (!st->done && st->beta == 0.0) { st->done = 1; st->breakdown = 1; }
    if (!isfinite(acc[0]) || !isfinite(acc[1]) || !isfinite(acc[2])) { st->done = 1; st->breakdown = KRYLOV_NONFINITE; }
  }
}

// ---- Jacobi-PCG

// Iteration it: convergence test / b from the previous update's partials; p <- z (it = 0) or
// p <- z + b p (grid-stride over the dofs)
__global__ void __launch_bounds__(VB) k_cg_p(long long N, const double *__restrict__ Z, double *__restrict__ P, KState *st,
                                             const double *upart, unsigned n_upart, int it) {
  __shared__ double lds[32];
  if (solve_ended(st)) return;
  double b = 0.0;
  if (it > 0 && !krylov_next(st, upart, n_upart, it, lds, b)) return;
  // VPT entries per thread strided by the grid, all loads issued before the stores
  const long long stride = (long long)gridDim.x * VB;
  for (long long r0 = blockIdx.x * (long long)VB + threadIdx.x; r0 < N; r0 += VPT * stride) {
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

// q = K p on the node-block matrix; block partials of p.q
template <typename VT>
__global__ void __launch_bounds__(VB) k_cg_spmv(MeshDims m, const VT *__restrict__ K, const double *__restrict__ P,
                                                double *__restrict__ Q, const KState *st, double *opart) {
  __shared__ double lds[16];
  if (solve_ended(st)) return;
  const long long n = blockIdx.x * (long long)VB + threadIdx.x;
  double acc[1] = {0.0};
  if (n < m.nnode) {
    int i, j, k;
    node_ijk32(m, n, i, j, k);
    double y[3] = {0, 0, 0};
    const long long nn = m.nnode;
    // nine slots in flight per lane for FP32 blocks (three for FP64, whose registers run out first)
    constexpr int UNR = sizeof(VT) == 4 ? 9 : 3;
#pragma unroll UNR
    for (int s = 0; s < 27; ++s) {
      bool ok;
      const long long mn = nb_masked(m, n, i, j, k, s, ok);
      // unconditional loads (mn is always a valid node), then the mask
      const double t0 = P[3 * mn], t1 = P[3 * mn + 1], t2 = P[3 * mn + 2];
      const double w = ok ? 1.0 : 0.0;
      block_mv(K + ((long long)s * nn + n) * 9, w * t0, w * t1, w * t2, y);
    }
    const double pc[3] = {P[3 * n], P[3 * n + 1], P[3 * n + 2]};
#pragma unroll
    for (int r = 0; r < 3; ++r) Q[3 * n + r] = y[r];
    acc[0] = y[0] * pc[0] + y[1] * pc[1] + y[2] * pc[2];
  }
  block_partials<1>(acc, opart, lds);
}

// r = B - K X for the last solve (B: the condensed residual, X: its solution), with block partials
// of r.r, (D^-1 r).(D^-1 r) and B.B (for pa_last_residual)
template <typename VT>
__global__ void __launch_bounds__(VB) k_true_residual(MeshDims m, const VT *__restrict__ K, const double *__restrict__ X,
                                                      const double *__restrict__ B, const double *__restrict__ dinv,
                                                      double *part) {
  __shared__ double lds[3 * 16];
  const long long n = blockIdx.x * (long long)VB + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (n < m.nnode) {
    int i, j, k;
    node_ijk32(m, n, i, j, k);
    double y[3] = {0, 0, 0};
    const long long nn = m.nnode;
    for (int s = 0; s < 27; ++s) {
      bool ok;
      const long long mn = nb_masked(m, n, i, j, k, s, ok);
      const double w = ok ? 1.0 : 0.0;
      block_mv(K + ((long long)s * nn + n) * 9, w * X[3 * mn], w * X[3 * mn + 1], w * X[3 * mn + 2], y);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double b = B[3 * n + c], r = b - y[c], z = dinv[3 * n + c] * r;
      acc[0] += r * r;
      acc[1] += z * z;
      acc[2] += b * b;
    }
  }
  block_partials<3>(acc, part, lds);
}

// a = (r.z)_it / (p.q); x += a p; r -= a q; z = D^-1 r; block partials of z.z and r.z.
// Each thread handles VPT entries, strided by the grid, with all loads issued before use.
__global__ void __launch_bounds__(VB) k_cg_update(long long N, double *__restrict__ X, const double *__restrict__ P,
                                                  double *__restrict__ R, const double *__restrict__ Q, double *__restrict__ Z,
                                                  const double *__restrict__ dinv, const KState *st, const double *opart,
                                                  unsigned n_opart, double *upart, int it) {
  __shared__ double lds[32];
  if (solve_ended(st)) return;
  double pq[1];
  sum_partials<1>(opart, n_opart, pq, lds);
  const double a = st->rzh[it & 1] / pq[0];
  const long long stride = (long long)gridDim.x * VB;
  double acc[2] = {0.0, 0.0};
  for (long long r0 = blockIdx.x * (long long)VB + threadIdx.x; r0 < N; r0 += VPT * stride) {
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
        const double z = rn * d[j];
        X[r] = x[j] + a * p[j];
        R[r] = rn;
        Z[r] = z;
        acc[0] += z * z;
        acc[1] += rn * z;
      }
    }
  }
  __syncthreads();
  block_partials<2>(acc, upart, lds);
}

template <typename KT>
__global__ void k_dinv(MeshDims m, const KT *__restrict__ K, double *__restrict__ dinv) {
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= 3 * m.nnode) return;
  const long long n = r / 3;
  const int c = (int)(r % 3);
  const double d = (double)K[(13 * m.nnode + n) * 9 + 4 * c];
  dinv[r] = (d == 0.0) ? 1.0 : 1.0 / d;
}

// constraints.distribute(du) and x += du (solveLinearSystem.cc:32-34)
__global__ void k_apply_du(long long N, double *__restrict__ u, const double *__restrict__ du, const unsigned char *__restrict__ con,
                           const double *__restrict__ delu, int iter0) {
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= N) return;
  u[r] += con[r] ? (iter0 ? delu[r] : 0.0) : du[r];
}

__global__ void k_norm2(long long N, const double *__restrict__ v, double *out, double *partials, unsigned *counter) {
  __shared__ double lds[16];
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc[1] = {0.0};
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < N; r += stride) acc[0] += v[r] * v[r];
  if (grid_sum<1>(acc, partials, counter, lds)) *out = sqrt(acc[0]);
}

__global__ void k_zero(long long N, double *__restrict__ v) {
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r < N) v[r] = 0.0;
}

// ---- post-processing projection (projectionFunctions.cc:143-223, updateAfterIncrement.cc:147-192):
// per qp von Mises stress sqrt(3/2)|dev T|, equivalent strain sqrt(2/3)|dev E| with E = (F^T F - I)/2,
// and the twinned flag; out[f * nq + Q]
__global__ void k_proj_values(MeshDims m, double hx, double hy, double hz, long long nq, const double *__restrict__ u,
                              const double *__restrict__ T, const int *__restrict__ var, double *__restrict__ out) {
  const long long Q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (Q >= nq) return;
  double F[9], C[9], t[9];
  const double hh[3] = {hx, hy, hz};
  qp_F(m, hh, u, Q, F);
  cp::m3_tmul(F, F, C);
  for (int i = 0; i < 9; ++i) t[i] = T[i * nq + Q];
  double E[9];
  for (int i = 0; i < 9; ++i) E[i] = 0.5 * (C[i] - ((i % 4 == 0) ? 1.0 : 0.0));
  const double trE = E[0] + E[4] + E[8], trT = t[0] + t[4] + t[8];
  double se = 0.0, st = 0.0;
  for (int i = 0; i < 9; ++i) {
    const double de = E[i] - ((i % 4 == 0) ? trE / 3 : 0.0), dt = t[i] - ((i % 4 == 0) ? trT / 3 : 0.0);
    se += de * de;
    st += dt * dt;
  }
  out[Q] = sqrt(3.0 / 2.0) * sqrt(st);
  out[nq + Q] = sqrt(2.0 / 3.0) * sqrt(se);
  out[2 * nq + Q] = (var && var[Q] != 0) ? 1.0 : 0.0;
}

// right-hand side of the L2 projection of field v: b_n = sum over the elements and Gauss points
// around node n of N_n(q) v(q) JxW, in component 0 of node n (components 1, 2 zero)
__global__ void k_proj_rhs(MeshDims m, double JxW, const double *__restrict__ v, double *__restrict__ b) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= m.nnode) return;
  const int i = (int)(n % m.nnx), j = (int)((n / m.nnx) % m.nny), k = (int)(n / ((long long)m.nnx * m.nny));
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  double acc = 0.0;
  for (int c = 0; c < 8; ++c) {  // element c: the node is its vertex 7 - c
    const int ex = i - 1 + (c & 1), ey = j - 1 + ((c >> 1) & 1), ez = k - 1 + ((c >> 2) & 1);
    if (ex < 0 || ey < 0 || ez < 0 || ex >= m.nx || ey >= m.ny || ez >= m.nz) continue;
    const int vtx = 7 - c;
    const long long e = ex + (long long)m.nx * (ey + (long long)m.ny * ez);
    for (int q = 0; q < 8; ++q) {
      double N = 1.0;
      for (int d = 0; d < 3; ++d) {
        const double xi = ((q >> d) & 1) ? g1 : g0;
        N *= ((vtx >> d) & 1) ? xi : 1.0 - xi;
      }
      acc += N * v[8 * e + q] * JxW;
    }
  }
  b[3 * n] = acc;
  b[3 * n + 1] = 0.0;
  b[3 * n + 2] = 0.0;
}

// the consistent Q1 mass matrix (projectionFunctions.cc:99-135) as node blocks M_nm I: slot s of node
// n sums N_n N_m JxW over the Gauss points of the elements holding both nodes
__global__ void k_mass_blocks(MeshDims m, double JxW, double *__restrict__ K) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= m.nnode) return;
  const int i = (int)(n % m.nnx), j = (int)((n / m.nnx) % m.nny), k = (int)(n / ((long long)m.nnx * m.nny));
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  for (int s = 0; s < 27; ++s) {
    const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
    double mv = 0.0;
    for (int c = 0; c < 8; ++c) {
      const int ex = i - 1 + (c & 1), ey = j - 1 + ((c >> 1) & 1), ez = k - 1 + ((c >> 2) & 1);
      if (ex < 0 || ey < 0 || ez < 0 || ex >= m.nx || ey >= m.ny || ez >= m.nz) continue;
      const int va = 7 - c;
      // the neighbour's vertex in this element
      const int bx = (va & 1) + dx, by = ((va >> 1) & 1) + dy, bz = ((va >> 2) & 1) + dz;
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
      const int vb = bx + 2 * by + 4 * bz;
      for (int q = 0; q < 8; ++q) {
        double Na = 1.0, Nb = 1.0;
        for (int d = 0; d < 3; ++d) {
          const double xi = ((q >> d) & 1) ? g1 : g0;
          Na *= ((va >> d) & 1) ? xi : 1.0 - xi;
          Nb *= ((vb >> d) & 1) ? xi : 1.0 - xi;
        }
        mv += Na * Nb * JxW;
      }
    }
    double *B = K + ((long long)s * m.nnode + n) * 9;
    for (int r = 0; r < 9; ++r) B[r] = (r % 4 == 0) ? mv : 0.0;
  }
}

#include "cp_mg.h"
#include "cp_pbc.h"

// ------------------------------------------------------------ host side
inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }
// grid of the reducing vector kernels: grid-stride loops over at most this many 256-lane blocks,
// which bounds the partial sums the last block adds up
constexpr unsigned VGRID_MAX = 256;
inline unsigned vgrid(long long n) { return std::min(nblk(n, 256), VGRID_MAX); }

}  // namespace

struct pa_context {
  pa_material mat;
  pa_problem pb;
  MeshDims m;
  double h[3], JxW, microvol = 0.0;
  int ns = 0;
  long long nq = 0, ndof = 0, nnz = 0;
  int method = PA_SOLVER_BICG, precision = PA_PREC_FP64;
  int qp_form = PA_QP_LANE;  // layout of the Newton-step qp passes (pa_set_qp_form)
  int asm_precision = -1;    // arithmetic form of the last assembly (-1: none yet)
  bool solved = false;       // a solve has run on the current assembly (pa_last_residual)
  double *resid_part = nullptr;  // block partials of k_true_residual (allocated on first use)
  KState last_krylov = {};   // solver state at the end of the last Krylov solve
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t kev[2] = {nullptr, nullptr};  // Krylov flag copies, one per chunk parity
  hipEvent_t qev[2] = {nullptr, nullptr};  // around the first qp pass of the latest assembly
  bool qev_set = false;
  KState *hks = nullptr;                   // pinned host copies of the solver state (2)
  double last_ms[3] = {0, 0, 0};
  bool grains_set = false;
  int n_grain_rows = 0;
  // device memory
  double *d_mat = nullptr, *d_gtab = nullptr;
  int gstride = 0;
  unsigned *d_egrain = nullptr;
  double *Fp_conv = nullptr, *Fp_iter = nullptr, *Fe_conv = nullptr, *Fe_iter = nullptr, *s_conv = nullptr, *s_iter = nullptr;
  double *slip_conv = nullptr, *slip_iter = nullptr, *rotnew = nullptr, *T = nullptr;
  bool kin = false;                                // kinematic hardening: backstress state below
  double *W_conv = nullptr, *W_iter = nullptr;     // [ns][nq], allocated only with kinematic hardening
  // twinning: ns counts slip + twin systems; nslip the slip systems, ntw the twin systems
  int nslip = 0, ntw = 0;
  double *tw_conv = nullptr, *tw_iter = nullptr;   // [ntw][nq] twin fractions
  int *var_conv = nullptr, *var_iter = nullptr;    // [nq] orientation variant (0: the grain's)
  double *d_vrod = nullptr;                        // [grain][1 + ntw][3] variant orientations
  std::vector<double> h_vrod;                      // the same on the host
  double F_r = 0.0, F_e = 0.0;                     // TwinRealVF / TwinMade of the last increment
  double *u = nullptr, *R = nullptr, *dinv = nullptr;
  void *K = nullptr;   // node-block Jacobian, double (FP64 form) or float (mixed form)
  void *Dq = nullptr;  // [81][nq] JxW dP/dF of the last assembly, double or float
  double *Pq = nullptr;  // [9][nq] JxW P
  unsigned *queue = nullptr;  // deferred qps of the pseudo-inverse pass
  SingStore sing = {};        // their singular systems and solutions
  double *ws = nullptr;       // 48-system variant: per-lane matrix workspace (WS_BLOCKS x 64 lanes)
  unsigned char *con = nullptr;
  double *delu = nullptr;
  double *kv[7] = {nullptr};  // Krylov vectors: X Rr Rl Zr Zl Pr Pl (CG: X R - Z Q P -)
  KState *kst = nullptr;
  double *partials = nullptr;
  unsigned *counters = nullptr;
  double *d_out = nullptr;
  long long *d_rp = nullptr;
  int bc_inc = -1;
  // periodic boundary conditions (pa_set_periodic): the closed constraint lines on the device and
  // the load scale of each increment's first Newton iteration
  // several phases (pa_create_phases): material of each phase, its block stride on the device and
  // the phase of each grain row; multiphase: Enable Multiphase (any phase count)
  int nph = 1;
  bool multiphase = false;
  pa_material phmat[4];
  int mstride = 0;
  std::vector<int> h_gphase;
  int *d_gphase = nullptr;
  bool pbc = false;
  PbcDev pbcd = {};
  std::vector<double> pbc_scale;
  double pbc_s = 0.0;  // load scale of the current assembly
  // multigrid preconditioner (PA_SOLVER_MGCG): levels 0..mg_levels, mg_levels the coarsest;
  // level 0 is the Jacobian with the Krylov residual as right-hand side and z as solution
  int mg_levels = 0;
  MeshDims mg_m[MG_MAX_LEVELS];
  void *mg_A[MG_MAX_LEVELS] = {nullptr};
  double *mg_dinv[MG_MAX_LEVELS] = {nullptr}, *mg_b[MG_MAX_LEVELS] = {nullptr}, *mg_x[MG_MAX_LEVELS] = {nullptr};
  double *mg_r[MG_MAX_LEVELS] = {nullptr}, *mg_d[MG_MAX_LEVELS] = {nullptr};
  double *mg_lam = nullptr, *mg_minv = nullptr, *mg_part = nullptr;
  // the V-cycle as a HIP graph, one per arithmetic form (its kernels take only level pointers,
  // sizes and device-side scalars, all fixed once the levels are allocated)
  hipGraph_t mg_graph[2] = {nullptr, nullptr};
  hipGraphExec_t mg_exec[2] = {nullptr, nullptr};
  std::vector<void *> allocs;
};

namespace {

template <typename T>
int dalloc(pa_context *c, T **p, size_t count) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, count * sizeof(T) + 16));
  c->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

int ns_template(int ns) { return ns <= 12 ? 12 : (ns <= 24 ? 24 : 48); }

// One-wave blocks of the one-lane qp passes of the 48-system variant (WS_BLOCKS): each lane of them
// owns a cp::LaneWs<48> array (80,640 bytes, lane-interleaved), 2.6 GB in all.
unsigned lane_grid(const pa_context *c, unsigned g) { return c->ws ? std::min(g, WS_BLOCKS) : g; }

bool mixed(const pa_context *c) { return c->precision == PA_PREC_MIXED; }

constexpr unsigned DEFER_GRID = 2048;  // 64-lane blocks of the replay / deferred passes (two waves per SIMD)

constexpr unsigned SVD_GRID = 8192;    // blocks of k_sing_svd (grid-stride over the entries)

template <int NS>
void sing_svd_round(pa_context *c, const unsigned *count) {
  hipLaunchKernelGGL(k_sing_svd<NS>, dim3(SVD_GRID), dim3(sing_per_block<NS>() * sing_group<NS>()), 0, c->stream, c->sing, count);
}

template <int NS, typename DT, bool KIN>
void qp_tangent_passes_k(pa_context *c, const QpArgs &a, bool main_only) {
  // one-wave blocks, a grid of 2048 (eight per CU) striding over the qps
  hipEventRecord(c->qev[0], c->stream);
  hipLaunchKernelGGL((k_qp_tangent<NS, DT, KIN>), dim3(lane_grid(c, std::min(2048u, nblk(c->nq, 64)))), dim3(64), 0, c->stream, a);
  hipEventRecord(c->qev[1], c->stream);
  c->qev_set = true;
  if (main_only) return;
  for (int r = 0; r < SING_ROUNDS; ++r) {
    sing_svd_round<NS>(c, a.qcount);
    hipLaunchKernelGGL((k_qp_tangent_replay<NS, DT, KIN>), dim3(lane_grid(c, DEFER_GRID)), dim3(64), 0, c->stream, a);
  }
  hipLaunchKernelGGL((k_qp_tangent_deferred<NS, DT, KIN>), dim3(lane_grid(c, DEFER_GRID)), dim3(64), 0, c->stream, a);
}
// Blocks of a persistent cooperative qp kernel: as many one-wave blocks as the device holds at
// once (from the kernel's occupancy), at most one per chunk of qps.
template <typename K>
unsigned resident_grid(K kernel, long long nq) {
  int dev = 0, ncu = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 64, 0) != hipSuccess)
    nb = 1, ncu = 256;
  const long long full = (long long)std::max(nb, 1) * std::max(ncu, 1);
  return (unsigned)std::max(1LL, std::min(full, (nq + QP_CHUNK - 1) / QP_CHUNK));
}
// cooperative form: one persistent launch on a work queue (counters 8: pseudo-inverse solves,
// 11: the queue)
template <int NS, typename DT, bool KIN>
void qp_tangent_coop_k(pa_context *c, const QpArgs &a) {
  static const unsigned grid_full = resident_grid(k_qp_tangent_coop<NS, DT, KIN>, 1LL << 40);
  const unsigned grid = (unsigned)std::min<long long>(grid_full, (c->nq + QP_CHUNK - 1) / QP_CHUNK);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, c->stream, c->counters + 11);
  hipEventRecord(c->qev[0], c->stream);
  hipLaunchKernelGGL((k_qp_tangent_coop<NS, DT, KIN>), dim3(grid), dim3(64), 0, c->stream, a);
  hipEventRecord(c->qev[1], c->stream);
  c->qev_set = true;
}
template <int NS, typename DT>
void qp_tangent_passes(pa_context *c, const QpArgs &a, bool main_only) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, c->stream, a.qcount);
  if (c->qp_form == PA_QP_COOP) {
    if (c->kin) qp_tangent_coop_k<NS, DT, true>(c, a);
    else qp_tangent_coop_k<NS, DT, false>(c, a);
    return;
  }
  if (c->kin) qp_tangent_passes_k<NS, DT, true>(c, a, main_only);
  else qp_tangent_passes_k<NS, DT, false>(c, a, main_only);
}

// twinning arguments of the qp passes of the current increment: F_T from the last increment's
// TwinRealVF (F_r) and TwinMade (F_e), calculatePlasticity.cc:969-985
TwinArgs twin_args(const pa_context *c) {
  TwinArgs t;
  std::memset(&t, 0, sizeof(t));
  if (!c->ntw) return t;
  t.on = 1;
  t.nslip = c->nslip;
  t.ntw = c->ntw;
  t.shear = c->mat.twin_shear;
  double F_T = c->F_r > 0 ? c->mat.twin_threshold + (c->mat.twin_saturation * c->F_e / c->F_r) : c->mat.twin_threshold;
  if (F_T > 1.0) F_T = 1.0;
  t.F_T = F_T;
  t.armed = c->F_r > 0;
  t.tw_conv = c->tw_conv; t.tw_iter = c->tw_iter;
  t.var_conv = c->var_conv; t.var_iter = c->var_iter;
  t.vrod = c->d_vrod;
  t.nvar = 1 + c->ntw;
  return t;
}

PhaseArgs phase_args(const pa_context *c) {
  PhaseArgs p;
  std::memset(&p, 0, sizeof(p));
  if (c->nph < 2) return p;
  p.gphase = c->d_gphase;
  p.mstride = c->mstride;
  for (int k = 0; k < c->nph; ++k) {
    const pa_material &m = c->phmat[k];
    p.nslip[k] = m.n_slip;
    p.ntw[k] = m.twinning ? m.n_twin : 0;
    p.ns[k] = p.nslip[k] + p.ntw[k];
    p.shear[k] = m.twinning ? m.twin_shear : 1.0;
    p.F_T[k] = m.twinning ? std::min(1.0, m.twin_threshold) : 2.0;
  }
  return p;
}

// main_only: pass 1 alone (every qp on the LU path; the singular ones are only queued)
int launch_qp_tangent(pa_context *c, bool main_only = false) {
  QpArgs a;
  a.m = c->m;
  for (int d = 0; d < 3; ++d) a.h[d] = c->h[d];
  a.JxW = c->JxW;
  a.u = c->u; a.egrain = c->d_egrain; a.gtab = c->d_gtab; a.gstride = c->gstride; a.mat = c->d_mat;
  a.ns = c->ns; a.max_solver_iter = c->mat.max_solver_iter; a.nq = c->nq;
  a.Fp_conv = c->Fp_conv; a.s_conv = c->s_conv; a.slip_conv = c->slip_conv; a.W_conv = c->W_conv;
  a.Fe_iter = c->Fe_iter; a.Fp_iter = c->Fp_iter; a.s_iter = c->s_iter; a.slip_iter = c->slip_iter; a.W_iter = c->W_iter;
  a.D = c->Dq; a.P = c->Pq;
  a.queue = c->queue; a.qcount = c->counters + 8; a.sing = c->sing;
  a.nsing = c->counters + 8; a.work = c->counters + 11;
  a.tw = twin_args(c);
  a.ph = phase_args(c);
  a.ws = c->ws;
  const int NS = ns_template(c->ns);
  const bool f = mixed(c);
  if (NS == 12) { if (f) qp_tangent_passes<12, float>(c, a, main_only); else qp_tangent_passes<12, double>(c, a, main_only); }
  else if (NS == 24) { if (f) qp_tangent_passes<24, float>(c, a, main_only); else qp_tangent_passes<24, double>(c, a, main_only); }
  else { if (f) qp_tangent_passes<48, float>(c, a, main_only); else qp_tangent_passes<48, double>(c, a, main_only); }
  HIPCHK(hipGetLastError());
  return 0;
}

template <int NS, bool KIN>
void post_passes_k(pa_context *c, const PostArgs &a, unsigned grid) {
  (void)grid;
  hipLaunchKernelGGL((k_post_qp<NS, KIN>), dim3(lane_grid(c, std::min(2048u, nblk(c->nq, 64)))), dim3(64), 0, c->stream, a);
  for (int r = 0; r < SING_ROUNDS; ++r) {
    sing_svd_round<NS>(c, a.qcount);
    hipLaunchKernelGGL((k_post_qp_replay<NS, KIN>), dim3(lane_grid(c, DEFER_GRID)), dim3(64), 0, c->stream, a);
  }
  hipLaunchKernelGGL((k_post_qp_deferred<NS, KIN>), dim3(lane_grid(c, DEFER_GRID)), dim3(64), 0, c->stream, a);
}
template <int NS, bool KIN>
void post_coop_k(pa_context *c, const PostArgs &a) {
  static const unsigned grid_full = resident_grid(k_post_qp_coop<NS, KIN>, 1LL << 40);
  const unsigned grid = (unsigned)std::min<long long>(grid_full, (c->nq + QP_CHUNK - 1) / QP_CHUNK);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, c->stream, c->counters + 12);
  hipLaunchKernelGGL((k_post_qp_coop<NS, KIN>), dim3(grid), dim3(64), 0, c->stream, a);
}
template <int NS>
void post_passes(pa_context *c, const PostArgs &a, unsigned grid) {
  if (c->qp_form == PA_QP_COOP) {
    if (c->kin) post_coop_k<NS, true>(c, a);
    else post_coop_k<NS, false>(c, a);
    return;
  }
  if (c->kin) post_passes_k<NS, true>(c, a, grid);
  else post_passes_k<NS, false>(c, a, grid);
}

int launch_elements(pa_context *c, int iteration) {
  ElemArgs a;
  a.m = c->m;
  for (int d = 0; d < 3; ++d) a.h[d] = c->h[d];
  a.nq = c->nq;
  a.D = c->Dq; a.P = c->Pq;
  a.con = c->con; a.delu = c->delu; a.iter0 = (iteration == 0);
  a.K = c->K; a.R = c->R;
  for (int col = 0; col < 8; ++col) {
    a.cx = col & 1; a.cy = (col >> 1) & 1; a.cz = (col >> 2) & 1;
    a.ncx = (c->m.nx - a.cx + 1) / 2; a.ncy = (c->m.ny - a.cy + 1) / 2; a.ncz = (c->m.nz - a.cz + 1) / 2;
    const long long ncol = (long long)a.ncx * a.ncy * a.ncz;
    if (ncol == 0) continue;
    const unsigned grid = nblk(ncol, ASM_EL);
    if (mixed(c)) hipLaunchKernelGGL(k_element<float>, dim3(grid), dim3(64), 0, c->stream, a);
    else hipLaunchKernelGGL(k_element<double>, dim3(grid), dim3(64), 0, c->stream, a);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int launch_assemble(pa_context *c, int iteration) {
  int rc = launch_qp_tangent(c);
  if (rc) return rc;
  return launch_elements(c, iteration);
}

// Partial-sum regions (doubles): [0, off) one per block of the operator kernels; [off, off2) two
// per block of the update kernels; [off2, ...) the once-per-solve reductions (init, norm; at most
// 3 per block of VGRID_MAX blocks).  The post-increment pass uses 13 per 256-lane block from 0.
long long spmv_partials_offset(const pa_context *c) { return ((long long)nblk(c->m.nnode, VB) + 64) & ~63LL; }
unsigned ugrid(long long n) { return nblk(n, (long long)VB * VPT); }
long long update_partials_end(const pa_context *c) { return (spmv_partials_offset(c) + 2LL * ugrid(c->ndof) + 64) & ~63LL; }

size_t partials_size(const pa_context *c) {
  const size_t vec = (size_t)update_partials_end(c) + 3 * VGRID_MAX + 64;
  const size_t post = 15 * (size_t)nblk(c->nq, 256) + 64;
  return std::max(vec, post);
}

size_t kbytes(const pa_context *c) { return (mixed(c) ? sizeof(float) : sizeof(double)) * 243 * (size_t)c->m.nnode; }

// periodic BCs: y (and y2) folded onto the unconstrained dofs (cp_pbc.h); st: skip once the Krylov
// solve has ended (null outside a solve)
void pbc_fold(pa_context *c, double *y, double *y2, const KState *st) {
  const PbcDev &p = c->pbcd;
  if (p.nshort) hipLaunchKernelGGL(k_pbc_fold_short, dim3(nblk(p.nshort, VB)), dim3(VB), 0, c->stream, p, y, y2, st);
  if (p.nlong) hipLaunchKernelGGL(k_pbc_fold_long, dim3(PBC_GL, p.nlong), dim3(VB), 0, c->stream, p, (const double *)y, (const double *)y2, st);
  hipLaunchKernelGGL(k_pbc_fold_final, dim3(std::max(1u, nblk(std::max(p.ncon, p.nlong), VB))), dim3(VB), 0, c->stream, p, y, y2, st);
}
void pbc_expand(pa_context *c, double *x, double *x2, double scale, const KState *st) {
  if (c->pbcd.ncon) hipLaunchKernelGGL(k_pbc_expand, dim3(nblk(c->pbcd.ncon, VB)), dim3(VB), 0, c->stream, c->pbcd, x, x2, scale, st);
}
// periodic BCs after the unconstrained assembly: R <- C^T (R - K s g), Jacobi entries of C^T K C
template <typename KT>
int pbc_assembled(pa_context *c, int inc, int it) {
  const PbcDev &p = c->pbcd;
  const double sc = (it == 0) ? c->pbc_scale[std::min<size_t>((size_t)inc, c->pbc_scale.size() - 1)] : 0.0;
  c->pbc_s = sc;
  const KT *K = (const KT *)c->K;
  if (sc != 0.0) {
    double *ge = c->kv[5];
    hipLaunchKernelGGL(k_zero, dim3(nblk(c->ndof, 256)), dim3(256), 0, c->stream, c->ndof, ge);
    pbc_expand(c, ge, nullptr, sc, nullptr);
    hipLaunchKernelGGL(k_pbc_rhs<KT>, dim3(nblk(c->m.nnode, VB)), dim3(VB), 0, c->stream, c->m, K, (const double *)ge, c->R);
  }
  pbc_fold(c, c->R, nullptr, nullptr);
  if (p.nshort) hipLaunchKernelGGL(k_pbc_diag_short<KT>, dim3(nblk(p.nshort, VB)), dim3(VB), 0, c->stream, p, c->m, K, c->dinv);
  if (p.nlong) hipLaunchKernelGGL(k_pbc_diag_long<KT>, dim3(PBC_GL, p.nlong), dim3(VB), 0, c->stream, p, c->m, K);
  hipLaunchKernelGGL(k_pbc_diag_final, dim3(std::max(1u, nblk(std::max(p.ncon, p.nlong), VB))), dim3(VB), 0, c->stream, p, c->dinv);
  HIPCHK(hipGetLastError());
  return 0;
}

int set_dirichlet(pa_context *c, int inc) {
  if (c->pbc) return 0;  // boundaryConditions.cc:455: the face rules are not applied with periodic BCs
  if (c->bc_inc == inc) return 0;
  BCArgs bc;
  std::memset(&bc, 0, sizeof(bc));
  for (int f = 0; f < 6; ++f)
    for (int d = 0; d < 3; ++d) {
      bc.fdc[f][d] = c->pb.face_dof_constrained[f][d];
      bc.dU[f][d] = c->pb.face_total_u[f][d] / c->pb.total_increments;  // deluConstraint, init.cc:341
    }
  if (c->pb.tabular) {  // boundaryConditions.cc:153-267: slope of the current table segment times delT
    int tc = 1;          // timeCounter (:162-170), re-tested at every boundary node of every increment
    for (int k = 0; k <= inc; ++k) {
      const double t = c->pb.delT * (k + 1);
      while (tc < c->pb.n_table - 1 && t > c->pb.table_time[tc]) ++tc;
    }
    for (int f = 0; f < 6; ++f)
      for (int d = 0; d < 3; ++d)
        bc.dU[f][d] = (c->pb.table_u[f][d][tc] - c->pb.table_u[f][d][tc - 1]) / (c->pb.table_time[tc] - c->pb.table_time[tc - 1]) *
                      c->pb.delT;
  }
  bc.cyclic = c->pb.cyclic;
  bc.csign = 1.0;
  if (c->pb.cyclic) {
    bc.cface = c->pb.cyclic_face; bc.cdof = c->pb.cyclic_dof;
    const int cf = bc.cface - 1, cd = bc.cdof - 1;
    bc.dU[cf][cd] = bc.dU[cf][cd] * c->pb.total_increments * c->pb.delT / c->pb.quarter_cycle_time;  // init.cc:343-345
    const double qt = c->pb.quarter_cycle_time, t = std::fmod(inc * c->pb.delT, 4 * qt);
    bc.csign = (t < qt) ? 1.0 : (t < 3 * qt ? -1.0 : 1.0);
  }
  bc.velgrad = c->pb.velgrad;
  if (c->pb.velgrad) {  // updateBeforeIncrement.cc:10-32: F_(k+1) = (I - L delT)^-1 F_k from F_0 = I
    double A[3][3], Minv[3][3], F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) A[a][b] = (a == b ? 1.0 : 0.0) - c->pb.vel_grad[a][b] * c->pb.delT;
    if (!host_inv3(A, Minv)) return fail(PA_ERR_ARG, "I - L*delT is singular");
    for (int k = 0; k <= inc; ++k) {
      double Fn[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Fn[a][b] = Minv[a][0] * F[0][b] + Minv[a][1] * F[1][b] + Minv[a][2] * F[2][b];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) { bc.dF[a][b] = Fn[a][b] - F[a][b]; F[a][b] = Fn[a][b]; }
    }
    for (int d = 0; d < 3; ++d) bc.h[d] = c->h[d];
  }
  hipLaunchKernelGGL(k_dirichlet, dim3(nblk(c->m.nnode, 256)), dim3(256), 0, c->stream, c->m, bc, c->con, c->delu);
  HIPCHK(hipGetLastError());
  c->bc_inc = inc;
  return 0;
}

int do_assemble(pa_context *c, int inc, int it, double *res_norm) {
  c->asm_precision = -1;
  c->solved = false;
  int rc = set_dirichlet(c, inc);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(c->K, 0, kbytes(c), c->stream));
  HIPCHK(hipMemsetAsync(c->R, 0, sizeof(double) * (size_t)c->ndof, c->stream));
  rc = launch_assemble(c, it);
  if (rc) return rc;
  if (mixed(c)) hipLaunchKernelGGL(k_dinv<float>, dim3(nblk(c->ndof, 256)), dim3(256), 0, c->stream, c->m, (const float *)c->K, c->dinv);
  else hipLaunchKernelGGL(k_dinv<double>, dim3(nblk(c->ndof, 256)), dim3(256), 0, c->stream, c->m, (const double *)c->K, c->dinv);
  HIPCHK(hipGetLastError());
  if (c->pbc) {
    rc = mixed(c) ? pbc_assembled<float>(c, inc, it) : pbc_assembled<double>(c, inc, it);
    if (rc) return rc;
  }
  if (res_norm) {
    hipLaunchKernelGGL(k_norm2, dim3(vgrid(c->ndof)), dim3(256), 0, c->stream, c->ndof, c->R, c->d_out + 16,
                       c->partials + update_partials_end(c),
                       c->counters + 3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(res_norm, c->d_out + 16, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->asm_precision = c->precision;
  if (res_norm && !std::isfinite(*res_norm))
    return fail(PA_ERR_NONFINITE, "assembled residual norm is " + std::to_string(*res_norm));
  return 0;
}

// ---- multigrid preconditioner, host side

// Levels: halve every axis while the level has more than MG_COARSE_NODES nodes.
int mg_plan(pa_context *c) {
  MeshDims m = c->m;
  int L = 0;
  c->mg_m[0] = m;
  while (m.nnode > MG_COARSE_NODES) {
    if ((m.nx & 1) || (m.ny & 1) || (m.nz & 1) || L + 1 >= MG_MAX_LEVELS)
      return fail(PA_ERR_UNSUPPORTED, "multigrid needs element counts that halve down to at most 2x2x2 elements");
    m.nx /= 2; m.ny /= 2; m.nz /= 2;
    m.nnx = m.nx + 1; m.nny = m.ny + 1; m.nnz = m.nz + 1;
    m.nnode = (long long)m.nnx * m.nny * m.nnz;
    m.nelem = (long long)m.nx * m.ny * m.nz;
    c->mg_m[++L] = m;
  }
  if (L == 0) return fail(PA_ERR_UNSUPPORTED, "multigrid needs a mesh with more than 2x2x2 elements");
  c->mg_levels = L;
  return 0;
}

// Level storage (once per context): operators sized for FP64 so that either form fits.
int mg_alloc(pa_context *c) {
  if (c->mg_lam) return 0;
  int rc = mg_plan(c);
  if (rc) return rc;
  const int L = c->mg_levels;
  c->mg_A[0] = c->K;
  c->mg_dinv[0] = c->dinv;
  c->mg_b[0] = c->kv[1];  // the Krylov residual r
  c->mg_x[0] = c->kv[3];  // z = M r
  c->mg_r[0] = c->kv[2];  // vectors the CG does not use
  c->mg_d[0] = c->kv[6];
  for (int l = 1; l <= L; ++l) {
    const long long nn = c->mg_m[l].nnode;
    if ((rc = dalloc(c, (double **)&c->mg_A[l], 243 * nn)) || (rc = dalloc(c, &c->mg_dinv[l], 3 * nn)) ||
        (rc = dalloc(c, &c->mg_b[l], 3 * nn)) || (rc = dalloc(c, &c->mg_x[l], 3 * nn)) || (rc = dalloc(c, &c->mg_r[l], 3 * nn)) ||
        (rc = dalloc(c, &c->mg_d[l], 3 * nn)))
      return rc;
  }
  if ((rc = dalloc(c, &c->mg_minv, (size_t)MG_COARSE_DOFS * MG_COARSE_DOFS)) || (rc = dalloc(c, &c->mg_part, nblk(c->m.nnode, 256) + 64)) ||
      (rc = dalloc(c, &c->mg_lam, MG_MAX_LEVELS)))
    return rc;
  return 0;
}

// Galerkin operators, diagonals, Gershgorin bounds and the coarsest inverse for the current Jacobian.
template <typename VT>
int mg_setup_t(pa_context *c) {
  const int L = c->mg_levels;
  c->mg_A[0] = c->K;
  for (int l = 1; l <= L; ++l)
    hipLaunchKernelGGL(k_mg_galerkin<VT>, dim3(nblk(27 * c->mg_m[l].nnode, 256)), dim3(256), 0, c->stream, c->mg_m[l - 1],
                       (const VT *)c->mg_A[l - 1], l == 1 ? (const unsigned char *)c->con : nullptr, c->mg_m[l], (VT *)c->mg_A[l]);
  for (int l = 0; l < L; ++l) {
    const unsigned g = nblk(c->mg_m[l].nnode, 256);
    hipLaunchKernelGGL(k_mg_diag<VT>, dim3(g), dim3(256), 0, c->stream, c->mg_m[l], (const VT *)c->mg_A[l],
                       l == 0 ? nullptr : c->mg_dinv[l], c->mg_part);
    hipLaunchKernelGGL(k_mg_lam, dim3(1), dim3(256), 0, c->stream, (const double *)c->mg_part, g, c->mg_lam + l);
  }
  hipLaunchKernelGGL(k_mg_factor<VT>, dim3(1), dim3(256), 0, c->stream, c->mg_m[L], (const VT *)c->mg_A[L], c->mg_minv);
  HIPCHK(hipGetLastError());
  return 0;
}
int mg_setup(pa_context *c) { return mixed(c) ? mg_setup_t<float>(c) : mg_setup_t<double>(c); }

// z = M r: one V-cycle; the last kernel leaves the block partials of r.z at the front of partials.
// Levels above lt run one kernel per step; levels lt..L run in the one-block k_mg_tail.
template <typename VT>
int mg_vcycle_t(pa_context *c) {
  const int L = c->mg_levels;
  const KState *st = c->kst;
  auto mask = [&](int l) { return l == 0 ? (const unsigned char *)c->con : nullptr; };
  int lt = 1;
  while (lt < L && c->mg_m[lt].nnode > MG_TAIL_NODES) ++lt;
  // one-wave blocks below level 0, so that the coarser levels spread over many CUs (level 0 keeps
  // VB-lane blocks: its block partials feed the Krylov kernels)
  auto bs = [&](int l) { return l == 0 ? (unsigned)VB : 64u; };
  for (int l = 0; l < lt; ++l) {
    const MeshDims &m = c->mg_m[l];
    const unsigned g = nblk(m.nnode, bs(l));
    const VT *A = (const VT *)c->mg_A[l];
    hipLaunchKernelGGL(k_mg_pre<VT>, dim3(g), dim3(bs(l)), 0, c->stream, m, A, (const double *)c->mg_dinv[l], (const double *)(c->mg_lam + l),
                       (const double *)c->mg_b[l], c->mg_x[l], c->mg_d[l], c->mg_r[l], st);
    hipLaunchKernelGGL(k_mg_resid<VT>, dim3(g), dim3(bs(l)), 0, c->stream, m, A, (const double *)c->mg_d[l], c->mg_r[l], st);
    if (l + 1 < lt)
      hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(c->mg_m[l + 1].nnode, 64)), dim3(64), 0, c->stream, m, (const double *)c->mg_r[l],
                         mask(l), c->mg_m[l + 1], c->mg_b[l + 1], st);
  }
  MgTail t;
  std::memset(&t, 0, sizeof(t));
  t.lt = lt;
  t.L = L;
  t.mf = c->mg_m[lt - 1];
  t.rf = c->mg_r[lt - 1];
  t.con = mask(lt - 1);
  for (int l = lt; l <= L; ++l) {
    t.m[l] = c->mg_m[l];
    t.A[l] = c->mg_A[l];
    t.dinv[l] = c->mg_dinv[l];
    t.b[l] = c->mg_b[l];
    t.x[l] = c->mg_x[l];
    t.r[l] = c->mg_r[l];
    t.d[l] = c->mg_d[l];
  }
  t.lam = c->mg_lam;
  t.minv = c->mg_minv;
  hipLaunchKernelGGL(k_mg_tail<VT>, dim3(1), dim3(MG_TAIL_THREADS), 0, c->stream, t, st);
  for (int l = lt - 1; l >= 0; --l) {
    const MeshDims &m = c->mg_m[l];
    const unsigned g = nblk(m.nnode, bs(l));
    const VT *A = (const VT *)c->mg_A[l];
    hipLaunchKernelGGL(k_mg_prolong, dim3(g), dim3(bs(l)), 0, c->stream, m, c->mg_x[l], mask(l), c->mg_m[l + 1],
                       (const double *)c->mg_x[l + 1], st);
    hipLaunchKernelGGL(k_mg_post1<VT>, dim3(g), dim3(bs(l)), 0, c->stream, m, A, (const double *)c->mg_dinv[l],
                       (const double *)(c->mg_lam + l), (const double *)c->mg_b[l], (const double *)c->mg_x[l], c->mg_d[l], c->mg_r[l], st);
    if (l == 0)
      hipLaunchKernelGGL((k_mg_post2<VT, true>), dim3(g), dim3(VB), 0, c->stream, m, A, (const double *)c->mg_dinv[l],
                         (const double *)(c->mg_lam + l), (const double *)c->mg_b[l], c->mg_x[l], (const double *)c->mg_d[l],
                         (const double *)c->mg_r[l], st, c->partials);
    else
      hipLaunchKernelGGL((k_mg_post2<VT, false>), dim3(g), dim3(bs(l)), 0, c->stream, m, A, (const double *)c->mg_dinv[l],
                         (const double *)(c->mg_lam + l), (const double *)c->mg_b[l], c->mg_x[l], (const double *)c->mg_d[l],
                         (const double *)c->mg_r[l], st, (double *)nullptr);
  }
  HIPCHK(hipGetLastError());
  return 0;
}
// The V-cycle's ~20 launches are captured once per form into a graph and replayed.
int mg_vcycle(pa_context *c) {
  const int f = mixed(c) ? 1 : 0;
  if (!c->mg_exec[f]) {
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = f ? mg_vcycle_t<float>(c) : mg_vcycle_t<double>(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail(PA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      return fail(PA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    c->mg_graph[f] = g;
    c->mg_exec[f] = ex;
  }
  HIPCHK(hipGraphLaunch(c->mg_exec[f], c->stream));
  return 0;
}

// One Krylov launch sequence for iteration k: p update, operator (CG: K p; BiCG: K p and K^T p~), update.
int launch_krylov_iteration(pa_context *c, int k) {
  const long long N = c->ndof;
  double *X = c->kv[0], *Rr = c->kv[1], *Rl = c->kv[2], *Zr = c->kv[3], *Zl = c->kv[4], *Pr = c->kv[5], *Pl = c->kv[6];
  const unsigned gu = ugrid(N), gn = nblk(c->m.nnode, VB);
  double *partA = c->partials, *partU = c->partials + spmv_partials_offset(c);
  if (c->method == PA_SOLVER_MGCG) {
    double *Q = Zl;
    hipLaunchKernelGGL(k_mgcg_p, dim3(gu), dim3(VB), 0, c->stream, N, (const double *)Zr, Pr, c->kst, (const double *)partU, gu,
                       (const double *)partA, gn, k);
    if (mixed(c))
      hipLaunchKernelGGL(k_cg_spmv<float>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const float *)c->K, (const double *)Pr, Q,
                         (const KState *)c->kst, partA);
    else
      hipLaunchKernelGGL(k_cg_spmv<double>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const double *)c->K, (const double *)Pr, Q,
                         (const KState *)c->kst, partA);
    hipLaunchKernelGGL(k_mgcg_update, dim3(gu), dim3(VB), 0, c->stream, N, X, (const double *)Pr, Rr, (const double *)Q,
                       (const double *)c->dinv, (const KState *)c->kst, (const double *)partA, gn, partU, k);
    HIPCHK(hipGetLastError());
    return mg_vcycle(c);
  } else if (c->method == PA_SOLVER_CG) {
    double *Q = Zl;
    hipLaunchKernelGGL(k_cg_p, dim3(gu), dim3(VB), 0, c->stream, N, (const double *)Zr, Pr, c->kst, (const double *)partU, gu, k);
    if (c->pbc) pbc_expand(c, Pr, nullptr, 0.0, c->kst);
    if (mixed(c))
      hipLaunchKernelGGL(k_cg_spmv<float>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const float *)c->K, (const double *)Pr, Q,
                         (const KState *)c->kst, partA);
    else
      hipLaunchKernelGGL(k_cg_spmv<double>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const double *)c->K, (const double *)Pr, Q,
                         (const KState *)c->kst, partA);
    if (c->pbc) pbc_fold(c, Q, nullptr, c->kst);
    hipLaunchKernelGGL(k_cg_update, dim3(gu), dim3(VB), 0, c->stream, N, X, (const double *)Pr, Rr, (const double *)Q, Zr,
                       (const double *)c->dinv, (const KState *)c->kst, (const double *)partA, gn, partU, k);
  } else {
    hipLaunchKernelGGL(k_bicg_p, dim3(gu), dim3(VB), 0, c->stream, N, (const double *)Zr, (const double *)Zl, Pr, Pl, c->kst,
                       (const double *)partU, gu, k);
    if (c->pbc) pbc_expand(c, Pr, Pl, 0.0, c->kst);
    if (mixed(c))
      hipLaunchKernelGGL(k_bicg_spmv<float>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const float *)c->K, (const double *)Pr,
                         (const double *)Pl, Zr, Zl, (const KState *)c->kst, partA);
    else
      hipLaunchKernelGGL(k_bicg_spmv<double>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const double *)c->K, (const double *)Pr,
                         (const double *)Pl, Zr, Zl, (const KState *)c->kst, partA);
    if (c->pbc) pbc_fold(c, Zr, Zl, c->kst);
    hipLaunchKernelGGL(k_bicg_update, dim3(gu), dim3(VB), 0, c->stream, N, X, (const double *)Pr, Rr, Rl, Zr, Zl,
                       (const double *)c->dinv, (const KState *)c->kst, (const double *)partA, gn, partU, k);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int krylov_run(pa_context *c, KState &hs);

int do_solve(pa_context *c, int it, int *iters) {
  if (c->asm_precision != c->precision)
    return fail(PA_ERR_STATE, "no assembly in the current arithmetic form: call pa_assemble after pa_set_solver");
  KState hs;
  if (int rc = krylov_run(c, hs)) return rc;
  if (iters) *iters = hs.it;
  c->last_krylov = hs;
  c->solved = true;
  if (hs.breakdown == KRYLOV_NONFINITE)  // as PETSc's KSP_DIVERGED_NANORINF: no correction is applied
    return fail(PA_ERR_NONFINITE, "Krylov solve stopped on a NaN/Inf residual norm at iteration " + std::to_string(hs.it) +
                                      " (right-hand side norm " + std::to_string(hs.bnorm) + ")");
  const long long N = c->ndof;
  const double *X = c->kv[0];
  // the CG iterate x is complete; BiCG likewise (x lives in kv[0] for both)
  if (c->pbc)
    hipLaunchKernelGGL(k_pbc_apply, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, c->u, X, c->pbcd, c->pbc_s);
  else
    hipLaunchKernelGGL(k_apply_du, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, c->u, X, (const unsigned char *)c->con,
                       (const double *)c->delu, (int)(it == 0));
  HIPCHK(hipGetLastError());
  return 0;
}

// the Krylov solve of K X = R (method and form of the context) into kv[0]; hs: the final state
int krylov_run(pa_context *c, KState &hs) {
  const long long N = c->ndof;
  double *X = c->kv[0], *Rr = c->kv[1], *Rl = c->kv[2], *Zr = c->kv[3], *Zl = c->kv[4];
  const bool cg = c->method != PA_SOLVER_BICG;
  const bool mg = c->method == PA_SOLVER_MGCG;
  if (mg) {
    const int rc = mg_setup(c);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_krylov_init, dim3(vgrid(N)), dim3(VB), 0, c->stream, N, c->R, X, Rr, cg ? nullptr : Rl, Zr, cg ? nullptr : Zl,
                     c->dinv, c->kst, c->pb.lin_tol, c->pb.lin_maxit, c->partials + update_partials_end(c), c->counters + 0);
  HIPCHK(hipGetLastError());
  if (mg) {
    const int rc = mg_vcycle(c);  // z_0 = M r_0 and the partials of r_0.z_0 for iteration 0
    if (rc) return rc;
  }
  // The kernels test convergence themselves and return at once after it.  The host reads the
  // flag every `chunk` iterations, one chunk behind: chunk j+1 is queued before the copy of the
  // flag after chunk j is waited on, so the stream never idles on the host round trip; at most
  // one chunk of launches runs after convergence, each returning at its first instruction.
  // (multigrid: a few dozen iterations of ~30 launches each, so a short chunk bounds the
  // launches that run after convergence)
  const int chunk = mg ? 4 : 16;
  for (int j = 0;; ++j) {
    for (int k = 0; k < chunk; ++k) {
      const int rc = launch_krylov_iteration(c, j * chunk + k);
      if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(&c->hks[j & 1], c->kst, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->kev[j & 1], c->stream));
    if (j == 0) continue;
    HIPCHK(hipEventSynchronize(c->kev[(j - 1) & 1]));
    hs = c->hks[(j - 1) & 1];
    if (hs.done) break;
  }
  return 0;
}

}  // namespace

namespace {
// The one-lane qp passes hold each lane's active-set and tangent matrices in private memory: 8.6 KB
// per lane for up to 12 systems, 25 KB for up to 24.  The 48-system variant would need 84 KB, a
// scratch reservation the runtime refuses on MI355X (HSA_STATUS_ERROR_OUT_OF_RESOURCES at the first
// dispatch), so its matrices live in the context's per-lane workspace (cp::LaneWs, 5-6 KB private
// left).  The batch entry point always runs cooperatively.
int check_qp_form(const pa_context *c) {
  if (c->qp_form != PA_QP_LANE) return fail(PA_ERR_UNSUPPORTED, "only the one-lane qp form takes Newton steps");
  return 0;
}
}  // namespace

extern "C" {

const char *pa_last_error(void) { return g_err.c_str(); }

int pa_create(const pa_material *mat, const pa_problem *pb, pa_context **out) {
  if (!mat || !pb || !out) return fail(PA_ERR_ARG, "null argument");
  *out = nullptr;
  if (mat->n_slip < 1 || mat->n_slip > PA_MAX_SLIP) return fail(PA_ERR_ARG, "n_slip out of range");
  if (mat->max_slip_search != 1) return fail(PA_ERR_UNSUPPORTED, "Max Slip Search Iterations must be 1");
  if (mat->max_solver_iter < 1) return fail(PA_ERR_ARG, "Max Solver Iterations must be >= 1");
  if (mat->twinning) {
    if (mat->n_twin < 1 || mat->n_slip + mat->n_twin > PA_MAX_SLIP)
      return fail(PA_ERR_ARG, "twinning: need 1 <= n_twin and n_slip + n_twin <= PA_MAX_SLIP");
    if (!(mat->twin_shear > 0)) return fail(PA_ERR_ARG, "twinning: Characteristic Twin Shear must be > 0");
  }
  if (pb->nx < 1 || pb->ny < 1 || pb->nz < 1) return fail(PA_ERR_ARG, "mesh size must be >= 1");
  if (pb->total_increments < 1 || pb->max_newton < 1 || pb->lin_maxit < 1) return fail(PA_ERR_ARG, "bad increment/solver counts");
  if (pb->tabular && (pb->n_table < 2 || pb->n_table > PA_MAX_TABLE)) return fail(PA_ERR_ARG, "tabular BCs need 2..PA_MAX_TABLE time points");
  if (pb->tabular && pb->cyclic) return fail(PA_ERR_UNSUPPORTED, "tabular and cyclic boundary conditions together");
  if (pb->velgrad) {
    double A[3][3], R[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) A[a][b] = (a == b ? 1.0 : 0.0) - pb->vel_grad[a][b] * pb->delT;
    if (!host_inv3(A, R)) return fail(PA_ERR_ARG, "velocity gradient BC: I - L*delT is singular");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PA_ERR_HIP, "no HIP device available");
  pa_context *c = new pa_context();
  c->mat = *mat;
  c->pb = *pb;
  c->kin = mat->kinematic_hardening != 0;
  c->nslip = mat->n_slip;
  c->ntw = mat->twinning ? mat->n_twin : 0;
  c->ns = c->nslip + c->ntw;
  MeshDims &m = c->m;
  m.nx = pb->nx; m.ny = pb->ny; m.nz = pb->nz;
  m.nnx = m.nx + 1; m.nny = m.ny + 1; m.nnz = m.nz + 1;
  m.nnode = (long long)m.nnx * m.nny * m.nnz;
  m.nelem = (long long)m.nx * m.ny * m.nz;
  c->nq = 8 * m.nelem;
  c->ndof = 3 * m.nnode;
  for (int d = 0; d < 3; ++d) c->h[d] = pb->span[d] / (d == 0 ? m.nx : d == 1 ? m.ny : m.nz);
  c->JxW = c->h[0] * c->h[1] * c->h[2] / 8.0;
  c->microvol = 0.0;  // microvol = the sum of JxW over every qp in order (updateAfterIncrement.cc:221)
  for (long long e = 0; e < 8 * c->m.nelem; ++e) c->microvol += c->JxW;
  auto cnt = [](int i, int nn) { return (long long)(std::min(i + 1, nn - 1) - std::max(i - 1, 0) + 1); };
  auto tot = [&](int nn) { long long s = 0; for (int i = 0; i < nn; ++i) s += cnt(i, nn); return s; };
  c->nnz = 9 * tot(m.nnx) * tot(m.nny) * tot(m.nnz);
  if (c->nnz >= (1LL << 31)) { delete c; return fail(PA_ERR_ARG, "mesh too large for int32 column indices"); }
  int rc = 0;
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return fail(PA_ERR_HIP, "hipStreamCreate failed"); }
  for (int i = 0; i < 4; ++i) hipEventCreate(&c->ev[i]);
  for (int i = 0; i < 2; ++i) hipEventCreateWithFlags(&c->kev[i], hipEventDisableTiming);
  for (int i = 0; i < 2; ++i) hipEventCreate(&c->qev[i]);
  if (hipHostMalloc((void **)&c->hks, 2 * sizeof(KState), hipHostMallocDefault) != hipSuccess) {
    pa_destroy(c);
    return fail(PA_ERR_HIP, "hipHostMalloc failed");
  }
  const int NS = ns_template(c->ns);
  // material block
  {
    // layout of cp::MatLayout<NS>: h0 ss pw q tol C_1 C_2
    const int tol = 3 * NS + NS * NS, c1 = tol + 1, c2 = c1 + NS;
    std::vector<double> mb(c2 + NS, 0.0);
    for (int a = 0; a < c->ns; ++a) {
      mb[a] = mat->h0[a]; mb[NS + a] = mat->ss[a]; mb[2 * NS + a] = mat->pw[a];
      for (int b = 0; b < c->ns; ++b) mb[3 * NS + a * NS + b] = mat->q[a][b];
      if (c->kin) { mb[c1 + a] = mat->C1[a]; mb[c2 + a] = mat->C2[a]; }
    }
    mb[tol] = mat->stress_tol;
    if ((rc = dalloc(c, &c->d_mat, mb.size()))) { pa_destroy(c); return rc; }
    if (hipMemcpy(c->d_mat, mb.data(), mb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { pa_destroy(c); return fail(PA_ERR_HIP, "copy material"); }
  }
  c->gstride = 9 * NS + 36 + 81 + 9;
  const long long nq = c->nq, nn = m.nnode, N = c->ndof;
  const int ns = c->ns;
  if ((rc = dalloc(c, &c->Fp_conv, 9 * nq)) || (rc = dalloc(c, &c->Fp_iter, 9 * nq)) || (rc = dalloc(c, &c->Fe_conv, 9 * nq)) ||
      (rc = dalloc(c, &c->Fe_iter, 9 * nq)) || (rc = dalloc(c, &c->s_conv, ns * nq)) || (rc = dalloc(c, &c->s_iter, ns * nq)) ||
      (rc = dalloc(c, &c->slip_conv, ns * nq)) || (rc = dalloc(c, &c->slip_iter, ns * nq)) || (rc = dalloc(c, &c->rotnew, 3 * nq)) ||
      (rc = dalloc(c, &c->T, 9 * nq)) || (rc = dalloc(c, &c->u, N)) || (rc = dalloc(c, &c->R, N)) ||
      (rc = dalloc(c, (double **)&c->K, 243 * nn)) || (rc = dalloc(c, (double **)&c->Dq, 81 * nq)) ||
      (rc = dalloc(c, &c->Pq, 9 * nq)) || (rc = dalloc(c, &c->queue, nq)) || (rc = dalloc(c, &c->dinv, N)) || (rc = dalloc(c, &c->con, N)) ||
      (rc = dalloc(c, &c->delu, N)) || (rc = dalloc(c, &c->d_egrain, m.nelem)) || (rc = dalloc(c, &c->kst, 1)) ||
      (rc = dalloc(c, &c->partials, partials_size(c))) || (rc = dalloc(c, &c->counters, 16)) ||
      (rc = dalloc(c, &c->d_out, 64))) {
    pa_destroy(c);
    return rc;
  }
  if (c->ntw) {  // twin fractions start at zero, every qp in its grain's orientation (init.cc:274-324)
    if ((rc = dalloc(c, &c->tw_conv, (long long)c->ntw * nq)) || (rc = dalloc(c, &c->tw_iter, (long long)c->ntw * nq)) ||
        (rc = dalloc(c, &c->var_conv, nq)) || (rc = dalloc(c, &c->var_iter, nq))) {
      pa_destroy(c);
      return rc;
    }
    const size_t bt = sizeof(double) * c->ntw * nq, bv = sizeof(int) * nq;
    if (hipMemset(c->tw_conv, 0, bt) != hipSuccess || hipMemset(c->tw_iter, 0, bt) != hipSuccess ||
        hipMemset(c->var_conv, 0, bv) != hipSuccess || hipMemset(c->var_iter, 0, bv) != hipSuccess) {
      pa_destroy(c);
      return fail(PA_ERR_HIP, "state init");
    }
  }
  for (int i = 0; i < 7; ++i)
    if ((rc = dalloc(c, &c->kv[i], N))) { pa_destroy(c); return rc; }
  if (cp::lane_ws<48>() && ns_template(c->ns) == 48 &&
      (rc = dalloc(c, &c->ws, (size_t)WS_BLOCKS * 64 * cp::LaneWs<48>::size))) {
    pa_destroy(c);
    return rc;
  }
  {  // singular-system store: as many entries as fit 768 MiB (4 GiB for 48 systems, whose entries
     // take 19 KB and whose qps meet singular systems more often; at most one per qp)
    const size_t ns_t = (size_t)ns_template(c->ns);
    const size_t per = sizeof(double) * (ns_t * ns_t + ns_t + SING_ROUNDS * ns_t) + SING_ROUNDS * 8 + 3 * sizeof(int);
    const size_t budget = ns_t == 48 ? ((size_t)4 << 30) : ((size_t)768 << 20);
    const size_t cap = std::min<size_t>((size_t)nq, std::max<size_t>(64, budget / per));
    c->sing.cap = (unsigned)cap;
    c->sing.miss = c->counters + 10;
    if ((rc = dalloc(c, &c->sing.A, cap * ns_t * ns_t)) || (rc = dalloc(c, &c->sing.b, cap * ns_t)) || (rc = dalloc(c, &c->sing.n, cap)) ||
        (rc = dalloc(c, &c->sing.X, cap * SING_ROUNDS * ns_t)) || (rc = dalloc(c, &c->sing.have, cap)) ||
        (rc = dalloc(c, &c->sing.done, cap)) || (rc = dalloc(c, &c->sing.key, cap * SING_ROUNDS))) {
      pa_destroy(c);
      return rc;
    }
  }
  // initial state (crystalPlasticity::init, init.cc:302-330): Fp = Fe = I, s = s0, slip = 0
  {
    std::vector<double> hbuf(9 * nq, 0.0);
    for (int i = 0; i < 3; ++i) std::fill(hbuf.begin() + (4 * i) * nq, hbuf.begin() + (4 * i + 1) * nq, 1.0);
    if (hipMemcpy(c->Fp_conv, hbuf.data(), 9 * nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->Fe_conv, hbuf.data(), 9 * nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->Fp_iter, hbuf.data(), 9 * nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->Fe_iter, hbuf.data(), 9 * nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      pa_destroy(c);
      return fail(PA_ERR_HIP, "state init copy");
    }
    std::vector<double> sb((size_t)ns * nq);
    for (int a = 0; a < ns; ++a) std::fill(sb.begin() + a * nq, sb.begin() + (a + 1) * nq, mat->s0[a]);
    if (hipMemcpy(c->s_conv, sb.data(), sb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->s_iter, sb.data(), sb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      pa_destroy(c);
      return fail(PA_ERR_HIP, "state init copy");
    }
  }
  if (c->kin) {  // backstress W_kh starts at zero (init.cc:302-330)
    if ((rc = dalloc(c, &c->W_conv, ns * nq)) || (rc = dalloc(c, &c->W_iter, ns * nq))) { pa_destroy(c); return rc; }
    if (hipMemset(c->W_conv, 0, sizeof(double) * ns * nq) != hipSuccess || hipMemset(c->W_iter, 0, sizeof(double) * ns * nq) != hipSuccess) {
      pa_destroy(c);
      return fail(PA_ERR_HIP, "state init");
    }
  }
  if (hipMemset(c->slip_conv, 0, sizeof(double) * ns * nq) != hipSuccess || hipMemset(c->slip_iter, 0, sizeof(double) * ns * nq) != hipSuccess ||
      hipMemset(c->u, 0, sizeof(double) * N) != hipSuccess || hipMemset(c->T, 0, sizeof(double) * 9 * nq) != hipSuccess ||
      hipMemset(c->counters, 0, sizeof(unsigned) * 16) != hipSuccess || hipMemset(c->kst, 0, sizeof(KState)) != hipSuccess) {
    pa_destroy(c);
    return fail(PA_ERR_HIP, "memset");
  }
  *out = c;
  return 0;
}

int pa_destroy(pa_context *c) {
  if (!c) return 0;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (int f = 0; f < 2; ++f) {
    if (c->mg_exec[f]) (void)hipGraphExecDestroy(c->mg_exec[f]);
    if (c->mg_graph[f]) (void)hipGraphDestroy(c->mg_graph[f]);
  }
  for (void *p : c->allocs) hipFree(p);
  if (c->d_gtab) hipFree(c->d_gtab);
  if (c->d_vrod) hipFree(c->d_vrod);
  if (c->d_gphase) hipFree(c->d_gphase);
  for (int i = 0; i < 4; ++i)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  for (int i = 0; i < 2; ++i)
    if (c->kev[i]) hipEventDestroy(c->kev[i]);
  for (int i = 0; i < 2; ++i)
    if (c->qev[i]) hipEventDestroy(c->qev[i]);
  if (c->hks) hipHostFree(c->hks);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int pa_sizes(pa_context *c, long long *nelem, long long *nqp, long long *ndof, long long *nnz) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  if (nelem) *nelem = c->m.nelem;
  if (nqp) *nqp = c->nq;
  if (ndof) *ndof = c->ndof;
  if (nnz) *nnz = c->nnz;
  return 0;
}

int pa_element_grains(pa_context *c, const unsigned *vox, const int nv[3], unsigned *elem_grain) {
  if (!c || !vox || !nv || !elem_grain) return fail(PA_ERR_ARG, "null argument");
  if (nv[0] < 1 || nv[1] < 1 || nv[2] < 1) return fail(PA_ERR_ARG, "voxel counts must be >= 1");
  const size_t nvox = (size_t)nv[0] * nv[1] * nv[2];
  unsigned *dv = nullptr, *de = nullptr;
  HIPCHK(hipMalloc(&dv, nvox * sizeof(unsigned)));
  if (hipMalloc(&de, c->m.nelem * sizeof(unsigned)) != hipSuccess) { hipFree(dv); return fail(PA_ERR_HIP, "hipMalloc"); }
  int rc = 0;
  if (hipMemcpy(dv, vox, nvox * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) rc = fail(PA_ERR_HIP, "copy voxels");
  if (!rc) {
    hipLaunchKernelGGL(k_element_grains, dim3(nblk(c->m.nelem, 256)), dim3(256), 0, c->stream, c->m, c->pb.span[0], c->pb.span[1],
                       c->pb.span[2], (const unsigned *)dv, nv[0], nv[1], nv[2], de);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(PA_ERR_HIP, "grain kernel");
  }
  if (!rc && hipMemcpy(elem_grain, de, c->m.nelem * sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(PA_ERR_HIP, "copy grains");
  hipFree(dv);
  hipFree(de);
  return rc;
}

namespace {
// grain tables of a context with several phases: each grain row with its phase's systems and
// elastic moduli (no twinning variants); s starts at each qp's phase's s0 (init.cc:1125-1175)
int set_grains_phases(pa_context *c, const unsigned *elem_grain, const double *grain_rod, int n_grain_rows) {
  const int NS = ns_template(c->ns);
  const int nvar = 1 + c->ntw;  // orientation rows per grain: the grain's, then its phase's twin variants
  std::vector<double> tab((size_t)n_grain_rows * nvar * c->gstride, 0.0);
  c->h_vrod.assign((size_t)n_grain_rows * nvar * 3, 0.0);
  for (int g = 0; g < n_grain_rows; ++g) {
    const pa_material &mt = c->phmat[c->h_gphase[g]];
    const int ntw = mt.twinning ? mt.n_twin : 0, nsys = mt.n_slip + ntw;
    std::vector<double> mv(3 * nsys), nv(3 * nsys);
    for (int a = 0; a < nsys; ++a) {  // normalised slip and twin vectors (init.cc:73-123)
      double mn = 0, nn = 0;
      for (int j = 0; j < 3; ++j) { mn += mt.m[a][j] * mt.m[a][j]; nn += mt.n[a][j] * mt.n[a][j]; }
      mn = std::sqrt(mn); nn = std::sqrt(nn);
      for (int j = 0; j < 3; ++j) { mv[3 * a + j] = mt.m[a][j] / mn; nv[3 * a + j] = mt.n[a][j] / nn; }
    }
    for (int v = 0; v < nvar; ++v) {
      double *r = &c->h_vrod[((size_t)g * nvar + v) * 3];
      if (v == 0 || v > ntw) for (int i = 0; i < 3; ++i) r[i] = grain_rod[3 * g + i];
      else twin_variant_rod(grain_rod + 3 * g, &nv[3 * (mt.n_slip + v - 1)], r);
      grain_table(mt, (const double(*)[3])mv.data(), (const double(*)[3])nv.data(), r, nsys, NS,
                  &tab[((size_t)g * nvar + v) * c->gstride]);
    }
  }
  if (c->ntw) {
    if (c->d_vrod) hipFree(c->d_vrod);
    c->d_vrod = nullptr;
    HIPCHK(hipMalloc(&c->d_vrod, c->h_vrod.size() * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_vrod, c->h_vrod.data(), c->h_vrod.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (c->d_gtab) hipFree(c->d_gtab);
  c->d_gtab = nullptr;
  HIPCHK(hipMalloc(&c->d_gtab, tab.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_gtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  if (c->d_gphase) hipFree(c->d_gphase);
  c->d_gphase = nullptr;
  HIPCHK(hipMalloc(&c->d_gphase, (size_t)n_grain_rows * sizeof(int)));
  HIPCHK(hipMemcpy(c->d_gphase, c->h_gphase.data(), (size_t)n_grain_rows * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_egrain, elem_grain, c->m.nelem * sizeof(unsigned), hipMemcpyHostToDevice));
  const long long nq = c->nq;
  std::vector<double> sb((size_t)c->ns * nq, 0.0), rn(3 * nq);
  for (long long e = 0; e < c->m.nelem; ++e) {
    const pa_material &mt = c->phmat[c->h_gphase[elem_grain[e]]];
    const int nsys = mt.n_slip + (mt.twinning ? mt.n_twin : 0);
    for (int q = 0; q < 8; ++q) {
      const long long Q = 8 * e + q;
      for (int a = 0; a < nsys; ++a) sb[(size_t)a * nq + Q] = mt.s0[a];
      for (int i = 0; i < 3; ++i) rn[i * nq + Q] = grain_rod[3 * elem_grain[e] + i];
    }
  }
  HIPCHK(hipMemcpy(c->s_conv, sb.data(), sb.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->s_iter, sb.data(), sb.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->rotnew, rn.data(), rn.size() * sizeof(double), hipMemcpyHostToDevice));
  c->n_grain_rows = n_grain_rows;
  c->grains_set = true;
  return 0;
}
}  // namespace

int pa_create_phases(const pa_material *mats, int n_phases, const pa_problem *pb, pa_context **out) {
  if (!mats || !pb || !out) return fail(PA_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_phases < 1 || n_phases > 4) return fail(PA_ERR_ARG, "1 to 4 phases");
  // the envelope: the largest system count, kinematic hardening if any phase has it; the
  // tolerance and local-loop settings are global prm keys, so the phases must agree on them
  pa_material env = mats[0];
  int tot = 0, ntw = 0;
  for (int k = 0; k < n_phases; ++k) {
    const pa_material &m = mats[k];
    const int mt = m.twinning ? m.n_twin : 0;
    if (m.n_slip < 1 || m.n_slip + mt > PA_MAX_SLIP) return fail(PA_ERR_ARG, "n_slip out of range");
    if (m.twinning && (m.n_twin < 1 || !(m.twin_shear > 0))) return fail(PA_ERR_ARG, "twinning: n_twin >= 1 and twin shear > 0");
    if (m.max_slip_search != 1) return fail(PA_ERR_UNSUPPORTED, "Max Slip Search Iterations must be 1");
    if (m.max_solver_iter != mats[0].max_solver_iter || m.stress_tol != mats[0].stress_tol)
      return fail(PA_ERR_ARG, "the phases must share Max Solver Iterations and Stress Tolerance");
    tot = std::max(tot, m.n_slip + mt);
    ntw = std::max(ntw, mt);
    env.kinematic_hardening |= m.kinematic_hardening;
  }
  // state for the most systems and the most twin systems of any phase
  env.twinning = ntw > 0;
  env.n_twin = ntw;
  env.n_slip = tot - ntw;
  if (env.twinning && !(env.twin_shear > 0)) env.twin_shear = 1.0;  // per phase in the kernels
  if (env.n_slip < 1) return fail(PA_ERR_ARG, "every phase needs slip systems");
  int rc = pa_create(&env, pb, out);
  if (rc) return rc;
  pa_context *c = *out;
  c->multiphase = true;
  if (n_phases == 1) { c->mat = mats[0]; return 0; }
  const int NS = ns_template(c->ns);
  const int tol = 3 * NS + NS * NS, c1 = tol + 1, c2 = c1 + NS;
  c->mstride = c2 + NS;  // cp::MatLayout<NS>::size
  std::vector<double> mb((size_t)n_phases * c->mstride, 0.0);
  for (int k = 0; k < n_phases; ++k) {
    const pa_material &m = mats[k];
    double *b = &mb[(size_t)k * c->mstride];
    const int nsys = m.n_slip + (m.twinning ? m.n_twin : 0);
    for (int a = 0; a < nsys; ++a) {
      b[a] = m.h0[a]; b[NS + a] = m.ss[a]; b[2 * NS + a] = m.pw[a];
      for (int q = 0; q < nsys; ++q) b[3 * NS + a * NS + q] = m.q[a][q];
      if (m.kinematic_hardening) { b[c1 + a] = m.C1[a]; b[c2 + a] = m.C2[a]; }
    }
    b[tol] = m.stress_tol;
    c->phmat[k] = m;
  }
  double *d = nullptr;
  if ((rc = dalloc(c, &d, mb.size())) || hipMemcpy(d, mb.data(), mb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    pa_destroy(c);
    *out = nullptr;
    return rc ? rc : fail(PA_ERR_HIP, "copy materials");
  }
  c->d_mat = d;
  c->nph = n_phases;
  return 0;
}

int pa_set_grain_phases(pa_context *c, const int *grain_phase, int n_grain_rows) {
  if (!c || !grain_phase || n_grain_rows < 1) return fail(PA_ERR_ARG, "null argument");
  if (c->nph < 2) return fail(PA_ERR_STATE, "the context has one phase (pa_create_phases makes several)");
  std::vector<int> gp(n_grain_rows);
  for (int g = 0; g < n_grain_rows; ++g) {
    if (grain_phase[g] < 1 || grain_phase[g] > c->nph) return fail(PA_ERR_ARG, "grain phase out of range");
    gp[g] = grain_phase[g] - 1;
  }
  c->h_gphase = gp;
  c->grains_set = false;  // the grain tables depend on the phases
  return 0;
}

int pa_set_grains(pa_context *c, const unsigned *elem_grain, const double *grain_rod, int n_grain_rows) {
  if (!c || !elem_grain || !grain_rod || n_grain_rows < 1) return fail(PA_ERR_ARG, "null argument");
  for (long long e = 0; e < c->m.nelem; ++e)
    if ((int)elem_grain[e] >= n_grain_rows) return fail(PA_ERR_ARG, "grain id without an orientation row");
  if (c->nph > 1 && (int)c->h_gphase.size() < n_grain_rows)
    return fail(PA_ERR_STATE, "several phases: pa_set_grain_phases must cover every grain row first");
  if (c->nph > 1) return set_grains_phases(c, elem_grain, grain_rod, n_grain_rows);
  const int NS = ns_template(c->ns);
  // normalised slip vectors (init.cc:73-123)
  std::vector<double> mv(3 * c->ns), nv(3 * c->ns);
  for (int a = 0; a < c->ns; ++a) {
    double mn = 0, nn = 0;
    for (int j = 0; j < 3; ++j) { mn += c->mat.m[a][j] * c->mat.m[a][j]; nn += c->mat.n[a][j] * c->mat.n[a][j]; }
    mn = std::sqrt(mn); nn = std::sqrt(nn);
    for (int j = 0; j < 3; ++j) { mv[3 * a + j] = c->mat.m[a][j] / mn; nv[3 * a + j] = c->mat.n[a][j] / nn; }
  }
  // orientation tables: one row per grain, or with twinning one per (grain, variant): the grain's
  // orientation and its 180-degree turns about each twin normal
  const int nvar = 1 + c->ntw;
  c->h_vrod.assign((size_t)n_grain_rows * nvar * 3, 0.0);
  for (int g = 0; g < n_grain_rows; ++g)
    for (int v = 0; v < nvar; ++v) {
      double *r = &c->h_vrod[((size_t)g * nvar + v) * 3];
      if (v == 0) for (int i = 0; i < 3; ++i) r[i] = grain_rod[3 * g + i];
      else twin_variant_rod(grain_rod + 3 * g, &nv[3 * (c->nslip + v - 1)], r);
    }
  std::vector<double> tab((size_t)n_grain_rows * nvar * c->gstride, 0.0);
  for (size_t row = 0; row < (size_t)n_grain_rows * nvar; ++row)
    grain_table(c->mat, (const double(*)[3])mv.data(), (const double(*)[3])nv.data(), &c->h_vrod[3 * row], c->ns, NS,
                &tab[row * c->gstride]);
  if (c->ntw) {
    if (c->d_vrod) hipFree(c->d_vrod);
    c->d_vrod = nullptr;
    HIPCHK(hipMalloc(&c->d_vrod, c->h_vrod.size() * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_vrod, c->h_vrod.data(), c->h_vrod.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (c->d_gtab) hipFree(c->d_gtab);
  c->d_gtab = nullptr;
  HIPCHK(hipMalloc(&c->d_gtab, tab.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_gtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_egrain, elem_grain, c->m.nelem * sizeof(unsigned), hipMemcpyHostToDevice));
  // rotnew starts at the grain orientation of each qp (init.cc:225-245)
  std::vector<double> rn(3 * c->nq);
  for (long long e = 0; e < c->m.nelem; ++e)
    for (int q = 0; q < 8; ++q)
      for (int i = 0; i < 3; ++i) rn[i * c->nq + 8 * e + q] = grain_rod[3 * elem_grain[e] + i];
  HIPCHK(hipMemcpy(c->rotnew, rn.data(), rn.size() * sizeof(double), hipMemcpyHostToDevice));
  c->n_grain_rows = n_grain_rows;
  c->grains_set = true;
  return 0;
}

int pa_set_solver(pa_context *c, int method, int precision) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  if (method != PA_SOLVER_BICG && method != PA_SOLVER_CG && method != PA_SOLVER_MGCG) return fail(PA_ERR_ARG, "unknown solver");
  if (precision != PA_PREC_FP64 && precision != PA_PREC_MIXED) return fail(PA_ERR_ARG, "unknown precision");
  if (method == PA_SOLVER_MGCG && c->pbc) return fail(PA_ERR_UNSUPPORTED, "multigrid-PCG does not take periodic boundary conditions");
  if (method == PA_SOLVER_MGCG) {
    const int rc = mg_alloc(c);
    if (rc) return rc;
  }
  c->method = method;
  c->precision = precision;
  return 0;
}

int pa_periodic_lines(const pa_problem *pb, const pa_periodic *per, int *constrained, long long *rowptr, long long *master,
                      double *weight, double *g, long long cap, long long *n_entries) {
  if (!pb || !per || !n_entries || cap < 0 || (cap > 0 && (!master || !weight))) return fail(PA_ERR_ARG, "null argument");
  PbcLines L;
  std::string err;
  const int rc = pbc_build(*pb, *per, L, err);
  if (rc) return fail(rc, err);
  const long long N = (long long)L.crow.size();
  long long k = 0;
  for (long long d = 0; d < N; ++d) {
    const int r = L.crow[d];
    if (constrained) constrained[d] = r >= 0;
    if (rowptr) rowptr[d] = k;
    if (g) g[d] = r >= 0 ? L.cg[r] : 0.0;
    if (r < 0) continue;
    for (int q = L.cptr[r]; q < L.cptr[r + 1]; ++q, ++k)
      if (k < cap) { master[k] = L.cidx[q]; weight[k] = L.cw[q]; }
  }
  if (rowptr) rowptr[N] = k;
  *n_entries = k;
  return 0;
}

int pa_set_periodic(pa_context *c, const pa_periodic *per) {
  if (!c || !per) return fail(PA_ERR_ARG, "null argument");
  if (c->pbc) return fail(PA_ERR_STATE, "periodic boundary conditions are already set");
  if (c->method == PA_SOLVER_MGCG) return fail(PA_ERR_UNSUPPORTED, "multigrid-PCG does not take periodic boundary conditions");
  PbcLines L;
  std::string err;
  int rc = pbc_build(c->pb, *per, L, err);
  if (rc) return fail(rc, err);
  PbcDev &p = c->pbcd;
  p.ncon = (int)L.cdof.size();
  p.nshort = (int)L.sdof.size();
  p.nlong = (int)L.ldof.size();
  auto up_i = [&](const std::vector<int> &v, const int **dst) -> int {
    int *d = nullptr;
    if (int r = dalloc(c, &d, std::max<size_t>(v.size(), 1))) return r;
    if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    *dst = d;
    return 0;
  };
  auto up_d = [&](const std::vector<double> &v, const double **dst) -> int {
    double *d = nullptr;
    if (int r = dalloc(c, &d, std::max<size_t>(v.size(), 1))) return r;
    if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    *dst = d;
    return 0;
  };
  if ((rc = up_i(L.crow, &p.crow)) || (rc = up_i(L.cdof, &p.cdof)) || (rc = up_i(L.cptr, &p.cptr)) || (rc = up_i(L.cidx, &p.cidx)) ||
      (rc = up_d(L.cw, &p.cw)) || (rc = up_d(L.cg, &p.cg)) || (rc = up_i(L.sdof, &p.sdof)) || (rc = up_i(L.sptr, &p.sptr)) ||
      (rc = up_i(L.sidx, &p.sidx)) || (rc = up_d(L.sw, &p.sw)) || (rc = up_i(L.ldof, &p.ldof)) || (rc = up_i(L.lptr, &p.lptr)) ||
      (rc = up_i(L.lidx, &p.lidx)) || (rc = up_d(L.lw, &p.lw)))
    return rc;
  if ((rc = dalloc(c, &p.part, std::max<size_t>(1, (size_t)p.nlong * PBC_GL * 2)))) return rc;
  // no Dirichlet lines: the face rules are not applied (boundaryConditions.cc:455)
  HIPCHK(hipMemsetAsync(c->con, 0, (size_t)c->ndof, c->stream));
  HIPCHK(hipMemsetAsync(c->delu, 0, sizeof(double) * (size_t)c->ndof, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->pbc_scale = L.scale;
  c->pbc = true;
  c->bc_inc = -1;
  return 0;
}

int pa_set_qp_form(pa_context *c, int form) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  if (form != PA_QP_LANE && form != PA_QP_COOP) return fail(PA_ERR_ARG, "unknown qp form");
  // The cooperative Newton-step passes took a GPU memory fault whose cause is not found; until it
  // is, they are refused (the batch entry point and the pseudo-inverse paths keep the cooperative
  // routine, which runs clean there).
  if (form == PA_QP_COOP)
    return fail(PA_ERR_UNSUPPORTED, "the cooperative Newton-step qp passes are switched off (an unexplained GPU memory "
                                    "fault); use PA_QP_LANE");
  c->qp_form = form;
  return 0;
}

int pa_assemble(pa_context *c, int inc, int it, double *res_norm) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
  if (int rc = check_qp_form(c)) return rc;
  HIPCHK(hipEventRecord(c->ev[0], c->stream));
  int rc = do_assemble(c, inc, it, res_norm);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev[1], c->stream));
  return 0;
}

int pa_linear_solve(pa_context *c, int inc, int it, int *iters) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  (void)inc;
  HIPCHK(hipEventRecord(c->ev[2], c->stream));
  int rc = do_solve(c, it, iters);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev[3], c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) c->last_ms[0] = ms;
  if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) c->last_ms[1] = ms;
  return 0;
}

int pa_newton_iteration(pa_context *c, int inc, int it, double *res_norm, int *iters) {
  int rc = pa_assemble(c, inc, it, res_norm);
  if (rc) return rc;
  return pa_linear_solve(c, inc, it, iters);
}

int pa_commit_increment(pa_context *c, double row[15]) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
  if (int rc = check_qp_form(c)) return rc;
  PostArgs a;
  a.m = c->m;
  for (int d = 0; d < 3; ++d) a.h[d] = c->h[d];
  a.JxW = c->JxW;
  a.u = c->u; a.egrain = c->d_egrain; a.gtab = c->d_gtab; a.gstride = c->gstride; a.mat = c->d_mat;
  a.ns = c->ns; a.max_solver_iter = c->mat.max_solver_iter; a.nq = c->nq;
  a.Fp_conv = c->Fp_conv; a.s_conv = c->s_conv; a.slip_conv = c->slip_conv; a.Fe_conv = c->Fe_conv; a.W_conv = c->W_conv;
  a.Fe_iter = c->Fe_iter; a.Fp_iter = c->Fp_iter; a.s_iter = c->s_iter; a.slip_iter = c->slip_iter; a.T = c->T; a.rotnew = c->rotnew;
  a.W_iter = c->W_iter;
  a.partials = c->partials; a.counter = c->counters + 4; a.out = c->d_out;
  a.queue = c->queue; a.qcount = c->counters + 9; a.sing = c->sing;
  a.nsing = c->counters + 9; a.work = c->counters + 12;
  a.tw = twin_args(c);
  a.ph = phase_args(c);
  a.ws = c->ws;
  HIPCHK(hipEventRecord(c->ev[0], c->stream));
  const int NS = ns_template(c->ns);
  const unsigned grid = nblk(c->nq, 256);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, c->stream, a.qcount);
  if (NS == 12) post_passes<12>(c, a, grid);
  else if (NS == 24) post_passes<24>(c, a, grid);
  else post_passes<48>(c, a, grid);
  hipLaunchKernelGGL(k_post_reduce, dim3(grid), dim3(256), 0, c->stream, a);
  HIPCHK(hipGetLastError());
  // _conv <- _iter (updateAfterIncrement.cc:298-305); _iter keeps its values
  const size_t b9 = 9 * c->nq * sizeof(double), bs = (size_t)c->ns * c->nq * sizeof(double);
  HIPCHK(hipMemcpyAsync(c->Fe_conv, c->Fe_iter, b9, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Fp_conv, c->Fp_iter, b9, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->s_conv, c->s_iter, bs, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->slip_conv, c->slip_iter, bs, hipMemcpyDeviceToDevice, c->stream));
  if (c->kin) HIPCHK(hipMemcpyAsync(c->W_conv, c->W_iter, bs, hipMemcpyDeviceToDevice, c->stream));
  if (c->ntw) {  // twinfraction_conv, twin_conv and rot_conv (as the variant) <- _iter (:302-305)
    HIPCHK(hipMemcpyAsync(c->tw_conv, c->tw_iter, sizeof(double) * c->ntw * c->nq, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->var_conv, c->var_iter, sizeof(int) * c->nq, hipMemcpyDeviceToDevice, c->stream));
  }
  double h[15];
  HIPCHK(hipMemcpyAsync(h, c->d_out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipEventRecord(c->ev[1], c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) c->last_ms[2] = ms;
  const double vol = c->microvol;
  // updateAfterIncrement.cc:719-725: the F_T of the next increment and its F_r > 0 reorientation
  // test use these; with Enable Multiphase the reference sets F_r, F_e and F_s to zero, so no qp
  // reorients by twinning and the three columns read 0
  if (c->ntw && !c->multiphase) {
    c->F_r = h[13] / vol;
    c->F_e = h[14] / vol;
  }
  if (row) {
    for (int k = 0; k < 12; ++k) row[k] = h[k] / vol;
    row[12] = c->ntw ? c->F_r : 0.0;  // TwinRealVF (without twinning no twin system is ever active)
    row[13] = c->ntw ? c->F_e : 0.0;  // TwinMade
    row[14] = c->multiphase ? 0.0 : h[12] / vol;
  }
  c->bc_inc = -1;
  for (int k = 0; k < 15; ++k)
    if (!std::isfinite(h[k])) return fail(PA_ERR_NONFINITE, "post-increment averages are not finite (NaN/Inf stress or strain)");
  return 0;
}

int pa_run_increment(pa_context *c, int inc, double row[15], int *lin_iters, double *res_norms) {
  if (!c) return fail(PA_ERR_ARG, "null context");
  double t_asm = 0, t_kry = 0;
  for (int it = 0; it < c->pb.max_newton; ++it) {
    double rn = 0;
    int its = 0;
    int rc = pa_newton_iteration(c, inc, it, res_norms ? &rn : nullptr, &its);
    if (rc) return rc;
    if (res_norms) res_norms[it] = rn;
    if (lin_iters) lin_iters[it] = its;
    t_asm += c->last_ms[0];
    t_kry += c->last_ms[1];
  }
  int rc = pa_commit_increment(c, row);
  c->last_ms[0] = t_asm;
  c->last_ms[1] = t_kry;
  return rc;
}

int pa_last_residual(pa_context *c, double out[3]) {
  if (!c || !out) return fail(PA_ERR_ARG, "null argument");
  if (c->pbc) return fail(PA_ERR_UNSUPPORTED, "the true residual is not formed under periodic boundary conditions");
  if (!c->solved || c->asm_precision != c->precision) return fail(PA_ERR_STATE, "no linear solve on the current assembly");
  const unsigned g = nblk(c->m.nnode, VB);
  if (!c->resid_part) {
    int rc = dalloc(c, &c->resid_part, 3 * (size_t)g);
    if (rc) return rc;
  }
  if (mixed(c))
    hipLaunchKernelGGL(k_true_residual<float>, dim3(g), dim3(VB), 0, c->stream, c->m, (const float *)c->K, (const double *)c->kv[0],
                       (const double *)c->R, (const double *)c->dinv, c->resid_part);
  else
    hipLaunchKernelGGL(k_true_residual<double>, dim3(g), dim3(VB), 0, c->stream, c->m, (const double *)c->K, (const double *)c->kv[0],
                       (const double *)c->R, (const double *)c->dinv, c->resid_part);
  HIPCHK(hipGetLastError());
  std::vector<double> h(3 * (size_t)g);
  HIPCHK(hipMemcpyAsync(h.data(), c->resid_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double t[3] = {0, 0, 0};
  for (unsigned b = 0; b < g; ++b)
    for (int k = 0; k < 3; ++k) t[k] += h[3 * (size_t)b + k];
  const double bn = std::sqrt(t[2]);
  out[0] = bn > 0 ? std::sqrt(t[0]) / bn : 0.0;
  out[1] = bn > 0 ? std::sqrt(t[1]) / bn : 0.0;
  out[2] = c->last_krylov.dp / (bn > 0 ? bn : 1.0);
  return 0;
}

int pa_last_timings(pa_context *c, double ms[3]) {
  if (!c || !ms) return fail(PA_ERR_ARG, "null argument");
  for (int i = 0; i < 3; ++i) ms[i] = c->last_ms[i];
  return 0;
}

int pa_last_qp_pass_ms(pa_context *c, double *ms) {
  if (!c || !ms) return fail(PA_ERR_ARG, "null argument");
  if (!c->qev_set) return fail(PA_ERR_STATE, "no assembly has run");
  HIPCHK(hipEventSynchronize(c->qev[1]));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, c->qev[0], c->qev[1]));
  *ms = t;
  return 0;
}

int pa_last_counts(pa_context *c, long long counts[3]) {
  if (!c || !counts) return fail(PA_ERR_ARG, "null argument");
  unsigned h[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(h, c->counters + 8, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 3; ++i) counts[i] = h[i];
  return 0;
}

int pa_get_state(pa_context *c, int which, double *host, long long n) {
  if (!c || !host) return fail(PA_ERR_ARG, "null argument");
  const long long nq = c->nq;
  const double *src = nullptr;
  int comps = 0;
  bool node = false;
  switch (which) {
    case 0: src = c->T; comps = 9; break;
    case 1: src = c->Fp_conv; comps = 9; break;
    case 2: src = c->Fe_conv; comps = 9; break;
    case 3: src = c->s_conv; comps = c->ns; break;
    case 4: src = c->rotnew; comps = 3; break;
    case 5: src = c->slip_conv; comps = c->nslip; break;
    case 6: src = c->u; comps = 3; node = true; break;
    case 7: src = c->R; comps = 3; node = true; break;
    case 8: src = c->Fp_iter; comps = 9; break;
    case 9: src = c->Fe_iter; comps = 9; break;
    case 10: src = c->s_iter; comps = c->ns; break;
    case 11: src = c->W_conv; comps = c->ns; break;
    case 12: src = c->W_iter; comps = c->ns; break;
    case 13: src = c->tw_conv; comps = c->ntw; break;
    case 14:    // twinned (0/1), and 15: the orientation of the constitutive update (rot_conv)
    case 15: {
      if (!c->ntw) return fail(PA_ERR_STATE, "no twin state: twinning is off");
      const int k = which == 14 ? 1 : 3;
      if (n < k * nq) return fail(PA_ERR_ARG, "host buffer too small");
      if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
      std::vector<int> var(nq);
      std::vector<unsigned> eg(c->m.nelem);
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipMemcpy(var.data(), c->var_conv, nq * sizeof(int), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(eg.data(), c->d_egrain, c->m.nelem * sizeof(unsigned), hipMemcpyDeviceToHost));
      for (long long Q = 0; Q < nq; ++Q) {
        if (which == 14) { host[Q] = var[Q] != 0 ? 1.0 : 0.0; continue; }
        const double *r = &c->h_vrod[((size_t)eg[Q >> 3] * (1 + c->ntw) + var[Q]) * 3];
        for (int i = 0; i < 3; ++i) host[3 * Q + i] = r[i];
      }
      return 0;
    }
    default: return fail(PA_ERR_ARG, "unknown state id");
  }
  if (!src) {
    if (which == 11 || which == 12) return fail(PA_ERR_STATE, "no backstress state: kinematic hardening is off");
    if (which == 13) return fail(PA_ERR_STATE, "no twin state: twinning is off");
    return fail(PA_ERR_STATE, "state not allocated");
  }
  const long long cnt = node ? c->ndof : comps * nq;
  if (n < cnt) return fail(PA_ERR_ARG, "host buffer too small");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (node) {
    HIPCHK(hipMemcpy(host, src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<double> tmp(cnt);
  HIPCHK(hipMemcpy(tmp.data(), src, cnt * sizeof(double), hipMemcpyDeviceToHost));
  for (long long Q = 0; Q < nq; ++Q)
    for (int k = 0; k < comps; ++k) host[Q * comps + k] = tmp[k * nq + Q];
  return 0;
}

int pa_get_csr(pa_context *c, long long *row_ptr, int *col_idx, double *values) {
  if (!c || !row_ptr || !col_idx) return fail(PA_ERR_ARG, "null argument");
  long long *drp = nullptr;
  int *dci = nullptr;
  double *dval = nullptr;
  HIPCHK(hipMalloc(&drp, (c->ndof + 1) * sizeof(long long)));
  int rc = 0;
  if (hipMalloc(&dci, c->nnz * sizeof(int)) != hipSuccess) rc = fail(PA_ERR_HIP, "hipMalloc");
  if (!rc && values && hipMalloc(&dval, c->nnz * sizeof(double)) != hipSuccess) rc = fail(PA_ERR_HIP, "hipMalloc");
  if (!rc) {
    hipLaunchKernelGGL(k_csr_rowptr, dim3(nblk(c->ndof + 1, 256)), dim3(256), 0, c->stream, c->m, drp);
    if (mixed(c))
      hipLaunchKernelGGL(k_csr_cols<float>, dim3(nblk(c->ndof, 4)), dim3(81, 4), 0, c->stream, c->m, (const long long *)drp, dci,
                         (const float *)c->K, dval);
    else
      hipLaunchKernelGGL(k_csr_cols<double>, dim3(nblk(c->ndof, 4)), dim3(81, 4), 0, c->stream, c->m, (const long long *)drp, dci,
                         (const double *)c->K, dval);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(PA_ERR_HIP, "csr kernels");
  }
  if (!rc && hipMemcpy(row_ptr, drp, (c->ndof + 1) * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(PA_ERR_HIP, "copy");
  if (!rc && hipMemcpy(col_idx, dci, c->nnz * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(PA_ERR_HIP, "copy");
  if (!rc && values && hipMemcpy(values, dval, c->nnz * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(PA_ERR_HIP, "copy");
  hipFree(drp);
  hipFree(dci);
  if (dval) hipFree(dval);
  return rc;
}

int pa_qp_update(pa_context *c, long long n, const double *F, const double *Fp_conv, const double *s_conv, const unsigned *grain,
                 int flag, double *P, double *T, double *dPdF, double *Fe, double *Fp, double *s_out, double *slip_inc, int *n_active,
                 const double *W_conv, double *W_out) {
  if (c && c->nph > 1) return fail(PA_ERR_UNSUPPORTED, "the batch qp entry points take one phase");
  if (!c || n < 0 || !F || !Fp_conv || !s_conv || !grain || !P || !T || !Fe || !Fp || !s_out || !slip_inc || !n_active)
    return fail(PA_ERR_ARG, "null argument");
  if (flag && !dPdF) return fail(PA_ERR_ARG, "dPdF required for flag 1");
  if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
  for (long long i = 0; i < n; ++i)
    if ((int)grain[i] >= c->n_grain_rows) return fail(PA_ERR_ARG, "grain index out of range");
  if (n == 0) return 0;
  const int ns = c->ns;
  const size_t b9 = 9 * n * sizeof(double), bs = (size_t)ns * n * sizeof(double);
  double *dF, *dFp, *ds, *dP, *dT, *dD = nullptr, *dFe, *dFpo, *dso, *dx, *dW = nullptr;
  unsigned *dg;
  int *dna;
  std::vector<void *> tmp;
  auto al = [&](void **p, size_t b) { if (hipMalloc(p, b) != hipSuccess) return false; tmp.push_back(*p); return true; };
  bool ok = al((void **)&dF, b9) && al((void **)&dFp, b9) && al((void **)&ds, bs) && al((void **)&dP, b9) && al((void **)&dT, b9) &&
            al((void **)&dFe, b9) && al((void **)&dFpo, b9) && al((void **)&dso, bs) && al((void **)&dx, bs) &&
            al((void **)&dg, n * sizeof(unsigned)) && al((void **)&dna, n * sizeof(int)) && (!flag || al((void **)&dD, 81 * n * sizeof(double))) &&
            (!c->kin || al((void **)&dW, bs));
  int rc = 0;
  if (!ok) rc = fail(PA_ERR_HIP, "hipMalloc");
  // orientation-table rows: with twinning each grain has 1 + n_twin rows; a batch call uses the
  // grain's own orientation (row 0 of its block)
  std::vector<unsigned> rows(grain, grain + n);
  if (c->ntw)
    for (auto &r : rows) r *= (unsigned)(1 + c->ntw);
  if (!rc && (hipMemcpy(dF, F, b9, hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(dFp, Fp_conv, b9, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(ds, s_conv, bs, hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(dg, rows.data(), n * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess))
    rc = fail(PA_ERR_HIP, "copy inputs");
  if (!rc && c->kin &&
      (W_conv ? hipMemcpy(dW, W_conv, bs, hipMemcpyHostToDevice) : hipMemset(dW, 0, bs)) != hipSuccess)
    rc = fail(PA_ERR_HIP, "copy inputs");
  if (!rc) {
    const int NS = ns_template(ns);
    const int msi = c->mat.max_solver_iter;
    BatchArgs b{n, dF, dFp, ds, dg, c->d_gtab, c->gstride, c->d_mat, ns, msi, flag, dP, dT, dD, dFe, dFpo, dso, dx, dna, dW};
    if (c->ntw) b.nslip = c->nslip;
    if (NS == 12) launch_qp_batch<12>(c->stream, c->kin, b);
    else if (NS == 24) launch_qp_batch<24>(c->stream, c->kin, b);
    else launch_qp_batch<48>(c->stream, c->kin, b);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(PA_ERR_HIP, "qp kernel");
  }
  if (!rc && (hipMemcpy(P, dP, b9, hipMemcpyDeviceToHost) != hipSuccess || hipMemcpy(T, dT, b9, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(Fe, dFe, b9, hipMemcpyDeviceToHost) != hipSuccess || hipMemcpy(Fp, dFpo, b9, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(s_out, dso, bs, hipMemcpyDeviceToHost) != hipSuccess || hipMemcpy(slip_inc, dx, bs, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(n_active, dna, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
              (flag && hipMemcpy(dPdF, dD, 81 * n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
              (c->kin && W_out && hipMemcpy(W_out, dW, bs, hipMemcpyDeviceToHost) != hipSuccess)))
    rc = fail(PA_ERR_HIP, "copy outputs");
  for (void *p : tmp) hipFree(p);
  return rc;
}

// ------------------------------------------------------------ warm-start network
struct pa_mlp {
  float *blob = nullptr, *m1 = nullptr, *m2 = nullptr, *part = nullptr, *lpart = nullptr, *grad = nullptr;
  unsigned short *wt = nullptr, *w = nullptr;
  double *lacc = nullptr;
  long long part_waves = 0;
  long long adam_steps = 0;
  hipStream_t stream = nullptr;
};

namespace {
mlp::Weights mlp_weights(const pa_mlp *m) { return mlp::Weights{m->wt, m->w, m->blob}; }

int mlp_part(pa_mlp *m, long long waves) {
  if (waves <= m->part_waves) return 0;
  if (m->part) hipFree(m->part);
  if (m->lpart) hipFree(m->lpart);
  m->part = nullptr; m->lpart = nullptr; m->part_waves = 0;
  HIPCHK(hipMalloc(&m->part, (size_t)waves * mlp::NPARAM * sizeof(float)));
  HIPCHK(hipMalloc(&m->lpart, (size_t)waves * sizeof(float)));
  m->part_waves = waves;
  return 0;
}

// inference on device buffers, on stream st
int mlp_predict_dev(const pa_mlp *m, hipStream_t st, long long n, const float *X, float *Y) {
  if (n == 0) return 0;
  const long long tiles = (n + 15) / 16;
  const unsigned grid = (unsigned)std::min<long long>((tiles + 3) / 4, 8192);
  hipLaunchKernelGGL(mlp::k_mlp_predict, dim3(grid), dim3(256), 0, st, mlp_weights(m), n, X, Y);
  HIPCHK(hipGetLastError());
  return 0;
}

// gradient of one batch of `count` samples (indices (mul*(base+k)+add) mod n); mode 0: into m->grad,
// mode 1: one Adam step at lr_t
int mlp_step(pa_mlp *m, long long n, const float *X, const float *Y, long long base, long long count, long long mul,
             long long add, int mode, float lr_t) {
  const long long blocks = (count + 63) / 64;
  if (int rc = mlp_part(m, blocks * 4)) return rc;
  hipLaunchKernelGGL(mlp::k_mlp_grad, dim3((unsigned)blocks), dim3(256), 0, m->stream, mlp_weights(m), n, X, Y, base, count, mul, add,
                     1.0f / (float)(count * mlp::NOUT), m->part, m->lpart);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(mlp::k_mlp_reduce, dim3(nblk(mlp::NPARAM, 256)), dim3(256), 0, m->stream, mode, (int)(blocks * 4),
                     (const float *)m->part, (const float *)m->lpart, m->grad, m->blob, m->m1, m->m2, m->wt, m->w, lr_t, 0.9f,
                     0.999f, 1e-8f, m->lacc);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace

int pa_mlp_create(pa_mlp **out) {
  if (!out) return fail(PA_ERR_ARG, "null argument");
  *out = nullptr;
  pa_mlp *m = new pa_mlp();
  auto bad = [&](const char *what) { pa_mlp_destroy(m); return fail(PA_ERR_HIP, what); };
  if (hipStreamCreate(&m->stream) != hipSuccess) return bad("hipStreamCreate");
  if (hipMalloc(&m->blob, mlp::NBLOB * sizeof(float)) != hipSuccess || hipMalloc(&m->m1, mlp::NPARAM * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->m2, mlp::NPARAM * sizeof(float)) != hipSuccess || hipMalloc(&m->grad, mlp::NPARAM * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->wt, mlp::NWBF * sizeof(unsigned short)) != hipSuccess ||
      hipMalloc(&m->w, mlp::NWBF * sizeof(unsigned short)) != hipSuccess || hipMalloc(&m->lacc, sizeof(double)) != hipSuccess)
    return bad("hipMalloc");
  if (hipMemset(m->blob, 0, mlp::NBLOB * sizeof(float)) != hipSuccess || hipMemset(m->m1, 0, mlp::NPARAM * sizeof(float)) != hipSuccess ||
      hipMemset(m->m2, 0, mlp::NPARAM * sizeof(float)) != hipSuccess || hipMemset(m->wt, 0, mlp::NWBF * 2) != hipSuccess ||
      hipMemset(m->w, 0, mlp::NWBF * 2) != hipSuccess)
    return bad("hipMemset");
  *out = m;
  return 0;
}

int pa_mlp_destroy(pa_mlp *m) {
  if (!m) return 0;
  for (void *p : {(void *)m->blob, (void *)m->m1, (void *)m->m2, (void *)m->part, (void *)m->lpart, (void *)m->grad, (void *)m->wt,
                  (void *)m->w, (void *)m->lacc})
    if (p) hipFree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int pa_mlp_set_params(pa_mlp *m, const float *blob) {
  if (!m || !blob) return fail(PA_ERR_ARG, "null argument");
  for (int i = 0; i < mlp::NBLOB; ++i)
    if (!std::isfinite(blob[i])) return fail(PA_ERR_ARG, "non-finite parameter");
  for (int i = 0; i < mlp::NOUT; ++i)
    if (blob[mlp::OFF_OSCALE + i] == 0.0f) return fail(PA_ERR_ARG, "zero output scale");
  HIPCHK(hipMemcpy(m->blob, blob, mlp::NBLOB * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemsetAsync(m->m1, 0, mlp::NPARAM * sizeof(float), m->stream));
  HIPCHK(hipMemsetAsync(m->m2, 0, mlp::NPARAM * sizeof(float), m->stream));
  m->adam_steps = 0;
  hipLaunchKernelGGL(mlp::k_mlp_cast, dim3(nblk(mlp::NPARAM, 256)), dim3(256), 0, m->stream, (const float *)m->blob, m->wt, m->w);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int pa_mlp_get_params(pa_mlp *m, float *blob) {
  if (!m || !blob) return fail(PA_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(blob, m->blob, mlp::NBLOB * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C++" {
namespace {
struct DevBuf {
  std::vector<void *> p;
  ~DevBuf() { for (void *q : p) (void)hipFree(q); }
  template <typename T> bool get(T **q, size_t bytes) {
    if (hipMalloc((void **)q, bytes ? bytes : 16) != hipSuccess) return false;
    p.push_back(*q);
    return true;
  }
};
}  // namespace
}  // extern "C++"

int pa_mlp_predict(pa_mlp *m, long long n, const float *X, float *Y) {
  if (!m || n < 0 || (n && (!X || !Y))) return fail(PA_ERR_ARG, "bad argument");
  if (n == 0) return 0;
  DevBuf b;
  float *dX, *dY;
  if (!b.get(&dX, n * mlp::NIN * sizeof(float)) || !b.get(&dY, n * mlp::NOUT * sizeof(float))) return fail(PA_ERR_HIP, "hipMalloc");
  HIPCHK(hipMemcpy(dX, X, n * mlp::NIN * sizeof(float), hipMemcpyHostToDevice));
  if (int rc = mlp_predict_dev(m, m->stream, n, dX, dY)) return rc;
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(Y, dY, n * mlp::NOUT * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pa_mlp_grad(pa_mlp *m, long long n, const float *X, const float *Y, float *grad, double *loss) {
  if (!m || n < 1 || n > 65536 || !X || !Y || !grad) return fail(PA_ERR_ARG, "bad argument (1 <= n <= 65536)");
  DevBuf b;
  float *dX, *dY;
  if (!b.get(&dX, n * mlp::NIN * sizeof(float)) || !b.get(&dY, n * mlp::NOUT * sizeof(float))) return fail(PA_ERR_HIP, "hipMalloc");
  HIPCHK(hipMemcpy(dX, X, n * mlp::NIN * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dY, Y, n * mlp::NOUT * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemsetAsync(m->lacc, 0, sizeof(double), m->stream));
  if (int rc = mlp_step(m, n, dX, dY, 0, n, 1, 0, 0, 0.0f)) return rc;
  double l = 0;
  HIPCHK(hipMemcpyAsync(grad, m->grad, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipMemcpyAsync(&l, m->lacc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return 0;
}

int pa_mlp_train(pa_mlp *m, long long n, const float *X, const float *Y, int epochs, int batch, float lr,
                 unsigned long long seed, double *epoch_loss) {
  if (!m || n < 1 || !X || !Y || epochs < 0 || batch < 1 || batch > 65536 || !(lr > 0.0f))
    return fail(PA_ERR_ARG, "bad argument (n >= 1, 1 <= batch <= 65536, lr > 0)");
  DevBuf b;
  float *dX, *dY;
  if (!b.get(&dX, n * mlp::NIN * sizeof(float)) || !b.get(&dY, n * mlp::NOUT * sizeof(float))) return fail(PA_ERR_HIP, "hipMalloc");
  HIPCHK(hipMemcpy(dX, X, n * mlp::NIN * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dY, Y, n * mlp::NOUT * sizeof(float), hipMemcpyHostToDevice));
  const long long steps_per_epoch = (n + batch - 1) / batch;
  const double total = (double)steps_per_epoch * epochs;
  unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto rnd = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
  auto gcd = [](long long a, long long c) { while (c) { const long long t = a % c; a = c; c = t; } return a; };
  long long k = 0;
  for (int e = 0; e < epochs; ++e) {
    // the epoch visits the samples in the order (mul * i + add) mod n, gcd(mul, n) = 1
    long long mul = n > 1 ? (long long)(rnd() % (unsigned long long)n) : 1, add = n > 1 ? (long long)(rnd() % (unsigned long long)n) : 0;
    while (n > 1 && (mul == 0 || gcd(mul, n) != 1)) mul = (mul + 1) % n;
    if (n == 1) mul = 1;
    HIPCHK(hipMemsetAsync(m->lacc, 0, sizeof(double), m->stream));
    for (long long s = 0; s < steps_per_epoch; ++s, ++k) {
      const long long base = s * batch, count = std::min<long long>(batch, n - base);
      // cosine decay from lr to lr/20 over the run, Adam bias corrections folded in
      const double lr_k = lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * k / std::max(1.0, total))));
      m->adam_steps++;
      const double t = (double)m->adam_steps;
      const float lr_t = (float)(lr_k * std::sqrt(1.0 - std::pow(0.999, t)) / (1.0 - std::pow(0.9, t)));
      if (int rc = mlp_step(m, n, dX, dY, base, count, mul, add, 1, lr_t)) return rc;
    }
    double l = 0;
    HIPCHK(hipMemcpyAsync(&l, m->lacc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (epoch_loss) epoch_loss[e] = l / (double)steps_per_epoch;
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int pa_qp_update_warm(pa_context *c, pa_mlp *net, long long n, const double *F, const double *Fp_conv, const double *Fe_conv,
                      const double *s_conv, const unsigned *grain, int flag, double *P, double *T, double *dPdF, double *Fe,
                      double *Fp, double *s_out, double *slip_inc, int *n_active, int *iters, float *feat, float *x0) {
  if (c && c->nph > 1) return fail(PA_ERR_UNSUPPORTED, "the batch qp entry points take one phase");
  if (!c || n < 0 || !F || !Fp_conv || !Fe_conv || !s_conv || !grain || !P || !T || !Fe || !Fp || !s_out || !slip_inc || !n_active || !iters)
    return fail(PA_ERR_ARG, "null argument");
  if (flag && !dPdF) return fail(PA_ERR_ARG, "dPdF required for flag 1");
  if (c->ns != mlp::NOUT) return fail(PA_ERR_UNSUPPORTED, "the warm start covers 12 slip systems (FCC)");
  if (c->kin) return fail(PA_ERR_UNSUPPORTED, "the warm start does not cover kinematic hardening");
  if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
  for (long long i = 0; i < n; ++i)
    if ((int)grain[i] >= c->n_grain_rows) return fail(PA_ERR_ARG, "grain index out of range");
  if (n == 0) return 0;
  const int ns = c->ns;
  const size_t b9 = 9 * n * sizeof(double), bs = (size_t)ns * n * sizeof(double);
  DevBuf b;
  double *dF, *dFp, *dFec, *ds, *dP, *dT, *dD = nullptr, *dFe, *dFpo, *dso, *dx;
  unsigned *dg;
  int *dna, *dit;
  float *dfeat, *dx0 = nullptr;
  if (!(b.get(&dF, b9) && b.get(&dFp, b9) && b.get(&dFec, b9) && b.get(&ds, bs) && b.get(&dP, b9) && b.get(&dT, b9) &&
        b.get(&dFe, b9) && b.get(&dFpo, b9) && b.get(&dso, bs) && b.get(&dx, bs) && b.get(&dg, n * sizeof(unsigned)) &&
        b.get(&dna, n * sizeof(int)) && b.get(&dit, n * sizeof(int)) && b.get(&dfeat, n * mlp::NIN * sizeof(float)) &&
        (!flag || b.get(&dD, 81 * n * sizeof(double))) && (!net || b.get(&dx0, n * mlp::NOUT * sizeof(float)))))
    return fail(PA_ERR_HIP, "hipMalloc");
  HIPCHK(hipMemcpy(dF, F, b9, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dFp, Fp_conv, b9, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dFec, Fe_conv, b9, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ds, s_conv, bs, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dg, grain, n * sizeof(unsigned), hipMemcpyHostToDevice));
  const unsigned grid = nblk(n, 64);
  hipLaunchKernelGGL(k_qp_features<12>, dim3(grid), dim3(64), 0, c->stream, n, (const double *)dF, (const double *)dFp,
                     (const double *)dFec, (const double *)ds, (const double *)nullptr, (const unsigned *)dg, (const double *)c->d_gtab,
                     c->gstride, ns, dfeat);
  HIPCHK(hipGetLastError());
  if (net) {
    HIPCHK(hipStreamSynchronize(net->stream));  // weights settled
    if (int rc = mlp_predict_dev(net, c->stream, n, dfeat, dx0)) return rc;
  }
  hipLaunchKernelGGL(k_qp_batch_warm<12>, dim3(grid), dim3(64), 0, c->stream, n, (const double *)dF, (const double *)dFp,
                     (const double *)ds, (const unsigned *)dg, (const double *)c->d_gtab, c->gstride, (const double *)c->d_mat, ns,
                     c->mat.max_solver_iter, flag, dP, dT, dD, dFe, dFpo, dso, dx, dna, (const float *)dx0, dit);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(P, dP, b9, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(T, dT, b9, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Fe, dFe, b9, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Fp, dFpo, b9, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(s_out, dso, bs, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(slip_inc, dx, bs, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(n_active, dna, n * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(iters, dit, n * sizeof(int), hipMemcpyDeviceToHost));
  if (flag) HIPCHK(hipMemcpy(dPdF, dD, 81 * n * sizeof(double), hipMemcpyDeviceToHost));
  if (feat) HIPCHK(hipMemcpy(feat, dfeat, n * mlp::NIN * sizeof(float), hipMemcpyDeviceToHost));
  if (x0 && net) HIPCHK(hipMemcpy(x0, dx0, n * mlp::NOUT * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pa_project_fields(pa_context *c, double *fields, int *iters) {
  if (!c || !fields) return fail(PA_ERR_ARG, "null argument");
  if (!c->grains_set) return fail(PA_ERR_STATE, "grains not set");
  const long long nq = c->nq, nn = c->m.nnode;
  // per-qp values into the staging of the assembly's P (9 per qp, 3 used)
  hipLaunchKernelGGL(k_proj_values, dim3(nblk(nq, 256)), dim3(256), 0, c->stream, c->m, c->h[0], c->h[1], c->h[2], nq,
                     (const double *)c->u, (const double *)c->T, (const int *)c->var_iter, c->Pq);
  hipLaunchKernelGGL(k_mass_blocks, dim3(nblk(nn, 256)), dim3(256), 0, c->stream, c->m, c->JxW, (double *)c->K);
  hipLaunchKernelGGL(k_dinv<double>, dim3(nblk(c->ndof, 256)), dim3(256), 0, c->stream, c->m, (const double *)c->K, c->dinv);
  HIPCHK(hipGetLastError());
  // solveLinearSystem2 (solveLinearSystem.cc:43-65): Jacobi-PCG in FP64, one field at a time, the
  // field in component 0 of every node (the other components stay zero)
  const int method = c->method, precision = c->precision;
  const bool pbc = c->pbc;
  c->method = PA_SOLVER_CG; c->precision = PA_PREC_FP64; c->pbc = false;
  c->asm_precision = -1;  // K now holds the mass matrix: the next Newton step assembles again
  c->solved = false;
  std::vector<double> x(3 * nn);
  int rc = 0;
  for (int f = 0; f < 3 && !rc; ++f) {
    hipLaunchKernelGGL(k_proj_rhs, dim3(nblk(nn, 256)), dim3(256), 0, c->stream, c->m, c->JxW, (const double *)(c->Pq + f * nq), c->R);
    KState hs;
    rc = krylov_run(c, hs);
    if (rc) break;
    if (iters) iters[f] = hs.it;
    if (hipMemcpyAsync(x.data(), c->kv[0], sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = fail(PA_ERR_HIP, "copy projected field");
      break;
    }
    for (long long n = 0; n < nn; ++n) fields[f * nn + n] = x[3 * n];
  }
  c->method = method; c->precision = precision; c->pbc = pbc;
  return rc;
}

int pa_grain_averages(pa_context *c, int n_grains, double *out) {
  if (!c || n_grains < 1 || n_grains > (1 << 20) || !out) return fail(PA_ERR_ARG, "bad argument (1 <= n_grains <= 2^20)");
  if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
  DevBuf b;
  double *d;
  if (!b.get(&d, (size_t)n_grains * 7 * sizeof(double))) return fail(PA_ERR_HIP, "hipMalloc");
  hipLaunchKernelGGL(k_grain_sums, dim3((unsigned)n_grains), dim3(256), 0, c->stream, (long long)c->m.nelem, (long long)c->nq,
                     (const unsigned *)c->d_egrain, (const double *)c->T, c->JxW, d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(out, d, (size_t)n_grains * 7 * sizeof(double), hipMemcpyDeviceToHost));
  for (int g = 0; g < n_grains; ++g)  // updateAfterIncrement.cc:676-682
    if (std::fabs(out[7 * g]) > 1e-15)
      for (int j = 1; j < 7; ++j) out[7 * g + j] = out[7 * g + j] / out[7 * g];
  return 0;
}

int pa_pinv_solve(int path, int n, long long count, const double *A, const double *b, double *x) {
  if (path < 0 || path > 2 || n < 1 || n > 48 || count < 0 || count > (1ll << 20) || (count && (!A || !b || !x)))
    return fail(PA_ERR_ARG, "bad argument");
  if (count == 0) return 0;
  const int NS = ns_template(n);
  const unsigned cnt = (unsigned)count;
  std::vector<double> hA((size_t)cnt * NS * NS, 0.0), hb((size_t)cnt * NS, 0.0), hx((size_t)cnt * NS, 0.0);
  for (size_t e = 0; e < cnt; ++e)
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) hA[(e * NS + i) * NS + j] = A[(e * n + i) * n + j];
      hb[e * NS + i] = b[e * n + i];
    }
  std::vector<void *> tmp;
  auto al = [&](void **p, size_t bytes) { if (hipMalloc(p, bytes) != hipSuccess) return false; tmp.push_back(*p); return true; };
  SingStore st{};
  unsigned *dcount = nullptr;
  int rc = 0;
  if (!(al((void **)&st.A, hA.size() * sizeof(double)) && al((void **)&st.b, hb.size() * sizeof(double)) &&
        al((void **)&st.X, hx.size() * sizeof(double)) && al((void **)&st.n, cnt * sizeof(int)) &&
        al((void **)&st.have, cnt * sizeof(int)) && al((void **)&st.done, cnt * sizeof(int)) && al((void **)&dcount, sizeof(unsigned))))
    rc = fail(PA_ERR_HIP, "hipMalloc");
  st.cap = cnt;
  const std::vector<int> hn(cnt, n);
  if (!rc && (hipMemcpy(st.A, hA.data(), hA.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(st.b, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(st.X, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(st.n, hn.data(), cnt * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(dcount, &cnt, sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemset(st.have, 0, cnt * sizeof(int)) != hipSuccess || hipMemset(st.done, 0, cnt * sizeof(int)) != hipSuccess))
    rc = fail(PA_ERR_HIP, "copy inputs");
  if (!rc) {
    if (path == 0) {
      if (NS == 12) hipLaunchKernelGGL(k_pinv_lane<12>, dim3(nblk(cnt, 64)), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
      else if (NS == 24) hipLaunchKernelGGL(k_pinv_lane<24>, dim3(nblk(cnt, 64)), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
      else hipLaunchKernelGGL(k_pinv_lane<48>, dim3(nblk(cnt, 64)), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
    } else if (path == 1) {
      if (NS == 12) hipLaunchKernelGGL(k_sing_svd<12>, dim3(SVD_GRID), dim3(sing_per_block<12>() * sing_group<12>()), 0, 0, st, (const unsigned *)dcount);
      else if (NS == 24) hipLaunchKernelGGL(k_sing_svd<24>, dim3(SVD_GRID), dim3(sing_per_block<24>() * sing_group<24>()), 0, 0, st, (const unsigned *)dcount);
      else hipLaunchKernelGGL(k_sing_svd<48>, dim3(SVD_GRID), dim3(sing_per_block<48>() * sing_group<48>()), 0, 0, st, (const unsigned *)dcount);
    } else {  // the cooperative qp kernels' routine, one group of lanes per system
      if (NS == 12) hipLaunchKernelGGL(k_pinv_group<12>, dim3(nblk(cnt, 64 / coop::group_size<12>())), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
      else if (NS == 24) hipLaunchKernelGGL(k_pinv_group<24>, dim3(nblk(cnt, 64 / coop::group_size<24>())), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
      else hipLaunchKernelGGL(k_pinv_group<48>, dim3(nblk(cnt, 64 / coop::group_size<48>())), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
    }
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = fail(PA_ERR_HIP, "pinv kernel");
  }
  if (!rc && hipMemcpy(hx.data(), st.X, hx.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(PA_ERR_HIP, "copy outputs");
  if (!rc)
    for (size_t e = 0; e < cnt; ++e)
      for (int i = 0; i < n; ++i) x[e * n + i] = hx[e * NS + i];
  for (void *p : tmp) hipFree(p);
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------ pa_pinv_solve without allocation
// Device memory of pa_pinv_solve_noalloc: one array per GPU, held until the process ends.  Host
// code never passes `scratch` itself to a kernel or a copy (in host code the name is a host-side
// stand-in, not a device address); it uses the address hipGetSymbolAddress returns for the
// current device, kept per device in g_scratch_dev.
static __device__ __align__(256) unsigned char scratch[PA_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
unsigned char *g_scratch_dev[kMaxDevices] = {};

// the current device's copy of `scratch` (looked up on the first call on each device)
int scratch_on_device(unsigned char **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return fail(PA_ERR_HIP, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return fail(PA_ERR_HIP, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = (unsigned char *)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

// Partition of the scratch for `count` systems of size n (leading dimension NS): compact staging
// of A, b, x as the caller lays them out, the padded store the pinv kernels read, and the store's
// per-system words; every segment starts on a 256-byte boundary.
struct NoallocLayout {
  size_t Ac, bc, xc, Ap, bp, Xp, nn, have, done, cnt, end;
};
NoallocLayout noalloc_layout(int n, int NS, size_t count) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  NoallocLayout L;
  size_t o = 0;
  L.Ac = o; o += up(count * n * n * sizeof(double));
  L.bc = o; o += up(count * n * sizeof(double));
  L.xc = o; o += up(count * n * sizeof(double));
  L.Ap = o; o += up(count * NS * NS * sizeof(double));
  L.bp = o; o += up(count * NS * sizeof(double));
  L.Xp = o; o += up(count * SING_ROUNDS * NS * sizeof(double));
  L.nn = o; o += up(count * sizeof(int));
  L.have = o; o += up(count * sizeof(int));
  L.done = o; o += up(count * sizeof(int));
  L.cnt = o; o += up(sizeof(unsigned));
  L.end = o;
  return L;
}

// compact (count x n x n, count x n) -> padded store (leading dimension NS), per-system words
// cleared (the array keeps the previous call's contents), queue length = count
__global__ void __launch_bounds__(256) k_noalloc_pack(unsigned count, int n, int NS, const double *Ac, const double *bc,
                                                      double *Ap, double *bp, double *Xp, int *nn, int *have, int *done,
                                                      unsigned *cnt) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)NS * NS;
  if (t < count * per) {
    const size_t e = t / per;
    const int i = (int)((t % per) / NS), j = (int)(t % NS);
    Ap[t] = (i < n && j < n) ? Ac[(e * n + i) * n + j] : 0.0;
  }
  if (t < (size_t)count * NS) {
    const size_t e = t / NS;
    const int i = (int)(t % NS);
    bp[t] = i < n ? bc[e * n + i] : 0.0;
  }
  if (t < (size_t)count * SING_ROUNDS * NS) Xp[t] = 0.0;
  if (t < count) { nn[t] = n; have[t] = 0; done[t] = 0; }
  if (t == 0) *cnt = count;
}

__global__ void __launch_bounds__(256) k_noalloc_unpack(unsigned count, int n, int NS, const double *Xp, double *xc) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)count * n) return;
  const size_t e = t / n;
  xc[t] = Xp[e * SING_ROUNDS * NS + t % n];
}
}  // namespace

extern "C" {

int pa_pinv_solve_noalloc(int path, int n, long long count, const double *A, const double *b, double *x) {
  if (path < 0 || path > 2 || n < 1 || n > 48 || count < 0 || (count && (!A || !b || !x)))
    return fail(PA_ERR_ARG, "bad argument");
  if (count == 0) return 0;
  const int NS = ns_template(n);
  if (count > (1ll << 20) || noalloc_layout(n, NS, (size_t)count).end > (size_t)PA_NOALLOC_SCRATCH_BYTES) {
    long long cap = 0;
    while (noalloc_layout(n, NS, (size_t)cap + 1).end <= (size_t)PA_NOALLOC_SCRATCH_BYTES) ++cap;
    return fail(PA_ERR_ARG, "pa_pinv_solve_noalloc: " + std::to_string(count) + " systems of size " + std::to_string(n) +
                                " exceed the scratch array (" + std::to_string(PA_NOALLOC_SCRATCH_BYTES) + " bytes: at most " +
                                std::to_string(cap) + " systems of this size)");
  }
  unsigned char *base = nullptr;
  if (int rc = scratch_on_device(&base)) return rc;
  const unsigned cnt = (unsigned)count;
  const NoallocLayout L = noalloc_layout(n, NS, cnt);
  double *Ac = (double *)(base + L.Ac), *bc = (double *)(base + L.bc), *xc = (double *)(base + L.xc);
  SingStore st{};
  st.cap = cnt;
  st.A = (double *)(base + L.Ap);
  st.b = (double *)(base + L.bp);
  st.X = (double *)(base + L.Xp);
  st.n = (int *)(base + L.nn);
  st.have = (int *)(base + L.have);
  st.done = (int *)(base + L.done);
  unsigned *dcount = (unsigned *)(base + L.cnt);
  if (hipMemcpy(Ac, A, (size_t)cnt * n * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(bc, b, (size_t)cnt * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(PA_ERR_HIP, "copy inputs");
  const size_t npack = std::max((size_t)cnt * NS * NS, (size_t)cnt * SING_ROUNDS * NS);
  hipLaunchKernelGGL(k_noalloc_pack, dim3((unsigned)((npack + 255) / 256)), dim3(256), 0, 0, cnt, n, NS, (const double *)Ac,
                     (const double *)bc, st.A, st.b, st.X, st.n, st.have, st.done, dcount);
  if (path == 0) {
    if (NS == 12) hipLaunchKernelGGL(k_pinv_lane<12>, dim3(nblk(cnt, 64)), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
    else if (NS == 24) hipLaunchKernelGGL(k_pinv_lane<24>, dim3(nblk(cnt, 64)), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
    else hipLaunchKernelGGL(k_pinv_lane<48>, dim3(nblk(cnt, 64)), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
  } else if (path == 1) {
    if (NS == 12) hipLaunchKernelGGL(k_sing_svd<12>, dim3(SVD_GRID), dim3(sing_per_block<12>() * sing_group<12>()), 0, 0, st, (const unsigned *)dcount);
    else if (NS == 24) hipLaunchKernelGGL(k_sing_svd<24>, dim3(SVD_GRID), dim3(sing_per_block<24>() * sing_group<24>()), 0, 0, st, (const unsigned *)dcount);
    else hipLaunchKernelGGL(k_sing_svd<48>, dim3(SVD_GRID), dim3(sing_per_block<48>() * sing_group<48>()), 0, 0, st, (const unsigned *)dcount);
  } else {
    if (NS == 12) hipLaunchKernelGGL(k_pinv_group<12>, dim3(nblk(cnt, 64 / coop::group_size<12>())), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
    else if (NS == 24) hipLaunchKernelGGL(k_pinv_group<24>, dim3(nblk(cnt, 64 / coop::group_size<24>())), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
    else hipLaunchKernelGGL(k_pinv_group<48>, dim3(nblk(cnt, 64 / coop::group_size<48>())), dim3(64), 0, 0, cnt, n, st.A, st.b, st.X);
  }
  hipLaunchKernelGGL(k_noalloc_unpack, dim3((unsigned)(((size_t)cnt * n + 255) / 256)), dim3(256), 0, 0, cnt, n, NS,
                     (const double *)st.X, xc);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return fail(PA_ERR_HIP, "pinv kernel");
  if (hipMemcpy(x, xc, (size_t)cnt * n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return fail(PA_ERR_HIP, "copy outputs");
  return 0;
}

int pa_time_kernel(pa_context *c, int which, int reps, double *avg_ms) {
  if (!c || !avg_ms || reps < 1) return fail(PA_ERR_ARG, "bad argument");
  if (!c->grains_set) return fail(PA_ERR_STATE, "pa_set_grains has not been called");
  if (int rc = check_qp_form(c)) return rc;
  const unsigned gn = nblk(c->m.nnode, VB), gu = ugrid(c->ndof);
  double *partA = c->partials, *partU = c->partials + spmv_partials_offset(c);
  // solver state that never ends (tol < 0, no iteration limit) with unit r.z history, and update
  // partials summing to z.z = r.z = 1, so the operator kernels run in full on the iteration >= 1
  // path (b = 1); operator partials summing to 1 for the update kernel
  KState ks;
  std::memset(&ks, 0, sizeof(ks));
  ks.tol = -1.0;
  ks.maxit = 0x7fffffff;
  ks.rzh[0] = ks.rzh[1] = 1.0;
  static const double ones[2] = {1.0, 1.0};
  HIPCHK(hipMemcpyAsync(c->kst, &ks, sizeof(ks), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(partU, 0, sizeof(double) * 2 * (size_t)gu, c->stream));
  HIPCHK(hipMemsetAsync(partA, 0, sizeof(double) * (size_t)gn, c->stream));
  HIPCHK(hipMemcpyAsync(partU, ones, sizeof(ones), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(partA, ones, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->ev[0], c->stream));
  for (int r = 0; r < reps; ++r) {
    int rc = 0;
    if (which == 0) {
      if (mixed(c))
        hipLaunchKernelGGL(k_bicg_spmv<float>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const float *)c->K, (const double *)c->kv[5],
                           (const double *)c->kv[6], c->kv[3], c->kv[4], (const KState *)c->kst, partA);
      else
        hipLaunchKernelGGL(k_bicg_spmv<double>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const double *)c->K, (const double *)c->kv[5],
                           (const double *)c->kv[6], c->kv[3], c->kv[4], (const KState *)c->kst, partA);
    } else if (which == 1) {
      rc = launch_assemble(c, 1);
    } else if (which == 2) {
      if (mixed(c))
        hipLaunchKernelGGL(k_cg_spmv<float>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const float *)c->K, (const double *)c->kv[5],
                           c->kv[4], (const KState *)c->kst, partA);
      else
        hipLaunchKernelGGL(k_cg_spmv<double>, dim3(gn), dim3(VB), 0, c->stream, c->m, (const double *)c->K, (const double *)c->kv[5],
                           c->kv[4], (const KState *)c->kst, partA);
    } else if (which == 3) {
      rc = launch_qp_tangent(c);
    } else if (which == 4) {
      rc = launch_elements(c, 1);
    } else if (which == 6) {
      hipLaunchKernelGGL(k_cg_p, dim3(gu), dim3(VB), 0, c->stream, c->ndof, (const double *)c->kv[3], c->kv[5], c->kst,
                         (const double *)partU, gu, 1);
    } else if (which == 7) {
      rc = launch_qp_tangent(c, true);
    } else if (which == 5) {
      hipLaunchKernelGGL(k_cg_update, dim3(gu), dim3(VB), 0, c->stream, c->ndof, c->kv[0], (const double *)c->kv[5], c->kv[1],
                         (const double *)c->kv[4], c->kv[3], (const double *)c->dinv, (const KState *)c->kst, (const double *)partA, gn,
                         partU, 1);
    } else {
      return fail(PA_ERR_ARG, "unknown kernel id");
    }
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[1], c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  *avg_ms = ms / reps;
  return 0;
}

}  // extern "C"
#include "cp_batch.h"

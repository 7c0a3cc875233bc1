// Test infrastructure: runs the wavefront-cooperative quadrature-point code of the HIP library
// (plasticity_amd/csrc/cp_coop.h, cp_passes.h) on the CPU, one std::thread per lane (the host
// stand-ins for the lane intrinsics are in tests/emu/hip/hip_runtime.h), so that its results can
// be compared with the oracle and its memory accesses checked (built with AddressSanitizer; every
// buffer, the LDS slices included, is an allocation of exactly its size).  C ABI for ctypes.
#include <hip/hip_runtime.h>

#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "../../include/plasticity_amd.h"
#include "../../plasticity_amd/csrc/cp_qp.h"
#include "../../plasticity_amd/csrc/cp_coop.h"
#include "../../plasticity_amd/csrc/cp_tables.h"

thread_local EmuDim3 threadIdx, blockIdx, blockDim, gridDim;
thread_local emu::Wave *emu::wave = nullptr;

namespace {
#include "../../plasticity_amd/csrc/cp_passes.h"

int ns_template(int ns) { return ns <= 12 ? 12 : (ns <= 24 ? 24 : 48); }

// material block in the layout of cp::MatLayout<NS> (as pa_create builds it)
std::vector<double> material_block(const pa_material &mat, int NS) {
  const int ns = mat.n_slip, tol = 3 * NS + NS * NS, c1 = tol + 1, c2 = c1 + NS;
  std::vector<double> mb(c2 + NS, 0.0);
  for (int a = 0; a < ns; ++a) {
    mb[a] = mat.h0[a]; mb[NS + a] = mat.ss[a]; mb[2 * NS + a] = mat.pw[a];
    for (int b = 0; b < ns; ++b) mb[3 * NS + a * NS + b] = mat.q[a][b];
    if (mat.kinematic_hardening) { mb[c1 + a] = mat.C1[a]; mb[c2 + a] = mat.C2[a]; }
  }
  mb[tol] = mat.stress_tol;
  return mb;
}

// per-grain tables (as pa_set_grains builds them)
std::vector<double> grain_tables(const pa_material &mat, const double *rod, int nrows, int NS, int &gstride) {
  const int ns = mat.n_slip;
  gstride = 9 * NS + 36 + 81 + 9;
  std::vector<double> mv(3 * ns), nv(3 * ns);
  for (int a = 0; a < ns; ++a) {
    double mn = 0, nn = 0;
    for (int j = 0; j < 3; ++j) { mn += mat.m[a][j] * mat.m[a][j]; nn += mat.n[a][j] * mat.n[a][j]; }
    mn = std::sqrt(mn); nn = std::sqrt(nn);
    for (int j = 0; j < 3; ++j) { mv[3 * a + j] = mat.m[a][j] / mn; nv[3 * a + j] = mat.n[a][j] / nn; }
  }
  std::vector<double> tab((size_t)nrows * gstride, 0.0);
  for (int g = 0; g < nrows; ++g)
    grain_table(mat, (const double(*)[3])mv.data(), (const double(*)[3])nv.data(), rod + 3 * g, ns, NS, &tab[(size_t)g * gstride]);
  return tab;
}

// Runs `body(lds)` as `grid` one-wavefront blocks (at most `concurrent` at a time), each block
// with its own LDS allocation of `lds_doubles`.
void launch(int G, unsigned grid, size_t lds_doubles, unsigned concurrent, const std::function<void(double *)> &body) {
  for (unsigned b0 = 0; b0 < grid; b0 += concurrent) {
    const unsigned nb = std::min(concurrent, grid - b0);
    std::vector<emu::Wave> waves(nb);
    std::vector<std::barrier<> *> bars;
    std::vector<double *> lds(nb);
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nb; ++w) {
      waves[w].G = G;
      waves[w].wbar = new std::barrier<>(64);
      for (int k = 0; k < 64 / G; ++k) waves[w].gbar[k] = new std::barrier<>(G);
      lds[w] = new double[lds_doubles];
      for (size_t i = 0; i < lds_doubles; ++i) lds[w][i] = std::nan("");  // nothing is read before it is written
    }
    for (unsigned w = 0; w < nb; ++w)
      for (int t = 0; t < 64; ++t)
        th.emplace_back([&, w, t]() {
          threadIdx.x = t;
          blockIdx.x = b0 + w;
          blockDim.x = 64;
          gridDim.x = grid;
          emu::wave = &waves[w];
          body(lds[w]);
        });
    for (auto &x : th) x.join();
    for (unsigned w = 0; w < nb; ++w) {
      delete waves[w].wbar;
      for (int k = 0; k < 64 / G; ++k) delete waves[w].gbar[k];
      delete[] lds[w];
    }
  }
}

template <int NS, bool KIN>
void batch_t(const BatchArgs &b) {
  constexpr int G = coop::group_size<NS>(), GPW = 64 / G;
  launch(G, (unsigned)((b.n + GPW - 1) / GPW), (size_t)GPW * coop::Slice<NS, KIN>::size, 8,
         [&](double *lds) { qp_batch_body<NS, KIN>(b, lds); });
}

template <int NS, typename DT, bool KIN>
void tangent_t(const QpArgs &a, unsigned grid) {
  constexpr int G = coop::group_size<NS>(), GPW = 64 / G;
  launch(G, grid, (size_t)GPW * coop::Slice<NS, KIN>::size, grid, [&](double *lds) { qp_tangent_coop_pass<NS, DT, KIN>(a, lds); });
}
template <int NS, bool KIN>
void post_t(const PostArgs &a, unsigned grid) {
  constexpr int G = coop::group_size<NS>(), GPW = 64 / G;
  launch(G, grid, (size_t)GPW * coop::Slice<NS, KIN>::size, grid, [&](double *lds) { post_qp_coop_pass<NS, KIN>(a, lds); });
}

// exact-size copy of a caller buffer (so that AddressSanitizer sees its bounds)
template <typename T>
std::vector<T> dup(const T *p, size_t n) { return std::vector<T>(p, p + n); }

}  // namespace

extern "C" {

// pa_qp_update's batch on the CPU: same arguments (W may be null without kinematic hardening).
// systems from g_nslip on are twin systems (one-sided shear); set by emu_set_nslip, default none
int g_nslip = 1 << 30;
void emu_set_nslip(int nslip) { g_nslip = nslip; }

int emu_qp_batch(const pa_material *mat, const double *rod, int nrows, long long n, const double *F, const double *Fp,
                 const double *s, const unsigned *grain, int flag, double *P, double *T, double *D, double *Fe, double *Fpo,
                 double *so, double *xo, int *na, double *W) {
  const int NS = ns_template(mat->n_slip), ns = mat->n_slip;
  int gstride = 0;
  std::vector<double> mb = material_block(*mat, NS), tab = grain_tables(*mat, rod, nrows, NS, gstride);
  std::vector<double> vF = dup(F, 9 * n), vFp = dup(Fp, 9 * n), vs = dup(s, ns * n);
  std::vector<unsigned> vg = dup(grain, n);
  std::vector<double> vP(9 * n), vT(9 * n), vD(flag ? 81 * n : 0), vFe(9 * n), vFpo(9 * n), vso(ns * n), vxo(ns * n);
  std::vector<int> vna(n);
  const bool kin = mat->kinematic_hardening != 0;
  std::vector<double> vW = kin ? (W ? dup(W, ns * n) : std::vector<double>(ns * n, 0.0)) : std::vector<double>();
  BatchArgs b{n, vF.data(), vFp.data(), vs.data(), vg.data(), tab.data(), gstride, mb.data(), ns, mat->max_solver_iter, flag,
              vP.data(), vT.data(), flag ? vD.data() : nullptr, vFe.data(), vFpo.data(), vso.data(), vxo.data(), vna.data(),
              kin ? vW.data() : nullptr};
  b.nslip = g_nslip;
  if (NS == 12) { if (kin) batch_t<12, true>(b); else batch_t<12, false>(b); }
  else if (NS == 24) { if (kin) batch_t<24, true>(b); else batch_t<24, false>(b); }
  else { if (kin) batch_t<48, true>(b); else batch_t<48, false>(b); }
  std::memcpy(P, vP.data(), 72 * n); std::memcpy(T, vT.data(), 72 * n); std::memcpy(Fe, vFe.data(), 72 * n);
  std::memcpy(Fpo, vFpo.data(), 72 * n); std::memcpy(so, vso.data(), 8 * ns * n); std::memcpy(xo, vxo.data(), 8 * ns * n);
  std::memcpy(na, vna.data(), 4 * n);
  if (flag) std::memcpy(D, vD.data(), 648 * n);
  if (kin && W) std::memcpy(W, vW.data(), 8 * ns * n);
  return 0;
}

// The one-lane routine (cp_qp.h qp_update, the form the warm-start batch runs) on the host, same
// arguments as emu_qp_batch: the reference against which the cooperative form is compared entry by
// entry.
int emu_qp_lane(const pa_material *mat, const double *rod, int nrows, long long n, const double *F, const double *Fp,
                const double *s, const unsigned *grain, int flag, double *P, double *T, double *D, double *Fe, double *Fpo,
                double *so, double *xo, int *na, double *W) {
  const int NS = ns_template(mat->n_slip), ns = mat->n_slip;
  const bool kin = mat->kinematic_hardening != 0;
  int gstride = 0;
  std::vector<double> mb = material_block(*mat, NS), tab = grain_tables(*mat, rod, nrows, NS, gstride);
  std::vector<double> ws(cp::LaneWs<48>::size);  // the 48-system variant's matrices (per-lane workspace)
  for (long long i = 0; i < n; ++i) {
    double s_[48], x_[48], W_[48];
    for (int a = 0; a < ns; ++a) { s_[a] = s[ns * i + a]; W_[a] = (kin && W) ? W[ns * i + a] : 0.0; }
    const double *gt = tab.data() + (size_t)grain[i] * gstride;
    auto run = [&](auto nsc, auto kinc) {
      constexpr int NSc = decltype(nsc)::value;
      constexpr bool K = decltype(kinc)::value;
      return cp::qp_update<NSc, double, K>(mb.data(), gt, ns, mat->max_solver_iter, F + 9 * i, Fp + 9 * i, s_, x_, flag, P + 9 * i,
                                           T + 9 * i, Fe + 9 * i, Fpo + 9 * i, flag ? D + 81 * i : nullptr, 1, nullptr, 1.0,
                                           cp::SvdInLane<NSc>{cp::LanePtr{ws.data() + cp::LaneWs<NSc>::UV, 1}}, W_, nullptr, 1,
                                           nullptr, g_nslip, cp::LanePtr{ws.data(), 1});
    };
    using I12 = std::integral_constant<int, 12>;
    using I24 = std::integral_constant<int, 24>;
    using I48 = std::integral_constant<int, 48>;
    using BT = std::true_type;
    using BF = std::false_type;
    int r;
    if (NS == 12) r = kin ? run(I12(), BT()) : run(I12(), BF());
    else if (NS == 24) r = kin ? run(I24(), BT()) : run(I24(), BF());
    else r = kin ? run(I48(), BT()) : run(I48(), BF());
    na[i] = r;
    for (int a = 0; a < ns; ++a) {
      so[ns * i + a] = s_[a];
      xo[ns * i + a] = x_[a];
      if (kin && W) W[ns * i + a] = W_[a];
    }
  }
  return 0;
}

// One qp pass over a structured mesh as the library runs it: flag 1 = assembly stage 1
// (k_qp_tangent: writes Fe/Fp/s/slip(/W)_iter, JxW dP/dF into D [81][nq] and JxW P into P [9][nq]),
// flag 0 = the post-increment pass (k_post_qp: the same history and T [9][nq]).  State arrays
// are component-major [c][nq]; u: 3 per node; egrain: 1 per element.  `grid` blocks pull from the
// work queue concurrently.  nsing receives the pseudo-inverse solves.
int emu_qp_pass(const pa_material *mat, const double *rod, int nrows, const int nel[3], const double h[3], const double *u,
                const unsigned *egrain, const double *Fp_conv, const double *s_conv, const double *slip_conv, const double *W_conv,
                int flag, int mixed, int grid, double *Fe_iter, double *Fp_iter, double *s_iter, double *slip_iter, double *W_iter,
                double *D, double *P, double *T, unsigned *nsing) {
  const int NS = ns_template(mat->n_slip), ns = mat->n_slip;
  const bool kin = mat->kinematic_hardening != 0;
  int gstride = 0;
  std::vector<double> mb = material_block(*mat, NS), tab = grain_tables(*mat, rod, nrows, NS, gstride);
  MeshDims m;
  m.nx = nel[0]; m.ny = nel[1]; m.nz = nel[2];
  m.nnx = m.nx + 1; m.nny = m.ny + 1; m.nnz = m.nz + 1;
  m.nnode = (long long)m.nnx * m.nny * m.nnz;
  m.nelem = (long long)m.nx * m.ny * m.nz;
  const long long nq = 8 * m.nelem;
  std::vector<double> vu = dup(u, 3 * m.nnode), vFp = dup(Fp_conv, 9 * nq), vs = dup(s_conv, ns * nq), vsl = dup(slip_conv, ns * nq);
  std::vector<double> vW = kin ? dup(W_conv, ns * nq) : std::vector<double>();
  std::vector<unsigned> veg = dup(egrain, m.nelem);
  std::vector<double> oFe(9 * nq), oFp(9 * nq), os(ns * nq), osl = dup(slip_conv, ns * nq), oW(kin ? ns * nq : 0), oP(9 * nq), oT(9 * nq);
  std::vector<double> oD(flag && !mixed ? 81 * nq : 0);
  std::vector<float> oDf(flag && mixed ? 81 * nq : 0);
  unsigned counters[2] = {0, 0};  // work queue, pseudo-inverse solves
  const double JxW = h[0] * h[1] * h[2] / 8.0;
  if (flag) {
    QpArgs a;
    a.m = m;
    for (int d = 0; d < 3; ++d) a.h[d] = h[d];
    a.JxW = JxW;
    a.u = vu.data(); a.egrain = veg.data(); a.gtab = tab.data(); a.gstride = gstride; a.mat = mb.data();
    a.ns = ns; a.max_solver_iter = mat->max_solver_iter; a.nq = nq;
    a.Fp_conv = vFp.data(); a.s_conv = vs.data(); a.slip_conv = vsl.data(); a.W_conv = kin ? vW.data() : nullptr;
    a.Fe_iter = oFe.data(); a.Fp_iter = oFp.data(); a.s_iter = os.data(); a.slip_iter = osl.data(); a.W_iter = kin ? oW.data() : nullptr;
    a.D = mixed ? (void *)oDf.data() : (void *)oD.data();
    a.P = oP.data();
    a.work = &counters[0]; a.nsing = &counters[1];
    if (NS == 12) {
      if (mixed) { if (kin) tangent_t<12, float, true>(a, grid); else tangent_t<12, float, false>(a, grid); }
      else { if (kin) tangent_t<12, double, true>(a, grid); else tangent_t<12, double, false>(a, grid); }
    } else if (NS == 24) {
      if (kin) tangent_t<24, double, true>(a, grid); else tangent_t<24, double, false>(a, grid);
    } else {
      if (kin) tangent_t<48, double, true>(a, grid); else tangent_t<48, double, false>(a, grid);
    }
  } else {
    PostArgs a;
    std::memset(&a, 0, sizeof(a));
    a.m = m;
    for (int d = 0; d < 3; ++d) a.h[d] = h[d];
    a.JxW = JxW;
    a.u = vu.data(); a.egrain = veg.data(); a.gtab = tab.data(); a.gstride = gstride; a.mat = mb.data();
    a.ns = ns; a.max_solver_iter = mat->max_solver_iter; a.nq = nq;
    a.Fp_conv = vFp.data(); a.s_conv = vs.data(); a.slip_conv = vsl.data(); a.W_conv = kin ? vW.data() : nullptr;
    a.Fe_iter = oFe.data(); a.Fp_iter = oFp.data(); a.s_iter = os.data(); a.slip_iter = osl.data(); a.W_iter = kin ? oW.data() : nullptr;
    a.T = oT.data();
    a.work = &counters[0]; a.nsing = &counters[1];
    if (NS == 12) { if (kin) post_t<12, true>(a, grid); else post_t<12, false>(a, grid); }
    else if (NS == 24) { if (kin) post_t<24, true>(a, grid); else post_t<24, false>(a, grid); }
    else { if (kin) post_t<48, true>(a, grid); else post_t<48, false>(a, grid); }
  }
  std::memcpy(Fe_iter, oFe.data(), 72 * nq); std::memcpy(Fp_iter, oFp.data(), 72 * nq);
  std::memcpy(s_iter, os.data(), 8 * ns * nq); std::memcpy(slip_iter, osl.data(), 8 * ns * nq);
  if (kin && W_iter) std::memcpy(W_iter, oW.data(), 8 * ns * nq);
  if (flag) {
    std::memcpy(P, oP.data(), 72 * nq);
    if (mixed) for (long long i = 0; i < 81 * nq; ++i) D[i] = oDf[i];
    else std::memcpy(D, oD.data(), 648 * nq);
  } else {
    std::memcpy(T, oT.data(), 72 * nq);
  }
  if (nsing) *nsing = counters[1];
  return 0;
}

}  // extern "C"

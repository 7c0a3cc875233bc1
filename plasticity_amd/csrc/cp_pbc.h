// Periodic boundary conditions (Use Periodic BCs, src/ellipticBVP/periodicBCs.cc): the affine
// constraints of the structured box, closed on the host, and the kernels that apply them to the
// Newton system without assembling a condensed matrix.
//
// A constrained dof d reads u_d = sum_k w_k u_{m_k} + s g_d over unconstrained masters m_k, where s
// is the load scale of the Newton iteration (1 at the first iteration of a loading increment, -1 of
// an unloading one of the tabular table, 0 otherwise: the reference's Inc0 / Inc0Neg / IncNot0
// constraint sets differ only in these inhomogeneities, which are linear in the vertex values).
// With C the matrix of the lines, the reference's condensed system (deal.II
// distribute_local_to_global over the element matrices) is C^T K C dx = C^T (R - K s g) on the
// unconstrained rows.  Here K and R stay the unconstrained node-block assembly, and
//   expand:  x_d = sum_k w_k x_{m_k} (+ s g_d) for constrained d, in place,
//   fold:    y_m += sum w y_d over the constrained images d of master m, then y_d = 0,
// bracket the node-block SpMV, so that a Krylov vector's unconstrained entries follow the
// condensed recurrences; constrained entries are zero in residuals and preconditioned residuals
// (Jacobi entry 0) and carry the expansion in search directions.  Masters with many images (the
// free corner dofs, whose images cover whole faces) are folded by block reductions.
//
// Included inside cp_api.hip's anonymous namespace after the Krylov helpers.

// ------------------------------------------------------------------ host: the closed lines
struct PbcLines {
  std::vector<int> crow;              // per dof: its constrained row, or -1
  std::vector<int> cdof, cptr, cidx;  // per row: dof; entries [cptr[r], cptr[r+1]) of masters cidx
  std::vector<double> cw, cg;         // weights; inhomogeneity at unit load scale
  std::vector<int> sdof, sptr, sidx;  // masters with a few images (no two of them neighbours)
  std::vector<double> sw;
  std::vector<int> ldof, lptr, lidx;  // masters with many or neighbouring images
  std::vector<double> lw;
  std::vector<double> scale;          // per increment: load scale of its first Newton iteration
};

// x86-64 conversion of a double to `unsigned int` (truncation to 64 bits, low 32 bits kept): the
// reference accumulates a derived vertex value in an unsigned int (periodicBCs.cc:56, 515-517).
inline double pbc_unsigned(double x) { return (double)(uint32_t)(int64_t)x; }

inline int pbc_build(const pa_problem &pb, const pa_periodic &P, PbcLines &L, std::string &err) {
  const int n = pb.nx;
  if (pb.ny != n || pb.nz != n || n < 2) { err = "periodic BCs need nx = ny = nz >= 2"; return PA_ERR_UNSUPPORTED; }
  if (P.n_vertex_lines < 0 || P.n_vertex_lines > PA_MAX_PBC_LINES || P.n_edge_lines < 0 || P.n_edge_lines > PA_MAX_PBC_LINES ||
      P.n_face_lines < 0 || P.n_face_lines > PA_MAX_PBC_LINES) {
    err = "periodic BCs: line counts out of range";
    return PA_ERR_ARG;
  }
  const long long nn = n + 1, N = 3 * nn * nn * nn;
  auto node = [&](long long i, long long j, long long k) { return i + nn * (j + nn * k); };
  auto vdof = [&](int r) {  // vertex row r: V000 V100 V010 V001 V110 V101 V011 V111
    static const int cor[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const int *c = cor[r / 3];
    return 3 * node(c[0] * n, c[1] * n, c[2] * n) + r % 3;
  };
  auto edof = [&](int r, int t) {  // edge row r, interior node t (ascending along the edge)
    const int e = r / 3, dir = e / 4, a = ((e % 4) >> 1) * n, b = (e & 1) * n;
    const long long nd = dir == 0 ? node(t + 1, a, b) : dir == 1 ? node(a, t + 1, b) : node(a, b, t + 1);
    return 3 * nd + r % 3;
  };
  // vertex values per increment (periodicBCs.cc:108-117) and the lines that fix a single unknown
  // vertex dof (:501-527)
  const double incs = P.tabular ? (double)(unsigned)(P.tabular_time / pb.delT) : (double)pb.total_increments;
  double val[24];
  int known[24];
  for (int i = 0; i < 24; ++i) { val[i] = P.vertex_total[i] / incs; known[i] = P.vertex_known[i] == 1; }
  const int nV = P.n_vertex_lines;
  for (int j = 0; j < nV; ++j) {
    const int cnt = P.vertex_dofs[j][0];
    if (cnt < 1 || cnt > 4) { err = "periodic BCs: vertex line entry count"; return PA_ERR_ARG; }
    for (int k = 1; k <= cnt; ++k)
      if (P.vertex_dofs[j][k] < 0 || P.vertex_dofs[j][k] >= 24) { err = "periodic BCs: vertex row out of range"; return PA_ERR_ARG; }
  }
  std::vector<char> line_done(nV, 0);
  for (int pass = 0; pass < nV; ++pass)
    for (int j = 0; j < nV; ++j) {
      const int cnt = P.vertex_dofs[j][0];
      int nk = 0, unk = 1;
      for (int k = 1; k <= cnt; ++k) {
        if (known[P.vertex_dofs[j][k]]) ++nk;
        else unk = k;
      }
      if (nk == cnt) line_done[j] = 1;
      if (nk == cnt - 1) {
        double acc = 0.0;
        for (int k = 1; k <= cnt; ++k)
          if (known[P.vertex_dofs[j][k]]) acc = pbc_unsigned(acc + val[P.vertex_dofs[j][k]] * P.vertex_coef[j][k]);
        const int d = P.vertex_dofs[j][unk];
        known[d] = 1;
        val[d] = acc * (-1.0 / P.vertex_coef[j][unk]);
        line_done[j] = 1;
      }
    }
  // raw lines in the reference's order: faces, edges, vertices (setPeriodicityConstraintsInc0)
  struct Raw {
    std::vector<std::pair<long long, double>> e;
    double g = 0.0;
    bool on = false;
  };
  std::vector<Raw> raw(N);
  auto entry = [&](long long d, long long m, double w) {  // an existing entry of the column stays
    for (const auto &q : raw[d].e) if (q.first == m) return;
    raw[d].e.push_back({m, w});
  };
  auto corners = [&](long long d, int v1, int v2, double c0, double c2, double c3) {  // :945-1003, 1133-1203
    if (known[v1] && known[v2]) raw[d].g = (-1 / c0) * c2 * val[v1] + (-1 / c0) * c3 * val[v2];
    else if (known[v1]) { raw[d].g = (-1 / c0) * c2 * val[v1]; entry(d, vdof(v2), (-1 / c0) * c3); }
    else if (known[v2]) { raw[d].g = (-1 / c0) * c3 * val[v2]; entry(d, vdof(v1), (-1 / c0) * c2); }
    else { entry(d, vdof(v2), (-1 / c0) * c3); entry(d, vdof(v1), (-1 / c0) * c2); }
  };
  for (int i = 0; i < P.n_face_lines; ++i) {
    const int *r = P.face_dofs[i];
    if (r[0] < 0 || r[0] >= 18 || ((r[0] / 3) & 1) == 0 || r[2] < 0 || r[2] >= 24 || r[3] < 0 || r[3] >= 24) {
      err = "periodic BCs: a face line must start with FXP, FYP or FZP and name vertex rows";
      return PA_ERR_ARG;
    }
  }
  // interior node (a, b) of the face normal to `dir` at coordinate index x
  auto fnode = [&](int dir, long long x, long long a, long long b) {
    return dir == 0 ? node(x, a, b) : dir == 1 ? node(a, x, b) : node(a, b, x);
  };
  for (int i = 0; i < P.n_face_lines; ++i) {  // make_periodicity_constraints (:1047-1075)
    const int dir = P.face_dofs[i][0] / 6, c = P.face_dofs[i][0] % 3;
    for (long long b = 1; b < n; ++b)
      for (long long a = 1; a < n; ++a) {
        const long long d = 3 * fnode(dir, n, a, b) + c;
        if (raw[d].on) continue;
        raw[d].on = true;
        entry(d, 3 * fnode(dir, 0, a, b) + c, 1.0);
      }
  }
  for (int i = 0; i < P.n_face_lines; ++i) {
    const int dir = P.face_dofs[i][0] / 6, c = P.face_dofs[i][0] % 3;
    for (long long b = 1; b < n; ++b)
      for (long long a = 1; a < n; ++a) {
        const long long d = 3 * fnode(dir, n, a, b) + c;
        if (raw[d].on)
          corners(d, P.face_dofs[i][2], P.face_dofs[i][3], P.face_coef[i][0], P.face_coef[i][2], P.face_coef[i][3]);
      }
  }
  for (int j = 0; j < P.n_edge_lines; ++j) {  // :922-1004
    const int *r = P.edge_dofs[j];
    const double *cf = P.edge_coef[j];
    if (r[0] < 0 || r[0] >= 36 || r[1] < 0 || r[1] >= 36 || r[2] < 0 || r[2] >= 24 || r[3] < 0 || r[3] >= 24) {
      err = "periodic BCs: edge line rows out of range";
      return PA_ERR_ARG;
    }
    for (int t = 0; t < n - 1; ++t) {
      const long long d = edof(r[0], t);
      raw[d].on = true;
      entry(d, edof(r[1], t), (-1 / cf[0]) * cf[1]);
      corners(d, r[2], r[3], cf[0], cf[2], cf[3]);
    }
  }
  for (int j = 0; j < nV; ++j) {  // :866-920
    if (line_done[j]) continue;
    const int cnt = P.vertex_dofs[j][0];
    if (known[P.vertex_dofs[j][1]]) { err = "periodic BCs: a vertex line constrains a known vertex dof"; return PA_ERR_UNSUPPORTED; }
    const long long d = vdof(P.vertex_dofs[j][1]);
    const double c1 = P.vertex_coef[j][1];
    raw[d].on = true;
    double inh = 0.0;
    for (int k = 2; k <= cnt; ++k) {
      const int rv = P.vertex_dofs[j][k];
      if (known[rv]) inh += P.vertex_coef[j][k] * val[rv];
      else entry(d, vdof(rv), (-1 / c1) * P.vertex_coef[j][k]);
    }
    raw[d].g = (-1 / c1) * inh;
  }
  for (int i = 0; i < 24; ++i)
    if (known[i]) { const long long d = vdof(i); raw[d].on = true; raw[d].g = val[i]; }
  // close (AffineConstraints::close): substitute constrained masters until none is left; a
  // substitution chain longer than the number of lines is a cycle
  std::vector<long long> rows;
  for (long long d = 0; d < N; ++d) if (raw[d].on) rows.push_back(d);
  for (size_t sweep = 0;; ++sweep) {
    if (sweep > rows.size() + 1) { err = "periodic BCs: the constraint lines form a cycle"; return PA_ERR_ARG; }
    bool changed = false;
    for (long long d : rows) {
      bool dep = false;
      for (const auto &q : raw[d].e) dep |= raw[q.first].on;
      if (!dep) continue;
      changed = true;
      std::vector<std::pair<long long, double>> out;
      auto add = [&](long long m, double w) {
        for (auto &q : out) if (q.first == m) { q.second += w; return; }
        out.push_back({m, w});
      };
      double g = raw[d].g;
      for (const auto &q : raw[d].e) {
        if (!raw[q.first].on) { add(q.first, q.second); continue; }
        g += q.second * raw[q.first].g;
        for (const auto &t : raw[q.first].e) add(t.first, q.second * t.second);
      }
      raw[d].e = out;
      raw[d].g = g;
    }
    if (!changed) break;
  }
  // device tables
  L.crow.assign(N, -1);
  L.cdof.clear(); L.cptr.assign(1, 0); L.cidx.clear(); L.cw.clear(); L.cg.clear();
  std::vector<std::vector<std::pair<int, double>>> img(N);
  for (long long d : rows) {
    std::sort(raw[d].e.begin(), raw[d].e.end());
    L.crow[d] = (int)L.cdof.size();
    L.cdof.push_back((int)d);
    for (const auto &q : raw[d].e) {
      L.cidx.push_back((int)q.first);
      L.cw.push_back(q.second);
      img[q.first].push_back({(int)d, q.second});
    }
    L.cptr.push_back((int)L.cidx.size());
    L.cg.push_back(raw[d].g);
  }
  auto near = [&](long long a, long long b) {  // nodes within one step on every axis (K couples them)
    const long long na = a / 3, nb = b / 3;
    const long long ia = na % nn, ja = (na / nn) % nn, ka = na / (nn * nn), ib = nb % nn, jb = (nb / nn) % nn, kb = nb / (nn * nn);
    return std::llabs(ia - ib) <= 1 && std::llabs(ja - jb) <= 1 && std::llabs(ka - kb) <= 1;
  };
  L.sdof.clear(); L.sptr.assign(1, 0); L.sidx.clear(); L.sw.clear();
  L.ldof.clear(); L.lptr.assign(1, 0); L.lidx.clear(); L.lw.clear();
  for (long long m = 0; m < N; ++m) {
    if (img[m].empty()) continue;
    bool isolated = img[m].size() <= 16;
    for (size_t a = 0; isolated && a < img[m].size(); ++a) {
      isolated &= !near(m, img[m][a].first);
      for (size_t b = a + 1; isolated && b < img[m].size(); ++b) isolated &= !near(img[m][a].first, img[m][b].first);
    }
    if (isolated) {
      L.sdof.push_back((int)m);
      for (const auto &q : img[m]) { L.sidx.push_back(q.first); L.sw.push_back(q.second); }
      L.sptr.push_back((int)L.sidx.size());
    } else {
      L.ldof.push_back((int)m);
      for (const auto &q : img[m]) { L.lidx.push_back(q.first); L.lw.push_back(q.second); }
      L.lptr.push_back((int)L.lidx.size());
    }
  }
  // load scale of each increment's first Newton iteration (setPeriodicityConstraints, :1276-1310)
  L.scale.assign(std::max(pb.total_increments, 1), 1.0);
  if (P.tabular) {
    if (P.n_table < 1 || P.n_table > PA_MAX_TABLE) { err = "periodic BCs: tabular table size"; return PA_ERR_ARG; }
    int tc = 1;
    for (int inc = 0; inc < pb.total_increments; ++inc) {
      const double t = pb.delT * (inc + 1) - pb.delT / 1000;
      if (tc < P.n_table && t > P.table_time[tc]) ++tc;
      L.scale[inc] = (double)P.table_coef[tc - 1];
    }
  }
  return 0;
}

// ------------------------------------------------------------------ device
constexpr int PBC_GL = 64;  // blocks per many-image master in the block reductions

struct PbcDev {
  int ncon, nshort, nlong;
  const int *crow, *cdof, *cptr, *cidx;
  const double *cw, *cg;
  const int *sdof, *sptr, *sidx;
  const double *sw;
  const int *ldof, *lptr, *lidx;
  const double *lw;
  double *part;  // [nlong][PBC_GL][2] partial sums
};

__device__ __forceinline__ bool pbc_skip(const KState *st) { return st != nullptr && solve_ended(st); }

// constrained entries of x (and x2) from their masters, plus s g (x only)
__global__ void __launch_bounds__(VB) k_pbc_expand(PbcDev p, double *x, double *x2, double s, const KState *st) {
  if (pbc_skip(st)) return;
  const int r = blockIdx.x * VB + threadIdx.x;
  if (r >= p.ncon) return;
  double v = s * p.cg[r], v2 = 0.0;
  for (int k = p.cptr[r]; k < p.cptr[r + 1]; ++k) {
    v += p.cw[k] * x[p.cidx[k]];
    if (x2) v2 += p.cw[k] * x2[p.cidx[k]];
  }
  x[p.cdof[r]] = v;
  if (x2) x2[p.cdof[r]] = v2;
}

// masters with a few images: y_m += sum w y_d
__global__ void __launch_bounds__(VB) k_pbc_fold_short(PbcDev p, double *y, double *y2, const KState *st) {
  if (pbc_skip(st)) return;
  const int r = blockIdx.x * VB + threadIdx.x;
  if (r >= p.nshort) return;
  const int m = p.sdof[r];
  double a = y[m], a2 = y2 ? y2[m] : 0.0;
  for (int k = p.sptr[r]; k < p.sptr[r + 1]; ++k) {
    a += p.sw[k] * y[p.sidx[k]];
    if (y2) a2 += p.sw[k] * y2[p.sidx[k]];
  }
  y[m] = a;
  if (y2) y2[m] = a2;
}

// masters with many images: PBC_GL partial sums of w y_d per master (grid PBC_GL x nlong)
__global__ void __launch_bounds__(VB) k_pbc_fold_long(PbcDev p, const double *y, const double *y2, const KState *st) {
  __shared__ double lds[2 * 16];
  if (pbc_skip(st)) return;
  const int v = blockIdx.y;
  double acc[2] = {0.0, 0.0};
  for (int k = p.lptr[v] + blockIdx.x * VB + threadIdx.x; k < p.lptr[v + 1]; k += PBC_GL * VB) {
    acc[0] += p.lw[k] * y[p.lidx[k]];
    if (y2) acc[1] += p.lw[k] * y2[p.lidx[k]];
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    p.part[((size_t)v * PBC_GL + blockIdx.x) * 2] = acc[0];
    p.part[((size_t)v * PBC_GL + blockIdx.x) * 2 + 1] = acc[1];
  }
}

// the many-image masters take their partial sums; then every constrained entry is zeroed
__global__ void __launch_bounds__(VB) k_pbc_fold_final(PbcDev p, double *y, double *y2, const KState *st) {
  if (pbc_skip(st)) return;
  const int t = blockIdx.x * VB + threadIdx.x;
  if (t < p.nlong) {
    double a = 0.0, a2 = 0.0;
    for (int b = 0; b < PBC_GL; ++b) {
      a += p.part[((size_t)t * PBC_GL + b) * 2];
      a2 += p.part[((size_t)t * PBC_GL + b) * 2 + 1];
    }
    y[p.ldof[t]] += a;
    if (y2) y2[p.ldof[t]] += a2;
  }
  for (int r = t; r < p.ncon; r += gridDim.x * VB) {
    y[p.cdof[r]] = 0.0;
    if (y2) y2[p.cdof[r]] = 0.0;
  }
}

template <typename KT>
__device__ __forceinline__ double pbc_kdiag(const MeshDims &m, const KT *K, long long d) {
  return (double)K[(13 * m.nnode + d / 3) * 9 + 4 * (d % 3)];
}

// Jacobi entries of the condensed matrix: masters with a few isolated images, diag = K_mm +
// sum w^2 K_dd (no coupling between the images)
template <typename KT>
__global__ void __launch_bounds__(VB) k_pbc_diag_short(PbcDev p, MeshDims m, const KT *__restrict__ K, double *dinv) {
  const int r = blockIdx.x * VB + threadIdx.x;
  if (r >= p.nshort) return;
  const int mm = p.sdof[r];
  double d = pbc_kdiag(m, K, mm);
  for (int k = p.sptr[r]; k < p.sptr[r + 1]; ++k) d += p.sw[k] * p.sw[k] * pbc_kdiag(m, K, p.sidx[k]);
  dinv[mm] = (d == 0.0) ? 1.0 : 1.0 / d;
}

// weight of master v in the line of dof b (1 for b = v itself, 0 for another unconstrained dof)
__device__ __forceinline__ double pbc_weight(const PbcDev &p, int b, int v) {
  if (b == v) return 1.0;
  const int r = p.crow[b];
  if (r < 0) return 0.0;
  for (int k = p.cptr[r]; k < p.cptr[r + 1]; ++k)
    if (p.cidx[k] == v) return p.cw[k];
  return 0.0;
}

// many-image masters: diag = sum_{a,b} w_a K_ab w_b over the master and its images, partial sums
// over the rows a (grid PBC_GL x nlong; entry -1 is the master itself)
template <typename KT>
__global__ void __launch_bounds__(VB) k_pbc_diag_long(PbcDev p, MeshDims m, const KT *__restrict__ K) {
  __shared__ double lds[16];
  const int v = blockIdx.y, mv = p.ldof[v];
  const int len = p.lptr[v + 1] - p.lptr[v];
  double acc[1] = {0.0};
  for (int k = (int)(blockIdx.x * VB + threadIdx.x) - 1; k < len; k += PBC_GL * VB) {
    const int a = k < 0 ? mv : p.lidx[p.lptr[v] + k];
    const double wa = k < 0 ? 1.0 : p.lw[p.lptr[v] + k];
    const long long na = a / 3;
    const int ra = a % 3;
    int i, j, kk;
    node_ijk32(m, na, i, j, kk);
    for (int s = 0; s < 27; ++s) {
      bool ok;
      const long long nb = nb_node(m, na, i, j, kk, s, ok);
      if (!ok) continue;
      for (int c = 0; c < 3; ++c) {
        const double wb = pbc_weight(p, (int)(3 * nb + c), mv);
        if (wb != 0.0) acc[0] += wa * (double)K[((long long)s * m.nnode + na) * 9 + 3 * ra + c] * wb;
      }
    }
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) p.part[((size_t)v * PBC_GL + blockIdx.x) * 2] = acc[0];
}

// many-image masters' Jacobi entries from their partial sums; constrained entries 0
__global__ void __launch_bounds__(VB) k_pbc_diag_final(PbcDev p, double *dinv) {
  const int t = blockIdx.x * VB + threadIdx.x;
  if (t < p.nlong) {
    double d = 0.0;
    for (int b = 0; b < PBC_GL; ++b) d += p.part[((size_t)t * PBC_GL + b) * 2];
    dinv[p.ldof[t]] = (d == 0.0) ? 1.0 : 1.0 / d;
  }
  for (int r = t; r < p.ncon; r += gridDim.x * VB) dinv[p.cdof[r]] = 0.0;
}

// R -= K ge (ge: the inhomogeneities s g at the constrained entries, zero elsewhere)
template <typename KT>
__global__ void __launch_bounds__(VB) k_pbc_rhs(MeshDims m, const KT *__restrict__ K, const double *__restrict__ ge,
                                                double *__restrict__ R) {
  const long long n = blockIdx.x * (long long)VB + threadIdx.x;
  if (n >= m.nnode) return;
  int i, j, k;
  node_ijk32(m, n, i, j, k);
  double y[3] = {0, 0, 0};
  for (int s = 0; s < 27; ++s) {
    bool ok;
    const long long mn = nb_masked(m, n, i, j, k, s, ok);
    const double w = ok ? 1.0 : 0.0;
    block_mv(K + ((long long)s * m.nnode + n) * 9, w * ge[3 * mn], w * ge[3 * mn + 1], w * ge[3 * mn + 2], y);
  }
  for (int c = 0; c < 3; ++c) R[3 * n + c] -= y[c];
}

// constraints.distribute(du) and x += du: u_d += X_d (unconstrained), sum w X_m + s g (constrained)
__global__ void k_pbc_apply(long long N, double *__restrict__ u, const double *__restrict__ X, PbcDev p, double s) {
  const long long d = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (d >= N) return;
  const int r = p.crow[d];
  if (r < 0) { u[d] += X[d]; return; }
  double v = s * p.cg[r];
  for (int k = p.cptr[r]; k < p.cptr[r + 1]; ++k) v += p.cw[k] * X[p.cidx[k]];
  u[d] += v;
}

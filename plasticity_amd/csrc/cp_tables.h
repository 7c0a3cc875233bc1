// Host-side construction of the per-grain constant table read by the qp kernels.
#pragma once
#include "../../include/plasticity_amd.h"

namespace {

// per-grain constants: rotated Schmid tensors, Dmat (shear columns doubled), TM, R
// (calculatePlasticity.cc:39-112, rotationOperations.cc:4-51, matrixOperations.cc:189-234)
inline void grain_table(const pa_material &mat, const double mv[][3], const double nv[][3], const double r[3], int ns, int NS,
                 double *out) {
  double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double t1 = 1.0 - rr, t2 = 1.0 + rr;
  double R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = (i == j ? t1 : 0.0) + 2.0 * r[i] * r[j];
  R[0][1] -= 2.0 * r[2]; R[0][2] += 2.0 * r[1]; R[1][2] -= 2.0 * r[0];
  R[1][0] += 2.0 * r[2]; R[2][0] -= 2.0 * r[1]; R[2][1] += 2.0 * r[0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] *= 1.0 / t2;
  double N[6][6];
  const int pi[3][2] = {{1, 2}, {2, 0}, {0, 1}};
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
  double Tm[6][6], D2[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { double s = 0; for (int k = 0; k < 6; ++k) s += N[i][k] * mat.C[k][j]; Tm[i][j] = s; }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { double s = 0; for (int k = 0; k < 6; ++k) s += Tm[i][k] * N[j][k]; D2[i][j] = s; }
  double *S = out, *D = out + 9 * NS, *TM = out + 9 * NS + 36;
  for (int a = 0; a < ns; ++a) {
    double mn[3][3], t[3][3];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) mn[j][k] = mv[a][j] * nv[a][k];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) { double s = 0; for (int k = 0; k < 3; ++k) s += R[i][k] * mn[k][j]; t[i][j] = s; }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) { double s = 0; for (int k = 0; k < 3; ++k) s += t[i][k] * R[j][k]; S[9 * a + 3 * i + j] = s; }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) D[6 * i + j] = (j >= 3 ? 2 * D2[i][j] : D2[i][j]);
  const int v2[9] = {0, 5, 4, 5, 1, 3, 4, 3, 2};
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) TM[9 * i + j] = D2[v2[i]][v2[j]];
  double *Rout = TM + 81;  // the rotation itself (lattice -> sample), row-major
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rout[3 * i + j] = R[i][j];
}

// Orientation of a grain after twinning on a system with unit normal nt (calculatePlasticity.cc:
// 995-1008 with updateAfterIncrement.cc:780-821): the quaternion of r, q = cos(phi/2) (1, r) with
// phi/2 = atan|r|, times (0, nt) as R(p) = R(q) R(0, nt), sign chosen so that p0 >= 0, back to
// r = p_vec / p0 (|p0| raised to 1e-4 when smaller).
inline void twin_variant_rod(const double r[3], const double nt[3], double out[3]) {
  const double c = std::cos(std::atan(std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])));
  const double q[4] = {c, c * r[0], c * r[1], c * r[2]};
  const double t[4] = {0.0, nt[0], nt[1], nt[2]};
  const double dot = t[1] * q[1] + t[2] * q[2] + t[3] * q[3];
  double p[4];
  p[0] = q[0] * t[0] - dot;
  p[1] = q[0] * t[1] + t[0] * q[1] + q[2] * t[3] - t[2] * q[3];
  p[2] = q[0] * t[2] + t[0] * q[2] - q[1] * t[3] + t[1] * q[3];
  p[3] = q[0] * t[3] + t[0] * q[3] + q[1] * t[2] - t[1] * q[2];
  if (p[0] < 0)
    for (int i = 0; i < 4; ++i) p[i] = -p[i];
  if (std::fabs(p[0]) < 1e-4) p[0] = 1e-4;
  const double inv = 1 / p[0];
  for (int i = 0; i < 3; ++i) out[i] = p[i + 1] * inv;
}

}  // namespace

// FP64 singular value decomposition, polar rotation, fixed-corotated Kirchhoff stress and
// the singular-value clamp plasticity, for one particle (device code).
//
// 3x3: symmetric Jacobi eigen-analysis of A = F^T F (cyclic sweeps with exact rotations,
// Golub & Van Loan, "Matrix Computations", 8.5, stopped at an off-diagonal residual of 1e-10)
// gives V; B = F V is then reduced to upper
// triangular form with three Givens rotations (the QR step of McAdams et al. 2011, "Computing
// the SVD of 3x3 matrices with minimal branching"), which gives U and the singular values
// as diag(R).  U and V are rotations, so the smallest singular value carries det F's sign
// (the rotation-variant SVD that makes U V^T the polar rotation).
// 2x2: closed form of Blinn, "Consider the lowly 2x2 matrix" (1996).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"

#ifndef AGIMP_HD
#define AGIMP_HD __host__ __device__
#endif

namespace agimp {

struct Mat3d { double m[3][3]; };

// 1 / sqrt(x) in FP64 (the device library's reciprocal square root: one operation instead
// of a square root and a division)
AGIMP_HD inline double rsqrt_d(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  return ::rsqrt(x);
#else
  return 1.0 / sqrt(x);
#endif
}

AGIMP_HD inline void jacobi_rotate(double A[3][3], double V[3][3], int p, int q) {
  double apq = A[p][q];
  // (below 1e-150, d^2 + e^2 could underflow; such an entry is already zero to FP64 precision
  // next to any diagonal entry the sweep criterion can still resolve)
  if (fabs(apq) < 1e-150) return;
  // The rotation angle phi has tan(phi) = t, the root of t^2 + 2 theta t - 1 = 0 with
  // |t| <= 1, theta = (aqq - app) / (2 apq).  With d = aqq - app, e = 2 apq and
  // r = sqrt(d^2 + e^2): cos(2 phi) = |d| / r and sin(2 phi) = sgn(d) e / r, so by the
  // half-angle formulas cos^2(phi) = (1 + |d|/r) / 2 and sin(phi) = sin(2 phi) / (2 cos(phi)).
  // Two reciprocal square roots, no division and no square root.
  const double d = A[q][q] - A[p][p], e = 2.0 * apq;
  const double inv_r = rsqrt_d(d * d + e * e);
  const double c2 = 0.5 + 0.5 * fabs(d) * inv_r;  // in [1/2, 1]
  const double ic = rsqrt_d(c2);
  double c = c2 * ic;
  double s = 0.5 * (d >= 0.0 ? e : -e) * inv_r * ic;
  double t = s * ic;
  // A' = J^T A J with J the rotation in the (p,q) plane
  int r = 3 - p - q;
  double app = A[p][p], aqq = A[q][q], arp = A[r][p], arq = A[r][q];
  A[p][p] = app - t * apq;
  A[q][q] = aqq + t * apq;
  A[p][q] = A[q][p] = 0.0;
  double nrp = c * arp - s * arq, nrq = s * arp + c * arq;
  A[r][p] = A[p][r] = nrp;
  A[r][q] = A[q][r] = nrq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double vp = V[i][p], vq = V[i][q];
    V[i][p] = c * vp - s * vq;
    V[i][q] = s * vp + c * vq;
  }
}

// Givens rotation that zeroes B[j][k] against B[i][k] acting on rows i, j; accumulates U.
AGIMP_HD inline void givens_rows(double B[3][3], double U[3][3], int i, int j, int k) {
  double a = B[i][k], b = B[j][k];
  double r2 = a * a + b * b;
  double c = 1.0, s = 0.0;
  if (r2 > 1e-300) {
    const double inv = rsqrt_d(r2);
    c = a * inv;
    s = b * inv;
  }
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    double bi = B[i][col], bj = B[j][col];
    B[i][col] = c * bi + s * bj;
    B[j][col] = -s * bi + c * bj;
  }
  // B_new = G B  =>  B = G^T B_new ; U accumulates G^T
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    double ui = U[row][i], uj = U[row][j];
    U[row][i] = c * ui + s * uj;
    U[row][j] = -s * ui + c * uj;
  }
}

AGIMP_HD inline void svd3(const double F[3][3], double U[3][3], double s[3], double V[3][3]) {
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      A[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
      V[i][j] = (i == j) ? 1.0 : 0.0;
    }
  // stop once the off-diagonal part is 1e-10 of the diagonal (squares: 1e-20): the
  // reconstruction U diag(s) V^T then differs from F by ~1e-10 relative, far below the FP32
  // rounding of every output; on deformation gradients of the 1M sphere scene this needs at
  // most 3 sweeps for 99 % of particles (1e-30: 4 sweeps for 31 %)
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= 1e-20 * dia) break;
    jacobi_rotate(A, V, 0, 1);
    jacobi_rotate(A, V, 0, 2);
    jacobi_rotate(A, V, 1, 2);
  }
  // sort eigenvalues descending, permuting V's columns
  double ev[3] = {A[0][0], A[1][1], A[2][2]};
#define AGIMP_SWAPCOL(p, q)                                                          \
  {                                                                                  \
    double te = ev[p]; ev[p] = ev[q]; ev[q] = te;                                    \
    for (int i_ = 0; i_ < 3; ++i_) { double tv = V[i_][p]; V[i_][p] = V[i_][q]; V[i_][q] = -tv; } \
  }
  if (ev[0] < ev[1]) AGIMP_SWAPCOL(0, 1);
  if (ev[0] < ev[2]) AGIMP_SWAPCOL(0, 2);
  if (ev[1] < ev[2]) AGIMP_SWAPCOL(1, 2);
#undef AGIMP_SWAPCOL
  // (column swaps negate one column, so det V stays +1)
  double B[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      B[i][j] = F[i][0] * V[0][j] + F[i][1] * V[1][j] + F[i][2] * V[2][j];
      U[i][j] = (i == j) ? 1.0 : 0.0;
    }
  givens_rows(B, U, 0, 1, 0);
  givens_rows(B, U, 0, 2, 0);
  givens_rows(B, U, 1, 2, 1);
  s[0] = B[0][0]; s[1] = B[1][1]; s[2] = B[2][2];
}

// Blinn 1996: F = R(phi) diag(s0, s1) R(theta)^T with s0 >= |s1|; s1 < 0 iff det F < 0.
AGIMP_HD inline void svd2(const double F[2][2], double U[2][2], double s[2], double V[2][2]) {
  double E = 0.5 * (F[0][0] + F[1][1]), Fh = 0.5 * (F[0][0] - F[1][1]);
  double G = 0.5 * (F[1][0] + F[0][1]), H = 0.5 * (F[1][0] - F[0][1]);
  double Q = sqrt(E * E + H * H), R = sqrt(Fh * Fh + G * G);
  s[0] = Q + R;
  s[1] = Q - R;
  double a1 = atan2(G, Fh), a2 = atan2(H, E);
  double theta = 0.5 * (a2 - a1), phi = 0.5 * (a2 + a1);
  double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
  U[0][0] = cp; U[0][1] = -sp; U[1][0] = sp; U[1][1] = cp;
  // F = R(phi) diag(s) R(theta), so V = R(theta)^T
  V[0][0] = ct; V[0][1] = st; V[1][0] = -st; V[1][1] = ct;
}

template <int D>
AGIMP_HD inline void svd_d(const double F[D][D], double U[D][D], double s[D], double V[D][D]);
template <>
AGIMP_HD inline void svd_d<2>(const double F[2][2], double U[2][2], double s[2], double V[2][2]) {
  svd2(F, U, s, V);
}
template <>
AGIMP_HD inline void svd_d<3>(const double F[3][3], double U[3][3], double s[3], double V[3][3]) {
  svd3(F, U, s, V);
}

template <int D>
AGIMP_HD inline double det_d(const double F[D][D]) {
  if constexpr (D == 2) return F[0][0] * F[1][1] - F[0][1] * F[1][0];
  else
    return F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
           F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
}

// Plasticity (optional) then Kirchhoff stress of the resulting elastic F (utils.py:11-16):
//   tau = 2 mu (F - U V^T) F^T + la J (J - 1) I,   mu, la scaled by exp(xi (1 - Jp)).
// With clamp and P.plastic, F and Jp are updated in place (Stomakhin et al. 2013).
template <int D>
AGIMP_HD inline void constitutive(double F[D][D], double& Jp, const Physics& P, double tau[D][D], bool clamp) {
  double U[D][D], s[D], V[D][D];
  svd_d<D>(F, U, s, V);
  if (clamp && P.plastic) {
    double lo = 1.0 - (double)P.theta_c, hi = 1.0 + (double)P.theta_s;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double ns = fmin(fmax(s[a], lo), hi);
      if (ns != s[a]) Jp *= s[a] / ns;  // s / ns is exactly 1 otherwise
      s[a] = ns;
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) acc += U[i][k] * s[k] * V[j][k];
        F[i][j] = acc;
      }
  }
  double h = (P.plastic && P.hardening != 0.0) ? exp((double)P.hardening * (1.0 - Jp)) : 1.0;
  double mu = (double)P.mu0 * h, la = (double)P.la0 * h;
  double J = det_d<D>(F);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double Rik = 0.0;
#pragma unroll
        for (int q = 0; q < D; ++q) Rik += U[i][q] * V[k][q];
        acc += (F[i][k] - Rik) * F[j][k];
      }
      tau[i][j] = 2.0 * mu * acc + ((i == j) ? la * J * (J - 1.0) : 0.0);
    }
}

}  // namespace agimp

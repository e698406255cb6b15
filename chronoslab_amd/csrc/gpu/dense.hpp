// Small dense FP64 linear algebra on the host for the Rayleigh-Ritz steps of LOBPCG
// (matrices of order <= 48): Cholesky, the symmetric-definite generalized eigenproblem by
// Cholesky reduction + cyclic Jacobi (Golub & Van Loan, Alg. 8.5.3), condition numbers.
// All matrices are column-major.
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace chrdense {

using Mat = std::vector<double>;  // column-major m x m

// In-place lower Cholesky (strict upper part zeroed).  Returns false when not SPD.
inline bool chol_lower(int m, double *K) {
  for (int j = 0; j < m; ++j) {
    double d = K[j + j * m];
    for (int k = 0; k < j; ++k) d -= K[j + k * m] * K[j + k * m];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    K[j + j * m] = d;
    for (int i = j + 1; i < m; ++i) {
      double s = K[i + j * m];
      for (int k = 0; k < j; ++k) s -= K[i + k * m] * K[j + k * m];
      K[i + j * m] = s / d;
    }
    for (int i = 0; i < j; ++i) K[i + j * m] = 0.0;
  }
  return true;
}

inline void symmetrize(int m, double *M) {
  for (int j = 0; j < m; ++j)
    for (int i = j + 1; i < m; ++i) {
      const double s = 0.5 * (M[i + j * m] + M[j + i * m]);
      M[i + j * m] = M[j + i * m] = s;
    }
}

// Cyclic Jacobi: S (symmetric) is destroyed; w = eigenvalues (unsorted), Q = eigenvectors.
inline void jacobi_eig(int m, double *S, double *w, double *Q) {
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) Q[i + j * m] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        const double v = S[i + j * m] * S[i + j * m];
        tot += v;
        if (i != j) off += v;
      }
    if (off == 0.0 || off <= 1e-32 * tot) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = S[p + q * m];
        if (apq == 0.0) continue;
        const double tau = (S[q + q * m] - S[p + p * m]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double skp = S[k + p * m], skq = S[k + q * m];
          S[k + p * m] = c * skp - s * skq;
          S[k + q * m] = s * skp + c * skq;
        }
        for (int k = 0; k < m; ++k) {
          const double spk = S[p + k * m], sqk = S[q + k * m];
          S[p + k * m] = c * spk - s * sqk;
          S[q + k * m] = s * spk + c * sqk;
        }
        for (int k = 0; k < m; ++k) {
          const double qkp = Q[k + p * m], qkq = Q[k + q * m];
          Q[k + p * m] = c * qkp - s * qkq;
          Q[k + q * m] = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < m; ++i) w[i] = S[i + i * m];
}

// A v = lambda B v, A symmetric, B SPD: ascending eigenvalues w, B-orthonormal V.
inline bool sygv(int m, const double *A, const double *B, double *w, double *V) {
  Mat L(B, B + (size_t)m * m), C(A, A + (size_t)m * m), Q((size_t)m * m), wt(m);
  if (!chol_lower(m, L.data())) return false;
  for (int j = 0; j < m; ++j)  // C = L^-1 A
    for (int i = 0; i < m; ++i) {
      double s = C[i + j * m];
      for (int k = 0; k < i; ++k) s -= L[i + k * m] * C[k + j * m];
      C[i + j * m] = s / L[i + i * m];
    }
  for (int i = 0; i < m; ++i)  // C = C L^-T
    for (int j = 0; j < m; ++j) {
      double s = C[i + j * m];
      for (int k = 0; k < j; ++k) s -= C[i + k * m] * L[j + k * m];
      C[i + j * m] = s / L[j + j * m];
    }
  symmetrize(m, C.data());
  jacobi_eig(m, C.data(), wt.data(), Q.data());
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return wt[a] < wt[b]; });
  for (int jj = 0; jj < m; ++jj) {
    const int j = ord[jj];
    w[jj] = wt[j];
    double *v = V + (size_t)jj * m;
    for (int i = 0; i < m; ++i) v[i] = Q[i + j * m];
    for (int i = m - 1; i >= 0; --i) {  // v = L^-T q
      double s = v[i];
      for (int k = i + 1; k < m; ++k) s -= L[k + i * m] * v[k];
      v[i] = s / L[i + i * m];
    }
  }
  return true;
}

// 2-norm condition number of a symmetric matrix (|eigenvalue| ratio).
inline double cond_sym(int m, const double *M) {
  Mat S(M, M + (size_t)m * m), Q((size_t)m * m), w(m);
  jacobi_eig(m, S.data(), w.data(), Q.data());
  double mx = 0.0, mn = 1e308;
  for (double x : w) {
    mx = std::max(mx, std::fabs(x));
    mn = std::min(mn, std::fabs(x));
  }
  return mn == 0.0 ? INFINITY : mx / mn;
}

// Inverse of an upper-triangular U (column-major).
inline Mat inv_upper(int m, const double *U) {
  Mat R((size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    R[j + j * m] = 1.0 / U[j + j * m];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += U[i + k * m] * R[k + j * m];
      R[i + j * m] = -s / U[i + i * m];
    }
  }
  return R;
}

}  // namespace chrdense

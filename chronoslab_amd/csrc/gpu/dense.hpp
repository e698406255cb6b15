// Small dense FP64 linear algebra on the host for the Rayleigh-Ritz steps of LOBPCG
// (matrices of order <= 48): Cholesky, the symmetric-definite generalized eigenproblem by
// Cholesky reduction + the symmetric QR algorithm, condition numbers.
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

// Symmetric eigen-decomposition: S (symmetric) is destroyed; w = eigenvalues (unsorted),
// Q = orthonormal eigenvectors.  Householder reduction to tridiagonal form, then implicit
// QR steps with the Wilkinson shift and Givens bulge chasing (Golub & Van Loan, Algs. 8.3.1
// and 8.3.3); both accumulate into Q.
inline void sym_eig(int m, double *S, double *w, double *Q) {
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) Q[i + j * m] = (i == j) ? 1.0 : 0.0;
  std::vector<double> v(m), y(m);
  for (int k = 0; k + 2 < m; ++k) {
    // Reflector H = I - beta v v^T on indices k+1..m-1 mapping S[k+1.., k] onto e_1.
    double sig = 0.0;
    for (int i = k + 2; i < m; ++i) sig += S[i + k * m] * S[i + k * m];
    if (sig == 0.0) continue;
    const double x0 = S[k + 1 + k * m], alpha = std::sqrt(x0 * x0 + sig);
    const double v0 = x0 + (x0 >= 0.0 ? alpha : -alpha);
    v[k + 1] = v0;
    for (int i = k + 2; i < m; ++i) v[i] = S[i + k * m];
    const double beta = 2.0 / (v0 * v0 + sig);
    // S <- S H (all rows), then S <- H S (all columns), Q <- Q H.
    for (int i = 0; i < m; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < m; ++j) t += S[i + j * m] * v[j];
      y[i] = beta * t;
    }
    for (int j = k + 1; j < m; ++j)
      for (int i = 0; i < m; ++i) S[i + j * m] -= y[i] * v[j];
    for (int j = 0; j < m; ++j) {
      double t = 0.0;
      for (int i = k + 1; i < m; ++i) t += v[i] * S[i + j * m];
      t *= beta;
      for (int i = k + 1; i < m; ++i) S[i + j * m] -= t * v[i];
    }
    for (int i = 0; i < m; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < m; ++j) t += Q[i + j * m] * v[j];
      y[i] = beta * t;
    }
    for (int j = k + 1; j < m; ++j)
      for (int i = 0; i < m; ++i) Q[i + j * m] -= y[i] * v[j];
  }
  // Tridiagonal T: diagonal d, off-diagonal e[i] = T(i+1, i).
  std::vector<double> d(m), e(m, 0.0);
  for (int i = 0; i < m; ++i) d[i] = S[i + i * m];
  for (int i = 0; i + 1 < m; ++i) e[i] = 0.5 * (S[i + 1 + i * m] + S[i + (i + 1) * m]);
  const double eps = 2.220446049250313e-16;
  int hi = m - 1, guard = 0;
  while (hi > 0 && guard < 60 * m) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
    if (lo > 0) e[lo - 1] = 0.0;
    ++guard;
    // Wilkinson shift from the trailing 2x2 block of T(lo..hi).
    const double dd = 0.5 * (d[hi - 1] - d[hi]), b = e[hi - 1];
    const double mu = d[hi] - b * b / (dd + (dd >= 0.0 ? 1.0 : -1.0) * std::hypot(dd, b));
    double x = d[lo] - mu, z = e[lo], bulge = 0.0;
    for (int k = lo; k < hi; ++k) {
      // Givens rotation on rows/columns k, k+1 zeroing z against x.
      const double r = std::hypot(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
      if (k > lo) e[k - 1] = r;
      // Rotate the 2x2 block [d_k e_k; e_k d_k1] and the entries beside it.
      const double dk = d[k], dk1 = d[k + 1], ek = e[k];
      d[k] = c * c * dk + 2.0 * c * s * ek + s * s * dk1;
      d[k + 1] = s * s * dk - 2.0 * c * s * ek + c * c * dk1;
      e[k] = c * s * (dk1 - dk) + (c * c - s * s) * ek;
      if (k + 1 < hi) {
        bulge = s * e[k + 1];
        e[k + 1] *= c;
        x = e[k];
        z = bulge;
      }
      for (int i = 0; i < m; ++i) {
        const double qk = Q[i + k * m], qk1 = Q[i + (k + 1) * m];
        Q[i + k * m] = c * qk + s * qk1;
        Q[i + (k + 1) * m] = -s * qk + c * qk1;
      }
    }
  }
  for (int i = 0; i < m; ++i) w[i] = d[i];
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
  sym_eig(m, C.data(), wt.data(), Q.data());
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
  sym_eig(m, S.data(), w.data(), Q.data());
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

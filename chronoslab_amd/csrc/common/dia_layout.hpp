// Symmetric diagonal layout of a square sparse matrix (host side, no HIP): built here,
// uploaded by the GPU library for its single-vector products (spmv_dia_kernel) and the PCG
// operator product (pcg_p_dia_kernel), and replayed on the CPU by the host library
// (chs_dia_apply) with the kernel's own index arithmetic.
//
// A candidate has its nonzeros on at most kMaxDiag distinct diagonals at offsets >= 0 and
// the mirrored ones below, values that equal their transposed entries bit for bit, no
// stored zeros, and sorted rows.  Only the main and upper diagonals are stored:
//   off[0] = 0 < off[1] < ... < off[m-1],   val[k * n + i] = A(i, i + off[k])
// (0 where that entry is absent or i + off[k] >= n).  Entry (i, i - off[k]) of the lower
// triangle is read as val[k * n + i - off[k]].  A 7-point stencil matrix keeps 4 values per
// row and no column indices, against 7 values and 7 codes in a row-indexed layout.
//
// Row i is summed in column order: the lower diagonals from the farthest, the main
// diagonal, the upper diagonals from the nearest, each product rounded before it is added
// and absent (zero) entries skipped.  That is the sequence of additions of the sequential
// CSR product of a sorted row, so the sums are the same bits.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace chrdia {

constexpr int kMaxDiag = 16;  // stored diagonals (main + upper)

struct Layout {
  int32_t n = 0, m = 0;  // m = 0: the matrix is not a candidate
  int32_t off[kMaxDiag] = {};
  std::vector<double> val;  // m * n
};

inline bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// Builds L from CSR; returns false (L.m = 0) when the matrix is not a candidate.
inline bool build(int32_t n, int32_t ncols, const int32_t *rp, const int32_t *ci, const double *v, Layout &L) {
  L = Layout();
  L.n = n;
  if (n <= 0 || ncols != n || rp[n] <= 0) return false;
  const int64_t nnz = rp[n];
  // Upper offsets (the main diagonal always has slot 0); rows sorted, values nonzero.
  std::vector<int32_t> offs{0};
  int64_t upper = 0, lower = 0;
  for (int32_t r = 0; r < n; ++r)
    for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
      if (e > rp[r] && ci[e] <= ci[e - 1]) return false;
      if (!(v[e] != 0.0) || std::isnan(v[e])) return false;
      const int32_t d = ci[e] - r;
      if (d > 0) {
        ++upper;
        if (std::find(offs.begin(), offs.end(), d) == offs.end()) {
          if ((int)offs.size() == kMaxDiag) return false;
          offs.push_back(d);
        }
      } else if (d < 0) {
        ++lower;
      }
    }
  if (upper != lower) return false;
  std::sort(offs.begin(), offs.end());
  const int m = (int)offs.size();
  // Stored values (half the matrix plus padding) at most the CSR's value count plus n.
  if ((int64_t)m * n > nnz + n) return false;
  std::vector<int> slot((size_t)offs.back() + 1, -1);
  for (int k = 0; k < m; ++k) slot[(size_t)offs[k]] = k;
  L.val.assign((size_t)m * n, 0.0);
  for (int32_t r = 0; r < n; ++r)
    for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
      const int32_t d = ci[e] - r;
      if (d >= 0) L.val[(size_t)slot[(size_t)d] * n + r] = v[e];
    }
  // Every lower entry must be the mirror of a stored upper entry, bit for bit; with equal
  // counts (and distinct positions) that makes the value matrix symmetric.
  for (int32_t r = 0; r < n; ++r)
    for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
      const int32_t d = r - ci[e];
      if (d <= 0) continue;
      if (d >= (int32_t)slot.size() || slot[(size_t)d] < 0 ||
          !same_bits(L.val[(size_t)slot[(size_t)d] * n + ci[e]], v[e])) {
        L = Layout();
        L.n = n;
        return false;
      }
    }
  L.m = m;
  for (int k = 0; k < m; ++k) L.off[k] = offs[(size_t)k];
  return true;
}

}  // namespace chrdia

// Bootstrap-AMG least-squares interpolation: every fine node is interpolated from coarse
// nodes of its strength neighbourhood with the weights that best reproduce the test vectors
// (the near-null space) at that node.
//
// Restates the reference's BAMG prolongation:
//   row driver          sources/Preconditioner/AMG/prolong/MEX_BAMG_Prol/BAMG.cpp:16-241
//   belt growth         add_new_neighs.cpp:5-28 (breadth-first over the strength pattern)
//   basis selection     maxVol.cpp:6-85 (pivoted Householder rank reveal + maximum volume)
//                       maxVol_inner.cpp:27-74, move_row_front.cpp, mk_HouHolVec.cpp,
//                       Apply_HouHol_Rot.cpp, backsolve.cpp
//   weight reduction    Reduce_IntSet.cpp:6-77 (RELAX_FAC = 1.5, BAMG_params.h:20)
//   gateway             cpt_Prolongation_BAMG.cpp:18-164 (explicit zero weights dropped)
// The least-squares fits are the reference's LAPACKE_dgels calls (BAMG.cpp:165, a Householder
// QR of the ntv x k system); here they are a Householder QR of our own, so weights agree
// with the reference's to rounding, not bit for bit.  Everything else (the order in which
// neighbours are met, pivots, swaps, acceptance tests) follows the reference step for step.
#include "bamg.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace chsb {
namespace {

constexpr double kRelax = 1.5;  // BAMG_params.h:20

double nrm2(int n, const double *v) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i] * v[i];
  return std::sqrt(s);
}

// Unit Householder vector h (length len) with (I - 2 h h^T) v = -+|v| e_1; h = 0 if v = 0.
void reflector(int len, const double *v, double *h) {
  double s2 = 0.0;
  for (int i = 0; i < len; ++i) {
    h[i] = v[i];
    s2 += v[i] * v[i];
  }
  const double s = std::sqrt(s2);
  if (s == 0.0) {
    for (int i = 0; i < len; ++i) h[i] = 0.0;
    return;
  }
  double hn;
  if (h[0] == 0.0) {
    h[0] = s;
    hn = std::sqrt(2.0 * s2);
  } else if (h[0] > 0.0) {
    hn = std::sqrt(2.0 * (h[0] * s + s2));
    h[0] += s;
  } else {
    hn = std::sqrt(2.0 * (-h[0] * s + s2));
    h[0] -= s;
  }
  const double inv = 1.0 / hn;
  for (int i = 0; i < len; ++i) h[i] *= inv;
}

void reflect(int len, const double *h, double *v) {
  double d = 0.0;
  for (int i = 0; i < len; ++i) d += h[i] * v[i];
  const double f = -2.0 * d;
  for (int i = 0; i < len; ++i) v[i] += f * h[i];
}

struct Work {
  std::vector<int32_t> nb;     // neighbourhood in the order met, nb[0] = the node itself
  std::vector<uint8_t> seen;   // membership of nb, cleared after each row
  std::vector<int32_t> cand;   // candidate coarse nodes (fine indices)
  std::vector<double> T;       // candidate test-vector rows, row-major, stride m
  std::vector<double> h, zc, zr;
  std::vector<double> M;       // least-squares matrix, column-major m x k
  std::vector<double> b;       // right-hand side, then solution + residual components
};

// Chooses at most kmax of the nrow rows of T (row-major, stride m, overwritten) spanning
// their row space with a near-maximal volume; the chosen candidates end up in
// cand[0..rank).  Stage 1 (maxVol.cpp:23-61): pivoted Householder reduction, the largest
// remaining row first, stopping when the norm ratio to the first pivot exceeds condmax.
// Stage 2 (maxVol.cpp:67-84, maxVol_inner.cpp:27-74): the other rows in the basis of the
// chosen ones, Z = B L^-1, and while some |Z_ij| > 1 + delta, swap row i into the basis in
// place of pivot j with the rank-1 update of Z.
int select_rows(int kmax, double condmax, int itmax, double delta, int nrow, int m, Work &w) {
  double *T = w.T.data();
  int32_t *cand = w.cand.data();
  const int maxrank = std::min(std::min(nrow, m), kmax);
  auto bring_front = [&](int r, double &best) {  // largest norm of T[i][r..m), i >= r
    int ip = r;
    best = nrm2(m - r, T + (size_t)r * m + r);
    for (int i = r + 1; i < nrow; ++i) {
      const double s = nrm2(m - r, T + (size_t)i * m + r);
      if (s > best) {
        best = s;
        ip = i;
      }
    }
    if (ip != r) std::swap_ranges(T + (size_t)ip * m, T + (size_t)ip * m + m, T + (size_t)r * m);
    std::swap(cand[r], cand[ip]);
  };
  double smax;
  bring_front(0, smax);
  reflector(m, T, w.h.data());
  for (int i = 0; i < nrow; ++i) reflect(m, w.h.data(), T + (size_t)i * m);
  int rank = 1;
  while (rank < maxrank) {
    const int r = rank;
    ++rank;
    double s;
    bring_front(r, s);
    const double cond = std::fabs(s) > 1e-15 ? smax / std::fabs(s) : condmax + 1.0;
    if (cond > condmax) {
      --rank;
      break;
    }
    reflector(m - r, T + (size_t)r * m + r, w.h.data());
    for (int i = r; i < nrow; ++i) reflect(m - r, w.h.data(), T + (size_t)i * m + r);
  }
  // Rows 0..rank-1 now hold a lower-triangular L (row i has entries in columns 0..i);
  // replace each further row's first rank entries b by z with z L = b.
  for (int k = rank; k < nrow; ++k) {
    double *z = T + (size_t)k * m;
    for (int i = rank - 1; i >= 0; --i) {
      const double *Li = T + (size_t)i * m;
      z[i] /= Li[i];
      for (int j = i - 1; j >= 0; --j) z[j] -= Li[j] * z[i];
    }
  }
  const int nz = nrow - rank;
  double *Z = T + (size_t)rank * m;  // nz x rank, stride m
  auto zmax = [&](int &ir, int &jc) {
    double best = 0.0;
    ir = jc = 0;
    for (int j = 0; j < rank; ++j)
      for (int i = 0; i < nz; ++i) {
        const double a = std::fabs(Z[(size_t)i * m + j]);
        if (a > best) {
          best = a;
          ir = i;
          jc = j;
        }
      }
    return best;
  };
  int ir, jc;
  double zm = zmax(ir, jc);
  for (int it = 0; zm > 1.0 + delta && it < itmax; ++it) {
    double *zc = w.zc.data(), *zr = w.zr.data();
    for (int i = 0; i < nz; ++i) zc[i] = Z[(size_t)i * m + jc];
    zc[ir] -= 1.0;
    const double f = 1.0 / Z[(size_t)ir * m + jc];
    for (int j = 0; j < rank; ++j) zr[j] = Z[(size_t)ir * m + j] * f;
    zr[jc] += f;
    for (int i = 0; i < nz; ++i)
      for (int j = 0; j < rank; ++j) Z[(size_t)i * m + j] -= zc[i] * zr[j];
    std::swap(cand[rank + ir], cand[jc]);
    zm = zmax(ir, jc);
  }
  return rank;
}

// min || M x - b || for M (m x k column-major, full column rank) by Householder QR; b is
// overwritten with x in b[0..k) and the residual's components in b[k..m) (the output
// convention of LAPACK's xGELS).  false on a zero pivot (rank-deficient M).
bool least_squares(int m, int k, double *M, double *b, double *h) {
  for (int c = 0; c < k; ++c) {
    double *col = M + (size_t)c * m;
    const int len = m - c;
    double s2 = 0.0;
    for (int i = c; i < m; ++i) s2 += col[i] * col[i];
    if (s2 == 0.0) return false;
    const double s = std::sqrt(s2);
    const double alpha = col[c] >= 0.0 ? -s : s;
    for (int i = 0; i < len; ++i) h[i] = col[c + i];
    h[0] -= alpha;
    double hh = 0.0;
    for (int i = 0; i < len; ++i) hh += h[i] * h[i];
    const double inv = hh > 0.0 ? 2.0 / hh : 0.0;
    for (int c2 = c + 1; c2 < k; ++c2) {
      double *v = M + (size_t)c2 * m + c;
      double d = 0.0;
      for (int i = 0; i < len; ++i) d += h[i] * v[i];
      d *= inv;
      for (int i = 0; i < len; ++i) v[i] -= d * h[i];
    }
    double d = 0.0;
    for (int i = 0; i < len; ++i) d += h[i] * b[c + i];
    d *= inv;
    for (int i = 0; i < len; ++i) b[c + i] -= d * h[i];
    col[c] = alpha;
  }
  for (int i = k - 1; i >= 0; --i) {
    double t = b[i];
    for (int j = i + 1; j < k; ++j) t -= M[(size_t)j * m + i] * b[j];
    b[i] = t / M[(size_t)i * m + i];
  }
  return true;
}

}  // namespace

int prolong(int32_t n, const int32_t *rp, const int32_t *ci, const int32_t *fc, int32_t ntv,
            const double *tv, const Params &prm, int nthreads, std::vector<int32_t> &prp,
            std::vector<int32_t> &pci, std::vector<double> &pv, int32_t *mark, std::string &err) {
  const int m = ntv;
  const int kcap = std::max(prm.mmax, ntv);
  std::vector<int32_t> cnt(n, 0), tcol((size_t)n * kcap);
  std::vector<double> tval((size_t)n * kcap);
  int failed = 0;
#pragma omp parallel num_threads(nthreads)
  {
    Work w;
    w.seen.assign(n, 0);
    w.nb.reserve(256);
    w.h.resize(m + 1);
    w.zr.resize(m + 1);
    w.b.resize(m);
    w.M.resize((size_t)m * m);
    auto grow = [&](size_t b0, size_t b1) {  // add_new_neighs.cpp:5-28
      for (size_t q = b0; q < b1; ++q) {
        const int32_t j = w.nb[q];
        for (int32_t e = rp[j]; e < rp[j + 1]; ++e) {
          const int32_t c = ci[e];
          if (!w.seen[c]) {
            w.seen[c] = 1;
            w.nb.push_back(c);
          }
        }
      }
    };
    // Gathers the coarse nodes of nb[1..) and fits the node's test vectors from at most kmax
    // of them; returns the rank, the weights in w.b[0..rank), the residual in w.b[rank..m).
    auto fit = [&](int32_t i, int kmax, bool &ok) -> int {
      w.cand.clear();
      for (size_t q = 1; q < w.nb.size(); ++q)
        if (fc[w.nb[q]] >= 0) w.cand.push_back(w.nb[q]);
      const int nc = (int)w.cand.size();
      if (nc == 0) return 0;
      w.T.resize((size_t)nc * m);
      if ((int)w.zc.size() < nc) w.zc.resize(nc);
      for (int q = 0; q < nc; ++q)
        std::copy(tv + (size_t)w.cand[q] * m, tv + (size_t)w.cand[q] * m + m, w.T.begin() + (size_t)q * m);
      const int rank = select_rows(kmax, prm.maxcond, prm.itmax_vol, prm.tol_vol, nc, m, w);
      std::sort(w.cand.begin(), w.cand.begin() + rank);
      for (int q = 0; q < rank; ++q)
        std::copy(tv + (size_t)w.cand[q] * m, tv + (size_t)w.cand[q] * m + m, w.M.begin() + (size_t)q * m);
      std::copy(tv + (size_t)i * m, tv + (size_t)i * m + m, w.b.begin());
      ok = least_squares(m, rank, w.M.data(), w.b.data(), w.h.data());
      return rank;
    };
#pragma omp for schedule(dynamic, 32)
    for (int32_t i = 0; i < n; ++i) {
      mark[i] = 0;
      int stop;
#pragma omp atomic read
      stop = failed;
      if (stop) continue;
      if (fc[i] >= 0) {
        tcol[(size_t)i * kcap] = fc[i];
        tval[(size_t)i * kcap] = 1.0;
        cnt[i] = 1;
        continue;
      }
      w.nb.clear();
      w.nb.push_back(i);
      w.seen[i] = 1;
      size_t b0 = 0, b1 = 1;
      int dist = 0;
      while (dist < prm.dist_min) {
        ++dist;
        grow(b0, b1);
        if (w.nb.size() == b1) break;
        b0 = b1;
        b1 = w.nb.size();
      }
      const double bnorm = nrm2(m, tv + (size_t)i * m);
      const double res_ok = prm.eps * bnorm;
      int rank = 0;
      double res = 0.0, wn = 0.0;
      bool ok = true;
      // Reduce_IntSet.cpp:14-76: fewer points until the weights' norm is within bound.
      auto reduce = [&](double bound) {
        while (wn > bound) {
          if (rank <= 1) {
            rank = 0;
            break;
          }
          rank = fit(i, rank - 1, ok);
          if (!ok) return;
          wn = nrm2(rank, w.b.data());
        }
      };
      for (;;) {
        const int nc_before = [&] {
          int c = 0;
          for (size_t q = 1; q < w.nb.size(); ++q) c += fc[w.nb[q]] >= 0;
          return c;
        }();
        if (nc_before > 0) {
          rank = fit(i, prm.mmax, ok);
          if (!ok) break;
          res = nrm2(m - rank, w.b.data() + rank);
          wn = nrm2(rank, w.b.data());
        } else {
          rank = 0;
          res = bnorm;
          wn = 0.0;
        }
        if (wn < prm.maxrownrm && (res < res_ok || rank == m)) break;  // accepted
        if (dist < prm.dist_max) {
          ++dist;
          const size_t before = w.nb.size();
          grow(b0, b1);
          if (w.nb.size() == before) {  // no new neighbours: keep this fit
            if (nc_before == 0) {
              mark[i] = 1;
            } else {
              if (wn > kRelax * prm.maxrownrm) reduce(kRelax * prm.maxrownrm);
              if (!ok) break;
              res = rank == 0 ? bnorm : nrm2(m - rank, w.b.data() + rank);
              if (!(res < res_ok)) mark[i] = 1;
            }
            break;
          }
          b0 = b1;
          b1 = w.nb.size();
        } else {
          if (wn > kRelax * prm.maxrownrm) reduce(kRelax * prm.maxrownrm);
          if (!ok) break;
          res = rank == 0 ? bnorm : nrm2(m - rank, w.b.data() + rank);
          if (res > res_ok) mark[i] = 1;
          break;
        }
      }
      for (int32_t q : w.nb) w.seen[q] = 0;
      if (!ok) {
#pragma omp atomic write
        failed = 1;
        continue;
      }
      int c = 0;
      for (int q = 0; q < rank; ++q) {
        if (w.b[q] == 0.0) continue;  // cpt_Prolongation_BAMG.cpp:109,141
        tcol[(size_t)i * kcap + c] = fc[w.cand[q]];
        tval[(size_t)i * kcap + c] = w.b[q];
        ++c;
      }
      cnt[i] = c;
    }
  }
  if (failed) {
    err = "bamg: singular least-squares fit";
    return 2;
  }
  prp.assign((size_t)n + 1, 0);
  for (int32_t i = 0; i < n; ++i) prp[i + 1] = prp[i] + cnt[i];
  pci.resize(prp[n]);
  pv.resize(prp[n]);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t q = 0; q < cnt[i]; ++q) {
      pci[prp[i] + q] = tcol[(size_t)i * kcap + q];
      pv[prp[i] + q] = tval[(size_t)i * kcap + q];
    }
  return 0;
}

int orthonormalize(int64_t n, int32_t k, double *V) {
  constexpr int64_t kBlk = 4096;
  const int64_t nb = (n + kBlk - 1) / kBlk;
  std::vector<double> part(nb);
  auto dot = [&](const double *x, const double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
      double s = 0.0;
      for (int64_t i = b * kBlk; i < std::min(n, (b + 1) * kBlk); ++i) s += x[i] * y[i];
      part[b] = s;
    }
    double s = 0.0;
    for (int64_t b = 0; b < nb; ++b) s += part[b];
    return s;
  };
  int deficient = 0;
  for (int32_t c = 0; c < k; ++c) {
    double *v = V + c * n;
    const double n0 = std::sqrt(dot(v, v));
    for (int pass = 0; pass < 2; ++pass)
      for (int32_t j = 0; j < c; ++j) {
        const double *q = V + j * n;
        const double r = dot(q, v);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) v[i] -= r * q[i];
      }
    const double nr = std::sqrt(dot(v, v));
    if (!(nr > 1e-12 * n0) || nr == 0.0) {
      deficient = 1;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) v[i] = 0.0;
      continue;
    }
    const double inv = 1.0 / nr;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) v[i] *= inv;
  }
  return deficient;
}

}  // namespace chsb

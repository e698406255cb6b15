// Energy-minimising prolongation: the pattern of MEX_cptProlPatt.cpp and the matrix-product
// (solv_emin = 4), Jacobi-preconditioned (prec_emin = 1) variant of the reference's EMIN
// for symmetric A:
//   sources/Preconditioner/AMG/prolong/EMIN/MEX_EMIN/MEX_cptProlPatt.cpp:239-430 (pattern)
//   .../EMIN/Cpp_EMIN/EMIN_matfree.cpp:61-300 (driver), load_Jacobi.cpp, copy_Prol.cpp,
//   gather_B_QR.cpp:30-480 (constraint blocks), gather_f.cpp, KP_spmat.cpp, Orth_Q.cpp,
//   DEFL_PCG_matfree.cpp:28-250 (projected CG), Prol_add_Cnodes.cpp.
// P minimises trace(P^T A P) over the entries of the fine rows' pattern, with every fine
// row i held to P_i TV_C = TV_i (TV_C: the test space at the coarse nodes).  The constraint
// of row i is the block B_i (its pattern's coarse test-vector rows); Q_i is an orthonormal
// basis of range(B_i), so corrections stay in the constraint's null space under
// x -> x - Q_i Q_i^T x.  Dot products are summed over fixed 4096-entry blocks, so the result
// does not depend on the thread count.
#include "emin.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>

namespace chse {
namespace {

double det_dot(int64_t n, const double *x, const double *y, int nthreads) {
  const int64_t kB = 4096, nb = (n + kB - 1) / kB;
  std::vector<double> part((size_t)std::max<int64_t>(nb, 1), 0.0);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    double s = 0.0;
    const int64_t e = std::min(n, (b + 1) * kB);
    for (int64_t i = b * kB; i < e; ++i) s += x[i] * y[i];
    part[(size_t)b] = s;
  }
  double s = 0.0;
  for (int64_t b = 0; b < nb; ++b) s += part[(size_t)b];
  return s;
}

// Householder QR of B (nr x m, column-major, nr > m) in place: R in the upper triangle,
// reflector tails below the diagonal, scalars in tau (LAPACK dgeqrf's conventions).
void house_qr(int nr, int m, double *B, double *tau) {
  for (int k = 0; k < m; ++k) {
    double *x = B + (size_t)k * nr + k;
    const int len = nr - k;
    double sig = 0.0;
    for (int i = 1; i < len; ++i) sig += x[i] * x[i];
    if (sig == 0.0) {
      tau[k] = 0.0;
      continue;
    }
    const double alpha = x[0], nrm = std::sqrt(alpha * alpha + sig);
    const double beta = alpha <= 0.0 ? nrm : -nrm;
    tau[k] = (beta - alpha) / beta;
    const double scal = 1.0 / (alpha - beta);
    for (int i = 1; i < len; ++i) x[i] *= scal;
    x[0] = beta;
    for (int j = k + 1; j < m; ++j) {
      double *y = B + (size_t)j * nr + k;
      double w = y[0];
      for (int i = 1; i < len; ++i) w += x[i] * y[i];
      w *= tau[k];
      y[0] -= w;
      for (int i = 1; i < len; ++i) y[i] -= w * x[i];
    }
  }
}

// Thin Q (nr x m) of house_qr's factors.
void house_q(int nr, int m, const double *B, const double *tau, double *Q) {
  std::fill(Q, Q + (size_t)nr * m, 0.0);
  for (int j = 0; j < m; ++j) Q[(size_t)j * nr + j] = 1.0;
  for (int k = m - 1; k >= 0; --k) {
    if (tau[k] == 0.0) continue;
    const double *v = B + (size_t)k * nr + k;  // v[0] = 1 implicitly
    const int len = nr - k;
    for (int j = 0; j < m; ++j) {
      double *y = Q + (size_t)j * nr + k;
      double w = y[0];
      for (int i = 1; i < len; ++i) w += v[i] * y[i];
      w *= tau[k];
      y[0] -= w;
      for (int i = 1; i < len; ++i) y[i] -= w * v[i];
    }
  }
}

// One-sided Jacobi SVD (Hestenes): W (nr x m, column-major) becomes U Sigma, V (m x m) the
// right singular vectors; sig[j] = column norms of W.
void jacobi_svd(int nr, int m, double *W, double *V, double *sig) {
  std::fill(V, V + (size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) V[(size_t)j * m + j] = 1.0;
  const double eps = 2.220446049250313e-16;
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        double *wp = W + (size_t)p * nr, *wq = W + (size_t)q * nr;
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = 0; i < nr; ++i) {
          a += wp[i] * wp[i];
          b += wq[i] * wq[i];
          g += wp[i] * wq[i];
        }
        if (g == 0.0 || std::fabs(g) <= eps * std::sqrt(a * b)) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < nr; ++i) {
          const double xp = wp[i], xq = wq[i];
          wp[i] = c * xp - s * xq;
          wq[i] = s * xp + c * xq;
        }
        double *vp = V + (size_t)p * m, *vq = V + (size_t)q * m;
        for (int i = 0; i < m; ++i) {
          const double xp = vp[i], xq = vq[i];
          vp[i] = c * xp - s * xq;
          vq[i] = s * xp + c * xq;
        }
      }
    if (!rotated) break;
  }
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int i = 0; i < nr; ++i) s += W[(size_t)j * nr + i] * W[(size_t)j * nr + i];
    sig[j] = std::sqrt(s);
  }
}

// Row i's constraint block (gather_B_QR.cpp:150-420): B = TV at the pattern's coarse
// nodes, g = TV_i - B^T p.  Corrects p by the minimum-norm step with B^T p = TV_i (exactly
// when B has full column rank, in least squares over its numerical range otherwise) and
// leaves in Q (nr x m) an orthonormal basis of that range, zero-padded.  Returns false
// when the factorisation breaks down.
bool constrain_row(int nr, int m, const double *B0, const double *tvi, double condmax,
                   double *p, double *Q) {
  std::vector<double> B(B0, B0 + (size_t)nr * m), g(tvi, tvi + m);
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int k = 0; k < nr; ++k) s += B[(size_t)j * nr + k] * p[k];
    g[j] -= s;
  }
  if (nr > m) {
    std::vector<double> tau(m);
    house_qr(nr, m, B.data(), tau.data());
    double mx = 0.0, mn = INFINITY;
    for (int k = 0; k < m; ++k) {
      mx = std::max(mx, std::fabs(B[(size_t)k * nr + k]));
      mn = std::min(mn, std::fabs(B[(size_t)k * nr + k]));
    }
    const double ratio = mx / mn;
    if (ratio != ratio) return false;  // 0/0: an all-zero block (dtrtrs would fail)
    if (!(ratio > condmax)) {
      // g <- R^-T g, then p += Q g.
      for (int i = 0; i < m; ++i) {
        double s = g[i];
        for (int k = 0; k < i; ++k) s -= B[(size_t)i * nr + k] * g[k];
        g[i] = s / B[(size_t)i * nr + i];
      }
      house_q(nr, m, B.data(), tau.data(), Q);
      for (int j = 0; j < m; ++j)
        for (int k = 0; k < nr; ++k) p[k] += Q[(size_t)j * nr + k] * g[j];
      return true;
    }
    B.assign(B0, B0 + (size_t)nr * m);
  }
  // Rank-revealing least squares by the SVD (gather_B_QR.cpp:290-360).
  std::vector<double> V((size_t)m * m), sig(m);
  jacobi_svd(nr, m, B.data(), V.data(), sig.data());
  std::vector<int> ord(m);
  for (int j = 0; j < m; ++j) ord[j] = j;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sig[a] > sig[b]; });
  const int mn = std::min(m, nr);
  int rank = 0;
  if (sig[ord[0]] > 1e-15) {
    rank = 1;
    if (nr > 1)
      while (rank < mn && sig[ord[0]] < condmax * std::fabs(sig[ord[rank]])) {
        ++rank;
        if (rank == mn) break;
      }
  }
  std::fill(Q, Q + (size_t)nr * m, 0.0);
  for (int r = 0; r < rank; ++r) {
    const int j = ord[r];
    double w = 0.0;
    for (int i = 0; i < m; ++i) w += V[(size_t)j * m + i] * g[i];
    w /= sig[j];
    for (int k = 0; k < nr; ++k) {
      const double u = B[(size_t)j * nr + k] / sig[j];
      Q[(size_t)r * nr + k] = u;
      p[k] += u * w;
    }
  }
  return true;
}

}  // namespace

void pattern(int32_t n, int32_t nc, const int32_t *arp, const int32_t *aci, const double *av,
             const int32_t *prp, const int32_t *pci, const double *pv, const int32_t *fc,
             int32_t ntv, const Params &prm, int nthreads, std::vector<int32_t> &trp,
             std::vector<int32_t> &tci) {
  // P without stored zeros (the MEX reads P through a zero filter).
  std::vector<int32_t> p0rp((size_t)n + 1, 0), p0ci;
  std::vector<double> p0v;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t k = prp[i]; k < prp[i + 1]; ++k)
      if (pv[k] != 0.0) {
        p0ci.push_back(pci[k]);
        p0v.push_back(pv[k]);
      }
    p0rp[i + 1] = (int32_t)p0ci.size();
  }
  trp.assign((size_t)n + 1, 0);
  tci.clear();
  if (prm.patt_pow <= 1) {
    for (int32_t i = 0; i < n; ++i) {
      if (fc[i] < 0) tci.insert(tci.end(), p0ci.begin() + p0rp[i], p0ci.begin() + p0rp[i + 1]);
      trp[i + 1] = (int32_t)tci.size();
    }
    return;
  }
  // S: DOM strength |a_ij| / sqrt(a_ii a_jj) of the entries above patt_tau (cpt_SoC_dom.m).
  std::vector<double> d(n, 0.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = arp[i]; k < arp[i + 1]; ++k)
      if (aci[k] == i) d[i] = av[k];
  std::vector<int32_t> srp((size_t)n + 1, 0), sci;
  std::vector<double> sv;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t k = arp[i]; k < arp[i + 1]; ++k) {
      if (av[k] == 0.0) continue;
      const double s = std::fabs(av[k]) / std::sqrt(d[i] * d[aci[k]]);
      if (s > prm.patt_tau) {
        sci.push_back(aci[k]);
        sv.push_back(s);
      }
    }
    srp[i + 1] = (int32_t)sci.size();
  }
  const int32_t nnzr = std::max(prm.nnzr_max, ntv);
  // Current operand of the product (P, then the previous power's pattern with its weights).
  std::vector<int32_t> crp = p0rp, cci = p0ci;
  std::vector<double> cv = p0v;
  std::vector<std::vector<int32_t>> rows_c((size_t)n);
  std::vector<std::vector<double>> rows_v((size_t)n);
  for (int32_t ipow = 2; ipow <= prm.patt_pow; ++ipow) {
    const bool last = ipow == prm.patt_pow;
#pragma omp parallel num_threads(nthreads)
    {
      std::vector<int32_t> marker((size_t)nc, 0);
      std::vector<double> acc((size_t)nc, 0.0);
      std::vector<int32_t> cand;
      struct Ent {
        int32_t col;
        double val;
      };
      std::vector<Ent> ext;
      int32_t stamp = 0;
#pragma omp for schedule(dynamic, 64)
      for (int32_t i = 0; i < n; ++i) {
        auto &oc = rows_c[i];
        auto &ov = rows_v[i];
        oc.clear();
        ov.clear();
        if (last && fc[i] >= 0) continue;
        const int32_t n0 = p0rp[i + 1] - p0rp[i];
        const int32_t budget = std::max(0, nnzr - n0);
        if (budget > 0) {
          stamp += 2;
          const int32_t st_p0 = stamp, st_seen = stamp + 1;
          for (int32_t k = p0rp[i]; k < p0rp[i + 1]; ++k) marker[p0ci[k]] = st_p0;
          cand.clear();
          for (int32_t ks = srp[i]; ks < srp[i + 1]; ++ks) {
            const int32_t j = sci[ks];
            const double s = sv[ks];
            for (int32_t kp = crp[j]; kp < crp[j + 1]; ++kp) {
              const int32_t col = cci[kp];
              if (marker[col] == st_seen) {
                acc[col] += s * cv[kp];
              } else if (marker[col] < st_p0) {
                marker[col] = st_seen;
                cand.push_back(col);
                acc[col] = s * cv[kp];
              }
            }
          }
          ext.clear();
          for (int32_t col : cand)
            if (acc[col] != 0.0) ext.push_back({col, acc[col]});
          if ((int64_t)ext.size() > budget) {
            // The budget largest |weight|, ties to the lower column.
            std::nth_element(ext.begin(), ext.begin() + budget, ext.end(), [](const Ent &a, const Ent &b) {
              const double x = std::fabs(a.val), y = std::fabs(b.val);
              return x != y ? x > y : a.col < b.col;
            });
            ext.resize((size_t)budget);
          }
        } else {
          ext.clear();
        }
        for (int32_t k = p0rp[i]; k < p0rp[i + 1]; ++k) ext.push_back({p0ci[k], 1.0});
        std::sort(ext.begin(), ext.end(), [](const Ent &a, const Ent &b) { return a.col < b.col; });
        for (const Ent &e : ext) {
          oc.push_back(e.col);
          ov.push_back(e.val);
        }
      }
    }
    if (!last) {
      crp.assign((size_t)n + 1, 0);
      cci.clear();
      cv.clear();
      for (int32_t i = 0; i < n; ++i) {
        cci.insert(cci.end(), rows_c[i].begin(), rows_c[i].end());
        cv.insert(cv.end(), rows_v[i].begin(), rows_v[i].end());
        crp[i + 1] = (int32_t)cci.size();
      }
    }
  }
  for (int32_t i = 0; i < n; ++i) {
    tci.insert(tci.end(), rows_c[i].begin(), rows_c[i].end());
    trp[i + 1] = (int32_t)tci.size();
  }
}

int improve(int32_t n, int32_t nc, const int32_t *arp, const int32_t *aci, const double *av,
            const int32_t *prp, const int32_t *pci, const double *pv, const int32_t *trp,
            const int32_t *tci, const int32_t *fc, int32_t ntv, const double *tv,
            const Params &prm, int nthreads, std::vector<int32_t> &orp, std::vector<int32_t> &oci,
            std::vector<double> &ov, int32_t &iter, std::string &err,
            const std::function<int(const double *, double *)> &kp_ext) {
  iter = 0;
  const int64_t nt = trp[n];
  std::vector<int32_t> c2g((size_t)nc, -1);
  for (int32_t i = 0; i < n; ++i)
    if (fc[i] >= 0) {
      if (fc[i] >= nc) {
        err = "emin: coarse index out of range";
        return 1;
      }
      c2g[fc[i]] = i;
    }
  for (int32_t c = 0; c < nc; ++c)
    if (c2g[c] < 0) {
      err = "emin: a coarse index has no node";
      return 1;
    }
  for (int32_t i = 0; i < n; ++i)
    if (fc[i] >= 0 && trp[i + 1] != trp[i]) {
      err = "emin: the pattern has entries in a coarse row";
      return 1;
    }
  // Jacobi scaling 1/a_ii of every pattern entry of row i (load_Jacobi.cpp).
  std::vector<double> dinv((size_t)nt), p0((size_t)nt, 0.0);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    double inv = 1.0;
    for (int32_t k = arp[i]; k < arp[i + 1]; ++k)
      if (aci[k] == i && std::fabs(av[k]) > 1e-15) inv = 1.0 / av[k];
    for (int32_t k = trp[i]; k < trp[i + 1]; ++k) dinv[k] = inv;
    // P's fine-row values at the pattern's positions (copy_Prol.cpp).
    if (fc[i] < 0) {
      int32_t jj = trp[i];
      for (int32_t ii = prp[i]; ii < prp[i + 1]; ++ii) {
        while (jj < trp[i + 1] && tci[jj] < pci[ii]) ++jj;
        if (jj < trp[i + 1] && tci[jj] == pci[ii]) p0[jj++] = pv[ii];
      }
    }
  }
  // Constraint blocks: correct P's rows onto the constraint, keep Q.
  std::vector<double> Q((size_t)nt * ntv, 0.0);
  int bad = 0;
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<double> B, tvi((size_t)ntv);
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < n; ++i) {
      const int nr = trp[i + 1] - trp[i];
      if (fc[i] >= 0 || nr == 0) continue;
      B.resize((size_t)nr * ntv);
      for (int k = 0; k < nr; ++k) {
        const int32_t g = c2g[tci[trp[i] + k]];
        for (int j = 0; j < ntv; ++j) B[(size_t)j * nr + k] = tv[(size_t)j * n + g];
      }
      for (int j = 0; j < ntv; ++j) tvi[j] = tv[(size_t)j * n + i];
      if (!constrain_row(nr, ntv, B.data(), tvi.data(), prm.condmax, p0.data() + trp[i],
                         Q.data() + (size_t)trp[i] * ntv)) {
#pragma omp atomic write
        bad = 1;
      }
    }
  }
  if (bad) {
    err = "emin: a constraint block has no usable factorisation";
    return 4;
  }
  // f = -A(c, i) at pattern entry (i, c) (gather_f.cpp), A symmetric.
  std::vector<double> f((size_t)nt, 0.0);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = trp[i]; k < trp[i + 1]; ++k) {
      const int32_t r = c2g[tci[k]];
      const int32_t *b = aci + arp[r], *e = aci + arp[r + 1];
      const int32_t *hit = std::lower_bound(b, e, i);
      if (hit != e && *hit == i) f[k] = -av[arp[r] + (hit - b)];
    }
  // y = (A X) on the pattern, X on the pattern (KP_spmat.cpp).
  int kp_fail = 0;
  auto kp_host = [&](const double *x, double *y) {
#pragma omp parallel num_threads(nthreads)
    {
      std::vector<int32_t> wn((size_t)nc, -1);
#pragma omp for schedule(dynamic, 64)
      for (int32_t i = 0; i < n; ++i) {
        const int32_t c0 = trp[i], c1 = trp[i + 1];
        if (c0 == c1) continue;
        for (int32_t k = c0; k < c1; ++k) {
          y[k] = 0.0;
          wn[tci[k]] = k - c0;
        }
        for (int32_t ka = arp[i]; ka < arp[i + 1]; ++ka) {
          const int32_t m = aci[ka];
          const double a = av[ka];
          for (int32_t kb = trp[m]; kb < trp[m + 1]; ++kb) {
            const int32_t pos = wn[tci[kb]];
            if (pos >= 0) y[c0 + pos] += a * x[kb];
          }
        }
        for (int32_t k = c0; k < c1; ++k) wn[tci[k]] = -1;
      }
    }
  };
  auto kp = [&](const double *x, double *y) {
    if (kp_ext) {
      if (!kp_fail && kp_ext(x, y)) kp_fail = 1;
    } else {
      kp_host(x, y);
    }
  };
  // y = (I - Q Q^T) x row by row (Orth_Q.cpp).
  auto orth = [&](const double *x, double *y) {
#pragma omp parallel num_threads(nthreads)
    {
      std::vector<double> a((size_t)ntv);
#pragma omp for schedule(static)
      for (int32_t i = 0; i < n; ++i) {
        const int nr = trp[i + 1] - trp[i];
        if (nr == 0) continue;
        const double *q = Q.data() + (size_t)trp[i] * ntv;
        const double *xi = x + trp[i];
        double *yi = y + trp[i];
        for (int j = 0; j < ntv; ++j) {
          double s = 0.0;
          for (int k = 0; k < nr; ++k) s += q[(size_t)j * nr + k] * xi[k];
          a[j] = s;
        }
        for (int k = 0; k < nr; ++k) {
          double t = 0.0;
          for (int j = 0; j < ntv; ++j) t += a[j] * q[(size_t)j * nr + k];
          yi[k] = xi[k] - t;
        }
      }
    }
  };
  // Projected, Jacobi-preconditioned CG on the energy (DEFL_PCG_matfree.cpp:90-240).
  std::vector<double> v((size_t)nt), res((size_t)nt), z((size_t)nt), pdir((size_t)nt),
      qkp((size_t)nt), dp((size_t)nt, 0.0);
  kp(p0.data(), v.data());
  for (int64_t k = 0; k < nt; ++k) v[k] -= f[k];
  orth(v.data(), res.data());
  const double res0 = std::sqrt(det_dot(nt, res.data(), res.data(), nthreads));
  bool done = kp_fail || prm.itmax <= 0 || res0 < 1e-15;
  double gamma_old = 0.0, de0 = 0.0;
  while (!done) {
    ++iter;
    for (int64_t k = 0; k < nt; ++k) v[k] = dinv[k] * res[k];
    orth(v.data(), z.data());
    const double gamma = det_dot(nt, res.data(), z.data(), nthreads);
    if (iter == 1) {
      pdir = z;
    } else {
      const double beta = gamma / gamma_old;
      for (int64_t k = 0; k < nt; ++k) pdir[k] = z[k] + beta * pdir[k];
    }
    gamma_old = gamma;
    kp(pdir.data(), v.data());
    if (kp_fail) break;  // the external product failed: the result is discarded below
    orth(v.data(), qkp.data());
    const double denom = det_dot(nt, qkp.data(), pdir.data(), nthreads);
    if (std::fabs(denom) < 1e-18) break;
    const double dek = gamma * gamma / denom, alpha = gamma / denom;
    for (int64_t k = 0; k < nt; ++k) {
      dp[k] += alpha * pdir[k];
      res[k] -= alpha * qkp[k];
    }
    if (iter == 1) de0 = dek;
    done = iter == prm.itmax || dek / de0 < prm.en_tol;
  }
  if (kp_fail) {
    err = "emin: the external masked product failed";
    return 5;
  }
  for (int64_t k = 0; k < nt; ++k) p0[k] -= dp[k];
  // Coarse rows back as e_c (Prol_add_Cnodes.cpp); exact zeros dropped as MATLAB's sparse()
  // does when the gateway's output is assembled.
  orp.assign((size_t)n + 1, 0);
  oci.clear();
  ov.clear();
  for (int32_t i = 0; i < n; ++i) {
    if (fc[i] >= 0) {
      oci.push_back(fc[i]);
      ov.push_back(1.0);
    } else {
      for (int32_t k = trp[i]; k < trp[i + 1]; ++k)
        if (p0[k] != 0.0) {
          oci.push_back(tci[k]);
          ov.push_back(p0[k]);
        }
    }
    orp[i + 1] = (int32_t)oci.size();
  }
  return 0;
}

}  // namespace chse

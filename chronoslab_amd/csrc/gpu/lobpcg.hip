// LOBPCG (sources/EigenSolver/lobpcg.m:21-467 with B_orthonormalize.m:1-25, prodB = []
// and no constraint block, as EigDriver.m:202-226 calls it) on the GPU.
//
// The n x k blocks live in HBM as row-major FP64 blocks.  Products with A are block SpMMs,
// the preconditioner is the block V-cycle, and the Gram matrices are tall-skinny products
// reduced in a fixed order.  The Rayleigh-Ritz steps (order <= 3*neig) run in FP64 on the
// host: Cholesky, the generalized symmetric-definite eigenproblem and condition numbers.
#include "chronos_amg.h"
#include "dense.hpp"
#include "runtime.hpp"

using namespace chrrt;

namespace {

struct Lob {
  chr_hier *h;       // preconditioner (may be null)
  const DevCsr *A;   // prodA
  const int *zero;   // stop flag that is never set
  hipStream_t s;
  int n;
  DevBuf part, cdev, gout, idx;
  int nparts = 1, rows_per_part = 1;

  int init() {
    nparts = std::max(1, std::min(1024, (n + 255) / 256));
    rows_per_part = (n + nparts - 1) / nparts;
    int rc;
    if ((rc = part.alloc(sizeof(double) * (size_t)nparts * chr::kBlock))) return rc;
    if ((rc = cdev.alloc(sizeof(double) * 48 * 48))) return rc;
    if ((rc = gout.alloc(sizeof(double) * chr::kBlock))) return rc;
    if ((rc = idx.alloc(sizeof(int) * 64))) return rc;
    return CHR_OK;
  }
  static unsigned grid(int64_t cnt) { return (unsigned)((cnt + chr::kBlock - 1) / chr::kBlock); }

  // out (column-major a x b, host) = X^T Y for row-major n x a and n x b blocks.
  int gram(const double *X, int a, const double *Y, int b, std::vector<double> &out) {
    if (a < 1 || b < 1 || a > 16 || b > 16) return fail(CHR_EARG, "lobpcg: Gram block wider than 16");
    hipLaunchKernelGGL(chr::gram_partial_kernel, dim3(nparts), dim3(chr::kBlock), 0, s, n, a, X, b, Y,
                       rows_per_part, part.as<double>());
    HIPC(hipGetLastError());
    const unsigned rgrid = (unsigned)((a * b * chr::kWave + chr::kBlock - 1) / chr::kBlock);
    hipLaunchKernelGGL(chr::gram_reduce_kernel, dim3(rgrid), dim3(chr::kBlock), 0, s, nparts, a * b,
                       (const double *)part.as<double>(), gout.as<double>());
    HIPC(hipGetLastError());
    std::vector<double> g((size_t)a * b);
    HIPC(hipMemcpyAsync(g.data(), gout.p, sizeof(double) * a * b, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    out.assign((size_t)a * b, 0.0);
    for (int p = 0; p < a; ++p)
      for (int q = 0; q < b; ++q) out[p + (size_t)q * a] = g[(size_t)p * b + q];
    return CHR_OK;
  }
  // Z (n x b) = beta Z + X (n x a) Cm (column-major a x b, host).
  // Small host operands are copied with the stream drained afterwards: the host arrays
  // are temporaries of the Rayleigh-Ritz step, and a copy from pageable memory that is
  // still pending when they change or go away would read the wrong data.
  int upload_small(void *dst, const void *src, size_t bytes) {
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
    return CHR_OK;
  }
  int gemm(const double *X, int a, const double *Cm, int b, double beta, double *Z) {
    if (int rc = upload_small(cdev.p, Cm, sizeof(double) * a * b)) return rc;
    if (a < 1 || b < 1 || a > 48 || b > 48) return fail(CHR_EARG, "lobpcg: block product wider than 48");
    const unsigned ggrid = (unsigned)std::min<int64_t>(2048, ((int64_t)n + chr::kGemmTile - 1) / chr::kGemmTile);
    hipLaunchKernelGGL(chr::block_gemm_kernel, dim3(ggrid), dim3(chr::kBlock), 0, s, n, a,
                       X, b, (const double *)cdev.as<double>(), beta, Z);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  int applyA(const double *X, int k, double *Y) {
    HIPC(launch_op(*A, X, k, chr::EpiStore<double>{Y}, zero, s));
    return CHR_OK;
  }
  // X <- M X (block V-cycle), tmp: n x k scratch.
  int applyM(double *X, int k, double *tmp) {
    if (!h) return CHR_OK;
    HIPC(enqueue_vcycle(*h, X, tmp, h->zero.as<int>(), k));
    HIPC(hipMemcpyAsync(X, tmp, sizeof(double) * (size_t)n * k, hipMemcpyDeviceToDevice, s));
    return CHR_OK;
  }
  int resid(const double *AX, const double *X, const double *lam_host, int k, double *R) {
    if (int rc = upload_small(cdev.p, lam_host, sizeof(double) * k)) return rc;
    hipLaunchKernelGGL(chr::block_resid_kernel, dim3(grid((int64_t)n * k)), dim3(chr::kBlock), 0, s, n, k,
                       AX, X, (const double *)cdev.as<double>(), R);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  int gather(const double *X, int k, const std::vector<int> &cols, double *Y) {
    const int m = (int)cols.size();
    if (int rc = upload_small(idx.p, cols.data(), sizeof(int) * m)) return rc;
    hipLaunchKernelGGL(chr::block_gather_kernel, dim3(grid((int64_t)n * m)), dim3(chr::kBlock), 0, s, n, k,
                       X, m, (const int *)idx.as<int>(), Y);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  int add(const double *X, const double *Y, double *Z, int k) {
    hipLaunchKernelGGL(chr::block_add_kernel, dim3(grid((int64_t)n * k)), dim3(chr::kBlock), 0, s,
                       (int64_t)n * k, X, Y, Z);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  int copy(double *dst, const double *src, int k) {
    HIPC(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n * k, hipMemcpyDeviceToDevice, s));
    return CHR_OK;
  }
  // B_orthonormalize.m with prodB = []: X <- X U^-1 with U^T U = X^T X; returns 1 when
  // cond(X^T X) > 1/(k eps) (the reference's failure), U (column-major upper) otherwise.
  int borth(double *X, int k, std::vector<double> &U, double *tmp, int &failed) {
    std::vector<double> M;
    int rc;
    if ((rc = gram(X, k, X, k, M))) return rc;
    failed = 0;
    if (chrdense::cond_sym(k, M.data()) > 1.0 / ((double)k * 2.220446049250313e-16) ||
        !chrdense::chol_lower(k, M.data())) {
      failed = 1;
      return CHR_OK;
    }
    U.assign((size_t)k * k, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i <= j; ++i) U[i + (size_t)j * k] = M[j + (size_t)i * k];
    std::vector<double> Ui = chrdense::inv_upper(k, U.data());
    if ((rc = gemm(X, k, Ui.data(), k, 0.0, tmp))) return rc;
    return copy(X, tmp, k);
  }
};

}  // namespace

extern "C" int chr_lobpcg(chr_hier *h, chr_mat *Aop, int32_t n, int32_t neig, double *X, int64_t ldx,
                          int32_t largest, int32_t reslambda_check, int32_t itmax, double tol,
                          int32_t restart_control, int32_t ptr_kind, int32_t *niter_out,
                          double *lambda_out, double *resnorm_hist, double *lambda_hist,
                          int32_t *ierr_out) {
  if (!X || neig < 1 || neig > kMaxBlockCols || itmax < 2 || ldx < n || (int64_t)n < 3 * (int64_t)neig)
    return fail(CHR_EARG, "chr_lobpcg: need 1 <= neig <= 16, itmax >= 2, n >= 3*neig, ldx >= n");
  if (!h && !Aop) return fail(CHR_EARG, "chr_lobpcg: need a hierarchy or an operator");
  const DevCsr *A = Aop ? &Aop->A : &h->outer.A;
  if (A->nrows != n || A->ncols != n || (h && h->lev[0].n != n))
    return fail(CHR_EARG, "chr_lobpcg: operator / hierarchy size differs from n");
  int rc;
  hipStream_t s = h ? h->stream : nullptr;
  if (h && (rc = ensure_block_capacity(*h, neig))) return rc;
  DevBuf zero;
  if ((rc = zero.alloc(sizeof(int)))) return rc;
  HIPC(hipMemsetAsync(zero.p, 0, sizeof(int), s));
  Lob L{h, A, zero.as<int>(), s, n};
  if ((rc = L.init())) return rc;
  const size_t blk = sizeof(double) * (size_t)n * neig;
  DevBuf bX, bAX, bP, bAP, bR, bAR, bAAR, bAPa, bAAP, bBest, bT, bT2;
  for (DevBuf *b : {&bX, &bAX, &bP, &bAP, &bR, &bAR, &bAAR, &bAPa, &bAAP, &bBest, &bT, &bT2})
    if ((rc = b->alloc(blk))) return rc;
  double *dX = bX.as<double>(), *dAX = bAX.as<double>(), *dP = bP.as<double>(), *dAP = bAP.as<double>();
  double *dR = bR.as<double>(), *dAR = bAR.as<double>(), *dAAR = bAAR.as<double>();
  double *dAPa = bAPa.as<double>(), *dAAP = bAAP.as<double>(), *dBest = bBest.as<double>();
  double *dT = bT.as<double>(), *dT2 = bT2.as<double>();
  HIPC(hipMemsetAsync(dP, 0, blk, s));
  HIPC(hipMemsetAsync(dAP, 0, blk, s));
  // X0 (column-major) -> row-major.
  const double *xsrc = X;
  int64_t lds = ldx;
  if (ptr_kind == CHR_PTR_HOST) {
    HIPC(copy_cols(dT, 8 * (size_t)n, X, 8 * (size_t)ldx, 8 * (size_t)n, neig, hipMemcpyHostToDevice, s));
    xsrc = dT;
    lds = n;
  }
  hipLaunchKernelGGL(chr::cm_to_rm_kernel, dim3(Lob::grid((int64_t)n * neig)), dim3(chr::kBlock), 0, s, n,
                     neig, xsrc, lds, dX);
  HIPC(hipGetLastError());
  const int64_t H = (int64_t)itmax + 1;
  auto hist = [&](double *hh, int it, int j, double v) {
    if (hh) hh[(int64_t)j * H + it] = v;
  };
  std::vector<double> lam(neig), lam_old(neig, 0.0), relDelta(neig, INFINITY), resn(neig), U, gA, gB, E, w;
  std::vector<int> act;
  int32_t ierr = 0, iter = 0, n_active = neig, failed = 0;
  double smallest = 1e308;
  int restart = 1, forced = 0;
  auto rayleigh_ritz_X = [&]() -> int {  // X, AX <- X E, AX E with E from (X'AX, X'X)
    std::vector<double> xax, xx;
    int r;
    if ((r = L.gram(dX, neig, dAX, neig, xax)) || (r = L.gram(dX, neig, dX, neig, xx))) return r;
    chrdense::symmetrize(neig, xax.data());
    chrdense::symmetrize(neig, xx.data());
    w.resize(neig);
    E.resize((size_t)neig * neig);
    if (!chrdense::sygv(neig, xax.data(), xx.data(), w.data(), E.data())) return 1;
    if ((r = L.gemm(dX, neig, E.data(), neig, 0.0, dT)) || (r = L.copy(dX, dT, neig))) return r;
    if ((r = L.gemm(dAX, neig, E.data(), neig, 0.0, dT)) || (r = L.copy(dAX, dT, neig))) return r;
    for (int j = 0; j < neig; ++j) lam[j] = w[j];
    return 0;
  };
  if ((rc = L.copy(dBest, dX, neig))) return rc;
  if ((rc = L.borth(dX, neig, U, dT, failed))) return rc;
  if (failed) {
    ierr = 1;
  } else {
    if ((rc = L.applyA(dX, neig, dAX))) return rc;
    int r = rayleigh_ritz_X();
    if (r < 0) return r;
    if (r == 1) ierr = 1;
    for (int j = 0; j < neig; ++j) hist(lambda_hist, 0, j, lam[j]);
  }
  while (!ierr && iter < itmax - 1) {
    ++iter;
    if ((rc = L.resid(dAX, dX, lam.data(), neig, dR))) return rc;
    std::vector<double> rr;
    if ((rc = L.gram(dR, neig, dR, neig, rr))) return rc;
    double avg = 0.0;
    for (int j = 0; j < neig; ++j) {
      resn[j] = std::sqrt(rr[j + (size_t)j * neig]) / lam[j];
      hist(resnorm_hist, iter - 1, j, resn[j]);
      avg += resn[j];
      relDelta[j] = std::fabs(lam_old[j] - lam[j]) / lam[j];
      lam_old[j] = lam[j];
    }
    avg /= neig;
    if (avg < smallest) {
      smallest = avg;
      if ((rc = L.copy(dBest, dX, neig))) return rc;
    } else if (avg > std::ldexp(1.0, restart_control) * smallest) {
      forced = 1;
      if ((rc = L.applyA(dX, neig, dAX))) return rc;
    }
    act.clear();
    for (int j = 0; j < neig; ++j) {
      const bool conv = reslambda_check ? (relDelta[j] < tol) : (resn[j] < tol);
      if (!conv) act.push_back(j);
    }
    n_active = (int)act.size();
    if (n_active == 0) break;
    const int na = n_active;
    if ((rc = L.gather(dR, neig, act, dAR))) return rc;
    if (iter > 1) {
      if ((rc = L.gather(dP, neig, act, dAPa)) || (rc = L.gather(dAP, neig, act, dAAP))) return rc;
    }
    if ((rc = L.applyM(dAR, na, dT))) return rc;
    // AR <- AR - X (X^T AR)
    std::vector<double> xar;
    if ((rc = L.gram(dX, neig, dAR, na, xar))) return rc;
    for (double &v : xar) v = -v;
    if ((rc = L.gemm(dX, neig, xar.data(), na, 1.0, dAR))) return rc;
    if ((rc = L.borth(dAR, na, U, dT, failed))) return rc;
    if (failed) {
      ierr = 2;
      break;
    }
    if ((rc = L.applyA(dAR, na, dAAR))) return rc;
    if (iter > 1) {
      if ((rc = L.borth(dAPa, na, U, dT, failed))) return rc;
      if (failed) {
        restart = 1;
      } else {
        std::vector<double> Ui = chrdense::inv_upper(na, U.data());
        if ((rc = L.gemm(dAAP, na, Ui.data(), na, 0.0, dT)) || (rc = L.copy(dAAP, dT, na))) return rc;
        restart = forced;
      }
    }
    const int m = restart ? neig + na : neig + 2 * na;
    gA.assign((size_t)m * m, 0.0);
    gB.assign((size_t)m * m, 0.0);
    auto put = [&](std::vector<double> &M, int r0, int c0, int rows, int cols, const std::vector<double> &src,
                   bool sym) {
      for (int jj = 0; jj < cols; ++jj)
        for (int ii = 0; ii < rows; ++ii) {
          const double v = src[ii + (size_t)jj * rows];
          M[(r0 + ii) + (size_t)(c0 + jj) * m] = v;
          if (sym) M[(c0 + jj) + (size_t)(r0 + ii) * m] = v;
        }
    };
    std::vector<double> g;
    if ((rc = L.gram(dX, neig, dAX, neig, g))) return rc;
    chrdense::symmetrize(neig, g.data());
    put(gA, 0, 0, neig, neig, g, false);
    if ((rc = L.gram(dX, neig, dAAR, na, g))) return rc;
    put(gA, 0, neig, neig, na, g, true);
    if ((rc = L.gram(dAR, na, dAAR, na, g))) return rc;
    chrdense::symmetrize(na, g.data());
    put(gA, neig, neig, na, na, g, false);
    if ((rc = L.gram(dX, neig, dX, neig, g))) return rc;
    chrdense::symmetrize(neig, g.data());
    put(gB, 0, 0, neig, neig, g, false);
    if ((rc = L.gram(dX, neig, dAR, na, g))) return rc;
    put(gB, 0, neig, neig, na, g, true);
    if ((rc = L.gram(dAR, na, dAR, na, g))) return rc;
    chrdense::symmetrize(na, g.data());
    put(gB, neig, neig, na, na, g, false);
    if (!restart) {
      const int c2 = neig + na;
      if ((rc = L.gram(dX, neig, dAAP, na, g))) return rc;
      put(gA, 0, c2, neig, na, g, true);
      if ((rc = L.gram(dAR, na, dAAP, na, g))) return rc;
      put(gA, neig, c2, na, na, g, true);
      if ((rc = L.gram(dAPa, na, dAAP, na, g))) return rc;
      chrdense::symmetrize(na, g.data());
      put(gA, c2, c2, na, na, g, false);
      if ((rc = L.gram(dX, neig, dAPa, na, g))) return rc;
      put(gB, 0, c2, neig, na, g, true);
      if ((rc = L.gram(dAR, na, dAPa, na, g))) return rc;
      put(gB, neig, c2, na, na, g, true);
      if ((rc = L.gram(dAPa, na, dAPa, na, g))) return rc;
      chrdense::symmetrize(na, g.data());
      put(gB, c2, c2, na, na, g, false);
    }
    w.resize(m);
    E.resize((size_t)m * m);
    if (!chrdense::sygv(m, gA.data(), gB.data(), w.data(), E.data())) {
      ierr = 2;
      break;
    }
    std::vector<double> EX((size_t)neig * neig), ER((size_t)na * neig), EP((size_t)na * neig);
    for (int j = 0; j < neig; ++j) {
      const int src = largest ? m - 1 - j : j;
      lam[j] = w[src];
      hist(lambda_hist, iter, j, lam[j]);
      const double *e = E.data() + (size_t)src * m;
      for (int i = 0; i < neig; ++i) EX[i + (size_t)j * neig] = e[i];
      for (int i = 0; i < na; ++i) ER[i + (size_t)j * na] = e[neig + i];
      if (!restart)
        for (int i = 0; i < na; ++i) EP[i + (size_t)j * na] = e[neig + na + i];
    }
    if ((rc = L.gemm(dAR, na, ER.data(), neig, 0.0, dP))) return rc;
    if ((rc = L.gemm(dAAR, na, ER.data(), neig, 0.0, dAP))) return rc;
    if (!restart) {
      if ((rc = L.gemm(dAPa, na, EP.data(), neig, 1.0, dP))) return rc;
      if ((rc = L.gemm(dAAP, na, EP.data(), neig, 1.0, dAP))) return rc;
    }
    if ((rc = L.gemm(dX, neig, EX.data(), neig, 0.0, dT)) || (rc = L.add(dP, dT, dX, neig))) return rc;
    if ((rc = L.gemm(dAX, neig, EX.data(), neig, 0.0, dT2)) || (rc = L.add(dAP, dT2, dAX, neig))) return rc;
  }
  if (ierr != 1) {
    if (n_active > 0) {
      ++iter;
      if ((rc = L.resid(dAX, dX, lam.data(), neig, dR))) return rc;
      std::vector<double> rr;
      if ((rc = L.gram(dR, neig, dR, neig, rr))) return rc;
      for (int j = 0; j < neig; ++j) {
        resn[j] = std::sqrt(rr[j + (size_t)j * neig]) / lam[j];
        hist(resnorm_hist, iter - 1, j, resn[j]);
        hist(lambda_hist, iter - 1, j, lam[j]);
      }
    }
    double mx = 0.0;
    for (int j = 0; j < neig; ++j) mx = std::max(mx, reslambda_check ? relDelta[j] : resn[j]);
    if (mx > tol) {
      ierr = -1;
      if ((rc = L.copy(dX, dBest, neig))) return rc;
    }
    // Final Rayleigh-Ritz without preconditioning (lobpcg.m:420-459).
    ++iter;
    if ((rc = L.applyA(dX, neig, dAX))) return rc;
    int r = rayleigh_ritz_X();
    if (r < 0) return r;
    if (r == 1) {
      ierr = 1;
    } else {
      if ((rc = L.resid(dAX, dX, lam.data(), neig, dR))) return rc;
      std::vector<double> rr;
      if ((rc = L.gram(dR, neig, dR, neig, rr))) return rc;
      for (int j = 0; j < neig; ++j) {
        resn[j] = std::sqrt(rr[j + (size_t)j * neig]) / lam[j];
        hist(resnorm_hist, iter - 1, j, resn[j]);
        hist(lambda_hist, iter - 1, j, lam[j]);
      }
    }
  }
  // Row-major X -> caller's column-major block.
  double *xdst = ptr_kind == CHR_PTR_HOST ? dT : X;
  const int64_t ldd = ptr_kind == CHR_PTR_HOST ? n : ldx;
  hipLaunchKernelGGL(chr::rm_to_cm_kernel, dim3(Lob::grid((int64_t)n * neig)), dim3(chr::kBlock), 0, s, n, neig,
                     (const double *)dX, xdst, ldd);
  HIPC(hipGetLastError());
  if (ptr_kind == CHR_PTR_HOST)
    HIPC(copy_cols(X, 8 * (size_t)ldx, dT, 8 * (size_t)n, 8 * (size_t)n, neig, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (lambda_out) std::copy(lam.begin(), lam.end(), lambda_out);
  if (niter_out) *niter_out = iter - 1;
  if (ierr_out) *ierr_out = ierr;
  return CHR_OK;
}

// Smoother damping of the adaptive FSAI smoother on the GPU:
//   omega = min(1, 1.9 / lambda_max(G A G^T))
// (sources/Preconditioner/AMG/smoother/smoother.m:72-81, where lambda_max comes from
// eigs(..., 'la') with tolerance 5e-4).  lambda_max is found by Lanczos with full
// reorthogonalisation, as the host setup does (chs_fsai_omega): start vector U[-1,1) from
// the shared counter RNG (seed, stream 7), normalised; per step w = G (A (G^T v)), alpha =
// w.v, two classical Gram-Schmidt passes against every Lanczos vector, beta = |w|; the
// tridiagonal's largest eigenvalue by Sturm bisection on the host; stop when it changes by
// at most tol relative after step 4, or when beta vanishes.  The host uses modified
// Gram-Schmidt, so the two lambdas agree to rounding (the tests bound the difference).
//
// Kernels: the three FP64 products use the solve phase's stream SpMV; every dot product is
// a fixed-partition partial sum per workgroup reduced in workgroup order (deterministic),
// and the Lanczos vectors live in one n x (maxit+1) column-major block on the device.
#include "chronos_amg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "rng.hpp"
#include "runtime.hpp"

namespace chr {

constexpr int kLzParts = 512;  // workgroups of every dot product (fixed: deterministic sums)

// part[c * kLzParts + b] = sum over rows of workgroup b of X[:, c] . w  (grid: kLzParts x k)
static __global__ __launch_bounds__(kBlock) void lz_dots_kernel(int n, const double *__restrict__ X,
                                                                 int64_t ldx, const double *__restrict__ w,
                                                                 double *__restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  const int c = blockIdx.y;
  const int per = (n + kLzParts - 1) / kLzParts;
  const int i0 = blockIdx.x * per, i1 = min(n, i0 + per);
  const double *x = X + (size_t)c * ldx;
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += kBlock) s += x[i] * w[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[(size_t)c * kLzParts + blockIdx.x] = s;
}

// d[c] = sum_b part[c * kLzParts + b] in workgroup order (one thread per column).
static __global__ __launch_bounds__(kBlock) void lz_reduce_kernel(int k, const double *__restrict__ part,
                                                                  double *__restrict__ d) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= k) return;
  double s = 0.0;
  for (int b = 0; b < kLzParts; ++b) s += part[(size_t)c * kLzParts + b];
  d[c] = s;
}

// w -= X[:, 0:k] d
static __global__ __launch_bounds__(kBlock) void lz_orth_kernel(int n, int k, const double *__restrict__ X,
                                                                int64_t ldx, const double *__restrict__ d,
                                                                double *__restrict__ w) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int c = 0; c < k; ++c) s -= d[c] * X[(size_t)c * ldx + i];
  w[i] = s;
}

// y = x / nrm
static __global__ __launch_bounds__(kBlock) void lz_scale_kernel(int n, const double *__restrict__ x,
                                                                 double nrm, double *__restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = x[i] / nrm;
}

}  // namespace chr

using chrrt::DevBuf;
using chrrt::DevCsr;
using chrrt::fail;

namespace {

int sturm_count(int k, const double *a, const double *b, double x) {
  int c = 0;
  double q = a[0] - x;
  if (q < 0) ++c;
  for (int i = 1; i < k; ++i) {
    const double qq = q == 0.0 ? 1e-300 : q;
    q = a[i] - x - b[i - 1] * b[i - 1] / qq;
    if (q < 0) ++c;
  }
  return c;
}

// Largest eigenvalue of the symmetric tridiagonal (a, b) by bisection on Sturm counts.
double tridiag_max(int k, const double *a, const double *b) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i < k - 1 ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (sturm_count(k, a, b, mid) >= k) hi = mid;
    else lo = mid;
  }
  return hi;
}

}  // namespace

extern "C" int chr_fsai_omega(const chr_csr *A, const chr_csr *G, uint64_t seed, double *omega,
                              double *lambda) {
  if (!A || !G || A->nrows != A->ncols || G->nrows != A->nrows || G->ncols != A->ncols)
    return fail(CHR_EARG, "chr_fsai_omega: A and G must be square of one size");
  int rc;
  if ((rc = chrrt::check_csr(*A, "chr_fsai_omega A")) || (rc = chrrt::check_csr(*G, "chr_fsai_omega G")))
    return rc;
  const int32_t n = A->nrows;
  if (n == 0) return fail(CHR_EARG, "chr_fsai_omega: empty matrix");
  const int maxit = std::max(2, std::min<int>(40, n));
  const double tol = 1e-7;
  hipStream_t s = nullptr;
  HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevCsr dA, dG, dGT;
  std::vector<int32_t> trp, tci;
  std::vector<double> tv;
  chrrt::transpose_host(*G, trp, tci, tv);
  if ((rc = dA.upload(*A, false)) || (rc = dG.upload(*G, false)) ||
      (rc = dGT.upload(n, n, trp.data(), tci.data(), tv.data(), false)))
    return rc;
  const int64_t ldv = n;
  DevBuf V, w, t1, t2, part, dd, zero;
  if ((rc = V.alloc(8 * (size_t)n * (maxit + 1))) || (rc = w.alloc(8 * (size_t)n)) ||
      (rc = t1.alloc(8 * (size_t)n)) || (rc = t2.alloc(8 * (size_t)n)) ||
      (rc = part.alloc(8 * (size_t)chr::kLzParts * (maxit + 1))) || (rc = dd.alloc(8 * (size_t)(maxit + 1))) ||
      (rc = zero.alloc(sizeof(int))))
    return rc;
  HIPC(hipMemsetAsync(zero.p, 0, sizeof(int), s));
  // Start vector, normalised on the host (as the host setup does).
  std::vector<double> v0(n);
  double nv = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    v0[i] = -1.0 + 2.0 * chrng::u01(seed, 7, (uint64_t)i);
    nv += v0[i] * v0[i];
  }
  nv = std::sqrt(nv);
  for (double &t : v0) t /= nv;
  HIPC(hipMemcpyAsync(V.p, v0.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s));
  const unsigned gn = (unsigned)((n + chr::kBlock - 1) / chr::kBlock);
  double *Vp = V.as<double>(), *wp = w.as<double>(), *pp = part.as<double>(), *dp = dd.as<double>();
  const int *zp = zero.as<int>();
  // dd[0..k) = X[:, 0:k] . y
  auto dots = [&](const double *X, int k, const double *y) -> hipError_t {
    hipLaunchKernelGGL(chr::lz_dots_kernel, dim3(chr::kLzParts, k), dim3(chr::kBlock), 0, s, n, X, ldv, y, pp);
    hipLaunchKernelGGL(chr::lz_reduce_kernel, dim3((k + chr::kBlock - 1) / chr::kBlock), dim3(chr::kBlock), 0, s,
                       k, (const double *)pp, dp);
    return hipGetLastError();
  };
  std::vector<double> a, b;
  double lam = 0.0, lam_old = 0.0;
  for (int j = 0; j < maxit; ++j) {
    const double *vj = Vp + (size_t)j * ldv;
    HIPC(chrrt::launch_spmv(dGT, vj, chr::EpiStore<double>{t1.as<double>()}, zp, s));
    HIPC(chrrt::launch_spmv(dA, t1.as<double>(), chr::EpiStore<double>{t2.as<double>()}, zp, s));
    HIPC(chrrt::launch_spmv(dG, t2.as<double>(), chr::EpiStore<double>{wp}, zp, s));
    double alpha = 0.0;
    HIPC(dots(vj, 1, wp));
    HIPC(hipMemcpyAsync(&alpha, dp, 8, hipMemcpyDeviceToHost, s));
    // Full reorthogonalisation, two classical Gram-Schmidt passes.
    for (int pass = 0; pass < 2; ++pass) {
      HIPC(dots(Vp, j + 1, wp));
      hipLaunchKernelGGL(chr::lz_orth_kernel, dim3(gn), dim3(chr::kBlock), 0, s, n, j + 1, (const double *)Vp, ldv,
                         (const double *)dp, wp);
      HIPC(hipGetLastError());
    }
    double bb = 0.0;
    HIPC(dots(wp, 1, wp));
    HIPC(hipMemcpyAsync(&bb, dp, 8, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    a.push_back(alpha);
    const double beta = std::sqrt(bb);
    lam = tridiag_max((int)a.size(), a.data(), b.data());
    if (j > 3 && std::fabs(lam - lam_old) <= tol * std::fabs(lam)) break;
    lam_old = lam;
    if (beta <= 1e-14 * std::fabs(lam)) break;
    b.push_back(beta);
    hipLaunchKernelGGL(chr::lz_scale_kernel, dim3(gn), dim3(chr::kBlock), 0, s, n, (const double *)wp, beta,
                       Vp + (size_t)(j + 1) * ldv);
    HIPC(hipGetLastError());
  }
  HIPC(hipStreamSynchronize(s));
  if (lambda) *lambda = lam;
  if (omega) *omega = std::min(1.0, 1.9 / lam);
  return CHR_OK;
}

// chs_omega_fn adapter (chronos_setup.h): lets chs_build take omega from the GPU.
extern "C" int chr_setup_omega(void *ctx, int32_t n, const int32_t *rowptr_A, const int32_t *col_A,
                               const double *val_A, const int32_t *rowptr_G, const int32_t *col_G,
                               const double *val_G, uint64_t seed, double *omega, double *lambda) {
  (void)ctx;
  const chr_csr A{n, n, rowptr_A, col_A, val_A}, G{n, n, rowptr_G, col_G, val_G};
  return chr_fsai_omega(&A, &G, seed, omega, lambda);
}

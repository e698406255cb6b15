// EMIN's CG operator on the GPU (set-up): y = (A X) restricted to the interpolation pattern,
// X and y on the pattern (the masked product of KP_spmat.cpp, sources/Preconditioner/AMG/
// prolong/EMIN/Cpp_EMIN/KP_spmat.cpp:1-60, which the host restatement follows).
//
// One lane per pattern entry (i, c): it walks row i of A in order and, for each a_im, looks
// c up in row m of the pattern (columns ascending, binary search) and adds a_im x_(m,c).  The
// host product accumulates every entry over A's row in the same order with the product
// rounded before the add, so both give the same bits (this library is built with
// -ffp-contract=off).  A and the pattern are uploaded once per level (open); each apply
// moves X in and y out.
#include "chronos_amg.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

#include "runtime.hpp"

using chrrt::fail;

namespace {

constexpr int kB = 256;

__global__ __launch_bounds__(kB) void emin_rows_kernel(int n, const int *__restrict__ trp, int *__restrict__ row) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  for (int k = trp[i]; k < trp[i + 1]; ++k) row[k] = i;
}

__global__ __launch_bounds__(kB) void emin_kp_kernel(int64_t nt, const int *__restrict__ row,
                                                     const int *__restrict__ tci, const int *__restrict__ trp,
                                                     const int *__restrict__ arp, const int *__restrict__ aci,
                                                     const double *__restrict__ av, const double *__restrict__ x,
                                                     double *__restrict__ y) {
  const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (k >= nt) return;
  const int i = row[k], c = tci[k];
  double s = 0.0;
  for (int ka = arp[i]; ka < arp[i + 1]; ++ka) {
    const int m = aci[ka];
    int lo = trp[m], hi = trp[m + 1];
    const int end = hi;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tci[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    if (lo < end && tci[lo] == c) s = s + av[ka] * x[lo];
  }
  y[k] = s;
}

struct EminKp {
  int32_t n = 0;
  int64_t nt = 0;
  hipStream_t s = nullptr;
  DevBuf arp, aci, av, trp, tci, row, x, y;
  ~EminKp() {
    if (s) (void)hipStreamDestroy(s);
  }
};

}  // namespace

extern "C" int chr_setup_emin_open(void *ctx, int32_t n, int32_t nc, const int32_t *rowptr_A,
                                   const int32_t *col_A, const double *val_A, const int32_t *rowptr_T,
                                   const int32_t *col_T, void **handle) {
  (void)ctx;
  (void)nc;
  if (!rowptr_A || !col_A || !val_A || !rowptr_T || !col_T || !handle || n < 0)
    return fail(CHR_EARG, "chr_setup_emin_open: arguments");
  auto h = std::make_unique<EminKp>();
  h->n = n;
  h->nt = rowptr_T[n];
  const int64_t nnzA = rowptr_A[n];
  HIPC(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  int rc;
  if ((rc = h->arp.alloc(4 * ((size_t)n + 1))) || (rc = h->aci.alloc(4 * (size_t)nnzA + 4)) ||
      (rc = h->av.alloc(8 * (size_t)nnzA + 8)) || (rc = h->trp.alloc(4 * ((size_t)n + 1))) ||
      (rc = h->tci.alloc(4 * (size_t)h->nt + 4)) || (rc = h->row.alloc(4 * (size_t)h->nt + 4)) ||
      (rc = h->x.alloc(8 * (size_t)h->nt + 8)) || (rc = h->y.alloc(8 * (size_t)h->nt + 8)))
    return rc;
  HIPC(hipMemcpyAsync(h->arp.p, rowptr_A, 4 * ((size_t)n + 1), hipMemcpyHostToDevice, h->s));
  HIPC(hipMemcpyAsync(h->aci.p, col_A, 4 * (size_t)nnzA, hipMemcpyHostToDevice, h->s));
  HIPC(hipMemcpyAsync(h->av.p, val_A, 8 * (size_t)nnzA, hipMemcpyHostToDevice, h->s));
  HIPC(hipMemcpyAsync(h->trp.p, rowptr_T, 4 * ((size_t)n + 1), hipMemcpyHostToDevice, h->s));
  HIPC(hipMemcpyAsync(h->tci.p, col_T, 4 * (size_t)h->nt, hipMemcpyHostToDevice, h->s));
  if (n > 0)
    hipLaunchKernelGGL(emin_rows_kernel, dim3((unsigned)((n + kB - 1) / kB)), dim3(kB), 0, h->s, n,
                       (const int *)h->trp.as<int>(), h->row.as<int>());
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(h->s));
  *handle = h.release();
  return CHR_OK;
}

extern "C" int chr_setup_emin_apply(void *handle, const double *x, double *y) {
  auto *h = static_cast<EminKp *>(handle);
  if (!h || !x || !y) return fail(CHR_EARG, "chr_setup_emin_apply: arguments");
  if (h->nt == 0) return CHR_OK;
  HIPC(hipMemcpyAsync(h->x.p, x, 8 * (size_t)h->nt, hipMemcpyHostToDevice, h->s));
  hipLaunchKernelGGL(emin_kp_kernel, dim3((unsigned)((h->nt + kB - 1) / kB)), dim3(kB), 0, h->s, h->nt,
                     (const int *)h->row.as<int>(), (const int *)h->tci.as<int>(), (const int *)h->trp.as<int>(),
                     (const int *)h->arp.as<int>(), (const int *)h->aci.as<int>(),
                     (const double *)h->av.as<double>(), (const double *)h->x.as<double>(), h->y.as<double>());
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(y, h->y.p, 8 * (size_t)h->nt, hipMemcpyDeviceToHost, h->s));
  HIPC(hipStreamSynchronize(h->s));
  return CHR_OK;
}

extern "C" void chr_setup_emin_close(void *handle) { delete static_cast<EminKp *>(handle); }

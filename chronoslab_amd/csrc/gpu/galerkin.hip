// Galerkin coarse operator A_c = P^T (A P) on the GPU, equal bit for bit to the host set-up's
// product (sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:277, `Anext = Pf'*(A*Pf)`).
//
// The host forms both products row by row (Gustavson): output row i accumulates, for each
// entry (i, k) of the left factor in column order, the products l_ik * r_kj of row k of the
// right factor into an accumulator per column j, the first product of a column stored and
// the later ones added, each product rounded before the add; the row is then sorted by
// column and exact zeros are dropped.  P^T is P's transpose with every row in ascending
// column order.
//
// Here one workgroup builds one output row in an LDS hash table (column -> accumulator).
// The entries (i, k) are taken one at a time, in column order; the entries of row k of the
// right factor are spread over the workgroup's lanes.  They are distinct columns, so no two
// lanes touch one accumulator in a round, and rounds are separated by a barrier, so every
// accumulator sees its products in the host's order.  The table is then compacted, ranked by
// column (a count of smaller keys) and written without its zeros.  A symbolic pass (the
// same hashing, keys only) sizes the rows first.  Tables have 128, 512, 2048 or 8192 slots
// and are filled to at most half: rows of at most 64 products start in the smallest, the
// others in the next; a row that does not fit moves to the next size, and a row that fits
// none makes the call decline (the host then forms the product).  Transposes are stable radix sorts of the column indices (hipCUB), so
// every transposed row lists its entries in ascending row order, as the host's does.
#include "chronos_amg.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"

using chrrt::DevBuf;
using chrrt::fail;

namespace chr {

__device__ __forceinline__ int gl_insert(int *key, int T, int j, bool *fresh) {
  unsigned h = ((unsigned)j * 2654435761u) & (unsigned)(T - 1);
  for (int probe = 0; probe < T; ++probe) {
    const int k = key[h];
    if (k == j) {
      *fresh = false;
      return (int)h;
    }
    if (k == -1) {
      const int old = atomicCAS(&key[h], -1, j);
      if (old == -1) {
        *fresh = true;
        return (int)h;
      }
      if (old == j) {
        *fresh = false;
        return (int)h;
      }
    }
    h = (h + 1) & (unsigned)(T - 1);
  }
  return -1;
}

// One output row of C = L R per workgroup of NT threads, for the rows listed in rows[].
// Symbolic (NUMERIC false): cnt[row] = number of distinct columns.  Numeric: the row's
// nonzeros sorted by column at out_col/out_val[off[row] ...], cnt[row] = their number.
// ovf[row] = 1 when the row has more than T/2 distinct columns (nothing else is written).
template <int T, int NT, bool NUMERIC>
__global__ __launch_bounds__(NT) void spgemm_row_kernel(
    int nrows_bin, const int *__restrict__ rows, const int *__restrict__ lrp, const int *__restrict__ lci,
    const double *__restrict__ lv, const int *__restrict__ rrp, const int *__restrict__ rci,
    const double *__restrict__ rv, const int64_t *__restrict__ off, int *__restrict__ out_col,
    double *__restrict__ out_val, int *__restrict__ cnt, int *__restrict__ ovf) {
  extern __shared__ __align__(16) char smem[];
  int *key = reinterpret_cast<int *>(smem);                 // T
  int *misc = key + T;                                      // [0] distinct, [1] overflow, [2] compacted
  double *acc = reinterpret_cast<double *>(misc + 4);       // T (numeric)
  int *ck = reinterpret_cast<int *>(acc + (NUMERIC ? T : 0));  // T/2 (numeric)
  double *cv = reinterpret_cast<double *>(ck + (NUMERIC ? T / 2 : 0));  // T/2 (numeric)
  if ((int)blockIdx.x >= nrows_bin) return;
  const int row = rows[blockIdx.x];
  const int t = threadIdx.x;
  for (int s = t; s < T; s += NT) key[s] = -1;
  if (t < 4) misc[t] = 0;
  __syncthreads();
  const int e0 = lrp[row], e1 = lrp[row + 1];
  for (int e = e0; e < e1; ++e) {
    const int k = lci[e];
    const double a = NUMERIC ? lv[e] : 0.0;
    const int f0 = rrp[k], f1 = rrp[k + 1];
    for (int base = f0; base < f1; base += NT) {
      const int f = base + t;
      if (f < f1) {
        bool fresh = false;
        const int slot = gl_insert(key, T, rci[f], &fresh);
        if (slot < 0) {
          misc[1] = 1;
        } else {
          if (NUMERIC) {
            const double p = a * rv[f];
            if (fresh) acc[slot] = p;
            else acc[slot] += p;
          }
          if (fresh) atomicAdd(&misc[0], 1);
        }
      }
      __syncthreads();
      const bool stop = misc[1] || misc[0] > T / 2;
      __syncthreads();
      if (stop) {
        if (t == 0) ovf[row] = 1;
        return;
      }
    }
  }
  const int u = misc[0];
  if (!NUMERIC) {
    if (t == 0) cnt[row] = u;
    return;
  }
  // Compaction (any order: the ranks below fix the output order).
  for (int s = t; s < T; s += NT)
    if (key[s] != -1) {
      const int q = atomicAdd(&misc[2], 1);
      ck[q] = key[s];
      cv[q] = acc[s];
    }
  __syncthreads();
  int nz = 0;
  for (int q = t; q < u; q += NT) {
    const int kq = ck[q];
    const double vq = cv[q];
    int rank = 0;
    for (int p = 0; p < u; ++p) rank += (ck[p] < kq && cv[p] != 0.0) ? 1 : 0;
    if (vq != 0.0) {
      out_col[off[row] + rank] = kq;
      out_val[off[row] + rank] = vq;
      ++nz;
    }
  }
  if (nz) atomicAdd(&misc[3], nz);
  __syncthreads();
  if (t == 0) cnt[row] = misc[3];
}

// prod[i] = number of products of output row i (sum of the right factor's row lengths).
__global__ __launch_bounds__(kBlock) void spgemm_products_kernel(int n, const int *__restrict__ lrp,
                                                                 const int *__restrict__ lci,
                                                                 const int *__restrict__ rrp,
                                                                 int64_t *__restrict__ prod) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  int64_t s = 0;
  for (int e = lrp[i]; e < lrp[i + 1]; ++e) s += rrp[lci[e] + 1] - rrp[lci[e]];
  prod[i] = s;
}

// Moves each row's cnt[row] entries from the padded offset off[row] to rp[row].
__global__ __launch_bounds__(kBlock) void spgemm_compact_kernel(int n, const int64_t *__restrict__ off,
                                                                const int *__restrict__ rp,
                                                                const int *__restrict__ pc,
                                                                const double *__restrict__ pv,
                                                                int *__restrict__ col,
                                                                double *__restrict__ val) {
  const int row = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x % kWave;
  if (row >= n) return;
  const int len = rp[row + 1] - rp[row];
  for (int q = lane; q < len; q += kWave) {
    col[rp[row] + q] = pc[off[row] + q];
    val[rp[row] + q] = pv[off[row] + q];
  }
}

__global__ __launch_bounds__(kBlock) void row_index_kernel(int n, const int *__restrict__ rp,
                                                           int *__restrict__ row_of) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  for (int e = rp[i]; e < rp[i + 1]; ++e) row_of[e] = i;
}

__global__ __launch_bounds__(kBlock) void iota_kernel(int64_t n, int *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = (int)i;
}

__global__ __launch_bounds__(kBlock) void transpose_gather_kernel(int64_t nnz, const int *__restrict__ perm,
                                                                  const int *__restrict__ row_of,
                                                                  const double *__restrict__ v,
                                                                  int *__restrict__ ciT,
                                                                  double *__restrict__ vT) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nnz) return;
  const int e = perm[t];
  ciT[t] = row_of[e];
  vT[t] = v[e];
}

__global__ __launch_bounds__(kBlock) void col_count_kernel(int64_t nnz, const int *__restrict__ ci,
                                                           int *__restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < nnz) atomicAdd(&cnt[ci[t]], 1);
}

}  // namespace chr

namespace {

// A CSR matrix in device memory.
struct DCsr {
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  DevBuf rp, ci, v;
};

unsigned grid_for(int64_t n) { return (unsigned)((n + chr::kBlock - 1) / chr::kBlock); }

int upload(const chr_csr &M, DCsr &D, hipStream_t s) {
  int rc;
  D.nrows = M.nrows;
  D.ncols = M.ncols;
  D.nnz = M.rowptr[M.nrows];
  if ((rc = D.rp.alloc(4 * ((size_t)M.nrows + 1))) || (rc = D.ci.alloc(4 * std::max<int64_t>(D.nnz, 1))) ||
      (rc = D.v.alloc(8 * std::max<int64_t>(D.nnz, 1))))
    return rc;
  HIPC(hipMemcpyAsync(D.rp.p, M.rowptr, 4 * ((size_t)M.nrows + 1), hipMemcpyHostToDevice, s));
  if (D.nnz) {
    HIPC(hipMemcpyAsync(D.ci.p, M.col, 4 * (size_t)D.nnz, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(D.v.p, M.val, 8 * (size_t)D.nnz, hipMemcpyHostToDevice, s));
  }
  return CHR_OK;
}

// Stable transpose: a radix sort of the column indices carrying the entry numbers.
int transpose(const DCsr &M, DCsr &T, hipStream_t s) {
  int rc;
  T.nrows = M.ncols;
  T.ncols = M.nrows;
  T.nnz = M.nnz;
  if (M.nnz > INT_MAX) return 1;
  const int nnz = (int)M.nnz;
  DevBuf keys_out, perm_in, perm_out, row_of, cnt, tmp;
  if ((rc = T.rp.alloc(4 * ((size_t)T.nrows + 1))) || (rc = T.ci.alloc(4 * std::max(nnz, 1))) ||
      (rc = T.v.alloc(8 * std::max(nnz, 1))) || (rc = keys_out.alloc(4 * std::max(nnz, 1))) ||
      (rc = perm_in.alloc(4 * std::max(nnz, 1))) || (rc = perm_out.alloc(4 * std::max(nnz, 1))) ||
      (rc = row_of.alloc(4 * std::max(nnz, 1))) || (rc = cnt.alloc(4 * ((size_t)T.nrows + 1))))
    return rc;
  HIPC(hipMemsetAsync(cnt.p, 0, cnt.bytes, s));
  if (nnz > 0) {
    hipLaunchKernelGGL(chr::iota_kernel, dim3(grid_for(nnz)), dim3(chr::kBlock), 0, s, (int64_t)nnz, perm_in.as<int>());
    hipLaunchKernelGGL(chr::row_index_kernel, dim3(grid_for(M.nrows)), dim3(chr::kBlock), 0, s, M.nrows,
                       M.rp.as<int>(), row_of.as<int>());
    hipLaunchKernelGGL(chr::col_count_kernel, dim3(grid_for(nnz)), dim3(chr::kBlock), 0, s, (int64_t)nnz,
                       M.ci.as<int>(), cnt.as<int>());
    HIPC(hipGetLastError());
    int bits = 1;
    while (bits < 31 && (1LL << bits) < (int64_t)M.ncols) ++bits;
    size_t tb = 0;
    HIPC(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, M.ci.as<int>(), keys_out.as<int>(), perm_in.as<int>(),
                                            perm_out.as<int>(), nnz, 0, bits, s));
    if ((rc = tmp.alloc(std::max<size_t>(tb, 1)))) return rc;
    HIPC(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, M.ci.as<int>(), keys_out.as<int>(), perm_in.as<int>(),
                                            perm_out.as<int>(), nnz, 0, bits, s));
    hipLaunchKernelGGL(chr::transpose_gather_kernel, dim3(grid_for(nnz)), dim3(chr::kBlock), 0, s, (int64_t)nnz,
                       perm_out.as<int>(), row_of.as<int>(), M.v.as<double>(), T.ci.as<int>(), T.v.as<double>());
    HIPC(hipGetLastError());
  }
  // Row pointers: exclusive scan of the column counts (host, T.nrows + 1 entries).
  std::vector<int32_t> c((size_t)T.nrows + 1, 0), rp((size_t)T.nrows + 1, 0);
  HIPC(hipMemcpyAsync(c.data(), cnt.p, 4 * (size_t)T.nrows, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  for (int32_t j = 0; j < T.nrows; ++j) rp[j + 1] = rp[j] + c[j];
  HIPC(hipMemcpyAsync(T.rp.p, rp.data(), 4 * rp.size(), hipMemcpyHostToDevice, s));
  HIPC(hipStreamSynchronize(s));
  return CHR_OK;
}

constexpr int kBins = 4;
constexpr int kBinT[kBins] = {128, 512, 2048, 8192};

size_t bin_lds(int b, bool numeric) {
  const size_t T = (size_t)kBinT[b];
  return 4 * T + 16 + (numeric ? 8 * T + 4 * (T / 2) + 8 * (T / 2) : 0);
}

template <int T, int NT, bool NUMERIC>
int launch_bin(int nb, const int *rows, const DCsr &L, const DCsr &R, const int64_t *off, int *oc, double *ov,
               int *cnt, int *ovf, size_t lds, hipStream_t s) {
  if (lds > 65536)
    HIPC(hipFuncSetAttribute((const void *)chr::spgemm_row_kernel<T, NT, NUMERIC>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  for (int r0 = 0; r0 < nb; r0 += (1 << 20)) {
    const int c = std::min(nb - r0, 1 << 20);
    hipLaunchKernelGGL((chr::spgemm_row_kernel<T, NT, NUMERIC>), dim3(c), dim3(NT), lds, s, c, rows + r0,
                       L.rp.as<int>(), L.ci.as<int>(), L.v.as<double>(), R.rp.as<int>(), R.ci.as<int>(),
                       R.v.as<double>(), off, oc, ov, cnt, ovf);
    HIPC(hipGetLastError());
  }
  return CHR_OK;
}

int launch_any(int b, bool numeric, int nb, const int *rows, const DCsr &L, const DCsr &R, const int64_t *off,
               int *oc, double *ov, int *cnt, int *ovf, hipStream_t s) {
  const size_t lds = bin_lds(b, numeric);
  switch (b * 2 + (numeric ? 1 : 0)) {
    case 0: return launch_bin<128, 64, false>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    case 1: return launch_bin<128, 64, true>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    case 2: return launch_bin<512, 64, false>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    case 3: return launch_bin<512, 64, true>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    case 4: return launch_bin<2048, 256, false>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    case 5: return launch_bin<2048, 256, true>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    case 6: return launch_bin<8192, 256, false>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
    default: return launch_bin<8192, 256, true>(nb, rows, L, R, off, oc, ov, cnt, ovf, lds, s);
  }
}

// C = L R on the device.  Returns 1 when a row has more distinct columns than the largest
// table holds or C's size exceeds 32-bit row pointers (the caller falls back to the host).
int spgemm(const DCsr &L, const DCsr &R, DCsr &C, hipStream_t s) {
  int rc;
  const int n = L.nrows;
  C.nrows = n;
  C.ncols = R.ncols;
  DevBuf prod, rows, cnt, ovf, off;
  if ((rc = prod.alloc(8 * ((size_t)n + 1))) || (rc = rows.alloc(4 * ((size_t)n + 1))) ||
      (rc = cnt.alloc(4 * ((size_t)n + 1))) || (rc = ovf.alloc(4 * ((size_t)n + 1))) ||
      (rc = off.alloc(8 * ((size_t)n + 1))))
    return rc;
  std::vector<int64_t> hprod(n);
  if (n > 0) {
    hipLaunchKernelGGL(chr::spgemm_products_kernel, dim3(grid_for(n)), dim3(chr::kBlock), 0, s, n, L.rp.as<int>(),
                       L.ci.as<int>(), R.rp.as<int>(), prod.as<int64_t>());
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(hprod.data(), prod.p, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
  }
  HIPC(hipStreamSynchronize(s));
  // Rows of at most 64 products (hence distinct columns) start in the smallest table, the
  // others in the next one; symbolic passes promote the rows that do not fit.
  std::vector<int> bin(n, 0);
  for (int i = 0; i < n; ++i) bin[i] = hprod[i] <= kBinT[0] / 2 ? 0 : 1;
  std::vector<int32_t> hcnt(n, 0), hovf(n, 0);
  std::vector<std::vector<int>> lists(kBins);
  for (int i = 0; i < n; ++i) lists[bin[i]].push_back(i);
  for (int b = 0; b < kBins; ++b) {
    auto &li = lists[b];
    if (li.empty()) continue;
    HIPC(hipMemcpyAsync(rows.p, li.data(), 4 * li.size(), hipMemcpyHostToDevice, s));
    HIPC(hipMemsetAsync(ovf.p, 0, ovf.bytes, s));
    if ((rc = launch_any(b, false, (int)li.size(), rows.as<int>(), L, R, nullptr, nullptr, nullptr, cnt.as<int>(),
                         ovf.as<int>(), s)))
      return rc;
    HIPC(hipMemcpyAsync(hovf.data(), ovf.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (int i : li)
      if (hovf[i]) {
        if (b == kBins - 1) return 1;
        bin[i] = b + 1;
        lists[b + 1].push_back(i);
      }
    if (b + 1 < kBins) std::sort(lists[b + 1].begin(), lists[b + 1].end());
  }
  HIPC(hipMemcpyAsync(hcnt.data(), cnt.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  std::vector<int64_t> hoff((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) hoff[i + 1] = hoff[i] + hcnt[i];
  const int64_t total = hoff[n];
  if (total > INT_MAX) return 1;
  DevBuf pc, pv;
  if ((rc = pc.alloc(4 * std::max<int64_t>(total, 1))) || (rc = pv.alloc(8 * std::max<int64_t>(total, 1)))) return rc;
  HIPC(hipMemcpyAsync(off.p, hoff.data(), 8 * hoff.size(), hipMemcpyHostToDevice, s));
  // Numeric passes, each row in the bin its symbolic pass settled in.
  std::vector<std::vector<int>> fin(kBins);
  for (int i = 0; i < n; ++i) fin[bin[i]].push_back(i);
  std::vector<DevBuf> keep(kBins);
  for (int b = 0; b < kBins; ++b) {
    auto &li = fin[b];
    if (li.empty()) continue;
    if ((rc = keep[b].alloc(4 * li.size()))) return rc;
    HIPC(hipMemcpyAsync(keep[b].p, li.data(), 4 * li.size(), hipMemcpyHostToDevice, s));
    if ((rc = launch_any(b, true, (int)li.size(), keep[b].as<int>(), L, R, off.as<int64_t>(), pc.as<int>(),
                         pv.as<double>(), cnt.as<int>(), ovf.as<int>(), s)))
      return rc;
  }
  HIPC(hipMemcpyAsync(hcnt.data(), cnt.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  std::vector<int32_t> hrp((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) hrp[i + 1] = hrp[i] + hcnt[i];
  C.nnz = hrp[n];
  if ((rc = C.rp.alloc(4 * ((size_t)n + 1)))) return rc;
  HIPC(hipMemcpyAsync(C.rp.p, hrp.data(), 4 * hrp.size(), hipMemcpyHostToDevice, s));
  if (C.nnz == total) {  // no zeros dropped: the padded rows are the result
    C.ci = std::move(pc);
    C.v = std::move(pv);
  } else {
    if ((rc = C.ci.alloc(4 * std::max<int64_t>(C.nnz, 1))) || (rc = C.v.alloc(8 * std::max<int64_t>(C.nnz, 1))))
      return rc;
    hipLaunchKernelGGL(chr::spgemm_compact_kernel, dim3(grid_for((int64_t)n * chr::kWave)), dim3(chr::kBlock), 0, s,
                       n, off.as<int64_t>(), C.rp.as<int>(), pc.as<int>(), pv.as<double>(), C.ci.as<int>(),
                       C.v.as<double>());
    HIPC(hipGetLastError());
  }
  HIPC(hipStreamSynchronize(s));
  return CHR_OK;
}

}  // namespace

extern "C" int chr_galerkin(const chr_csr *A, const chr_csr *P, int32_t *rowptr_C, void *sink,
                            chr_reserve_fn reserve) {
  if (!A || !P || !rowptr_C || !reserve) return fail(CHR_EARG, "chr_galerkin: arguments");
  int rc;
  if ((rc = chrrt::check_csr(*A, "chr_galerkin A")) || (rc = chrrt::check_csr(*P, "chr_galerkin P"))) return rc;
  if (A->nrows != A->ncols || P->nrows != A->ncols)
    return fail(CHR_EARG, "chr_galerkin: A must be square and P must have A's row count");
  hipStream_t s = nullptr;
  HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DCsr dA, dP, dAP, dPT, dC;
  if ((rc = upload(*A, dA, s)) || (rc = upload(*P, dP, s))) return rc;
  if ((rc = spgemm(dA, dP, dAP, s))) return rc;  // 1: declined
  dA = DCsr();
  if ((rc = transpose(dP, dPT, s))) return rc;
  dP = DCsr();
  if ((rc = spgemm(dPT, dAP, dC, s))) return rc;
  int32_t *col = nullptr;
  double *val = nullptr;
  if ((rc = reserve(sink, dC.nnz, &col, &val)) || (dC.nnz > 0 && (!col || !val)))
    return fail(CHR_EALLOC, "chr_galerkin: output storage");
  HIPC(hipMemcpyAsync(rowptr_C, dC.rp.p, 4 * ((size_t)dC.nrows + 1), hipMemcpyDeviceToHost, s));
  if (dC.nnz) {
    HIPC(hipMemcpyAsync(col, dC.ci.p, 4 * (size_t)dC.nnz, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(val, dC.v.p, 8 * (size_t)dC.nnz, hipMemcpyDeviceToHost, s));
  }
  HIPC(hipStreamSynchronize(s));
  return CHR_OK;
}

// chs_galerkin_fn adapter (chronos_setup.h).  ctx is unused.
extern "C" int chr_setup_galerkin(void *ctx, int32_t n, const int32_t *rowptr_A, const int32_t *col_A,
                                  const double *val_A, int32_t nc, const int32_t *rowptr_P,
                                  const int32_t *col_P, const double *val_P, int32_t *rowptr_C, void *sink,
                                  chr_reserve_fn reserve) {
  (void)ctx;
  const chr_csr A{n, n, rowptr_A, col_A, val_A};
  const chr_csr P{n, nc, rowptr_P, col_P, val_P};
  return chr_galerkin(&A, &P, rowptr_C, sink, reserve);
}

// Adaptive symmetric FSAI factor on the GPU: every row's greedy pattern growth and its
// dense SPD solves, for all rows of a level in one launch.
//
// Row i of G grows a pattern P_i of columns j < i over at most nstep steps
// (sources/Preconditioner/AMG/smoother/MEX_SYM_AFSAI/cpt_afsai_coef.cpp:316-554, driven
// by compute_local_fsai.cpp:6-116).  One step:
//   1. candidates: the columns j < i, j not in P_i, of row i of A and of the rows of A
//      indexed by P_i, listed in order of first appearance (row i first, then the pattern
//      rows in pattern order, each in column order), with the gradient
//      g_j = a_ij + sum_{k in P_i} x_k a_kj accumulated in that same order
//      (cpt_afsai_coef.cpp:227-312, kap_grad);
//   2. the step_size candidates of largest |g_j| join P_i, chosen by the partial
//      partition of cpt_afsai_coef.cpp:40-89 (ri_sortsplit), whose tie order is kept;
//   3. K = A_PP and w = -a_Pi are gathered (gather_fullsys, :144-223), K is factored by
//      Cholesky and x = K^-1 w (dpotrf/dpotrs, :446,454);
//   4. growth stops after nstep steps, when the pattern stops growing, or when the Kaporin
//      quantity a_Pi^T A_PP^-1 a_Pi changes by less than eps relative (:460-468).
// The row is s*[x; 1] with s = 1/sqrt(a_ii - a_Pi^T x), diagonal last (:480-534).
//
// Layout of the work: one wave (64 lanes) per row, one workgroup per row.  The candidate
// table is an open-addressing hash in LDS (column -> list index, -1 for pattern columns);
// each source row is streamed with one lane per nonzero (the columns of one row are
// distinct, so the lanes of a round never collide on a key), and new candidates take list
// indices in lane order, which is the order of first appearance.  Gradient contributions
// are added round by round, i.e. in source order, with each product rounded before the
// add (the library is built without FP contraction).  The Cholesky factor is computed with
// one lane per row of K, every dot product summed in index order; the triangular solves
// and the Kaporin sums run on lane 0 in index order.  The arithmetic is therefore the same
// sequence of IEEE operations as the CPU setup's restatement of the reference, and the
// factor matches it bit for bit.
//
// No matrix cores: the systems are at most nstep x nstep (15 x 15 at the configured
// parameters) and are refactored after every pattern step, so there is no GEMM-shaped
// work in them; the kernel is bound by the latency of the dependent gathers and of the
// small factorisations, which the many rows in flight per CU hide.
#include "chronos_amg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace chr {

constexpr int kFsaiLanes = 64;
constexpr int kFsaiMaxM = 63;  // pattern entries per row (nstep * step_size)

struct FsaiShape {
  int tab;   // hash slots (power of two)
  int cap;   // candidate list capacity
  int ld;    // leading dimension of K (>= nstep * step_size)
  size_t lds_bytes() const {
    return (size_t)tab * 8 + (size_t)cap * 12 + (size_t)ld * ld * 8 + (size_t)ld * 8 * 2 +
           (size_t)ld * 4 + 64;
  }
};

__device__ __forceinline__ unsigned fsai_hash(int j, int tab) {
  return ((unsigned)j * 2654435761u) & (unsigned)(tab - 1);
}

// Finds column j in the table, inserting it when absent.  Returns the slot; *fresh tells
// whether this call inserted it.
__device__ __forceinline__ int fsai_probe(int *hkey, int tab, int j, bool *fresh) {
  unsigned h = fsai_hash(j, tab);
  for (;;) {
    const int k = hkey[h];
    if (k == j) {
      *fresh = false;
      return (int)h;
    }
    if (k == -1) {
      const int old = atomicCAS(&hkey[h], -1, j);
      if (old == -1) {
        *fresh = true;
        return (int)h;
      }
      if (old == j) {
        *fresh = false;
        return (int)h;
      }
    }
    h = (h + 1) & (unsigned)(tab - 1);
  }
}

__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  const int lane = threadIdx.x;
  return lane == 0 ? 0 : __popcll(mask & ((~0ull) >> (64 - lane)));
}

// Partial partition of (val, idx)[0, n): afterwards the first ncut entries hold the ncut
// largest |val|; pivot = first element of the active range, strict '>' (ri_sortsplit).
__device__ void fsai_select_seq(int n, int ncut, double *val, int *idx) {
  int first = 0, last = n - 1;
  if (ncut < 1 || ncut > n) return;
  const int target = ncut - 1;
  for (;;) {
    int mid = first;
    const double pivot = fabs(val[first]);
    for (int j = first + 1; j <= last; ++j)
      if (fabs(val[j]) > pivot) {
        ++mid;
        const double tv = val[mid];
        val[mid] = val[j];
        val[j] = tv;
        const int ti = idx[mid];
        idx[mid] = idx[j];
        idx[j] = ti;
      }
    {
      const double tv = val[mid];
      val[mid] = val[first];
      val[first] = tv;
      const int ti = idx[mid];
      idx[mid] = idx[first];
      idx[first] = ti;
    }
    if (mid == target) return;
    if (mid > target)
      last = mid - 1;
    else
      first = mid + 1;
  }
}

// K (row-major, leading dimension ld: K[q*ld + c], q >= c) = A_PP and w = -a_iP, gathered
// from row pat[c] of A for the columns pat[q], q >= c, and from row i.  Entries not stored
// in A are zero.  Returns whether row i has an entry left of its diagonal.
__device__ bool fsai_gather(int i, int m, const int *__restrict__ rp, const int *__restrict__ ci,
                            const double *__restrict__ v, const int *pat, double *K, int ld, double *w) {
  const int lane = threadIdx.x;
  const int a0 = rp[i], a1 = rp[i + 1];
  const bool has_rhs = a0 < a1 && ci[a0] < i;
  for (int t = lane; t < m * ld; t += kFsaiLanes) K[t] = 0.0;
  if (has_rhs && lane < m) w[lane] = 0.0;
  __syncthreads();
  auto find = [&](int j) -> int {  // index of j in the sorted pattern, or -1
    int lo = 0, hi = m - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int pm = pat[mid];
      if (pm == j) return mid;
      if (pm < j) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
  };
  for (int c = 0; c < m; ++c) {
    const int r = pat[c];
    const int e1 = rp[r + 1];
    for (int e = rp[r] + lane; e < e1; e += kFsaiLanes) {
      const int j = ci[e];
      if (j < r) continue;
      const int q = find(j);
      if (q >= c) K[q * ld + c] = v[e];
    }
  }
  if (has_rhs)
    for (int e = a0 + lane; e < a1; e += kFsaiLanes) {
      const int j = ci[e];
      if (j >= i) break;
      const int q = find(j);
      if (q >= 0) w[q] = -v[e];
    }
  __syncthreads();
  return has_rhs;
}

// In-place lower Cholesky of the leading m x m block of K (row-major).  Column j: lane q
// (j <= q < m) forms K[q][j] - sum_{k<j} K[q][k] K[j][k] in index order; the diagonal is
// its square root and the rest is divided by it.  Returns false (uniformly) if a pivot is
// not positive.
__device__ bool fsai_chol(int m, double *K, int ld, int *flag) {
  const int lane = threadIdx.x;
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    if (lane >= j && lane < m) {
      s = K[lane * ld + j];
      for (int k = 0; k < j; ++k) s -= K[lane * ld + k] * K[j * ld + k];
    }
    if (lane == j) *flag = (s > 0.0) ? 1 : 0;
    __syncthreads();
    if (!*flag) return false;
    const double d = sqrt(__shfl(s, j, kFsaiLanes));
    if (lane == j) K[j * ld + j] = d;
    else if (lane > j && lane < m) K[lane * ld + j] = s / d;
    __syncthreads();
  }
  return true;
}

// b <- K^-1 b with K = L L^T (lane 0; forward then backward substitution, index order).
__device__ void fsai_chol_solve(int m, const double *L, int ld, double *b) {
  for (int i = 0; i < m; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[i * ld + k] * b[k];
    b[i] = s / L[i * ld + i];
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < m; ++k) s -= L[k * ld + i] * b[k];
    b[i] = s / L[i * ld + i];
  }
}

// One row of G per workgroup (one wave).  Outputs: gcol/gval[row * kmax + q] for q <
// rowlen[row] (pattern in ascending order, diagonal last), and *notposdef set to 1 when a
// block is not positive definite or a row's Schur complement is not positive.
__global__ __launch_bounds__(kFsaiLanes) void afsai_row_kernel(
    int row0, int nrows, const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ v, int nstep, int step_size, double eps, FsaiShape sh, int kmax,
    int *__restrict__ gcol, double *__restrict__ gval, int *__restrict__ rowlen,
    int *__restrict__ notposdef, int *__restrict__ overflow) {
  extern __shared__ __align__(16) char smem[];
  const int tab = sh.tab, cap = sh.cap, ld = sh.ld;
  double *cgrad = reinterpret_cast<double *>(smem);
  double *K = cgrad + cap;
  double *w = K + (size_t)ld * ld;
  double *wsav = w + ld;
  int *hkey = reinterpret_cast<int *>(wsav + ld);
  int *hval = hkey + tab;
  int *ccol = hval + tab;
  int *pat = ccol + cap;
  int *misc = pat + ld;  // [0] candidates, [1] pivot flag, [2] chosen column
  const int lane = threadIdx.x;
  const int i = row0 + blockIdx.x;
  if (i >= nrows) return;

  if (lane < ld) wsav[lane] = 0.0;
  int m = 0;
  double dk_old = 0.0;
  bool refine = nstep >= 1, posdef = true;
  int istep = 0;
  const int ri0 = rp[i], ri1 = rp[i + 1];

  while (refine) {
    ++istep;
    const int m_old = m;
    // 1. Candidate list and gradients.
    for (int t = lane; t < tab; t += kFsaiLanes) hkey[t] = -1;
    if (lane == 0) misc[0] = 0;
    __syncthreads();
    if (lane < m) {
      bool fresh;
      const int h = fsai_probe(hkey, tab, pat[lane], &fresh);
      hval[h] = -1;
    }
    __syncthreads();
    // Source s = 0 is row i (values a_ij); s > 0 is row pat[s-1] (values x_{s-1} a_kj).
    for (int s = 0; s <= m; ++s) {
      const int r = s == 0 ? i : pat[s > 0 ? s - 1 : 0];
      const double xk = s == 0 ? 1.0 : w[s > 0 ? s - 1 : 0];
      const int e0 = s == 0 ? ri0 : rp[r], e1 = s == 0 ? ri1 : rp[r + 1];
      for (int base = e0; base < e1; base += kFsaiLanes) {
        const int e = base + lane;
        int j = INT_MAX;
        double a = 0.0;
        if (e < e1) {
          j = ci[e];
          a = v[e];
        }
        const bool live = j < i;
        int slot = -1;
        bool fresh = false;
        if (live) slot = fsai_probe(hkey, tab, j, &fresh);
        const unsigned long long newmask = __ballot(live && fresh);
        const int nc0 = misc[0];
        if (live) {
          if (fresh) {
            const int idx = nc0 + lanes_below(newmask);
            if (idx < cap) {
              hval[slot] = idx;
              ccol[idx] = j;
              cgrad[idx] = s == 0 ? a : xk * a;
            } else {
              hval[slot] = -2;
              atomicOr(overflow, 1);
            }
          } else {
            const int idx = hval[slot];
            if (idx >= 0) cgrad[idx] += xk * a;
          }
        }
        __syncthreads();
        if (lane == 0) misc[0] = min(cap, nc0 + __popcll(newmask));
        __syncthreads();
      }
    }
    const int ncand = misc[0];
    const int ncut = min(step_size, ncand);
    // 2. Selection.
    if (ncut > 0) {
      bool seq = ncut != 1;
      if (!seq) {
        // Unique largest |g|: that candidate is the one the partition leaves in front.
        // Ties and NaNs go to the sequential procedure, which fixes their order.
        double mx = -1.0;
        bool nan = false;
        for (int t = lane; t < ncand; t += kFsaiLanes) {
          const double a = fabs(cgrad[t]);
          nan |= !(a == a);
          mx = a > mx ? a : mx;
        }
        if (__any(nan)) {
          seq = true;
        } else {
          for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, kFsaiLanes));
          int cnt = 0, who = INT_MAX;
          for (int t = lane; t < ncand; t += kFsaiLanes)
            if (fabs(cgrad[t]) == mx) {
              ++cnt;
              who = min(who, t);
            }
          for (int off = 32; off > 0; off >>= 1) {
            cnt += __shfl_xor(cnt, off, kFsaiLanes);
            who = min(who, __shfl_xor(who, off, kFsaiLanes));
          }
          if (cnt == 1) {
            if (lane == 0) misc[2] = ccol[who];
          } else {
            seq = true;
          }
        }
      }
      if (seq && lane == 0) {
        fsai_select_seq(ncand, ncut, cgrad, ccol);
        misc[2] = ccol[0];
      }
      __syncthreads();
      // Append the chosen columns and sort the pattern (values are distinct: rank = count
      // of smaller entries).
      if (lane < ncut) pat[m + lane] = ncut == 1 ? misc[2] : ccol[lane];
      m += ncut;
      __syncthreads();
      int mine = 0, rank = 0;
      if (lane < m) {
        mine = pat[lane];
        for (int q = 0; q < m; ++q) rank += pat[q] < mine;
      }
      __syncthreads();
      if (lane < m) pat[rank] = mine;
      __syncthreads();
    }
    // 3. Solve on the grown pattern.
    if (m > m_old) {
      const bool has_rhs = fsai_gather(i, m, rp, ci, v, pat, K, ld, w);
      if (has_rhs && m > 0) {
        if (!fsai_chol(m, K, ld, &misc[1])) {
          posdef = false;
        } else {
          if (lane < m) wsav[lane] = w[lane];
          __syncthreads();
          if (lane == 0) fsai_chol_solve(m, K, ld, w);
        }
      }
      __syncthreads();
      double dk_new = 0.0;
      for (int q = 0; q < m; ++q) dk_new += w[q] * wsav[q];
      if (istep == nstep) {
        refine = false;
      } else {
        refine = (fabs(dk_new - dk_old) >= eps * dk_old) && (dk_new != 0.0);
        dk_old = dk_new;
      }
    } else {
      refine = false;
    }
    __syncthreads();
  }
  // Scaling from the diagonal entry (every lane walks row i to it).
  double aii = 0.0;
  bool found = false;
  for (int e = ri0; e < ri1 && ci[e] <= i; ++e)
    if (ci[e] == i) {
      aii = v[e];
      found = true;
      break;
    }
  double sc = 0.0;
  if (!found) {
    posdef = false;
  } else {
    double d = 0.0;
    for (int q = 0; q < m; ++q) d += wsav[q] * w[q];
    sc = aii - d;
  }
  if (sc < 0.0) {
    posdef = false;
    if (m > 0 && found) {
      __syncthreads();
      fsai_gather(i, m, rp, ci, v, pat, K, ld, w);
      if (fsai_chol(m, K, ld, &misc[1])) {
        if (lane < m) wsav[lane] = w[lane];
        __syncthreads();
        if (lane == 0) fsai_chol_solve(m, K, ld, w);
        __syncthreads();
        double d = 0.0;
        for (int q = 0; q < m; ++q) d += wsav[q] * w[q];
        sc = aii - d;
      }
    }
  }
  if (sc <= 0.0) {
    posdef = false;
    sc = 1.0;
  }
  sc = 1.0 / sqrt(sc);
  if (lane < m) {
    gcol[(size_t)i * kmax + lane] = pat[lane];
    gval[(size_t)i * kmax + lane] = sc * w[lane];
  }
  if (lane == 0) {
    gcol[(size_t)i * kmax + m] = i;
    gval[(size_t)i * kmax + m] = sc;
    rowlen[i] = m + 1;
    if (!posdef) atomicOr(notposdef, 1);
  }
}

// Compaction of the padded rows into CSR (row pointers from an exclusive scan on the host).
__global__ __launch_bounds__(kBlock) void afsai_compact_kernel(int n, int kmax, const int *__restrict__ rowptr,
                                                               const int *__restrict__ gcol,
                                                               const double *__restrict__ gval,
                                                               int *__restrict__ col, double *__restrict__ val) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * kmax) return;
  const int r = (int)(t / kmax), q = (int)(t % kmax);
  const int len = rowptr[r + 1] - rowptr[r];
  if (q < len) {
    col[rowptr[r] + q] = gcol[t];
    val[rowptr[r] + q] = gval[t];
  }
}

}  // namespace chr

using chrrt::DevBuf;
using chrrt::fail;

namespace {

// Host sizing of the per-row work: the candidate list holds at most the lower entries of
// row i plus those of its pattern rows, i.e. (m + 1) * (longest row) entries.
chr::FsaiShape fsai_shape(const chr_csr &A, int mmax) {
  int maxrow = 0;
  for (int32_t r = 0; r < A.nrows; ++r) maxrow = std::max(maxrow, A.rowptr[r + 1] - A.rowptr[r]);
  chr::FsaiShape sh{};
  sh.cap = std::max(64, (mmax + 1) * maxrow);
  int tab = 64;
  while (tab < (4 * (sh.cap + mmax)) / 3 + 1) tab <<= 1;
  sh.tab = tab;
  sh.ld = std::max(1, mmax);
  return sh;
}

}  // namespace

extern "C" int chr_afsai(const chr_csr *A, int32_t nstep, int32_t step_size, double eps,
                         int32_t *rowptr_G, int32_t *col_G, double *val_G, int64_t *nnz_G,
                         int32_t *is_posdef) {
  if (!A || !rowptr_G || !col_G || !val_G || nstep < 0 || step_size < 1)
    return fail(CHR_EARG, "chr_afsai: arguments");
  if (A->nrows != A->ncols) return fail(CHR_EARG, "chr_afsai: A must be square");
  int rc;
  if ((rc = chrrt::check_csr(*A, "chr_afsai A"))) return rc;
  const int32_t n = A->nrows;
  const int mmax = std::max(1, nstep * step_size);
  if (nstep * step_size > chr::kFsaiMaxM)
    return fail(CHR_EARG, "chr_afsai: nstep * step_size must not exceed 63");
  const int kmax = mmax + 1;
  if (n == 0) {
    rowptr_G[0] = 0;
    if (nnz_G) *nnz_G = 0;
    if (is_posdef) *is_posdef = 1;
    return CHR_OK;
  }
  chr::FsaiShape sh = fsai_shape(*A, mmax);
  const size_t lds = sh.lds_bytes();
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  int max_lds = 0;
  HIPC(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (lds > (size_t)max_lds)
    return fail(CHR_EARG, "chr_afsai: rows too long for the per-row work space (" + std::to_string(lds) +
                              " bytes of LDS needed, " + std::to_string(max_lds) + " available)");
  if (lds > 65536) HIPC(hipFuncSetAttribute((const void *)chr::afsai_row_kernel,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int64_t nnz = A->rowptr[n];
  hipStream_t s = nullptr;
  HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf drp, dci, dv, gcol, gval, rowlen, flags, drowptr, dcol, dval;
  if ((rc = drp.alloc(4 * (size_t)(n + 1))) || (rc = dci.alloc(4 * (size_t)std::max<int64_t>(nnz, 1))) ||
      (rc = dv.alloc(8 * (size_t)std::max<int64_t>(nnz, 1))) || (rc = gcol.alloc(4 * (size_t)n * kmax)) ||
      (rc = gval.alloc(8 * (size_t)n * kmax)) || (rc = rowlen.alloc(4 * (size_t)n)) ||
      (rc = flags.alloc(2 * sizeof(int))))
    return rc;
  HIPC(hipMemcpyAsync(drp.p, A->rowptr, 4 * (size_t)(n + 1), hipMemcpyHostToDevice, s));
  if (nnz) {
    HIPC(hipMemcpyAsync(dci.p, A->col, 4 * (size_t)nnz, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dv.p, A->val, 8 * (size_t)nnz, hipMemcpyHostToDevice, s));
  }
  HIPC(hipMemsetAsync(flags.p, 0, flags.bytes, s));
  // Rows in launches of at most 2^20 workgroups.
  for (int32_t r0 = 0; r0 < n; r0 += (1 << 20)) {
    const int32_t cnt = std::min<int32_t>(n - r0, 1 << 20);
    hipLaunchKernelGGL(chr::afsai_row_kernel, dim3(cnt), dim3(chr::kFsaiLanes), lds, s, r0, n,
                       drp.as<int>(), dci.as<int>(), dv.as<double>(), nstep, step_size, eps, sh, kmax,
                       gcol.as<int>(), gval.as<double>(), rowlen.as<int>(), flags.as<int>(),
                       flags.as<int>() + 1);
    HIPC(hipGetLastError());
  }
  std::vector<int32_t> len(n);
  int hflags[2] = {0, 0};
  HIPC(hipMemcpyAsync(len.data(), rowlen.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(hflags, flags.p, sizeof(hflags), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (hflags[1]) return fail(CHR_EHIP, "chr_afsai: candidate table overflow (internal sizing error)");
  rowptr_G[0] = 0;
  for (int32_t r = 0; r < n; ++r) rowptr_G[r + 1] = rowptr_G[r] + len[r];
  const int64_t nnzG = rowptr_G[n];
  if ((rc = drowptr.alloc(4 * (size_t)(n + 1))) || (rc = dcol.alloc(4 * (size_t)nnzG)) ||
      (rc = dval.alloc(8 * (size_t)nnzG)))
    return rc;
  HIPC(hipMemcpyAsync(drowptr.p, rowptr_G, 4 * (size_t)(n + 1), hipMemcpyHostToDevice, s));
  const unsigned grid = (unsigned)(((int64_t)n * kmax + chr::kBlock - 1) / chr::kBlock);
  hipLaunchKernelGGL(chr::afsai_compact_kernel, dim3(grid), dim3(chr::kBlock), 0, s, n, kmax,
                     drowptr.as<int>(), gcol.as<int>(), gval.as<double>(), dcol.as<int>(), dval.as<double>());
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(col_G, dcol.p, 4 * (size_t)nnzG, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(val_G, dval.p, 8 * (size_t)nnzG, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (nnz_G) *nnz_G = nnzG;
  if (is_posdef) *is_posdef = hflags[0] ? 0 : 1;
  return CHR_OK;
}

// chs_afsai_fn adapter (chronos_setup.h): lets the host setup (chs_build) take its aFSAI
// factors from the GPU kernel.  ctx is unused.
extern "C" int chr_setup_afsai(void *ctx, int32_t n, const int32_t *rowptr, const int32_t *col,
                               const double *val, int32_t nstep, int32_t step_size, double eps,
                               int32_t *rowptr_G, int32_t *col_G, double *val_G, int64_t *nnz_G,
                               int32_t *is_posdef) {
  (void)ctx;
  const chr_csr A{n, n, rowptr, col, val};
  // Rows too long for one workgroup's LDS work space: the host set-up factors this level.
  if (nstep >= 0 && step_size >= 1 && nstep * step_size <= chr::kFsaiMaxM && n > 0 && rowptr) {
    int dev = 0, max_lds = 0;
    HIPC(hipGetDevice(&dev));
    HIPC(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (fsai_shape(A, std::max(1, nstep * step_size)).lds_bytes() > (size_t)max_lds) return 1;
  }
  return chr_afsai(&A, nstep, step_size, eps, rowptr_G, col_G, val_G, nnz_G, is_posdef);
}

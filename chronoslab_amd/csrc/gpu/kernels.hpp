// Device kernels of the AMG solve phase (gfx950, wave64).
//
// Every product hands its row sums to an epilogue functor, which fuses the vector operation
// that follows the product in the V-cycle (scaling by omega, residual, correction) so that
// no intermediate vector makes an extra trip through HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chr {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// Sum over the LANES lanes of a row group (xor butterfly, every lane gets the total).
template <int LANES, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
  return v;
}

// Block-wide sum of one double per thread; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = group_sum<kWave>(v);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / kWave; ++i) s += lds[i];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------------------
// Epilogues: called once per row by lane 0 of the row group with the row's product acc.
// ---------------------------------------------------------------------------------------
template <typename TO>
struct EpiStore {  // out = acc
  TO *out;
  __device__ void operator()(int i, double acc) const { out[i] = (TO)acc; }
};
template <typename TO>
struct EpiScale {  // out = w * acc
  TO *out;
  double w;
  __device__ void operator()(int i, double acc) const { out[i] = (TO)(w * acc); }
};
template <typename TB, typename TO>
struct EpiResidual {  // out = b - acc
  const TB *b;
  TO *out;
  __device__ void operator()(int i, double acc) const { out[i] = (TO)((double)b[i] - acc); }
};
template <typename TX, typename TO>
struct EpiAxpy {  // out = x + w * acc   (x and out may alias: same row, same lane)
  const TX *x;
  TO *out;
  double w;
  __device__ void operator()(int i, double acc) const { out[i] = (TO)((double)x[i] + w * acc); }
};

// Accumulator type: float products for the FP32 cycle, double otherwise.
template <typename TV>
struct Acc {
  using type = double;
};
template <>
struct Acc<float> {
  using type = float;
};

// ---------------------------------------------------------------------------------------
// CSR-stream SpMV.  The rows are cut on the host into blocks of whole consecutive rows
// holding at most kStreamCap nonzeros and at most kBlock rows (a row longer than kStreamCap
// is a block of its own).  A workgroup streams its block's values and column indices with
// coalesced loads, kStreamPer independent loads in flight per thread, stores the products
// in LDS, and then each thread sums one row in column order (deterministic) and hands the
// sum to the epilogue.  No padding and no dependence on the row-length distribution.
// ---------------------------------------------------------------------------------------
constexpr int kStreamCap = 2048;
constexpr int kStreamPer = kStreamCap / kBlock;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8).
// Map b to a logical block so that each XCD owns one contiguous range of rows and the
// gathers of neighbouring rows hit that XCD's L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7;
  const int x = b & 7, j = b >> 3;
  return x * per + min(x, rem) + j;
}

template <typename T>
__device__ __forceinline__ T block_sum_t(T v, T *lds) {
  v = group_sum<kWave>(v);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) lds[w] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / kWave; ++i) s += lds[i];
  __syncthreads();
  return s;
}

// Gathers kStreamPer products of the block's nonzeros into prod[]; `colval` maps a column
// index to the multiplied operand.
template <typename TA, typename TV, typename Xf>
__device__ __forceinline__ void stream_products(int e0, int e1, const int *__restrict__ ci,
                                                const TV *__restrict__ val, Xf colval, TA *prod) {
  int c[kStreamPer];
  TV v[kStreamPer];
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kStreamPer; ++k) {
    const int e = e0 + k * kBlock + tid;
    const bool in = e < e1;
    c[k] = in ? ci[e] : 0;
    v[k] = in ? val[e] : TV(0);
  }
#pragma unroll
  for (int k = 0; k < kStreamPer; ++k) {
    const int e = e0 + k * kBlock + tid;
    if (e < e1) prod[k * kBlock + tid] = (TA)v[k] * colval(c[k]);
  }
}

template <typename TA, typename TX>
struct ColLoad {
  const TX *__restrict__ x;
  __device__ __forceinline__ TA operator()(int j) const { return (TA)x[j]; }
};

template <typename TV, typename TX, typename Epi>
__global__ __launch_bounds__(kBlock) void spmv_stream_kernel(int nblk,
                                                             const int *__restrict__ blk_rows,
                                                             const int *__restrict__ rp,
                                                             const int *__restrict__ ci,
                                                             const TV *__restrict__ val,
                                                             const TX *__restrict__ x, Epi epi,
                                                             const int *__restrict__ stop) {
  if (stop && *stop) return;
  using TA = typename Acc<TV>::type;
  __shared__ TA prod[kStreamCap];
  const int b = xcd_block(blockIdx.x, nblk);
  const int r0 = blk_rows[b], r1 = blk_rows[b + 1];
  const int e0 = rp[r0], e1 = rp[r1];
  const int tid = threadIdx.x;
  if (e1 - e0 > kStreamCap) {  // one long row: strided sum + block reduction
    TA s = 0;
    for (int e = e0 + tid; e < e1; e += kBlock) s += (TA)val[e] * (TA)x[ci[e]];
    s = block_sum_t<TA>(s, prod);
    if (tid == 0) epi(r0, (double)s);
    return;
  }
  stream_products<TA>(e0, e1, ci, val, ColLoad<TA, TX>{x}, prod);
  __syncthreads();
  const int r = r0 + tid;
  if (r < r1) {
    const int a = rp[r] - e0, z = rp[r + 1] - e0;
    TA s = 0;
    for (int q = a; q < z; ++q) s += prod[q];
    epi(r, (double)s);
  }
}

// Coarsest solve x = Ainv b with the FP64 dense inverse (row-major n x n): one wave per
// row, FP64 arithmetic whatever the cycle precision.
template <typename TX>
__global__ __launch_bounds__(kBlock) void coarse_gemv_kernel(int n, const double *__restrict__ Ainv,
                                                             const TX *__restrict__ b,
                                                             TX *__restrict__ x,
                                                             const int *__restrict__ stop) {
  if (stop && *stop) return;
  const int row = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x % kWave;
  if (row >= n) return;
  const double *a = Ainv + (size_t)row * n;
  double acc = 0.0;
  for (int j = lane; j < n; j += kWave) acc += a[j] * (double)b[j];
  acc = group_sum<kWave>(acc);
  if (lane == 0) x[row] = (TX)acc;
}

// Strided block column copy with type conversion: dst[i] = src[i] (one column).
template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void convert_kernel(int n, const TS *__restrict__ src,
                                                         TD *__restrict__ dst) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = (TD)src[i];
}

// ---------------------------------------------------------------------------------------
// PCG_CJ kernels (FP64).  Reductions are deterministic: every block of a fixed grid writes
// its partial sum; the last block to finish (atomic ticket) adds the partials in block
// order and computes the scalar update, then re-arms the ticket.
// ---------------------------------------------------------------------------------------
struct PcgState {
  double bnorm, resiter, ptap, pr, zap, alpha, beta, tol;
  int iter, itmax, stop, breakdown;
  unsigned int ticket[4];
};

__device__ __forceinline__ bool last_block(unsigned int *ticket, unsigned int nblocks,
                                           int *flag_lds) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = atomicAdd(ticket, 1u);
    *flag_lds = (t == nblocks - 1);
  }
  __syncthreads();
  const bool last = *flag_lds != 0;
  if (last) __threadfence();  // acquire: the other blocks' partials (other XCDs' L2s)
  return last;
}

// Deterministic sum of `count` partials by the calling block (thread 0 holds the result).
__device__ __forceinline__ double sum_partials(const double *part, int count, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += kBlock) s += part[i];
  return block_sum(s, lds);
}

// r = b - A x ; partials of b.b and r.r ; last block: bnorm, resvec[0], stop when b = 0.
template <int LANES>
__global__ __launch_bounds__(kBlock) void pcg_init_kernel(int n, const int *__restrict__ rp,
                                                          const int *__restrict__ ci,
                                                          const double *__restrict__ val,
                                                          const double *__restrict__ b,
                                                          const double *__restrict__ x,
                                                          double *__restrict__ r,
                                                          double *__restrict__ part,
                                                          PcgState *st, double *resvec) {
  __shared__ double lds[kBlock / kWave];
  __shared__ int flag;
  const int groups_per_block = kBlock / LANES;
  const int lane = threadIdx.x % LANES;
  double bb = 0.0, rr = 0.0;
  for (int row = blockIdx.x * groups_per_block + threadIdx.x / LANES; row < n;
       row += gridDim.x * groups_per_block) {
    double acc = 0.0;
    for (int e = rp[row] + lane; e < rp[row + 1]; e += LANES) acc += val[e] * x[ci[e]];
    acc = group_sum<LANES>(acc);
    if (lane == 0) {
      const double bi = b[row], ri = bi - acc;
      r[row] = ri;
      bb += bi * bi;
      rr += ri * ri;
    }
  }
  bb = block_sum(bb, lds);
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bb;
    part[gridDim.x + blockIdx.x] = rr;
  }
  if (last_block(&st->ticket[0], gridDim.x, &flag)) {
    double sbb = sum_partials(part, gridDim.x, lds);
    double srr = sum_partials(part + gridDim.x, gridDim.x, lds);
    if (threadIdx.x == 0) {
      st->bnorm = sqrt(sbb);
      st->iter = 0;
      st->breakdown = 0;
      if (sbb == 0.0) {
        st->resiter = 0.0;
        st->stop = 1;
      } else {
        st->resiter = sqrt(srr) / st->bnorm;
        st->stop = st->itmax <= 0 ? 1 : 0;
      }
      resvec[0] = st->resiter;
      st->ticket[0] = 0;
    }
  }
}

// zap = z . Ap_old ; last block: beta = -zap / ptap (iteration > 1).
__global__ __launch_bounds__(kBlock) void pcg_beta_kernel(int n, const double *__restrict__ z,
                                                          const double *__restrict__ ap,
                                                          double *__restrict__ part,
                                                          PcgState *st) {
  if (st->stop) return;
  __shared__ double lds[kBlock / kWave];
  __shared__ int flag;
  double s = 0.0;
  if (st->iter > 0)
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += z[i] * ap[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[2 * gridDim.x + blockIdx.x] = s;
  if (last_block(&st->ticket[1], gridDim.x, &flag)) {
    double t = sum_partials(part + 2 * gridDim.x, gridDim.x, lds);
    if (threadIdx.x == 0) {
      st->zap = t;
      st->beta = st->iter > 0 ? -t / st->ptap : 0.0;
      st->ticket[1] = 0;
    }
  }
}

// p_new = z + beta p_old (formed on the fly at every gathered column), Ap = A p_new,
// partials of p.Ap and p.r over the row blocks of A (as in spmv_stream_kernel); last block:
// ptap, alpha = p.r / ptap.
struct PColVal {
  const double *__restrict__ z;
  const double *__restrict__ p;
  double beta;
  __device__ __forceinline__ double operator()(int j) const { return z[j] + beta * p[j]; }
};

__global__ __launch_bounds__(kBlock) void pcg_p_stream_kernel(
    int nblk, const int *__restrict__ blk_rows, const int *__restrict__ rp,
    const int *__restrict__ ci, const double *__restrict__ val, const double *__restrict__ z,
    const double *__restrict__ p_old, double *__restrict__ p_new, double *__restrict__ ap,
    const double *__restrict__ r, double *__restrict__ part, PcgState *st) {
  if (st->stop) return;
  __shared__ double prod[kStreamCap];
  __shared__ int flag;
  const double beta = st->beta;
  const int b = xcd_block(blockIdx.x, nblk);
  const int r0 = blk_rows[b], r1 = blk_rows[b + 1];
  const int e0 = rp[r0], e1 = rp[r1];
  const int tid = threadIdx.x;
  PColVal cv{z, p_old, beta};
  double pap = 0.0, prr = 0.0;
  if (e1 - e0 > kStreamCap) {
    double s = 0.0;
    for (int e = e0 + tid; e < e1; e += kBlock) s += val[e] * cv(ci[e]);
    s = block_sum_t<double>(s, prod);
    if (tid == 0) {
      const double pi = cv(r0);
      p_new[r0] = pi;
      ap[r0] = s;
      pap = pi * s;
      prr = pi * r[r0];
    }
  } else {
    stream_products<double>(e0, e1, ci, val, cv, prod);
    __syncthreads();
    const int row = r0 + tid;
    if (row < r1) {
      const int a = rp[row] - e0, q1 = rp[row + 1] - e0;
      double s = 0.0;
      for (int q = a; q < q1; ++q) s += prod[q];
      const double pi = cv(row);
      p_new[row] = pi;
      ap[row] = s;
      pap = pi * s;
      prr = pi * r[row];
    }
    __syncthreads();
  }
  pap = block_sum_t<double>(pap, prod);
  prr = block_sum_t<double>(prr, prod);
  if (tid == 0) {
    part[b] = pap;
    part[nblk + b] = prr;
  }
  if (last_block(&st->ticket[2], gridDim.x, &flag)) {
    double a = sum_partials(part, nblk, prod);
    double c = sum_partials(part + nblk, nblk, prod);
    if (tid == 0) {
      st->ptap = a;
      st->pr = c;
      st->alpha = c / a;
      if (!(a != 0.0) || !isfinite(a)) {
        st->breakdown = 1;
        st->stop = 1;
      }
      st->ticket[2] = 0;
    }
  }
}

// x += alpha p ; r -= alpha Ap ; partial r.r ; last block: resiter, resvec, stop test.
__global__ __launch_bounds__(kBlock) void pcg_update_kernel(int n, const double *__restrict__ p,
                                                            const double *__restrict__ ap,
                                                            double *__restrict__ x,
                                                            double *__restrict__ r,
                                                            double *__restrict__ part,
                                                            PcgState *st, double *resvec) {
  if (st->stop) return;
  __shared__ double lds[kBlock / kWave];
  __shared__ int flag;
  const double alpha = st->alpha;
  double rr = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    rr += ri * ri;
  }
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) part[5 * gridDim.x + blockIdx.x] = rr;
  if (last_block(&st->ticket[3], gridDim.x, &flag)) {
    double s = sum_partials(part + 5 * gridDim.x, gridDim.x, lds);
    if (threadIdx.x == 0) {
      const int it = st->iter + 1;
      st->iter = it;
      st->resiter = sqrt(s) / st->bnorm;
      resvec[it] = st->resiter;
      if (st->resiter < st->tol || it >= st->itmax) st->stop = 1;
      st->ticket[3] = 0;
    }
  }
}

}  // namespace chr

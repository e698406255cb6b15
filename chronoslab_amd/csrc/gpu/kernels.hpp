// Device kernels of the AMG solve phase (gfx950, wave64).
//
// CSR SpMV: a group of LANES consecutive lanes owns one row (LANES in {2,...,64}, chosen
// per matrix from its mean row length), strides over the row's entries, and reduces with
// cross-lane shuffles inside the group.  The result is handed to an epilogue functor, which
// fuses the vector operation that follows the product in the V-cycle (scaling by omega,
// residual, correction) so that no intermediate vector makes an extra trip through HBM.
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

template <int LANES, typename TV, typename TX, typename Epi>
__global__ __launch_bounds__(kBlock) void spmv_kernel(int nrows, const int *__restrict__ rp,
                                                      const int *__restrict__ ci,
                                                      const TV *__restrict__ val,
                                                      const TX *__restrict__ x, Epi epi,
                                                      const int *__restrict__ stop) {
  if (stop && *stop) return;
  using TA = typename Acc<TV>::type;
  const int row = (blockIdx.x * kBlock + threadIdx.x) / LANES;
  const int lane = threadIdx.x % LANES;
  if (row >= nrows) return;  // whole row groups leave together
  const int e0 = rp[row], e1 = rp[row + 1];
  TA acc = 0;
  for (int e = e0 + lane; e < e1; e += LANES) acc += (TA)val[e] * (TA)x[ci[e]];
  acc = group_sum<LANES>(acc);
  if (lane == 0) epi(row, (double)acc);
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
// partials of p.Ap and p.r ; last block: ptap, alpha = p.r / ptap.
template <int LANES>
__global__ __launch_bounds__(kBlock) void pcg_p_spmv_kernel(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
    const double *__restrict__ z, const double *__restrict__ p_old, double *__restrict__ p_new,
    double *__restrict__ ap, const double *__restrict__ r, double *__restrict__ part,
    PcgState *st) {
  if (st->stop) return;
  __shared__ double lds[kBlock / kWave];
  __shared__ int flag;
  const double beta = st->beta;
  const int groups_per_block = kBlock / LANES;
  const int lane = threadIdx.x % LANES;
  double pap = 0.0, prr = 0.0;
  for (int row = blockIdx.x * groups_per_block + threadIdx.x / LANES; row < n;
       row += gridDim.x * groups_per_block) {
    double acc = 0.0;
    for (int e = rp[row] + lane; e < rp[row + 1]; e += LANES) {
      const int j = ci[e];
      acc += val[e] * (z[j] + beta * p_old[j]);
    }
    acc = group_sum<LANES>(acc);
    if (lane == 0) {
      const double pi = z[row] + beta * p_old[row];
      p_new[row] = pi;
      ap[row] = acc;
      pap += pi * acc;
      prr += pi * r[row];
    }
  }
  pap = block_sum(pap, lds);
  prr = block_sum(prr, lds);
  if (threadIdx.x == 0) {
    part[3 * gridDim.x + blockIdx.x] = pap;
    part[4 * gridDim.x + blockIdx.x] = prr;
  }
  if (last_block(&st->ticket[2], gridDim.x, &flag)) {
    double a = sum_partials(part + 3 * gridDim.x, gridDim.x, lds);
    double c = sum_partials(part + 4 * gridDim.x, gridDim.x, lds);
    if (threadIdx.x == 0) {
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

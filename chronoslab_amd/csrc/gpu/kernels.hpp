// Device kernels of the AMG solve phase (gfx950, wave64).
//
// Every product hands its row sums to an epilogue functor, which fuses the vector operation
// that follows the product in the V-cycle (scaling by omega, residual, correction) so that
// no intermediate vector makes an extra trip through HBM.
#pragma once
#include <type_traits>

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
// Epilogues: called once per row with the row's product acc.  pre(i) loads the row's own
// operand (issued at kernel start, before the product is known); operator() finishes.
// ---------------------------------------------------------------------------------------
template <typename TO>
struct EpiStore {  // out = acc
  TO *out;
  __device__ __forceinline__ double pre(int) const { return 0.0; }
  __device__ __forceinline__ void operator()(int i, double acc, double) const { out[i] = (TO)acc; }
};
// Same as EpiStore; a distinct type so that the timing probe's kernel has a name of its own
// in profiler output (its average duration then covers the probe's launches only).
template <typename TO, char W>
struct EpiProbe : EpiStore<TO> {};
// wv (optional): a per-row weight (the smoother damping of the row's system in a batched
// hierarchy); row = flat index / k.
template <typename TO>
struct EpiScale {  // out = w * acc
  TO *out;
  double w;
  const double *wv = nullptr;
  int k = 1;
  __device__ __forceinline__ double pre(int i) const { return wv ? wv[i / k] : w; }
  __device__ __forceinline__ void operator()(int i, double acc, double ww) const { out[i] = (TO)(ww * acc); }
};
template <typename TB, typename TO>
struct EpiResidual {  // out = b - acc
  const TB *b;
  TO *out;
  __device__ __forceinline__ double pre(int i) const { return (double)b[i]; }
  __device__ __forceinline__ void operator()(int i, double acc, double bi) const { out[i] = (TO)(bi - acc); }
};
template <typename TX, typename TO>
struct EpiAxpy {  // out = x + w * acc   (x and out may alias: same row, same thread)
  const TX *x;
  TO *out;
  double w;
  const double *wv = nullptr;
  int k = 1;
  __device__ __forceinline__ double pre(int i) const { return (double)x[i]; }
  __device__ __forceinline__ void operator()(int i, double acc, double xi) const {
    out[i] = (TO)(xi + (wv ? wv[i / k] : w) * acc);
  }
};

// Products rounded before they are added (the build also disables contraction), so that
// every layout and the block kernel sum a row to the same bits.
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }

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
// Column indices are stored either as int32 or, when every row block's columns fit in a
// 65536-wide window, as uint16 offsets from a per-block base column (6 instead of 8 bytes
// per FP32 nonzero, 10 instead of 12 per FP64 nonzero).  The block table holds, per block,
// int4 {first row, first nonzero, base column, 0}; entry nblk closes the last block.
// ---------------------------------------------------------------------------------------
// CSR-stream SpMV.  The rows are cut on the host into blocks of whole consecutive rows
// holding at most PER * kBlock nonzeros and at most kBlock rows (a longer row is a block of
// its own).  A workgroup streams its block's values and column indices with coalesced
// loads, PER independent loads in flight per thread, stores the products
// in LDS, and then each thread sums one row in column order (deterministic) and hands the
// sum to the epilogue.  No padding and no dependence on the row-length distribution.
// ---------------------------------------------------------------------------------------
// PER = nonzeros staged per thread; the block capacity is PER * kBlock nonzeros.  PER is
// chosen per matrix at upload by timing the candidates.

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

// Sum of prod[ra..rz) in index order.  The LDS reads are issued 16 at a time ahead of the
// additions, so a long row waits for one LDS round trip per 16 elements instead of one per
// element; the additions themselves stay sequential (same bits as a plain loop).
template <typename T>
__device__ __forceinline__ T row_sum(const T *prod, int ra, int rz) {
  T s = 0;
  int q = ra;
  for (; q + 16 <= rz; q += 16) {  // full groups: a plain chain of adds
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = prod[q + k];
#pragma unroll
    for (int k = 0; k < 16; ++k) s += v[k];
  }
  if (q < rz) {
    const int n = rz - q;
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = prod[q + min(k, n - 1)];
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < n) s += v[k];
  }
  return s;
}

// Strided partial sum of one long row over the workgroup (thread t takes e0 + t, e0 + t +
// kBlock, ...), each thread adding its elements in that order; the loads of four steps are
// issued before their adds.
template <typename TA, typename Load>
__device__ __forceinline__ TA strided_row_sum(int e0, int e1, Load load) {
  TA s = 0;
  int e = e0 + (int)threadIdx.x;
  for (; e + 3 * kBlock < e1; e += 4 * kBlock) {
    TA v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = load(e + k * kBlock);
#pragma unroll
    for (int k = 0; k < 4; ++k) s += v[k];
  }
  for (; e < e1; e += kBlock) s += load(e);
  return s;
}

// Stages the products of the block's nonzeros [e0, e1) in prod[].  Indices are clamped to
// the last nonzero so that every load is unconditional: all 2*PER index/value loads issue
// back to back, then the PER gathers, then the LDS stores (masked).
template <int PER, typename TA, typename TV, typename TI, typename Xf>
__device__ __forceinline__ void stream_products(int e0, int e1, int base, const TI *__restrict__ ci,
                                                const TV *__restrict__ val, Xf colval, TA *prod) {
  if (e1 <= e0) return;
  int c[PER];
  TV v[PER];
  TA xv[PER];
  const int tid = threadIdx.x, last = e1 - 1;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = min(e0 + k * kBlock + tid, last);
    c[k] = base + (int)ci[e];
    v[k] = val[e];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) xv[k] = colval(c[k]);
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int o = k * kBlock + tid;
    if (e0 + o < e1) prod[o] = (TA)v[k] * xv[k];
  }
}

// Four consecutive elements of an index or value array at a 4-element-aligned position, in
// one 4-, 8- or 16-byte vector load (two for FP64).
template <typename T>
struct __align__(4 * sizeof(T)) Run4 {
  T v[4];
};
template <typename T>
__device__ __forceinline__ Run4<T> load_run4(const T *__restrict__ p) {
  Run4<T> q;
  if constexpr (sizeof(T) == 1) {
    const uchar4 u = *reinterpret_cast<const uchar4 *>(p);
    q.v[0] = (T)u.x, q.v[1] = (T)u.y, q.v[2] = (T)u.z, q.v[3] = (T)u.w;
  } else if constexpr (sizeof(T) == 2) {
    const ushort4 u = *reinterpret_cast<const ushort4 *>(p);
    q.v[0] = (T)u.x, q.v[1] = (T)u.y, q.v[2] = (T)u.z, q.v[3] = (T)u.w;
  } else if constexpr (std::is_same<T, float>::value) {
    const float4 u = *reinterpret_cast<const float4 *>(p);
    q.v[0] = u.x, q.v[1] = u.y, q.v[2] = u.z, q.v[3] = u.w;
  } else if constexpr (sizeof(T) == 4) {
    const int4 u = *reinterpret_cast<const int4 *>(p);
    q.v[0] = (T)u.x, q.v[1] = (T)u.y, q.v[2] = (T)u.z, q.v[3] = (T)u.w;
  } else {
    const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
    q.v[0] = a.x, q.v[1] = a.y, q.v[2] = b.x, q.v[3] = b.y;
  }
  return q;
}

// load_run4 with the non-temporal hint (the loads' nt bit), for data a product reads once.
// The lanes of the loaded vector are copied to scalars before the bit casts: a bit cast taken
// straight from a lane expression (w.y) yields lane 0 with this compiler.
template <typename T>
__device__ __forceinline__ Run4<T> load_run4_nt(const T *__restrict__ p) {
  typedef unsigned int u1;
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  Run4<T> q;
  if constexpr (sizeof(T) == 1) {
    const u1 w = __builtin_nontemporal_load(reinterpret_cast<const u1 *>(p));
#pragma unroll
    for (int i = 0; i < 4; ++i) q.v[i] = (T)((w >> (8 * i)) & 0xFFu);
  } else if constexpr (sizeof(T) == 2) {
    const u2 w = __builtin_nontemporal_load(reinterpret_cast<const u2 *>(p));
    const unsigned int w0 = w.x, w1 = w.y;
    q.v[0] = (T)(w0 & 0xFFFFu), q.v[1] = (T)(w0 >> 16), q.v[2] = (T)(w1 & 0xFFFFu), q.v[3] = (T)(w1 >> 16);
  } else if constexpr (sizeof(T) == 4) {
    const u4 w = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(p));
    const unsigned int w0 = w.x, w1 = w.y, w2 = w.z, w3 = w.w;
    q.v[0] = __builtin_bit_cast(T, w0), q.v[1] = __builtin_bit_cast(T, w1);
    q.v[2] = __builtin_bit_cast(T, w2), q.v[3] = __builtin_bit_cast(T, w3);
  } else {
    const u4 a = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(p));
    const u4 b = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(p) + 1);
    const unsigned long long a0 = a.x, a1 = a.y, a2 = a.z, a3 = a.w, b0 = b.x, b1 = b.y, b2 = b.z, b3 = b.w;
    q.v[0] = __builtin_bit_cast(T, a0 | (a1 << 32));
    q.v[1] = __builtin_bit_cast(T, a2 | (a3 << 32));
    q.v[2] = __builtin_bit_cast(T, b0 | (b1 << 32));
    q.v[3] = __builtin_bit_cast(T, b2 | (b3 << 32));
  }
  return q;
}

// Vector-load form of stream_products: lane tid takes groups of four consecutive nonzeros,
// group g covering [a + 4 g, a + 4 g + 4) with a = e0 rounded down to a multiple of 4, so each
// lane issues one index and one value load per group instead of four of each.  Index and
// value arrays are allocated to a multiple of four elements, so every group is in bounds;
// elements outside [e0, e1) are not gathered.  col(c) maps an index word to its column;
// before_gather() runs (block-uniformly) between the loads and the gathers.  The products
// land in LDS at prod[e - a] (prod 16-byte aligned, room for PER * kBlock + 8 elements), one
// aligned four-element store per group (zeros outside [e0, e1)), so a row's products start
// at prod[(e0 - a) + row start - e0]; they are stream_products' products, so the row sums
// are the same.
template <int PER, typename TA, typename TV, typename TI, typename Col, typename Xf, typename Pre>
__device__ __forceinline__ void stream_products4(int e0, int e1, const TI *__restrict__ ci,
                                                 const TV *__restrict__ val, Col col, Xf colval,
                                                 Pre before_gather, TA *prod) {
  if (e1 <= e0) return;
  constexpr int NG = PER / 4 + 1;  // groups per lane: (PER kBlock + 3) / 4 + 1 groups at most
  const int tid = threadIdx.x, a = e0 & ~3, last = ((e1 - 1 - a) >> 2);
  Run4<TI> c[NG];
  Run4<TV> v[NG];
  // With FP32 values the indices and values carry the non-temporal hint (read once per
  // product): the cache-cold level-0 G^T product in the packed layout 65.6 -> 59.8 us.
  constexpr bool kNT = sizeof(TV) == 4;
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    const int g = min(j * kBlock + tid, last);
    c[j] = kNT ? load_run4_nt(ci + a + 4 * g) : load_run4(ci + a + 4 * g);
    v[j] = kNT ? load_run4_nt(val + a + 4 * g) : load_run4(val + a + 4 * g);
  }
  before_gather();
  TA xv[NG][4];
#pragma unroll
  for (int j = 0; j < NG; ++j)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = a + 4 * (j * kBlock + tid) + q;
      xv[j][q] = (e >= e0 && e < e1) ? colval(col(c[j].v[q])) : (TA)0;
    }
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    const int g = j * kBlock + tid;
    if (g <= last) {
      Run4<TA> pq;
#pragma unroll
      for (int q = 0; q < 4; ++q) pq.v[q] = (TA)v[j].v[q] * xv[j][q];
      *reinterpret_cast<Run4<TA> *>(prod + 4 * g) = pq;
    }
  }
}

// The operand of a product: one vector (XOne), or two for a composite operator (XTwo) whose
// even columns 2j index row j of vector a and odd columns 2j + 1 row j of vector b (the two
// column ranges interleaved, so that a row block of the operator spans twice the columns of
// either part and keeps 16-bit column offsets where the parts do).  Both vectors are
// row-major n x k blocks; row(j, k) is the address of operand row j.
template <typename TX>
struct XOne {
  using elem = TX;
  const TX *a;
  __device__ __forceinline__ const TX *row(int j, int k) const { return a + (size_t)j * k; }
};
// b is held as a byte offset from a: the address stays one kernel-argument pointer plus an
// integer, which the compiler keeps in global (not flat) loads; a select between two
// pointers would turn every gather into a flat load.
template <typename TX>
struct XTwo {
  using elem = TX;
  const TX *a;
  int64_t db;  // (char *)b - (char *)a
  __host__ __device__ XTwo(const TX *a_, const TX *b_)
      : a(a_), db(reinterpret_cast<const char *>(b_) - reinterpret_cast<const char *>(a_)) {}
  __device__ __forceinline__ const TX *row(int j, int k) const {
    const char *base = reinterpret_cast<const char *>(a) + ((j & 1) ? db : 0);
    return reinterpret_cast<const TX *>(base) + (size_t)(j >> 1) * k;
  }
};

template <typename TA, typename XS>
struct ColLoad {
  XS x;
  __device__ __forceinline__ TA operator()(int j) const { return (TA)*x.row(j, 1); }
};

// stop: device flag (never null); the kernel does nothing once it is set.  It is read
// together with the block table so that it adds no latency of its own.
template <int PER, typename TV, typename TI, typename XS, typename Epi, bool V4 = false>
__global__ __launch_bounds__(kBlock) void spmv_stream_kernel(int nblk,
                                                             const int4 *__restrict__ blk,
                                                             const int *__restrict__ rp,
                                                             const TI *__restrict__ ci,
                                                             const TV *__restrict__ val,
                                                             XS x, Epi epi,
                                                             const int *__restrict__ stop) {
  using TA = typename Acc<TV>::type;
  __shared__ __align__(16) TA prod[PER * kBlock + 8];
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  if (*stop) return;
  const int r0 = lo.x, r1 = hi.x, e0 = lo.y, e1 = hi.y, base = lo.z;
  const int tid = threadIdx.x;
  if (e1 - e0 > PER * kBlock) {  // one long row: strided sum + block reduction
    TA s = strided_row_sum<TA>(e0, e1, [&](int e) { return (TA)val[e] * (TA)*x.row(base + (int)ci[e], 1); });
    s = block_sum_t<TA>(s, prod);
    if (tid == 0) epi(r0, (double)s, epi.pre(r0));
    return;
  }
  // Row-phase operands first, so their latency overlaps the product phase.
  const int r = r0 + tid;
  const int rc = min(r, r1 - 1);
  const int off = V4 ? e0 & 3 : 0;  // LDS position of the block's first product
  const int ra = rp[rc] - e0 + off, rz = rp[rc + 1] - e0 + off;
  const double pv = epi.pre(rc);
  if constexpr (V4)
    stream_products4<PER, TA>(
        e0, e1, ci, val, [&](TI c) { return base + (int)c; }, ColLoad<TA, XS>{x}, [] {}, prod);
  else
    stream_products<PER, TA>(e0, e1, base, ci, val, ColLoad<TA, XS>{x}, prod);
  __syncthreads();
  if (r < r1) epi(r, (double)row_sum<TA>(prod, ra, rz), pv);
}

// ---------------------------------------------------------------------------------------
// CSR-stream SpMV on the packed-code layout (pack_layout.hpp): each nonzero's 16-bit code
// holds its row's offset from the block's first row (high byte) and an index into the
// window's dictionary of row-relative offsets (low byte), so col = r0 + (code >> 8) -
// dict[code & 255].  Same block cut, staging, summation order and rounding as
// spmv_stream_kernel; the block table's third word is the block's dictionary window, whose
// 256 offsets are loaded into LDS beside the block's values and codes.
// ---------------------------------------------------------------------------------------
template <int PER, typename TV, typename TX, typename Epi, bool V4 = false>
__global__ __launch_bounds__(kBlock) void spmv_pack_kernel(int nblk, const int4 *__restrict__ blk,
                                                           const int *__restrict__ rp,
                                                           const uint16_t *__restrict__ code,
                                                           const TV *__restrict__ val,
                                                           const int *__restrict__ dict,
                                                           const TX *__restrict__ x, Epi epi,
                                                           const int *__restrict__ stop) {
  using TA = typename Acc<TV>::type;
  __shared__ __align__(16) TA prod[PER * kBlock + 8];
  __shared__ int ld[256];
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  if (*stop) return;
  const int r0 = lo.x, r1 = hi.x, e0 = lo.y, e1 = hi.y;
  const int tid = threadIdx.x;
  ld[tid] = dict[lo.z * 256 + tid];
  if (e1 - e0 > PER * kBlock) {  // one long row: strided sum + block reduction
    __syncthreads();
    TA s = strided_row_sum<TA>(e0, e1, [&](int e) {
      const uint32_t c = code[e];
      return (TA)val[e] * (TA)x[r0 + (int)(c >> 8) - ld[c & 255u]];
    });
    s = block_sum_t<TA>(s, prod);
    if (tid == 0) epi(r0, (double)s, epi.pre(r0));
    return;
  }
  const int r = r0 + tid;
  const int rc = min(r, r1 - 1);
  const int off = V4 ? e0 & 3 : 0;  // LDS position of the block's first product
  const int ra = rp[rc] - e0 + off, rz = rp[rc + 1] - e0 + off;
  const double pv = epi.pre(rc);
  if constexpr (V4) {
    stream_products4<PER, TA>(
        e0, e1, code, val, [&](uint16_t c) { return r0 + (int)(c >> 8) - ld[c & 255u]; },
        [&](int j) { return (TA)x[j]; }, [] { __syncthreads(); }, prod);  // sync: dictionary in LDS
  } else if (e1 > e0) {
    uint32_t c[PER];
    TV v[PER];
    TA xv[PER];
    const int last = e1 - 1;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = min(e0 + k * kBlock + tid, last);
      c[k] = code[e];
      v[k] = val[e];
    }
    __syncthreads();  // dictionary in LDS
#pragma unroll
    for (int k = 0; k < PER; ++k) xv[k] = (TA)x[r0 + (int)(c[k] >> 8) - ld[c[k] & 255u]];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int o = k * kBlock + tid;
      if (e0 + o < e1) prod[o] = (TA)v[k] * xv[k];
    }
  }
  __syncthreads();
  if (r < r1) epi(r, (double)row_sum<TA>(prod, ra, rz), pv);
}

// ---------------------------------------------------------------------------------------
// Sliced-ELL SpMV for square matrices with short rows (the smoother factors G, G^T and
// the fine-level operators).  Rows are taken in slices of 64 slots, one wave per slice and
// one row per lane; slice s holds len elements per lane (len a multiple of 4) in groups of
// four, element (k, lane) at sp[s] + 64 (k & ~3) + 4 lane + (k & 3): a lane loads a group
// with one vector load, and the wave's load is one contiguous 256-element segment.  The
// row sum stays in a register (no LDS staging, no row pointers).  A slice is as long
// as its longest row; shorter rows are padded with a PAD code.  Within sort windows of
// 2^sshift slots the rows may be ordered by decreasing length to cut padding: perm then
// holds the row of each slot as an offset from the sort window's first row (0xFFFF: empty
// slot).  Dictionaries (CB = 8) belong to windows of 2^wshift >= 2^sshift slots.
//
// Column codes, as row-relative offsets d = row - col:
//   CB = 8:  an index into the window's dictionary of at most 255 offsets (dict[w*256 +
//            code]; code 255 = PAD, whose dictionary entry is 0);
//   CB = 16: d - dmin as uint16 (0xFFFF = PAD);
//   CB = 32: the column itself (-1 = PAD).
// Each lane sums its row in column order with the product rounded before the add, which
// is the stream kernel's order and rounding: both layouts give the same bits.
// ---------------------------------------------------------------------------------------
template <int CB>
struct SellCode;
template <>
struct SellCode<8> {
  using type = uint8_t;
  static constexpr uint32_t pad = 0xFF;
};
template <>
struct SellCode<16> {
  using type = uint16_t;
  static constexpr uint32_t pad = 0xFFFF;
};
template <>
struct SellCode<32> {
  using type = int32_t;
  static constexpr uint32_t pad = 0xFFFFFFFFu;
};

// U = elements per chunk: every load of a chunk issues before its gathers; 16 when the
// layout has slices longer than 8 (one round trip for the FSAI factors' rows of <= 16).
// ulen > 0: every slice is ulen elements long (uniform ELL, unsorted), so slice offsets
// are computed and the value/code loads need no table read first.  All loads that do not
// depend on another (stop flag, slice offsets, row permutation, dictionary, and with ulen
// the first chunk) are issued before the first wait.
template <int CB, int U, typename TV, typename TX, typename Epi>
__global__ __launch_bounds__(kBlock) void spmv_sell_kernel(
    int nslices, int nrows, int wshift, int sshift, int ulen, const int *__restrict__ sp,
    const uint16_t *__restrict__ perm, const TV *__restrict__ val,
    const typename SellCode<CB>::type *__restrict__ code, const int *__restrict__ dict, int dmin,
    const TX *__restrict__ x, Epi epi, const int *__restrict__ stop) {
  using TA = typename Acc<TV>::type;
  using CT = typename SellCode<CB>::type;
  constexpr int kSlices = kBlock / kWave;
  __shared__ int ld[CB == 8 ? 256 : 1];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int s = b * kSlices + threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int sc = min(s, nslices - 1);  // clamped: every load below stays in bounds
  const int slot = sc * kWave + lane;
  int e0, e1;
  if (ulen > 0) {
    e0 = sc * ulen * kWave;
    e1 = e0 + ulen * kWave;
  } else {
    e0 = sp[sc];
    e1 = sp[sc + 1];
  }
  const int q = perm ? (int)perm[slot] : 0;
  int dv = 0;
  if constexpr (CB == 8) dv = dict[(((b * kSlices) * kWave) >> wshift) * 256 + threadIdx.x];
  const int len = (e1 - e0) / kWave;
  // A chunk's values and codes, a group of four per vector load (U is a multiple of 4),
  // the groups past the lane's last clamped to it (an empty slice reads element 0, which
  // every layout has, and uses none of it); the surplus products are dropped below.
  // FP32 values and their codes are loaded with the non-temporal hint: measured on the 128^3
  // diffusion hierarchy, it takes the cache-cold level-0 G product from 56 to 48 us and the
  // FP32 V-cycle from 0.530 to 0.510 ms, while with FP64 values it slowed the cycle by 2%.
  constexpr bool kNT = sizeof(TV) == 4;
  auto load_chunk = [&](int ch, TV(&vv)[U], CT(&cc)[U]) {
#pragma unroll
    for (int g = 0; g < U / 4; ++g) {
      const int at = len > 0 ? e0 + kWave * min(ch * U + 4 * g, len - 4) + 4 * lane : 0;
      const Run4<TV> a = kNT ? load_run4_nt(val + at) : load_run4(val + at);
      const Run4<CT> b = kNT ? load_run4_nt(code + at) : load_run4(code + at);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vv[4 * g + q] = a.v[q];
        cc[4 * g + q] = b.v[q];
      }
    }
  };
  TV v[U];
  CT c[U];
  load_chunk(0, v, c);
  if (*stop) return;
  if constexpr (CB == 8) {
    // The block's slices lie in one window (windows hold a multiple of kBlock slots).
    ld[threadIdx.x] = dv;
    __syncthreads();
  }
  if (s >= nslices) return;
  int row = slot;
  bool valid = slot < nrows;
  if (perm) {
    valid = q != 0xFFFF;
    row = ((slot >> sshift) << sshift) + q;
  }
  const int rc = valid ? row : 0;  // PAD elements read x[rc] and drop the product
  const double pv = epi.pre(rc);
  auto col_of = [&](CT cc) -> int {
    if constexpr (CB == 8) return rc - ld[cc];
    else if constexpr (CB == 16) return rc - (cc == (CT)SellCode<16>::pad ? 0 : (int)cc + dmin);
    else return cc < 0 ? rc : (int)cc;
  };
  // Chunks of U elements; the last chunk clamps its loads to the row's last element and
  // drops the surplus products.
  // FP64 values: the next chunk's values and codes are requested right after this chunk's
  // gathers, so a multi-chunk slice keeps one chunk of loads in flight while it waits (with
  // FP32 values the extra registers cost more occupancy than the overlap gains, measured).
  // Either way the products are added chunk by chunk in element order.
  constexpr bool kAhead = sizeof(TV) == 8;
  TA acc = 0;
  const int nch = (len + U - 1) / U;
  for (int ch = 0; ch < nch; ++ch) {
    if (!kAhead && ch > 0) load_chunk(ch, v, c);
    TA xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = (TA)x[col_of(c[u])];
    const bool more = kAhead && ch + 1 < nch;  // uniform over the wave (one slice)
    TV vn[U];
    CT cn[U];
    if (more) load_chunk(ch + 1, vn, cn);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const TA p = mul_rn((TA)v[u], xv[u]);
      const bool live = ch * U + u < len && ((uint32_t)c[u] & SellCode<CB>::pad) != SellCode<CB>::pad;
      acc = live ? acc + p : acc;
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        v[u] = vn[u];
        c[u] = cn[u];
      }
    }
  }
  if (valid) epi(row, (double)acc, pv);
}

// ---------------------------------------------------------------------------------------
// Symmetric diagonal layout (dia_layout.hpp): m stored diagonals at offsets off[0] = 0 <
// off[1] < ... < off[m-1], diagonal k at val + k n holding A(i, i + off[k]); the lower
// triangle is read from the mirrored upper entry, A(i, i - off[k]) = val[k n + i - off[k]].
// One row per thread; every load is a contiguous run of the wave's rows (no column
// indices, no gathers).  Row i is summed in column order (lower diagonals from the
// farthest, the main diagonal, upper diagonals from the nearest) with the product rounded
// before the add and zero (absent) entries skipped: the sequential CSR order, so the same
// bits as the other layouts.
// ---------------------------------------------------------------------------------------
struct DiaOffs {
  int m;
  int off[16];
};

// Loads of row i (clamped into [0, n)): lower/upper values and operands of diagonals
// 1..MAXM-1 (those >= m and those outside the matrix read as zero values), and the main
// diagonal.
template <int MAXM, typename TV, typename TA, typename XF>
struct DiaRow {
  TV lo[MAXM], up[MAXM], dg;
  TA xl[MAXM], xu[MAXM], xd;
  __device__ __forceinline__ void load(int i, int n, const DiaOffs &d, const TV *__restrict__ val, XF xf) {
#pragma unroll
    for (int k = 1; k < MAXM; ++k) {
      lo[k] = up[k] = TV(0);
      xl[k] = xu[k] = TA(0);
      if (k < d.m) {
        const int jl = i - d.off[k], ju = i + d.off[k];
        const int jlc = max(jl, 0), juc = min(ju, n - 1);
        const TV a = val[(size_t)k * n + jlc];
        lo[k] = jl >= 0 ? a : TV(0);
        xl[k] = (TA)xf(jlc);
        up[k] = val[(size_t)k * n + i];
        xu[k] = (TA)xf(juc);
      }
    }
    dg = val[i];
    xd = (TA)xf(i);
  }
  __device__ __forceinline__ TA sum() const {
    TA s = 0;
#pragma unroll
    for (int k = MAXM - 1; k >= 1; --k)
      if (lo[k] != TV(0)) s = s + mul_rn((TA)lo[k], xl[k]);
    if (dg != TV(0)) s = s + mul_rn((TA)dg, xd);
#pragma unroll
    for (int k = 1; k < MAXM; ++k)
      if (up[k] != TV(0)) s = s + mul_rn((TA)up[k], xu[k]);
    return s;
  }
};

template <typename TX>
struct VecLoad {
  const TX *__restrict__ x;
  __device__ __forceinline__ TX operator()(int j) const { return x[j]; }
};

template <int MAXM, typename TV, typename TX, typename Epi>
__global__ __launch_bounds__(kBlock) void spmv_dia_kernel(int n, DiaOffs d, const TV *__restrict__ val,
                                                          const TX *__restrict__ x, Epi epi,
                                                          const int *__restrict__ stop) {
  using TA = typename Acc<TV>::type;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i = b * kBlock + (int)threadIdx.x;
  const int ic = min(i, n - 1);
  DiaRow<MAXM, TV, TA, VecLoad<TX>> row;
  row.load(ic, n, d, val, VecLoad<TX>{x});
  const double pv = epi.pre(ic);
  if (*stop || i >= n) return;
  epi(i, (double)row.sum(), pv);
}

// ---------------------------------------------------------------------------------------
// Block product Y = M X for row-major n x k blocks (element (i, c) at i*k + c, k <= KL), on
// the stream layout's row blocks.  The block's values and column indices are staged in LDS
// with coalesced loads (cap = the layout's PER * kBlock nonzeros, in dynamic LDS); then each
// (row, column) pair is one lane, KL lanes per row, so the KL lanes of a row gather one
// contiguous k-vector of X per nonzero.  A lane reads 16 staged nonzeros and issues their 16
// gathers before it adds the products, in nonzero order and each rounded before the add:
// column c is bit for bit the single-vector stream kernel's result on column c.  A block
// holding one row longer than cap (the stream layout's long-row blocks) is summed by its KL
// lanes straight from memory, in the same order.  Epilogues see flat indices i*k + c.
// ---------------------------------------------------------------------------------------

// Copies a stream block's values and column indices [e0, e1), and its row starts rp[r0..r1]
// as offsets from e0, into LDS: the loads of eight strides are issued before their stores,
// so a block of up to 8 * kBlock nonzeros costs one memory round trip (a loop of single loads
// and stores would wait once per stride).
template <typename TV, typename TI>
__device__ __forceinline__ void stage_block(int r0, int r1, int e0, int e1, const int *__restrict__ rp,
                                            const TV *__restrict__ val, const TI *__restrict__ ci, TV *sv,
                                            TI *sc, int *srp) {
  const int tid = threadIdx.x;
  const int nr = r1 - r0;  // <= kBlock rows: nr + 1 row starts
  const int ra = rp[r0 + min(tid, nr)];
  const int rb = (tid == 0 && nr >= kBlock) ? rp[r1] : 0;
  const int cnt = e1 - e0;
  for (int t0 = 0; t0 < cnt; t0 += 8 * kBlock) {
    TV v[8];
    TI c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = min(t0 + u * kBlock + (int)threadIdx.x, cnt - 1);
      v[u] = val[e0 + t];
      c[u] = ci[e0 + t];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u * kBlock + (int)threadIdx.x;
      if (t < cnt) {
        sv[t] = v[u];
        sc[t] = c[u];
      }
    }
  }
  if (tid <= nr) srp[tid] = ra - e0;
  if (tid == 0 && nr >= kBlock) srp[nr] = rb - e0;
}

// Bytes of LDS a block product needs beyond its staged values and indices (the row starts).
constexpr int kStageRpBytes = (kBlock + 1) * 4;

template <int KL, typename TV, typename TI, typename XS, typename Epi>
__global__ __launch_bounds__(kBlock) void spmm_kernel(int nblk, const int4 *__restrict__ blk,
                                                      const int *__restrict__ rp,
                                                      const TI *__restrict__ ci,
                                                      const TV *__restrict__ val,
                                                      XS x, int k, int cap, Epi epi,
                                                      const int *__restrict__ stop) {
  using TA = typename Acc<TV>::type;
  extern __shared__ __align__(16) char smem[];
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  if (*stop) return;
  const int r0 = lo.x, r1 = hi.x, e0 = lo.y, e1 = hi.y, base = lo.z;
  const int tid = threadIdx.x;
  if (e1 - e0 > cap) {  // one long row
    const int c = tid % KL;
    if (tid >= KL || c >= k) return;  // no cross-lane operations below
    const int i = r0 * k + c;
    const double pv = epi.pre(i);
    TA s = 0;
    for (int e = rp[r0]; e < rp[r0 + 1]; ++e) s = s + mul_rn((TA)val[e], (TA)x.row(base + (int)ci[e], k)[c]);
    epi(i, (double)s, pv);
    return;
  }
  TV *sv = reinterpret_cast<TV *>(smem);
  TI *sc = reinterpret_cast<TI *>(smem + (size_t)cap * sizeof(TV));
  int *srp = reinterpret_cast<int *>(smem + (size_t)cap * (sizeof(TV) + sizeof(TI)));
  stage_block(r0, r1, e0, e1, rp, val, ci, sv, sc, srp);
  __syncthreads();
  const int npair = (r1 - r0) * KL;
  for (int p = tid; p < npair; p += kBlock) {
    const int r = r0 + p / KL, c = p % KL;
    if (c >= k) continue;
    const int i = r * k + c;
    const double pv = epi.pre(i);
    const int ra = srp[r - r0], rz = srp[r - r0 + 1];
    TA s = 0;
    for (int q = ra; q < rz; q += 16) {
      const int n = min(16, rz - q);
      TV vv[16];
      TA xv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int qq = q + min(u, n - 1);
        vv[u] = sv[qq];
        xv[u] = (TA)x.row(base + (int)sc[qq], k)[c];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (u < n) s = s + mul_rn((TA)vv[u], xv[u]);
    }
    epi(i, (double)s, pv);
  }
}

// Block product for k a multiple of 4 with X 16-byte aligned: as spmm_kernel, but each lane
// takes four adjacent columns of its row (KL / 4 lanes per row) with one 16-byte load per
// nonzero (two for FP64) and one read of the staged value and column for all four.  Every
// column is still summed in nonzero order with each product rounded before the add, so the
// results equal spmm_kernel's bit for bit.
template <typename T>
struct Quad;
template <>
struct Quad<float> {
  static constexpr int kUnroll = 16;
  __device__ __forceinline__ static void load(const float *p, float (&o)[4]) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
};
template <>
struct Quad<double> {
  static constexpr int kUnroll = 8;
  __device__ __forceinline__ static void load(const double *p, double (&o)[4]) {
    const double2 a = *reinterpret_cast<const double2 *>(p), b = *reinterpret_cast<const double2 *>(p + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
  }
};

template <int KL, typename TV, typename TI, typename XS, typename Epi>
__global__ __launch_bounds__(kBlock) void spmm4_kernel(int nblk, const int4 *__restrict__ blk,
                                                       const int *__restrict__ rp,
                                                       const TI *__restrict__ ci,
                                                       const TV *__restrict__ val,
                                                       XS x, int k, int cap, Epi epi,
                                                       const int *__restrict__ stop) {
  static_assert(KL % 4 == 0, "four columns per lane");
  using TX = typename XS::elem;
  using TA = typename Acc<TV>::type;
  constexpr int LR = KL / 4, U = Quad<TX>::kUnroll;
  extern __shared__ __align__(16) char smem[];
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  if (*stop) return;
  const int r0 = lo.x, r1 = hi.x, e0 = lo.y, e1 = hi.y, base = lo.z;
  const int tid = threadIdx.x;
  if (e1 - e0 > cap) {  // one long row: KL lanes, one column each (as spmm_kernel)
    const int c = tid % KL;
    if (tid >= KL || c >= k) return;  // no cross-lane operations below
    const int i = r0 * k + c;
    const double pv = epi.pre(i);
    TA s = 0;
    for (int e = rp[r0]; e < rp[r0 + 1]; ++e) s = s + mul_rn((TA)val[e], (TA)x.row(base + (int)ci[e], k)[c]);
    epi(i, (double)s, pv);
    return;
  }
  TV *sv = reinterpret_cast<TV *>(smem);
  TI *sc = reinterpret_cast<TI *>(smem + (size_t)cap * sizeof(TV));
  int *srp = reinterpret_cast<int *>(smem + (size_t)cap * (sizeof(TV) + sizeof(TI)));
  stage_block(r0, r1, e0, e1, rp, val, ci, sv, sc, srp);
  __syncthreads();
  const int npair = (r1 - r0) * LR;
  for (int p = tid; p < npair; p += kBlock) {
    const int r = r0 + p / LR, c0 = 4 * (p % LR);
    if (c0 >= k) continue;
    const int i0 = r * k + c0;
    double pv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) pv[j] = epi.pre(i0 + j);
    const int ra = srp[r - r0], rz = srp[r - r0 + 1];
    TA s[4] = {0, 0, 0, 0};
    for (int q = ra; q < rz; q += U) {
      const int n = min(U, rz - q);
      TV vv[U];
      TX xv[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int qq = q + min(u, n - 1);
        vv[u] = sv[qq];
        Quad<TX>::load(x.row(base + (int)sc[qq], k) + c0, xv[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (u < n) {
#pragma unroll
          for (int j = 0; j < 4; ++j) s[j] = s[j] + mul_rn((TA)vv[u], (TA)xv[u][j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) epi(i0 + j, (double)s[j], pv[j]);
  }
}

// Direct solves on a level's coarse rows [nd, nd + nrows): each belongs to a dense block
// (a system whose coarsest level this is) with its FP64 inverse (row-major m x m) packed in
// Ainv.  cmap[t] = {first row of the block, m, offset of the block's inverse, 0} for coarse
// row t.  One wave per row, FP64 arithmetic whatever the cycle precision; column blockIdx.y
// of a row-major n x k block.
template <typename TX>
__global__ __launch_bounds__(kBlock) void coarse_gemv_kernel(int nrows, int nd,
                                                             const int4 *__restrict__ cmap,
                                                             const double *__restrict__ Ainv,
                                                             const TX *__restrict__ b,
                                                             TX *__restrict__ x,
                                                             const int *__restrict__ stop, int k) {
  const int t = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x % kWave;
  const int c = blockIdx.y;
  if (t >= nrows) return;
  const int4 info = cmap[t];
  if (stop && *stop) return;
  const int row = nd + t, m = info.y, r0 = info.x;
  const double *a = Ainv + (size_t)info.z + (size_t)(row - r0) * m;
  // Each lane adds a[j] b[j] for j = lane, lane + 64, ... in order; the loads of sixteen
  // steps are issued before their adds (one memory round trip per 1024 columns).
  constexpr int kAhead = 16;
  double acc = 0.0;
  int j = lane;
  for (; j + (kAhead - 1) * kWave < m; j += kAhead * kWave) {
    double av[kAhead], bv[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      av[u] = a[j + u * kWave];
      bv[u] = (double)b[(size_t)(r0 + j + u * kWave) * k + c];
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u) acc += av[u] * bv[u];
  }
  if (j < m) {
    double av[kAhead], bv[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      const int jj = min(j + u * kWave, m - 1);
      av[u] = a[jj];
      bv[u] = (double)b[(size_t)(r0 + jj) * k + c];
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
      if (j + u * kWave < m) acc += av[u] * bv[u];
  }
  acc = group_sum<kWave>(acc);
  if (lane == 0) x[(size_t)row * k + c] = (TX)acc;
}

// ---------------------------------------------------------------------------------------
// Dense block helpers (FP64, row-major blocks with leading dimension = column count).
// ---------------------------------------------------------------------------------------
// dst (row-major n x k) <- src (column-major, leading dimension ld)
static __global__ __launch_bounds__(kBlock) void cm_to_rm_kernel(int n, int k, const double *__restrict__ src,
                                                          int64_t ld, double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  const int64_t i = t / k;
  const int c = (int)(t % k);
  dst[t] = src[c * ld + i];
}
// dst (column-major, leading dimension ld) <- src (row-major n x k)
static __global__ __launch_bounds__(kBlock) void rm_to_cm_kernel(int n, int k, const double *__restrict__ src,
                                                          double *__restrict__ dst, int64_t ld) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  const int64_t i = t % n;  // consecutive threads walk down a column of dst
  const int c = (int)(t / n);
  dst[c * ld + i] = src[i * k + c];
}

// ---------------------------------------------------------------------------------------
// Tall-skinny FP64 products of LOBPCG on the matrix cores (v_mfma_f64_16x16x4f64).  Lane
// layout of that instruction (checked on gfx950): A[i][k] at lane i + 16 k and B[k][j] at
// lane j + 16 k, one double each; D[i][j] at lane j + 16 (i % 4), element i / 4.
// ---------------------------------------------------------------------------------------
typedef double mf64x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mf64x4 mfma_f64_16x16x4(double a, double b, mf64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Gram products X_k^T Y_k of row-major n x w blocks (w <= 16).  The operands are numbered
// in order of first appearance; three fixed tables of (i, j) pairs, one nibble per pair:
//   table 0: one product, blocks 0 and 1;
//   table 1: the six products of a Rayleigh-Ritz step after a restart, over the blocks
//            X, AX, AAR, AR: X'AX, X'AAR, AR'AAR, X'X, X'AR, AR'AR;
//   table 2: table 1 and the six products with AAP and APa: X'AAP, AR'AAP, APa'AAP,
//            X'APa, AR'APa, APa'APa;
//   table 3: R^T R of the residual block R = AX - X diag(lam), formed on the fly from
//            blocks 0 (AX) and 1 (X) and never stored (the product's operand is block 0);
//   tables 4 and 5: the Rayleigh-Ritz products when the narrow blocks (AR, AAR, APa, AAP,
//            na <= 8 columns each) fit side by side in 16 columns: blocks 0 (X) and 1 (AX)
//            and packed operands of up to four narrow blocks in slots of sw = 4 or 8 columns
//            (columns [s sw, s sw + na) hold segment s, the rest are zero).  Table 4, one
//            packed block (2): X'AX, X'X, X'P, P'P; table 5, two (2 = [AR | AAR],
//            3 = [APa | AAP]): X'AX, X'X, X'P2, X'P3, P2'P2, P2'P3, P3'P3.  An entry of a
//            packed product is the same matrix-instruction sum as in the unpacked one.
constexpr int kGramMaxBlocks = 6, kGramMaxPairs = 12, kGramMaxOut = kGramMaxPairs * 256;
constexpr uint64_t kGramPi = 0x530530300300ull, kGramPj = 0x555444330221ull;
constexpr uint64_t kGramPi4 = 0x2000ull, kGramPj4 = 0x2201ull, kGramPi5 = 0x3220000ull, kGramPj5 = 0x3323201ull;
__host__ __device__ constexpr int gram_npairs(int t) {
  return t == 0 || t == 3 ? 1 : t == 1 ? 6 : t == 2 ? 12 : t == 4 ? 4 : 7;
}
__host__ __device__ constexpr int gram_nblocks(int t) {
  return t == 0 || t == 3 ? 2 : t == 1 ? 4 : t == 2 ? 6 : t == 4 ? 3 : 4;
}
__host__ __device__ constexpr int gram_pi(int t, int k) {
  return t == 0 || t == 3 ? 0 : (int)(((t == 4 ? kGramPi4 : t == 5 ? kGramPi5 : kGramPi) >> (4 * k)) & 15);
}
__host__ __device__ constexpr int gram_pj(int t, int k) {
  return t == 0 ? 1 : t == 3 ? 0 : (int)(((t == 4 ? kGramPj4 : t == 5 ? kGramPj5 : kGramPj) >> (4 * k)) & 15);
}
struct GramSet {
  const double *blk[kGramMaxBlocks];
  int w[kGramMaxBlocks];      // widths (packed blocks: 16 or the slots used)
  int po[kGramMaxPairs + 1];  // first output entry of each product (row-major w_i x w_j)
  const double *lam;          // table 3: the k eigenvalues
  const double *seg[2][4];    // tables 4, 5: segments of packed blocks 2 and 3
  int nseg[2], sw, na;        // segments per packed block, slot width, segment width
};

// Per-workgroup partial Gram products: wave v of the grid takes rows [v R, v R + R) (R a
// multiple of 4, fixed by the launch), four rows per matrix instruction, one accumulator
// per product; the workgroup's four waves are then added in wave order and written to
// part[g][po[k] + p w_j + q].  The row split and the order of the additions are fixed, so
// the partials do not depend on timing, and a product's partials are the same whichever
// table it is computed in.
template <int T>
__global__ __launch_bounds__(kBlock) void gram_mfma_kernel(int n, GramSet gs, int rows_per_wave,
                                                           double *__restrict__ part) {
  constexpr int NP = gram_npairs(T), NB = gram_nblocks(T), NW = kBlock / kWave;
  constexpr int UNR = NP == 1 ? 8 : NP <= 6 ? 4 : 2;  // row steps whose loads are issued together
  __shared__ double red[NW][kWave * 4];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int64_t v = (int64_t)blockIdx.x * NW + wv;
  const int i0 = (int)min<int64_t>(n, v * rows_per_wave), i1 = min(n, i0 + rows_per_wave);
  const int rsub = lane / 16, csub = lane % 16;
  mf64x4 acc[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) acc[k] = mf64x4{0.0, 0.0, 0.0, 0.0};
  const double *src[NB];
  int wd[NB];
  bool col_ok[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (T >= 4 && b >= 2) {  // packed: segment csub / sw, its column csub % sw
      const int sg = csub / gs.sw, c = csub % gs.sw;
      col_ok[b] = sg < gs.nseg[b - 2] && c < gs.na;
      wd[b] = gs.na;
      src[b] = col_ok[b] ? gs.seg[b - 2][sg] + c : gs.blk[0];
    } else {
      wd[b] = gs.w[b];
      col_ok[b] = csub < wd[b];
      src[b] = gs.blk[b] + csub;
    }
  }
  const double lam_c = (T == 3 && col_ok[0]) ? gs.lam[csub] : 0.0;
  // UNR row steps per round: all their loads first, then their instructions in row order
  // (steps past the wave's rows load zeros, which leave the accumulators unchanged).
  for (int r = i0; r < i1; r += 4 * UNR) {
    double f[UNR][NB];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int row = r + 4 * u + rsub;
#pragma unroll
      for (int b = 0; b < NB; ++b) f[u][b] = (row < i1 && col_ok[b]) ? src[b][(size_t)row * wd[b]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      if constexpr (T == 3) f[u][0] = f[u][0] - f[u][1] * lam_c;  // R = AX - X diag(lam), as block_resid_gather_kernel
#pragma unroll
      for (int k = 0; k < NP; ++k) acc[k] = mfma_f64_16x16x4(f[u][gram_pi(T, k)], f[u][gram_pj(T, k)], acc[k]);
    }
  }
  const int total = gs.po[NP];
  const int t = threadIdx.x, tl = t / 4, te = t % 4;
  const int p = tl / 16 + 4 * te, q = tl % 16;  // entry held by element te of lane tl
#pragma unroll
  for (int k = 0; k < NP; ++k) {
#pragma unroll
    for (int e = 0; e < 4; ++e) red[wv][lane * 4 + e] = acc[k][e];
    __syncthreads();
    double s = red[0][t];
#pragma unroll
    for (int u = 1; u < NW; ++u) s += red[u][t];
    const int wa = gs.w[gram_pi(T, k)], wb = gs.w[gram_pj(T, k)];
    if (p < wa && q < wb) part[(size_t)blockIdx.x * total + gs.po[k] + p * wb + q] = s;
    __syncthreads();
  }
}

// out[t] = sum over the nparts partials of entry t: one wave per entry, lane l adding
// partials l, l + 64, ... in order, then a fixed butterfly across the wave.
static __global__ __launch_bounds__(kBlock) void gram_reduce_kernel(int nparts, int ab,
                                                             const double *__restrict__ part,
                                                             double *__restrict__ out) {
  const int t = (blockIdx.x * kBlock + threadIdx.x) / kWave, lane = threadIdx.x % kWave;
  if (t >= ab) return;
  double s = 0.0;
  for (int g = lane; g < nparts; g += kWave) s += part[(size_t)g * ab + t];
  s = group_sum<kWave>(s);
  if (lane == 0) out[t] = s;  // row-major a x b: out[p*b + q]
}

// Z (n x b) = beta * Z + X (n x a) * C with C(p, q) = cscale * Cm[p * csp + q * csq]
// (column-major a x b: csp = 1, csq = a; row-major: csp = b, csq = 1), a <= 4 KS,
// b <= 16 CT, Z not aliasing X.  Each wave takes 16-row tiles; the coefficient fragments of
// the KS k-steps and CT column tiles stay in registers.  (LOBPCG's blocks are at most 16
// wide: KS = 4, CT = 1.)
template <int KS, int CT>
__global__ __launch_bounds__(kBlock) void tsgemm_mfma_kernel(int n, int a, const double *__restrict__ X, int b,
                                                             const double *__restrict__ Cm, int csp, int csq,
                                                             double cscale, double beta, double *__restrict__ Z) {
  const int lane = threadIdx.x % kWave, hi = lane / 16, lo = lane % 16;
  double cf[KS][CT];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int kk = 4 * s + hi, j = 16 * c + lo;
      cf[s][c] = (kk < a && j < b) ? cscale * Cm[(size_t)kk * csp + (size_t)j * csq] : 0.0;
    }
  const int64_t ntile = ((int64_t)n + 15) / 16;
  const int64_t nwave = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t tile = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; tile < ntile; tile += nwave) {
    const int64_t r0 = tile * 16;
    const int64_t ar = r0 + lo;  // the A-fragment row of this lane
    mf64x4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = mf64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int kk = 4 * s + hi;
      const double av = (ar < n && kk < a) ? X[ar * a + kk] : 0.0;
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[c] = mfma_f64_16x16x4(av, cf[s][c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int j = 16 * c + lo;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = r0 + hi + 4 * e;
        if (row < n && j < b) {
          const size_t zi = (size_t)row * b + j;
          Z[zi] = (beta != 0.0 ? beta * Z[zi] : 0.0) + acc[c][e];
        }
      }
    }
  }
}

// The closing update of a LOBPCG iteration (lobpcg.m:389-400) in one pass, for the
// search-direction side (R, Pa, X) or the operator side (AR, APa, AX):
//   Pout = R ER (+ Pa EP),   Xout = Pout + X EX,
// R and Pa n x na, X n x k, ER and EP column-major na x k, EX k x k (na, k <= 16), on the
// matrix cores: R ER and Pa EP accumulate in one accumulator, X EX in another.  X and Xout
// may alias: a wave reads all of its 16-row tile before writing it, and no two waves share
// a tile.
static __global__ __launch_bounds__(kBlock) void lob_update_mfma_kernel(int n, int na, int k, const double *__restrict__ R,
                                                                 const double *__restrict__ Pa, const double *X,
                                                                 const double *__restrict__ ER,
                                                                 const double *__restrict__ EP,
                                                                 const double *__restrict__ EX,
                                                                 double *__restrict__ Pout, double *Xout) {
  const int lane = threadIdx.x % kWave, hi = lane / 16, lo = lane % 16;
  double cr[4], cp[4], cx[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kk = 4 * s + hi;
    cr[s] = (kk < na && lo < k) ? ER[kk + (size_t)lo * na] : 0.0;
    cp[s] = (Pa && kk < na && lo < k) ? EP[kk + (size_t)lo * na] : 0.0;
    cx[s] = (kk < k && lo < k) ? EX[kk + (size_t)lo * k] : 0.0;
  }
  const int sr = (na + 3) / 4, sx = (k + 3) / 4;
  const int64_t ntile = ((int64_t)n + 15) / 16;
  const int64_t nwave = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t tile = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; tile < ntile; tile += nwave) {
    const int64_t r0 = tile * 16, ar = r0 + lo;
    const bool rok = ar < n;
    double fr[4], fp[4], fx[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int kk = 4 * s + hi;
      fr[s] = (s < sr && rok && kk < na) ? R[ar * na + kk] : 0.0;
      fp[s] = (Pa && s < sr && rok && kk < na) ? Pa[ar * na + kk] : 0.0;
      fx[s] = (s < sx && rok && kk < k) ? X[ar * k + kk] : 0.0;
    }
    mf64x4 p = {0.0, 0.0, 0.0, 0.0}, x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < sr) p = mfma_f64_16x16x4(fr[s], cr[s], p);
    if (Pa)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < sr) p = mfma_f64_16x16x4(fp[s], cp[s], p);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < sx) x = mfma_f64_16x16x4(fx[s], cx[s], x);
    if (lo < k)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = r0 + hi + 4 * e;
        if (row < n) {
          const size_t zi = (size_t)row * k + lo;
          Pout[zi] = p[e];
          Xout[zi] = p[e] + x[e];
        }
      }
  }
}

// Y (n x m) = columns idx[0..m) of the residual block R = AX - X diag(lam) (n x k): the
// product rounded, then the difference (as gram_mfma_kernel<3> forms R).
static __global__ __launch_bounds__(kBlock) void block_resid_gather_kernel(int n, int k, const double *__restrict__ AX,
                                                                    const double *__restrict__ X,
                                                                    const double *__restrict__ lam, int m,
                                                                    const int *__restrict__ idx,
                                                                    double *__restrict__ Y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * m) return;
  const int64_t i = t / m;
  const int c = idx[t % m];
  Y[t] = AX[i * k + c] - X[i * k + c] * lam[c];
}

// Y (n x m) = columns idx[0..m) of X (n x k).
static __global__ __launch_bounds__(kBlock) void block_gather_kernel(int n, int k, const double *__restrict__ X,
                                                              int m, const int *__restrict__ idx,
                                                              double *__restrict__ Y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * m) return;
  const int64_t i = t / m;
  Y[t] = X[i * k + idx[t % m]];
}

// Y (n x kp) = [X | 0] for X (n x k), k <= kp: the block widened with zero columns.
static __global__ __launch_bounds__(kBlock) void block_pad_kernel(int n, int k, int kp, const double *__restrict__ X,
                                                           double *__restrict__ Y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * kp) return;
  const int64_t i = t / kp;
  const int c = (int)(t % kp);
  Y[t] = c < k ? X[i * k + c] : 0.0;
}

// X (n x k) = the first k columns of Y (n x kp).
static __global__ __launch_bounds__(kBlock) void block_unpad_kernel(int n, int k, int kp, const double *__restrict__ Y,
                                                             double *__restrict__ X) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  X[t] = Y[(t / k) * kp + t % k];
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
// its partial sums with plain stores; a one-workgroup finalize kernel adds them in block
// order and performs the scalar update (the kernel boundary makes the partials visible,
// so no device-scope fence is needed per block).
// ---------------------------------------------------------------------------------------
struct PcgState {
  double bnorm, resiter, ptap, pr, zap, alpha, beta, tol;
  int iter, itmax, stop, breakdown;
};

constexpr int kFinBlock = 1024;

// Partials of b.b and r.r (r = b - A x is formed by the stream SpMV beforehand).
static __global__ __launch_bounds__(kBlock) void pcg_norms_kernel(int n, const double *__restrict__ b,
                                                           const double *__restrict__ r,
                                                           double *__restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  double bb = 0.0, rr = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    bb += b[i] * b[i];
    rr += r[i] * r[i];
  }
  bb = block_sum(bb, lds);
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = bb;
    part[gridDim.x + blockIdx.x] = rr;
  }
}

// Partials of z . Ap_old (iterations > 1).
static __global__ __launch_bounds__(kBlock) void pcg_zap_kernel(int n, const double *__restrict__ z,
                                                         const double *__restrict__ ap,
                                                         double *__restrict__ part,
                                                         const PcgState *st) {
  if (st->stop || st->iter == 0) return;
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += z[i] * ap[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// p_new = z + beta p_old, Ap = A p_new, and per-block partials of p.Ap and p.r over the row
// blocks of A (as in spmv_stream_kernel).  FORMED = false: p_new is formed on the fly at
// every gathered column (z and p_old gathered) and written for the block's rows; FORMED =
// true: p_new was formed by pcg_pnew_kernel and is the one gathered operand.  Both forms
// compute every p_i by the same expression, so they give the same bits.
struct PColVal {
  const double *__restrict__ z;
  const double *__restrict__ p;
  double beta;
  __device__ __forceinline__ double operator()(int j) const { return z[j] + beta * p[j]; }
};
struct PColFormed {
  const double *__restrict__ p;
  __device__ __forceinline__ double operator()(int j) const { return p[j]; }
};

// p_new = z + beta p_old (PCG_CJ.m:45), formed before the product so that the product
// gathers one operand per nonzero.
static __global__ __launch_bounds__(kBlock) void pcg_pnew_kernel(int n, const double *__restrict__ z,
                                                          const double *__restrict__ p_old,
                                                          double *__restrict__ p_new,
                                                          const PcgState *st) {
  if (st->stop) return;
  const double beta = st->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p_new[i] = z[i] + beta * p_old[i];
}

template <int PER, typename TI, bool FORMED>
__global__ __launch_bounds__(kBlock) void pcg_p_stream_kernel(
    int nblk, const int4 *__restrict__ blk, const int *__restrict__ rp, const TI *__restrict__ ci,
    const double *__restrict__ val, const double *__restrict__ z, const double *__restrict__ p_old,
    double *__restrict__ p_new, double *__restrict__ ap, const double *__restrict__ r,
    double *__restrict__ part, const PcgState *st) {
  __shared__ double prod[PER * kBlock];
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  if (st->stop) return;
  const double beta = st->beta;
  const int r0 = lo.x, r1 = hi.x, e0 = lo.y, e1 = hi.y, base = lo.z;
  const int tid = threadIdx.x;
  using CV = typename std::conditional<FORMED, PColFormed, PColVal>::type;
  CV cv;
  if constexpr (FORMED) cv = CV{p_new};
  else cv = CV{z, p_old, beta};
  double pap = 0.0, prr = 0.0;
  if (e1 - e0 > PER * kBlock) {
    double s = strided_row_sum<double>(e0, e1, [&](int e) { return val[e] * cv(base + (int)ci[e]); });
    s = block_sum_t<double>(s, prod);
    if (tid == 0) {
      const double pi = cv(r0);
      if constexpr (!FORMED) p_new[r0] = pi;
      ap[r0] = s;
      pap = pi * s;
      prr = pi * r[r0];
    }
  } else {
    const int row = r0 + tid;
    const int rc = min(row, r1 - 1);
    const int ra = rp[rc] - e0, rz = rp[rc + 1] - e0;
    const double pi = cv(rc), ri = r[rc];
    stream_products<PER, double>(e0, e1, base, ci, val, cv, prod);
    __syncthreads();
    if (row < r1) {
      const double s = row_sum<double>(prod, ra, rz);
      if constexpr (!FORMED) p_new[row] = pi;
      ap[row] = s;
      pap = pi * s;
      prr = pi * ri;
    }
    __syncthreads();
  }
  pap = block_sum_t<double>(pap, prod);
  prr = block_sum_t<double>(prr, prod);
  if (tid == 0) {
    part[b] = pap;
    part[nblk + b] = prr;
  }
}

// Ap = A p in the symmetric diagonal layout (spmv_dia_kernel's loads and summation order)
// and per-block partials of p.Ap and p.r, one block of kBlock rows per workgroup (grid =
// the blocks, partials at part[b] and part[nblk + b] for the row block b).  FORMED: p_new
// was formed by pcg_pnew_kernel and is the operand; otherwise the kernel reads z and p_old
// at the same positions, forms p = z + beta p_old there with pcg_pnew_kernel's expression
// (same bits) and writes p_new for its own rows.
template <int MAXM, bool FORMED>
__global__ __launch_bounds__(kBlock) void pcg_p_dia_kernel(int n, DiaOffs d, const double *__restrict__ val,
                                                           const double *__restrict__ z,
                                                           const double *__restrict__ p_old,
                                                           double *__restrict__ p_new, double *__restrict__ ap,
                                                           const double *__restrict__ r,
                                                           double *__restrict__ part, const PcgState *st) {
  __shared__ double lds[kBlock / kWave];
  const int nblk = (int)gridDim.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int i = b * kBlock + (int)threadIdx.x;
  const int ic = min(i, n - 1);
  DiaRow<MAXM, double, double, VecLoad<double>> row;
  row.load(ic, n, d, val, VecLoad<double>{FORMED ? p_new : z});
  double pl[MAXM], pu[MAXM], pd = 0.0;
  if constexpr (!FORMED) {
#pragma unroll
    for (int k = 1; k < MAXM; ++k) {
      pl[k] = pu[k] = 0.0;
      if (k < d.m) {
        pl[k] = p_old[max(ic - d.off[k], 0)];
        pu[k] = p_old[min(ic + d.off[k], n - 1)];
      }
    }
    pd = p_old[ic];
  }
  const double ri = r[ic];
  if (st->stop) return;
  if constexpr (!FORMED) {
    const double beta = st->beta;
#pragma unroll
    for (int k = 1; k < MAXM; ++k) {
      row.xl[k] = row.xl[k] + beta * pl[k];
      row.xu[k] = row.xu[k] + beta * pu[k];
    }
    row.xd = row.xd + beta * pd;
  }
  double pap = 0.0, prr = 0.0;
  if (i < n) {
    const double s = row.sum();
    if constexpr (!FORMED) p_new[i] = row.xd;
    ap[i] = s;
    pap = row.xd * s;
    prr = row.xd * ri;
  }
  pap = block_sum_t<double>(pap, lds);
  prr = block_sum_t<double>(prr, lds);
  if (threadIdx.x == 0) {
    part[b] = pap;
    part[nblk + b] = prr;
  }
}

// x += alpha p ; r -= alpha Ap ; partials of r.r; r32 (when not null) = (float) r.
static __global__ __launch_bounds__(kBlock) void pcg_update_kernel(int n, const double *__restrict__ p,
                                                            const double *__restrict__ ap,
                                                            double *__restrict__ x,
                                                            double *__restrict__ r,
                                                            float *__restrict__ r32,
                                                            double *__restrict__ part,
                                                            const PcgState *st) {
  if (st->stop) return;
  __shared__ double lds[kBlock / kWave];
  const double alpha = st->alpha;
  double rr = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    if (r32) r32[i] = (float)ri;
    rr += ri * ri;
  }
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// Deterministic sums of N arrays of count partials (array k at part + k * count) by one
// workgroup of kFinBlock threads; the results land in t[] of thread 0.  Each thread adds
// its strided partials of every array in order, with the loads of sixteen steps of all N
// arrays issued before their adds; then a fixed butterfly per wave and the waves in order.
// The order of each sum does not depend on N or on the look-ahead.
template <int N>
__device__ __forceinline__ void fin_sum(const double *part, int count, double *lds, double (&t)[N]) {
  constexpr int kAhead = 16;
  double s[N];
#pragma unroll
  for (int k = 0; k < N; ++k) s[k] = 0.0;
  int i = threadIdx.x;
  for (; i + (kAhead - 1) * kFinBlock < count; i += kAhead * kFinBlock) {
    double v[N][kAhead];
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int a = 0; a < kAhead; ++a) v[k][a] = part[(size_t)k * count + i + a * kFinBlock];
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int a = 0; a < kAhead; ++a) s[k] += v[k][a];
  }
  for (; i < count; i += kFinBlock)
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] += part[(size_t)k * count + i];
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    s[k] = group_sum<kWave>(s[k]);
    if (l == 0) lds[k * (kFinBlock / kWave) + w] = s[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    t[k] = 0.0;
    if (threadIdx.x == 0)
      for (int j = 0; j < kFinBlock / kWave; ++j) t[k] += lds[k * (kFinBlock / kWave) + j];
  }
  __syncthreads();
}

enum { kFinInit = 0, kFinBeta = 1, kFinAlpha = 2, kFinUpdate = 3 };

// Scalar updates of PCG_CJ.m after each reduction (one workgroup).
// host_out (may be null): host-mapped pinned memory that receives the state when the kernel
// ends (the iteration's closing finalisation), stopped or not, so that no copy follows.
static __global__ __launch_bounds__(kFinBlock) void pcg_finalize_kernel(int kind, int count,
                                                                 const double *__restrict__ part,
                                                                 PcgState *st,
                                                                 double *__restrict__ resvec,
                                                                 PcgState *host_out = nullptr) {
  __shared__ double lds[2 * (kFinBlock / kWave)];
  if (kind != kFinInit && st->stop) {
    if (host_out && threadIdx.x == 0) {
      *host_out = *st;
      __threadfence_system();
    }
    return;
  }
  if (kind == kFinBeta && st->iter == 0) {
    if (threadIdx.x == 0) st->beta = 0.0;  // first iteration: p = z (PCG_CJ.m:95-96)
    return;
  }
  double s0, s1 = 0.0;
  if (kind == kFinInit || kind == kFinAlpha) {
    double t[2];
    fin_sum<2>(part, count, lds, t);
    s0 = t[0];
    s1 = t[1];
  } else {
    double t[1];
    fin_sum<1>(part, count, lds, t);
    s0 = t[0];
  }
  if (threadIdx.x != 0) return;
  switch (kind) {
    case kFinInit:
      st->bnorm = sqrt(s0);
      st->iter = 0;
      st->breakdown = 0;
      if (s0 == 0.0) {
        st->resiter = 0.0;
        st->stop = 1;
      } else {
        st->resiter = sqrt(s1) / st->bnorm;
        st->stop = st->itmax <= 0 ? 1 : 0;
      }
      resvec[0] = st->resiter;
      break;
    case kFinBeta:
      st->zap = s0;
      st->beta = -s0 / st->ptap;
      break;
    case kFinAlpha:
      st->ptap = s0;
      st->pr = s1;
      st->alpha = s1 / s0;
      if (!(s0 != 0.0) || !isfinite(s0)) {
        st->breakdown = 1;
        st->stop = 1;
      }
      break;
    default: {
      const int it = st->iter + 1;
      st->iter = it;
      st->resiter = sqrt(s0) / st->bnorm;
      resvec[it] = st->resiter;
      if (st->resiter < st->tol || it >= st->itmax) st->stop = 1;
      if (host_out) {
        *host_out = *st;
        __threadfence_system();
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Adjoint gradient for the finite-volume diffusion operator with harmonic transmissibilities
// T_ij = 2 k_i k_j / (k_i + k_j) = -a_ij and Dirichlet terms proportional to k_i (their sum
// over the row is the row sum of A).  With A lam = dL/dx:
//   dL/d ln k_i = -( sum_{j != i} T_ij k_j/(k_i+k_j) (lam_i - lam_j)(x_i - x_j) + rowsum_i lam_i x_i )
// One thread per row over A's pattern (row blocks of the stream layout).
// ---------------------------------------------------------------------------------------
template <typename TI>
__global__ __launch_bounds__(kBlock) void fv_grad_kernel(int nblk, const int4 *__restrict__ blk,
                                                         const int *__restrict__ rp,
                                                         const TI *__restrict__ ci,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ kk,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ lam,
                                                         double *__restrict__ g) {
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  const int base = lo.z;
  for (int r = lo.x + threadIdx.x; r < hi.x; r += kBlock) {
    const double ki = kk[r], xi = x[r], li = lam[r];
    double s = 0.0, rowsum = 0.0;
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
      const int j = base + (int)ci[e];
      const double a = val[e];
      rowsum += a;
      if (j != r) {
        const double kj = kk[j];
        s += (-a) * kj / (ki + kj) * (li - lam[j]) * (xi - x[j]);
      }
    }
    g[r] = -(s + rowsum * li * xi);
  }
}

}  // namespace chr

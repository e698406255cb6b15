// Shared host runtime of the HIP solve phase: device buffers, the device CSR layout,
// kernel launchers, the opaque handle types and the V-cycle sequence.
#pragma once

#include "chronos_amg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.hpp"
#include "sell_layout.hpp"

namespace chrrt {

extern thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPC(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(CHR_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

// Owned device allocation.
// Element count of a CSR value or index array as allocated: a multiple of four (at least
// four), so that the products' four-element vector loads stay inside the array.
inline int64_t quad_len(int64_t nnz) { return std::max<int64_t>(4, (nnz + 3) & ~(int64_t)3); }

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t b) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = b;
    if (b == 0) return CHR_OK;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(CHR_EALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return CHR_OK;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

inline int check_csr(const chr_csr &m, const char *what) {
  if (m.nrows < 0 || m.ncols < 0 || !m.rowptr || (m.rowptr[m.nrows] > 0 && (!m.col || !m.val)))
    return fail(CHR_EARG, std::string(what) + ": null pointer or negative size");
  if (m.rowptr[0] != 0) return fail(CHR_EARG, std::string(what) + ": rowptr[0] != 0");
  for (int32_t i = 0; i < m.nrows; ++i) {
    if (m.rowptr[i + 1] < m.rowptr[i])
      return fail(CHR_EARG, std::string(what) + ": rowptr not monotone");
    for (int32_t e = m.rowptr[i]; e < m.rowptr[i + 1]; ++e)
      if (m.col[e] < 0 || m.col[e] >= m.ncols)
        return fail(CHR_EARG, std::string(what) + ": column index out of range");
  }
  return CHR_OK;
}

// Host CSR transpose (columns of the result come out sorted).
inline void transpose_host(const chr_csr &M, std::vector<int32_t> &rp, std::vector<int32_t> &ci,
                    std::vector<double> &v) {
  const int64_t nnz = M.rowptr[M.nrows];
  rp.assign((size_t)M.ncols + 1, 0);
  ci.resize(nnz);
  v.resize(nnz);
  for (int64_t e = 0; e < nnz; ++e) rp[M.col[e] + 1]++;
  for (int32_t i = 0; i < M.ncols; ++i) rp[i + 1] += rp[i];
  std::vector<int32_t> pos(rp.begin(), rp.end() - 1);
  for (int32_t i = 0; i < M.nrows; ++i)
    for (int32_t e = M.rowptr[i]; e < M.rowptr[i + 1]; ++e) {
      int32_t p = pos[M.col[e]]++;
      ci[p] = i;
      v[p] = M.val[e];
    }
}

// Copies ncols columns of `bytes` bytes each between column-major blocks with leading
// dimensions (in bytes) ldd and lds: one copy when both are packed, else one per column.
inline hipError_t copy_cols(void *dst, size_t ldd, const void *src, size_t lds, size_t bytes, int ncols,
                            hipMemcpyKind kind, hipStream_t s) {
  if (ldd == bytes && lds == bytes) return hipMemcpyAsync(dst, src, bytes * (size_t)ncols, kind, s);
  for (int c = 0; c < ncols; ++c) {
    hipError_t e = hipMemcpyAsync(static_cast<char *>(dst) + ldd * c, static_cast<const char *>(src) + lds * c,
                                  bytes, kind, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Candidate nonzeros-per-thread of the stream kernel (block capacity PER * 256).
constexpr int kPers[] = {4, 8, 16};

// Sliced-ELL copy of a square matrix for single-vector products (spmv_sell_kernel): 64-row
// slices stored column by column, rows optionally ordered by decreasing length inside
// sort windows of 2^sshift rows, column codes as described at the kernel.
struct DevSell {
  int cb = 0;  // code bits (8, 16, 32); 0 = not built
  int32_t nrows = 0, nslices = 0, wshift = 14, sshift = 14, dmin = 0;
  int64_t elems = 0;  // stored elements, padding included
  bool sorted = false;
  int unroll = 8;     // elements per chunk of the kernel (16 when a slice is longer than 8)
  int ulen = 0;       // > 0: uniform ELL, every slice ulen elements long
  DevBuf sp, perm, val, code, dict;
  // Builds the layout from host CSR (chrsell::build) and uploads it; leaves cb = 0 (and
  // returns CHR_OK) when the matrix is not a candidate.
  int build(int32_t n, int32_t ncols, const int32_t *hrp, const int32_t *hci, const double *hv,
            bool fp32, int sort_shift = -1);
  // Uploads a layout built on the host (cb = 0: not a candidate, nothing uploaded).
  int adopt(const chrsell::Layout &L, bool fp32);
  void release() {
    sp = DevBuf();
    perm = DevBuf();
    val = DevBuf();
    code = DevBuf();
    dict = DevBuf();
    cb = 0;
  }
  double bytes(bool fp32) const {
    return (double)elems * ((fp32 ? 4 : 8) + cb / 8) + 4.0 * (nslices + 1) +
           (double)(perm.bytes + dict.bytes);
  }
};

// Symmetric diagonal copy of a square matrix (dia_layout.hpp) for single-vector products
// (spmv_dia_kernel) and the PCG operator product (pcg_p_dia_kernel).
struct DevDia {
  int32_t n = 0, m = 0;  // m = 0: not built
  chr::DiaOffs offs{};
  DevBuf val;
  // Builds the layout from host CSR and uploads it; leaves m = 0 (and returns CHR_OK) when
  // the matrix is not a candidate.
  int build(int32_t n, int32_t ncols, const int32_t *hrp, const int32_t *hci, const double *hv, bool fp32);
  void release() {
    val = DevBuf();
    m = 0;
  }
  double bytes(bool fp32) const { return (double)m * n * (fp32 ? 4 : 8); }
  unsigned grid() const { return (unsigned)((n + chr::kBlock - 1) / chr::kBlock); }
};

// Packed-code index copy for single-vector stream products (spmv_pack_kernel): block table
// with dictionary windows, one 16-bit code per nonzero, per-window dictionaries.  Values
// and row pointers are the DevCsr's own (same nonzero order).
struct DevPack {
  int per = 8;
  int32_t nblk = 0;
  bool built = false;
  DevBuf blk, code, dict;
  int build(int32_t n, int32_t ncols, const int32_t *hrp, const int32_t *hci, int per_);
  void release() {
    blk = DevBuf();
    code = DevBuf();
    dict = DevBuf();
    built = false;
  }
};

// Device CSR in FP32 or FP64, laid out for the stream kernel.
struct DevCsr {
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  bool fp32 = false;
  bool idx16 = false;  // uint16 column offsets from a per-block base
  int per = 8;         // nonzeros per thread (block capacity per * kBlock)
  int32_t nblk = 0;    // row blocks
  DevBuf rp, val, blk;
  DevBuf ci;           // uint16 offsets when idx16, else int32 columns
  DevSell sell;        // sliced-ELL copy, used for single-vector products when use_sell
  bool use_sell = false;
  int sell_arg = -2;   // the sort argument the sliced-ELL copy was built with (-2: none)
  DevPack pack;        // packed-code index copy, used for single-vector products when use_pack
  bool use_pack = false;
  DevDia dia;          // symmetric diagonal copy, used for single-vector products when use_dia
  bool use_dia = false;
  bool dia_pcg = false;  // the PCG operator product runs on dia (a fixed rule, never timed)
  // FP32 stream and packed-code single-vector products stage their nonzeros with
  // four-element vector loads (else one element per load; FP64 always); same products and
  // sums either way.  Set by the autotuner.
  bool vec4 = false;
  int upload(int32_t nr, int32_t nc, const int32_t *hrp, const int32_t *hci, const double *hv,
             bool as_fp32, int per_ = 8) {
    nrows = nr;
    ncols = nc;
    nnz = hrp[nr];
    fp32 = as_fp32;
    int rc;
    if ((rc = rp.alloc(sizeof(int32_t) * (nr + 1)))) return rc;
    if ((rc = val.alloc((as_fp32 ? 4 : 8) * (size_t)quad_len(nnz)))) return rc;
    HIPC(hipMemcpy(rp.p, hrp, sizeof(int32_t) * (nr + 1), hipMemcpyHostToDevice));
    if (nnz) {
      if (as_fp32) {
        std::vector<float> f(hv, hv + nnz);
        HIPC(hipMemcpy(val.p, f.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
      } else {
        HIPC(hipMemcpy(val.p, hv, sizeof(double) * nnz, hipMemcpyHostToDevice));
      }
    }
    return layout(per_, hrp, hci);
  }
  // Row blocks for a given per: whole rows, <= per * kBlock nonzeros, <= kBlock rows, as
  // int4 {first row, first nonzero, base column, 0}, nblk + 1 of them; column indices as
  // uint16 offsets from the block's smallest column when every block spans < 65536 columns.
  int layout(int per_, const int32_t *hrp, const int32_t *hci) {
    per = per_;
    const int64_t cap = (int64_t)per * chr::kBlock;
    std::vector<int4> br;
    br.push_back(make_int4(0, 0, 0, 0));
    int32_t rows = 0;
    int64_t cnt = 0;
    for (int32_t r = 0; r < nrows; ++r) {
      const int64_t len = hrp[r + 1] - hrp[r];
      if (rows > 0 && (rows == chr::kBlock || cnt + len > cap)) {
        br.push_back(make_int4(r, hrp[r], 0, 0));
        rows = 0;
        cnt = 0;
      }
      ++rows;
      cnt += len;
    }
    if (nrows > 0) br.push_back(make_int4(nrows, hrp[nrows], 0, 0));
    nblk = (int32_t)br.size() - 1;
    idx16 = nnz > 0;
    for (int32_t b = 0; b < nblk && idx16; ++b) {
      int32_t lo = INT32_MAX, hi = -1;
      for (int64_t e = br[b].y; e < br[b + 1].y; ++e) {
        lo = std::min(lo, hci[e]);
        hi = std::max(hi, hci[e]);
      }
      if (hi >= 0 && hi - lo > 65535) idx16 = false;
      br[b].z = hi >= 0 ? lo : 0;
    }
    if (!idx16)
      for (auto &q : br) q.z = 0;
    int rc;
    if ((rc = blk.alloc(sizeof(int4) * br.size()))) return rc;
    HIPC(hipMemcpy(blk.p, br.data(), sizeof(int4) * br.size(), hipMemcpyHostToDevice));
    if (idx16) {
      std::vector<uint16_t> d(nnz);
      for (int32_t b = 0; b < nblk; ++b)
        for (int64_t e = br[b].y; e < br[b + 1].y; ++e) d[e] = (uint16_t)(hci[e] - br[b].z);
      if ((rc = ci.alloc(sizeof(uint16_t) * (size_t)quad_len(nnz)))) return rc;
      HIPC(hipMemcpy(ci.p, d.data(), sizeof(uint16_t) * nnz, hipMemcpyHostToDevice));
    } else {
      if ((rc = ci.alloc(sizeof(int32_t) * (size_t)quad_len(nnz)))) return rc;
      if (nnz) HIPC(hipMemcpy(ci.p, hci, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    }
    return CHR_OK;
  }
  int upload(const chr_csr &m, bool as_fp32) {
    return upload(m.nrows, m.ncols, m.rowptr, m.col, m.val, as_fp32);
  }
  // Bytes of one single-vector product as stored: values + column codes + row pointers and
  // block table (stream layout) or slice offsets, permutation and dictionaries (sliced ELL).
  double stream_bytes() const {
    if (use_dia) return dia.bytes(fp32);
    if (use_sell) return sell.bytes(fp32);
    if (use_pack)
      return (double)nnz * ((fp32 ? 4 : 8) + 2) + 4.0 * (nrows + 1) + 16.0 * (pack.nblk + 1) + (double)pack.dict.bytes;
    return (double)nnz * ((fp32 ? 4 : 8) + (idx16 ? 2 : 4)) + 4.0 * (nrows + 1) + 16.0 * (nblk + 1);
  }
  // The same with 4-byte column indices and no block table (the reference's CSR format).
  double csr_bytes() const { return (double)nnz * ((fp32 ? 4 : 8) + 4) + 4.0 * (nrows + 1); }
  double device_bytes() const {
    return (double)(rp.bytes + ci.bytes + val.bytes + blk.bytes + sell.sp.bytes + sell.perm.bytes +
                    sell.val.bytes + sell.code.bytes + sell.dict.bytes + pack.blk.bytes + pack.code.bytes +
                    pack.dict.bytes + dia.val.bytes);
  }
};

template <int MAXM, typename TV, typename TX, typename Epi>
hipError_t launch_dia_m(const DevDia &D, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  hipLaunchKernelGGL((chr::spmv_dia_kernel<MAXM, TV, TX, Epi>), dim3(D.grid()), dim3(chr::kBlock), 0, s, D.n,
                     D.offs, D.val.as<TV>(), x, epi, stop);
  return hipGetLastError();
}

template <typename TV, typename TX, typename Epi>
hipError_t launch_dia(const DevDia &D, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  if (D.m <= 4) return launch_dia_m<4, TV>(D, x, epi, stop, s);
  if (D.m <= 8) return launch_dia_m<8, TV>(D, x, epi, stop, s);
  return launch_dia_m<16, TV>(D, x, epi, stop, s);
}

template <int CB, int U, typename TV, typename TX, typename Epi>
hipError_t launch_sell_cb(const DevSell &S, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  const unsigned grid = (unsigned)((S.nslices + chr::kBlock / chr::kWave - 1) / (chr::kBlock / chr::kWave));
  hipLaunchKernelGGL((chr::spmv_sell_kernel<CB, U, TV, TX, Epi>), dim3(grid), dim3(chr::kBlock), 0, s,
                     S.nslices, S.nrows, S.wshift, S.sshift, S.ulen, S.sp.as<int>(),
                     S.sorted ? S.perm.as<uint16_t>() : nullptr, S.val.as<TV>(),
                     S.code.as<typename chr::SellCode<CB>::type>(), S.dict.as<int>(), S.dmin, x, epi,
                     stop);
  return hipGetLastError();
}

template <int U, typename TV, typename TX, typename Epi>
hipError_t launch_sell_u(const DevSell &S, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  switch (S.cb) {
    case 8: return launch_sell_cb<8, U, TV>(S, x, epi, stop, s);
    case 16: return launch_sell_cb<16, U, TV>(S, x, epi, stop, s);
    default: return launch_sell_cb<32, U, TV>(S, x, epi, stop, s);
  }
}

template <typename TV, typename TX, typename Epi>
hipError_t launch_sell(const DevSell &S, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  return S.unroll == 16 ? launch_sell_u<16, TV>(S, x, epi, stop, s) : launch_sell_u<8, TV>(S, x, epi, stop, s);
}

template <int PER, typename TV, typename TX, typename Epi, bool V4>
hipError_t launch_pack_v(const DevCsr &M, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  hipLaunchKernelGGL((chr::spmv_pack_kernel<PER, TV, TX, Epi, V4>), dim3(M.pack.nblk), dim3(chr::kBlock), 0, s,
                     M.pack.nblk, M.pack.blk.as<int4>(), M.rp.as<int>(), M.pack.code.as<uint16_t>(),
                     M.val.as<TV>(), M.pack.dict.as<int>(), x, epi, stop);
  return hipGetLastError();
}

template <int PER, typename TV, typename TX, typename Epi>
hipError_t launch_pack_p(const DevCsr &M, const TX *x, Epi epi, const int *stop, hipStream_t s) {
  if constexpr (std::is_same<TV, float>::value)
    if (M.vec4) return launch_pack_v<PER, TV, TX, Epi, true>(M, x, epi, stop, s);
  return launch_pack_v<PER, TV, TX, Epi, false>(M, x, epi, stop, s);
}

template <int PER, typename TV, typename XS, typename Epi, bool V4>
hipError_t launch_spmv_v(const DevCsr &M, XS x, Epi epi, const int *stop, hipStream_t s) {
  if (M.idx16)
    hipLaunchKernelGGL((chr::spmv_stream_kernel<PER, TV, uint16_t, XS, Epi, V4>), dim3(M.nblk),
                       dim3(chr::kBlock), 0, s, M.nblk, M.blk.as<int4>(), M.rp.as<int>(),
                       M.ci.as<uint16_t>(), M.val.as<TV>(), x, epi, stop);
  else
    hipLaunchKernelGGL((chr::spmv_stream_kernel<PER, TV, int, XS, Epi, V4>), dim3(M.nblk),
                       dim3(chr::kBlock), 0, s, M.nblk, M.blk.as<int4>(), M.rp.as<int>(),
                       M.ci.as<int>(), M.val.as<TV>(), x, epi, stop);
  return hipGetLastError();
}

template <int PER, typename TV, typename XS, typename Epi>
hipError_t launch_spmv_p(const DevCsr &M, XS x, Epi epi, const int *stop, hipStream_t s) {
  if constexpr (std::is_same<TV, float>::value)
    if (M.vec4) return launch_spmv_v<PER, TV, XS, Epi, true>(M, x, epi, stop, s);
  return launch_spmv_v<PER, TV, XS, Epi, false>(M, x, epi, stop, s);
}

template <typename XS>
struct is_x_one : std::false_type {};
template <typename TX>
struct is_x_one<chr::XOne<TX>> : std::true_type {};

template <typename TV, typename XS, typename Epi>
hipError_t launch_spmv_t(const DevCsr &M, XS x, Epi epi, const int *stop, hipStream_t s) {
  if (M.nrows == 0) return hipSuccess;
  if constexpr (is_x_one<XS>::value) {
    if (M.use_dia) return launch_dia<TV>(M.dia, x.a, epi, stop, s);
    if (M.use_sell) return launch_sell<TV>(M.sell, x.a, epi, stop, s);
    if (M.use_pack) {
      switch (M.pack.per) {
        case 4: return launch_pack_p<4, TV>(M, x.a, epi, stop, s);
        case 16: return launch_pack_p<16, TV>(M, x.a, epi, stop, s);
        default: return launch_pack_p<8, TV>(M, x.a, epi, stop, s);
      }
    }
  } else {
    // Composite operators (two operand vectors) are kept in the stream layout only.
    if (M.use_sell || M.use_pack || M.use_dia) return hipErrorInvalidValue;
  }
  switch (M.per) {
    case 4: return launch_spmv_p<4, TV>(M, x, epi, stop, s);
    case 16: return launch_spmv_p<16, TV>(M, x, epi, stop, s);
    default: return launch_spmv_p<8, TV>(M, x, epi, stop, s);
  }
}

template <typename TX>
chr::XOne<TX> xsrc(const TX *x) {
  return chr::XOne<TX>{x};
}
template <typename TX>
chr::XTwo<TX> xsrc(chr::XTwo<TX> x) {
  return x;
}

template <typename XP, typename Epi>
hipError_t launch_spmv(const DevCsr &M, XP xp, Epi epi, const int *stop, hipStream_t s) {
  const auto x = xsrc(xp);
  return M.fp32 ? launch_spmv_t<float>(M, x, epi, stop, s)
                : launch_spmv_t<double>(M, x, epi, stop, s);
}

template <typename TX>
bool aligned16(const chr::XOne<TX> &x) {
  return reinterpret_cast<uintptr_t>(x.a) % 16 == 0;
}
template <typename TX>
bool aligned16(const chr::XTwo<TX> &x) {
  return reinterpret_cast<uintptr_t>(x.a) % 16 == 0 && x.db % 16 == 0;
}

template <int KL, typename TV, typename XS, typename Epi>
hipError_t launch_spmm_kl(const DevCsr &M, XS x, int k, Epi epi, const int *stop,
                          hipStream_t s) {
  const int cap = M.per * chr::kBlock;  // nonzeros of a stream block (longer: one long row)
  if constexpr (KL >= 4) {
    // Four columns per lane when k allows it and X takes 16-byte loads (same bits).
    if (k % 4 == 0 && aligned16(x) && !std::getenv("CHRONOS_SPMM_SCALAR")) {
      if (M.idx16)
        hipLaunchKernelGGL((chr::spmm4_kernel<KL, TV, uint16_t, XS, Epi>), dim3(M.nblk), dim3(chr::kBlock),
                           (size_t)cap * (sizeof(TV) + sizeof(uint16_t)) + chr::kStageRpBytes, s, M.nblk, M.blk.as<int4>(),
                           M.rp.as<int>(), M.ci.as<uint16_t>(), M.val.as<TV>(), x, k, cap, epi, stop);
      else
        hipLaunchKernelGGL((chr::spmm4_kernel<KL, TV, int, XS, Epi>), dim3(M.nblk), dim3(chr::kBlock),
                           (size_t)cap * (sizeof(TV) + sizeof(int)) + chr::kStageRpBytes, s, M.nblk, M.blk.as<int4>(),
                           M.rp.as<int>(), M.ci.as<int>(), M.val.as<TV>(), x, k, cap, epi, stop);
      return hipGetLastError();
    }
  }
  if (M.idx16)
    hipLaunchKernelGGL((chr::spmm_kernel<KL, TV, uint16_t, XS, Epi>), dim3(M.nblk), dim3(chr::kBlock),
                       (size_t)cap * (sizeof(TV) + sizeof(uint16_t)) + chr::kStageRpBytes, s, M.nblk, M.blk.as<int4>(),
                       M.rp.as<int>(), M.ci.as<uint16_t>(), M.val.as<TV>(), x, k, cap, epi, stop);
  else
    hipLaunchKernelGGL((chr::spmm_kernel<KL, TV, int, XS, Epi>), dim3(M.nblk), dim3(chr::kBlock),
                       (size_t)cap * (sizeof(TV) + sizeof(int)) + chr::kStageRpBytes, s, M.nblk, M.blk.as<int4>(),
                       M.rp.as<int>(), M.ci.as<int>(), M.val.as<TV>(), x, k, cap, epi, stop);
  return hipGetLastError();
}

template <typename TV, typename XS, typename Epi>
hipError_t launch_spmm_t(const DevCsr &M, XS x, int k, Epi epi, const int *stop,
                         hipStream_t s) {
  if (k <= 2) return launch_spmm_kl<2, TV>(M, x, k, epi, stop, s);
  if (k <= 4) return launch_spmm_kl<4, TV>(M, x, k, epi, stop, s);
  if (k <= 8) return launch_spmm_kl<8, TV>(M, x, k, epi, stop, s);
  return launch_spmm_kl<16, TV>(M, x, k, epi, stop, s);
}

// Y = M X on row-major n x k blocks (k = 1: the stream kernel; 2 <= k <= 16: the block
// kernel, whose columns equal the stream kernel's results bit for bit).  x: a vector
// pointer, or chr::XTwo for a composite operator over two vectors.
template <typename XP, typename Epi>
hipError_t launch_op(const DevCsr &M, XP xp, int k, Epi epi, const int *stop, hipStream_t s) {
  if (k == 1) return launch_spmv(M, xp, epi, stop, s);
  if (M.nrows == 0) return hipSuccess;
  const auto x = xsrc(xp);
  return M.fp32 ? launch_spmm_t<float>(M, x, k, epi, stop, s)
                : launch_spmm_t<double>(M, x, k, epi, stop, s);
}

constexpr int kMaxBlockCols = 16;


// flush: a device buffer of 1 GiB or so whose halves are copied onto each other to evict
// the Infinity Cache before cache-cold timings (NULL: every candidate timed back to back).
// tune_k > 1: the matrix serves block products of that many columns, so the stream blocks'
// PER is timed with those and the single-vector layouts (packed, sliced ELL) are not built.
// stream_only: only the stream layout's PER is tuned (composite operators, whose products
// read two operand vectors, run in the stream layout).
int tune_layout(DevCsr &M, const int32_t *hrp, const int32_t *hci, const double *hv, hipStream_t s,
                const int *zero, bool tune_per = true, const DevBuf *flush = nullptr, int tune_k = 1,
                bool stream_only = false);

uint64_t next_uid();

}  // namespace chrrt

using chrrt::DevBuf;
using chrrt::DevCsr;

// ---------------------------------------------------------------------------------------
// Opaque objects
// ---------------------------------------------------------------------------------------
struct chr_mat {
  // Unique per object (never reused): keys the PCG graphs captured with this operator, so a
  // new matrix that happens to get a freed one's address never reuses its graphs.
  uint64_t uid = chrrt::next_uid();
  DevCsr A;
  DevBuf zero;  // an int 0: the stop flag of products outside a PCG solve
};

// The device buffers of one PCG_CJ solve, as plain pointers (the solver's kernels take these,
// whoever owns the memory).
struct PcgPtrs {
  int32_t n = 0;
  unsigned nblk = 0;  // blocks of the vector kernels (partials in part: 2 per block)
  double *x = nullptr, *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *ap = nullptr;
  double *part = nullptr, *part_p = nullptr, *resvec = nullptr;
  float *r32 = nullptr;  // FP32 copy of r (FP32 cycle only)
  chr::PcgState *st = nullptr;
};

struct PcgWork {
  int32_t n = 0, itmax_cap = 0;
  DevBuf x, r, z, p0, p1, ap, part, part_p, state, resvec;
  DevBuf r32;  // FP32 copy of r (FP32 cycle): the V-cycle's first smoother operand
  unsigned nblk = 0;
  chr::PcgState *pinned = nullptr;      // two slots of host-mapped pinned memory
  chr::PcgState *pinned_dev = nullptr;  // their device address (written by the finalize kernel)
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  uint64_t graph_uid = 0;  // uid of the outer operator the cached graphs were built for
  size_t part_p_cap = 0;   // doubles in part_p
  PcgPtrs ptrs(bool fp32) const {
    PcgPtrs q;
    q.n = n;
    q.nblk = nblk;
    q.x = x.as<double>();
    q.r = r.as<double>();
    q.z = z.as<double>();
    q.p0 = p0.as<double>();
    q.p1 = p1.as<double>();
    q.ap = ap.as<double>();
    q.part = part.as<double>();
    q.part_p = part_p.as<double>();
    q.resvec = resvec.as<double>();
    q.r32 = fp32 ? r32.as<float>() : nullptr;
    q.st = state.as<chr::PcgState>();
    return q;
  }
  void drop_graphs() {     // the graphs hold buffer addresses: drop them before any realloc
    for (auto &g : exec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    graph_uid = 0;
  }
  ~PcgWork() {
    for (auto &g : exec)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (pinned) (void)hipHostFree(pinned);
  }
};

struct chr_hier {
  // Level l holds n rows.  Rows [0, nd) go through smoothing and the transfers (A, G, G^T
  // have nd rows, P is nd x n_{l+1}); rows [nd, n) are the coarsest levels of the systems
  // that end here and are solved directly with dense FP64 inverses.  A single system has
  // nd = n on every level but the last, where nd = 0.  A batch of systems merged into one
  // block-diagonal hierarchy has systems ending at different levels.
  struct Level {
    int32_t n = 0, nd = 0;
    DevCsr A, G, GT, P, PT;
    double omega = 1.0;
    DevBuf omega_rows;  // per-row damping (batched hierarchies with differing omegas)
    int32_t nupre = 1, nupost = 1;
    DevBuf b, x, r, y;  // cycle vectors, n x kcap (b unused on level 0)
    DevBuf cmap;        // int4 per direct row: {block first row, m, inverse offset, 0}
    int32_t ncoarse_blocks = 0;
    int64_t coarse_m2 = 0;  // sum of m^2 over the level's direct blocks
    // Composite operators of a fused level (levels >= 1 of a single-system hierarchy with
    // one pre- and one post-sweep), formed on the host at upload.  [M1 | M2] applied to
    // [u; v] is one operator over two vectors, its columns interleaved (M1's column j is 2j,
    // M2's is 2j + 1; chr::XTwo):
    //   R = [-P^T A | P^T]          b_c = R [x; b]  (residual + restriction)
    //   U = [A P | A]               r = b - U [x_c; x]  (prolongation + residual; x itself is
    //                               updated by W's epilogue)
    //   W = [P | omega G^T]         out = x + W [x_c; G r]
    //    or [P | S]                 with S below:  out = x + W [x_c; r]
    //   S = omega G^T G             x = S b  (pre-smoothing in one product)
    // They are the same linear maps as the unfused sequence; only the rounding differs.
    DevCsr R, U, W, S;
    bool fuse_r = false, fuse_up = false, fuse_s = false;
  };
  int32_t precision = CHR_FP64;
  int device = 0;  // HIP device the hierarchy was created on
  std::vector<Level> lev;
  int32_t nc = 0;  // rows solved directly, over all levels
  DevBuf Ainv;     // packed FP64 dense inverses of the direct blocks (row-major each)
  int32_t kcap = 1;   // columns the cycle vectors are allocated for
  // Incremented whenever the cycle vectors are reallocated: graphs captured by other objects
  // (the batch's) embed their addresses and must be rebuilt when it changes.
  uint64_t buf_gen = 0;
  DevBuf io_b, io_x;  // FP64 staging for host-pointer and block calls (n_0 x kcap)
  chr_mat outer;     // level-0 operator in FP64 (default outer operator of PCG)
  DevBuf zero;       // an int 0: the stop flag of V-cycles outside a PCG solve
  hipStream_t stream = nullptr;
  std::unique_ptr<PcgWork> pcg;
  chr_hier_stats stats{};
  ~chr_hier() {
    pcg.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }
  bool fp32() const { return precision == CHR_FP32_CYCLE; }
};

namespace chrrt {

// ---------------------------------------------------------------------------------------
// V-cycle enqueue.  T is the cycle type (float or double); level 0 reads b0 and writes
// out0 in FP64.
// ---------------------------------------------------------------------------------------
// The cycle runs on row-major n_l x k blocks (k = 1: single vectors).
template <typename T>
struct Cycle {
  chr_hier &h;
  const int *stop;
  hipStream_t s;
  int k = 1;
  // Optional FP32 copy of the level-0 input, read by the first smoother product instead of
  // the FP64 input (the product rounds its operand to FP32 either way: same bits, half the
  // gathered bytes).
  const float *b0f = nullptr;

  template <typename XP, typename Epi>
  hipError_t op(const DevCsr &M, XP x, Epi epi) {
    return launch_op(M, x, k, epi, stop, s);
  }

  // Direct solves of rows [nd, n) of level L.
  template <typename TB, typename TO>
  hipError_t coarse(const chr_hier::Level &L, const TB *b, TO *x) {
    const int nr = L.n - L.nd;
    if (nr <= 0) return hipSuccess;
    static_assert(std::is_same<TB, TO>::value, "direct solve reads and writes one type");
    const dim3 grid((unsigned)(((int64_t)nr * chr::kWave + chr::kBlock - 1) / chr::kBlock), (unsigned)k);
    hipLaunchKernelGGL((chr::coarse_gemv_kernel<TO>), grid, dim3(chr::kBlock), 0, s, nr, L.nd,
                       L.cmap.as<int4>(), h.Ainv.as<double>(), b, x, stop, k);
    return hipGetLastError();
  }

  // x <- x + omega G^T (G (b - A x)) sweep, with the output written to out (may be x).
  template <typename TB, typename TO>
  hipError_t post_sweep(chr_hier::Level &L, const TB *b, TO *out) {
    hipError_t e;
    T *x = L.x.as<T>(), *r = L.r.as<T>(), *y = L.y.as<T>();
    if ((e = op(L.A, x, chr::EpiResidual<TB, T>{b, r}))) return e;
    if ((e = op(L.G, r, chr::EpiStore<T>{y}))) return e;
    return op(L.GT, y, chr::EpiAxpy<T, TO>{x, out, L.omega, L.omega_rows.as<double>(), k});
  }

  hipError_t fused_level(int l, const T *b, T *out);

  template <typename TB, typename TO>
  hipError_t level(int l, const TB *b, TO *out) {
    chr_hier::Level &L = h.lev[l];
    hipError_t e;
    // Systems ending on this level: direct solves of rows [nd, n).
    if (L.nd < L.n) {
      if constexpr (std::is_same<TB, TO>::value) {
        if ((e = coarse(L, b, out))) return e;
      } else {
        // Level 0 with TB != TO never happens: level-0 input and output are both FP64.
        return hipErrorInvalidValue;
      }
    }
    if (L.nd == 0) return hipSuccess;
    T *x = L.x.as<T>(), *r = L.r.as<T>(), *y = L.y.as<T>();
    if constexpr (std::is_same<TB, T>::value && std::is_same<TO, T>::value) {
      if (L.fuse_r) return fused_level(l, b, out);
    }
    // Pre-smoothing from a zero initial guess, then nupre-1 further sweeps.
    if (l == 0 && b0f && k == 1 && std::is_same<T, float>::value)
      e = op(L.G, b0f, chr::EpiStore<T>{y});
    else
      e = op(L.G, b, chr::EpiStore<T>{y});
    if (e) return e;
    if ((e = op(L.GT, y, chr::EpiScale<T>{x, L.omega, L.omega_rows.as<double>(), k}))) return e;
    for (int q = 2; q <= L.nupre; ++q)
      if ((e = post_sweep(L, b, x))) return e;
    // Residual and restriction.
    if ((e = op(L.A, x, chr::EpiResidual<TB, T>{b, r}))) return e;
    chr_hier::Level &C = h.lev[l + 1];
    T *bc = C.b.as<T>(), *xc = C.x.as<T>();
    if ((e = op(L.PT, r, chr::EpiStore<T>{bc}))) return e;
    // Coarse correction.
    if ((e = level<T, T>(l + 1, bc, xc))) return e;
    if ((e = op(L.P, xc, chr::EpiAxpy<T, T>{x, x, 1.0}))) return e;
    // Post-smoothing; the last sweep writes the level's output.
    if (L.nupost == 0) {
      const unsigned grid = (unsigned)(((int64_t)L.nd * k + chr::kBlock - 1) / chr::kBlock);
      hipLaunchKernelGGL((chr::convert_kernel<T, TO>), dim3(grid), dim3(chr::kBlock), 0, s, L.nd * k,
                         (const T *)x, out);
      return hipGetLastError();
    }
    for (int q = 1; q < L.nupost; ++q)
      if ((e = post_sweep(L, b, x))) return e;
    return post_sweep(L, b, out);
  }
};

// The fused sequence of a level with composite operators (see chr_hier::Level), one pre- and
// one post-sweep, zero initial guess: 4 products with S, 5 without, instead of 8.
template <typename T>
hipError_t Cycle<T>::fused_level(int l, const T *b, T *out) {
  chr_hier::Level &L = h.lev[l];
  chr_hier::Level &C = h.lev[l + 1];
  T *x = L.x.as<T>(), *r = L.r.as<T>(), *y = L.y.as<T>();
  T *bc = C.b.as<T>(), *xc = C.x.as<T>();
  hipError_t e;
  if (L.fuse_s) {
    e = op(L.S, b, chr::EpiStore<T>{x});
  } else {
    if ((e = op(L.G, b, chr::EpiStore<T>{y}))) return e;
    e = op(L.GT, y, chr::EpiScale<T>{x, L.omega, nullptr, k});
  }
  if (e) return e;
  if ((e = op(L.R, chr::XTwo<T>{x, b}, chr::EpiStore<T>{bc}))) return e;
  if ((e = level<T, T>(l + 1, bc, xc))) return e;
  if (!L.fuse_up) {
    if ((e = op(L.P, xc, chr::EpiAxpy<T, T>{x, x, 1.0}))) return e;
    return post_sweep(L, b, out);
  }
  if ((e = op(L.U, chr::XTwo<T>{xc, x}, chr::EpiResidual<T, T>{b, r}))) return e;
  if (L.fuse_s) return op(L.W, chr::XTwo<T>{xc, r}, chr::EpiAxpy<T, T>{x, out, 1.0});
  if ((e = op(L.G, r, chr::EpiStore<T>{y}))) return e;
  return op(L.W, chr::XTwo<T>{xc, y}, chr::EpiAxpy<T, T>{x, out, 1.0});
}

// Host description of one level for build_hier (views into caller-owned CSR arrays).
struct LevelIn {
  int32_t n = 0, nd = 0;
  chr_csr A{}, G{}, P{};            // nd-row matrices (used when nd > 0)
  double omega = 1.0;
  std::vector<double> omega_rows;   // per-row damping of rows [0, nd), or empty
  int32_t nupre = 1, nupost = 1;
  struct Direct {
    int32_t start;  // first row of the block on this level
    chr_csr A;      // the block's operator (square, dense-invertible)
  };
  std::vector<Direct> direct;       // rows [nd, n), in order
};
// tune_k: block width the level matrices are tuned for (1 for single systems; the batch
// tunes for its block products).
int build_hier(std::vector<LevelIn> &lv, int32_t precision, const chr_csr &outer, chr_hier **out,
               int tune_k = 1);

// One V-cycle on FP64 row-major n_0 x k blocks (k <= kcap).
// b0f (FP32 cycle, k = 1): an FP32 copy of b for the first smoother product.
hipError_t enqueue_vcycle(chr_hier &h, const double *b, double *out, const int *stop, int k = 1,
                          const float *b0f = nullptr);
// Grows the cycle vectors to k columns (synchronises the handle's stream).
int ensure_block_capacity(chr_hier &h, int k);

}  // namespace chrrt

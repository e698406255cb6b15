// Host replays of the GPU library's sliced-ELL product (spmv_sell_kernel), packed-code
// stream product (spmv_pack_kernel) and symmetric diagonal product (spmv_dia_kernel): each
// builds the layout with the same builder (sell_layout.hpp, pack_layout.hpp,
// dia_layout.hpp) and walks it with the kernel's own index arithmetic,
// workgroup by workgroup and lane by lane, checking every load against the buffers the
// kernel would read.  It lets the layout and the kernel's addressing be
// verified without a GPU; the product it returns has the kernel's summation order and
// rounding (FP32 or FP64 products and sums).
#include <stdint.h>

#include <cmath>
#include <cstdlib>
#include <vector>

#include "chronos_setup.h"
#include "dense.hpp"
#include "dia_layout.hpp"
#include "pack_layout.hpp"
#include "sell_layout.hpp"

namespace {

constexpr int kBlock = 256;

template <typename TA>
int replay(const chrsell::Layout &L, const double *x, double *y) {
  const int32_t n = L.nrows;
  const int64_t nslots = (int64_t)L.nslices * chrsell::kSlice;
  const int nblocks = (L.nslices + kBlock / chrsell::kSlice - 1) / (kBlock / chrsell::kSlice);
  const int64_t ndict = (int64_t)L.dict.size();
  std::vector<int> written((size_t)n, 0);
  for (int b = 0; b < nblocks; ++b) {
    const int64_t wi = ((int64_t)b * kBlock) >> L.wshift;  // the block's dictionary window
    if (L.cb == 8 && (wi + 1) * 256 > ndict) return -10;
    for (int t = 0; t < kBlock; ++t) {
      const int64_t s = (int64_t)b * (kBlock / chrsell::kSlice) + t / chrsell::kSlice;
      const int lane = t % chrsell::kSlice;
      if (s >= L.nslices) continue;
      const int64_t slot = s * chrsell::kSlice + lane;
      int64_t row = slot;
      bool valid = slot < n;
      if (L.sorted) {
        if (slot >= (int64_t)L.perm.size()) return -11;
        const int q = L.perm[(size_t)slot];
        valid = q != 0xFFFF;
        row = ((slot >> L.sshift) << L.sshift) + q;
      }
      if (valid && (row < 0 || row >= n)) return -12;
      const int64_t rc = valid ? row : 0;
      const int64_t d0 = L.sp[(size_t)s], d1 = L.sp[(size_t)s + 1];
      if (L.ulen > 0 && (d0 != s * L.ulen * chrsell::kSlice || d1 - d0 != L.ulen * chrsell::kSlice)) return -17;
      if ((d1 - d0) % chrsell::kSlice) return -13;
      const int64_t len = (d1 - d0) / chrsell::kSlice;
      TA acc = 0;
      const int kUnroll = L.unroll;
      const int64_t nch = (len + kUnroll - 1) / kUnroll;
      for (int64_t ch = 0; ch < nch; ++ch)
        for (int u = 0; u < kUnroll; ++u) {
          const int64_t k = std::min<int64_t>(ch * kUnroll + u, len - 1);
          // The kernel loads each group of four with one vector load: the group of chunk
          // element u, clamped to the lane's last group, starts at d0 + 64 k0 + 4 lane.
          if (len % 4) return -18;
          const int64_t k0 = std::min<int64_t>(ch * kUnroll + (u & ~3), len - 4);
          const int64_t g = len > 0 ? d0 + chrsell::kSlice * k0 + 4 * lane : 0;
          if (g % 4 != 0 || g + 3 >= L.elems) return -19;
          const int64_t e = len > 0 ? g + (u & 3) : 0;
          if (len > 0 && e != d0 + chrsell::elem_at((int32_t)k, lane) && ch * kUnroll + u < len) return -20;
          if (e < 0 || e >= L.elems) return -14;
          bool pad;
          int64_t col;
          if (L.cb == 8) {
            const uint8_t c = L.c8[(size_t)e];
            pad = c == 0xFF;
            col = rc - L.dict[(size_t)(wi * 256 + c)];
          } else if (L.cb == 16) {
            const uint16_t c = L.c16[(size_t)e];
            pad = c == 0xFFFF;
            col = rc - (pad ? 0 : (int64_t)c + L.dmin);
          } else {
            const int32_t c = L.c32[(size_t)e];
            pad = c < 0;
            col = pad ? rc : c;
          }
          if (col < 0 || col >= n) return -15;
          const TA p = (TA)L.val[(size_t)e] * (TA)x[col];
          if (ch * kUnroll + u < len && !pad) acc = acc + p;
        }
      if (valid) {
        y[row] = (double)acc;
        written[(size_t)row]++;
      }
    }
  }
  for (int32_t r = 0; r < n; ++r)
    if (written[(size_t)r] != 1) return -16;
  (void)nslots;
  return 0;
}

}  // namespace

extern "C" int chs_sell_apply(int32_t n, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                              const double *val, int32_t fp32, const double *x, double *y,
                              int32_t *info, int64_t *elems) {
  chrsell::Layout L;
  // CHRONOS_SELL_SORT fixes the sort window exactly as it does for the GPU library.
  const char *env_sort = std::getenv("CHRONOS_SELL_SORT");
  if (!chrsell::build(n, ncols, rowptr, col, val, fp32 ? 4 : 8, L, env_sort ? std::atoi(env_sort) : -1))
    return 1;
  if (info) {
    info[0] = L.cb;
    info[1] = L.sorted ? 1 : 0;
    info[2] = L.wshift;
    info[3] = L.nslices;
    info[4] = L.sshift;
    info[5] = L.ulen;
  }
  if (elems) *elems = L.elems;
  if (fp32) {
    // The kernel reads FP32 values and an FP32 vector.
    chrsell::Layout &M = L;
    for (auto &v : M.val) v = (double)(float)v;
    std::vector<double> xf((size_t)n);
    for (int32_t i = 0; i < n; ++i) xf[(size_t)i] = (double)(float)x[i];
    return replay<float>(M, xf.data(), y);
  }
  return replay<double>(L, x, y);
}

namespace {

template <typename TA>
int replay_pack(const chrpack::Layout &L, const int32_t *rp, const double *val, int64_t nnz, const double *x,
                double *y) {
  const int32_t n = L.nrows;
  const int cap = L.per * kBlock;
  const int64_t ndict = (int64_t)L.dict.size();
  std::vector<int> written((size_t)n, 0);
  std::vector<TA> prod((size_t)cap);
  auto col_of = [&](int64_t e, int32_t r0, int64_t w, int64_t &col) -> int {
    if (e < 0 || e >= nnz) return -21;
    const uint32_t c = L.code[(size_t)e];
    const int64_t di = w * 256 + (c & 255u);
    if (di < 0 || di >= ndict) return -22;
    col = (int64_t)r0 + (c >> 8) - L.dict[(size_t)di];
    if (col < 0 || col >= n) return -23;
    return 0;
  };
  for (int32_t b = 0; b < L.nblk; ++b) {
    const int32_t r0 = L.blk[4 * b], e0 = L.blk[4 * b + 1], w = L.blk[4 * b + 2];
    const int32_t r1 = L.blk[4 * (b + 1)], e1 = L.blk[4 * (b + 1) + 1];
    if (r1 - r0 > kBlock || r1 <= r0) return -24;
    if (e1 - e0 > cap) {  // long row: strided partial sums, then the block reduction
      if (r1 - r0 != 1) return -25;
      TA part[kBlock];
      for (int t = 0; t < kBlock; ++t) {
        TA s = 0;
        for (int64_t e = e0 + t; e < e1; e += kBlock) {
          int64_t col;
          if (int rc = col_of(e, r0, w, col)) return rc;
          s += (TA)val[e] * (TA)x[col];
        }
        part[t] = s;
      }
      // block_sum_t: xor butterfly inside each wave of 64, then waves 0..3 in order
      TA tot = 0;
      for (int wv = 0; wv < kBlock / 64; ++wv) {
        TA v[64];
        for (int l = 0; l < 64; ++l) v[l] = part[wv * 64 + l];
        for (int off = 32; off > 0; off >>= 1) {
          TA nv[64];
          for (int l = 0; l < 64; ++l) nv[l] = v[l] + v[l ^ off];
          for (int l = 0; l < 64; ++l) v[l] = nv[l];
        }
        tot += v[0];
      }
      y[r0] = (double)tot;
      written[(size_t)r0]++;
      continue;
    }
    for (int64_t o = 0; o < e1 - e0; ++o) {
      int64_t col;
      if (int rc = col_of(e0 + o, r0, w, col)) return rc;
      prod[(size_t)o] = (TA)val[e0 + o] * (TA)x[col];
    }
    for (int32_t r = r0; r < r1; ++r) {
      const int64_t ra = rp[r] - e0, rz = rp[r + 1] - e0;
      if (ra < 0 || rz > e1 - e0) return -26;
      TA s = 0;
      for (int64_t q = ra; q < rz; ++q) s += prod[(size_t)q];
      y[r] = (double)s;
      written[(size_t)r]++;
    }
  }
  for (int32_t r = 0; r < n; ++r)
    if (written[(size_t)r] != 1) return -27;
  return 0;
}

}  // namespace

extern "C" int chs_pack_apply(int32_t n, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                              const double *val, int32_t per, int32_t fp32, const double *x, double *y,
                              int32_t *nblk) {
  chrpack::Layout L;
  if (!chrpack::build(n, ncols, rowptr, col, per, L)) return 1;
  if (nblk) *nblk = L.nblk;
  const int64_t nnz = rowptr[n];
  if (fp32) {
    std::vector<double> vf((size_t)nnz), xf((size_t)n);
    for (int64_t e = 0; e < nnz; ++e) vf[(size_t)e] = (double)(float)val[e];
    for (int32_t i = 0; i < n; ++i) xf[(size_t)i] = (double)(float)x[i];
    return replay_pack<float>(L, rowptr, vf.data(), nnz, xf.data(), y);
  }
  return replay_pack<double>(L, rowptr, val, nnz, x, y);
}

namespace {

// spmv_dia_kernel row by row: the loads of DiaRow::load (indices clamped into [0, n), a
// lower entry outside the matrix read as zero), then DiaRow::sum's order and rounding.
template <typename TA>
int replay_dia(const chrdia::Layout &L, const std::vector<double> &val, const double *x, double *y) {
  const int32_t n = L.n;
  const int64_t nv = (int64_t)val.size();
  if (L.m < 1 || L.m > chrdia::kMaxDiag || L.off[0] != 0) return -31;
  for (int32_t i = 0; i < n; ++i) {
    TA lo[chrdia::kMaxDiag], up[chrdia::kMaxDiag], xl[chrdia::kMaxDiag], xu[chrdia::kMaxDiag];
    for (int k = 1; k < L.m; ++k) {
      if (L.off[k] <= L.off[k - 1]) return -32;
      const int64_t jl = (int64_t)i - L.off[k], ju = (int64_t)i + L.off[k];
      const int64_t jlc = std::max<int64_t>(jl, 0), juc = std::min<int64_t>(ju, n - 1);
      const int64_t el = (int64_t)k * n + jlc, eu = (int64_t)k * n + i;
      if (el < 0 || el >= nv || eu >= nv) return -33;
      lo[k] = jl >= 0 ? (TA)val[(size_t)el] : TA(0);
      up[k] = (TA)val[(size_t)eu];
      xl[k] = (TA)x[jlc];
      xu[k] = (TA)x[juc];
      if (ju >= n && up[k] != TA(0)) return -34;  // an upper entry past the last column
    }
    TA s = 0;
    for (int k = L.m - 1; k >= 1; --k)
      if (lo[k] != TA(0)) s = s + lo[k] * xl[k];
    if ((TA)val[(size_t)i] != TA(0)) s = s + (TA)val[(size_t)i] * (TA)x[i];
    for (int k = 1; k < L.m; ++k)
      if (up[k] != TA(0)) s = s + up[k] * xu[k];
    y[i] = (double)s;
  }
  return 0;
}

}  // namespace

extern "C" int chs_dia_apply(int32_t n, int32_t ncols, const int32_t *rowptr, const int32_t *col,
                             const double *val, int32_t fp32, const double *x, double *y, int32_t *ndiag,
                             int32_t *offsets) {
  chrdia::Layout L;
  if (!chrdia::build(n, ncols, rowptr, col, val, L)) return 1;
  if (ndiag) *ndiag = L.m;
  if (offsets)
    for (int k = 0; k < L.m; ++k) offsets[k] = L.off[k];
  if (fp32) {
    // The kernel reads FP32 values and an FP32 vector.
    std::vector<double> vf(L.val.size()), xf((size_t)n);
    for (size_t e = 0; e < vf.size(); ++e) vf[e] = (double)(float)L.val[e];
    for (int32_t i = 0; i < n; ++i) xf[(size_t)i] = (double)(float)x[i];
    return replay_dia<float>(L, vf, xf.data(), y);
  }
  return replay_dia<double>(L, L.val, x, y);
}

// The host dense routines of the GPU library's LOBPCG (dense.hpp), exported for testing.
extern "C" int chs_sygv(int32_t m, const double *A, const double *B, double *w, double *V) {
  if (m < 1 || !A || !B || !w || !V) return -1;
  return chrdense::sygv(m, A, B, w, V) ? 0 : 1;
}

extern "C" double chs_cond_sym(int32_t m, const double *M) {
  if (m < 1 || !M) return -1.0;
  return chrdense::cond_sym(m, M);
}

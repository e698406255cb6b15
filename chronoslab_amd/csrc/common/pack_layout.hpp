// Packed-code CSR-stream layout (host side, no HIP): built here, uploaded by the GPU library
// for its single-vector products (spmv_pack_kernel), and replayed on the CPU by the host
// library (chs_pack_apply) with the kernel's own index arithmetic.
//
// The rows are cut into blocks of whole consecutive rows as for the CSR-stream kernel (at
// most per * 256 nonzeros and 256 rows; a longer row is a block of its own), and no block
// straddles a dictionary window of 2^wshift rows.  Each nonzero stores one 16-bit code:
// the high byte is its row's offset from the block's first row, the low byte indexes the
// window's dictionary of at most 256 row-relative offsets d = row - col.  A nonzero thus
// costs 2 bytes of index where plain columns cost 4, also when the block's columns span
// more than 65536 (the FSAI factor transposes, whose rows reach several grid planes away).
// Block table entries: int4 {first row, first nonzero, window, 0}, nblk + 1 of them.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <climits>
#include <vector>

namespace chrpack {

constexpr int kRows = 256;  // rows per block at most (one workgroup)

struct Layout {
  bool ok = false;
  int per = 8;
  int wshift = 14;
  int32_t nrows = 0, nblk = 0;
  std::vector<int32_t> blk;  // 4 * (nblk + 1)
  std::vector<uint16_t> code;
  std::vector<int32_t> dict;  // 256 per window
};

// Returns false (L.ok = false) when the matrix is not square or some window holds more
// than 256 distinct row-relative offsets.
inline bool build(int32_t n, int32_t ncols, const int32_t *rp, const int32_t *ci, int per, Layout &L) {
  L = Layout();
  L.nrows = n;
  L.per = per;
  if (n <= 0 || ncols != n || rp[n] == 0) return false;
  int ws = 8;
  while (ws < 14 && ((int64_t)1 << ws) < n) ++ws;
  L.wshift = ws;
  const int64_t nwin = ((int64_t)n + ((int64_t)1 << ws) - 1) >> ws;
  // Dictionaries.
  int64_t dlo = INT64_MAX, dhi = INT64_MIN;
  for (int32_t r = 0; r < n; ++r)
    for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
      dlo = std::min<int64_t>(dlo, (int64_t)r - ci[e]);
      dhi = std::max<int64_t>(dhi, (int64_t)r - ci[e]);
    }
  if (dhi - dlo >= ((int64_t)1 << 26)) return false;
  std::vector<int32_t> lut((size_t)(dhi - dlo + 1), -1), stamp((size_t)(dhi - dlo + 1), -1);
  L.dict.assign((size_t)nwin * 256, 0);
  std::vector<int32_t> wcount((size_t)nwin, 0);
  for (int64_t w = 0; w < nwin; ++w) {
    const int64_t r1 = std::min<int64_t>((w + 1) << ws, n);
    int32_t cnt = 0;
    for (int64_t r = w << ws; r < r1; ++r)
      for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
        const size_t k = (size_t)(r - ci[e] - dlo);
        if (stamp[k] == (int32_t)w) continue;
        if (cnt == 256) return false;
        stamp[k] = (int32_t)w;
        L.dict[(size_t)w * 256 + cnt++] = (int32_t)(r - ci[e]);
      }
    wcount[(size_t)w] = cnt;
  }
  // Blocks (the CSR-stream cutting rule, plus window boundaries).
  const int64_t cap = (int64_t)per * 256;
  L.blk = {0, 0, 0, 0};
  int32_t rows = 0;
  int64_t cnt = 0;
  for (int32_t r = 0; r < n; ++r) {
    const int64_t len = rp[r + 1] - rp[r];
    const bool new_window = r > 0 && (r >> ws) != ((r - 1) >> ws);
    if (rows > 0 && (rows == kRows || cnt + len > cap || new_window)) {
      L.blk.insert(L.blk.end(), {r, rp[r], r >> ws, 0});
      rows = 0;
      cnt = 0;
    }
    ++rows;
    cnt += len;
  }
  L.blk.insert(L.blk.end(), {n, rp[n], 0, 0});
  L.nblk = (int32_t)(L.blk.size() / 4) - 1;
  // Codes.
  L.code.assign((size_t)rp[n], 0);
  int64_t lut_w = -1;
  for (int32_t b = 0; b < L.nblk; ++b) {
    const int32_t r0 = L.blk[4 * b], r1 = L.blk[4 * (b + 1)], w = L.blk[4 * b + 2];
    if (w != lut_w) {
      for (int c = 0; c < wcount[(size_t)w]; ++c) lut[(size_t)(L.dict[(size_t)w * 256 + c] - dlo)] = c;
      lut_w = w;
    }
    for (int32_t r = r0; r < r1; ++r)
      for (int32_t e = rp[r]; e < rp[r + 1]; ++e)
        L.code[(size_t)e] = (uint16_t)(((r - r0) << 8) | lut[(size_t)((int64_t)r - ci[e] - dlo)]);
  }
  L.ok = true;
  return true;
}

}  // namespace chrpack

// Sliced-ELL layout of a square sparse matrix (host side, no HIP): built here, uploaded by
// the GPU library for its single-vector products (spmv_sell_kernel), and replayed on the
// CPU by the host library (chs_sell_apply) with the kernel's own index arithmetic.
//
// Layout:
//   * slots are taken in slices of 64 (one wave, one row per lane) of len_s elements per
//     lane, len_s = the longest row of the slice (of the whole matrix when `ulen` > 0:
//     uniform ELL, sp[s] = 64 ulen s, unsorted) rounded up to a multiple of 4.  A lane's
//     elements come in groups of four consecutive ones, so that the kernel loads a group
//     with one vector load: element (k, lane) sits at sp[s] + 64 (k & ~3) + 4 lane + (k & 3)
//     (elem_at below);
//   * dictionary windows of 2^wshift slots (256 <= 2^wshift <= 16384); when `sorted`, the
//     rows of each sort window of 2^sshift slots (256 <= 2^sshift <= 2^wshift) sit in its
//     slots by decreasing length (stable) and perm[slot] is the row's offset from the sort
//     window's first row (0xFFFF: empty slot); otherwise slot = row;
//   * column codes are row-relative offsets d = row - col: cb = 8, an index into the
//     window's dictionary dict[w*256 + code] of at most 255 offsets (code 255 = padding,
//     dictionary entry 0); cb = 16, d - dmin (0xFFFF = padding); cb = 32, the column
//     itself (-1 = padding).  Padding elements have value 0.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <climits>
#include <vector>

namespace chrsell {

constexpr int kSlice = 64;
constexpr int kMaxRow = 256;

// Offset of element k of a lane's row from the slice start.
inline int64_t elem_at(int32_t k, int lane) { return (int64_t)kSlice * (k & ~3) + 4 * lane + (k & 3); }

// Elements per lane of a slice whose longest row has len elements.
inline int32_t slice_len(int32_t len) { return (len + 3) & ~3; }

struct Layout {
  int cb = 0;  // 0: the matrix is not a candidate
  int32_t nrows = 0, nslices = 0, wshift = 8, sshift = 8, dmin = 0;
  int unroll = 8;  // kernel chunk: 16 when some slice is longer than 8 elements
  int ulen = 0;    // > 0: uniform ELL (unsorted, every slice ulen elements long)
  int64_t elems = 0;
  bool sorted = false;
  std::vector<int32_t> sp, dict;
  std::vector<uint16_t> perm;
  std::vector<double> val;
  std::vector<uint8_t> c8;
  std::vector<uint16_t> c16;
  std::vector<int32_t> c32;
};

// value_bytes: 4 or 8, the stored value size (used to weigh sorting against its perm).
// Returns false (L.cb = 0) when the matrix is not square, empty, has a row longer than
// kMaxRow, or would be padded by more than 50%.
inline bool build(int32_t n, int32_t ncols, const int32_t *rp, const int32_t *ci, const double *v,
                  int value_bytes, Layout &L, int sort_shift = -1) {
  L = Layout();
  L.nrows = n;
  if (n <= 0 || ncols != n || rp[n] == 0) return false;
  int32_t maxlen = 0;
  for (int32_t r = 0; r < n; ++r) maxlen = std::max(maxlen, rp[r + 1] - rp[r]);
  if (maxlen > kMaxRow) return false;
  const int64_t nnz = rp[n];
  int ws = 8;
  while (ws < 14 && ((int64_t)1 << ws) < n) ++ws;
  const int64_t W = (int64_t)1 << ws;
  const int32_t nslices = (int32_t)((n + kSlice - 1) / kSlice);
  const int64_t nslots = (int64_t)nslices * kSlice;
  auto rlen = [&](int32_t r) { return r < 0 ? 0 : rp[r + 1] - rp[r]; };

  // Candidate slot -> row maps: identity, and rows by decreasing length inside sort windows
  // of 2^ss slots (ss = 8, 10, 12, ..., ws).  Small windows keep a wave's rows (and so its
  // gathers and stores) close together; large ones cut the padding.
  std::vector<int32_t> ident((size_t)nslots, -1);
  for (int32_t r = 0; r < n; ++r) ident[(size_t)r] = r;
  // Rows of each sort window ordered by decreasing length, ties in row order (a stable
  // counting sort: lengths are at most kMaxRow).
  auto sorted_map = [&](int ss) {
    std::vector<int32_t> m((size_t)nslots, -1);
    const int64_t SW = (int64_t)1 << ss;
    std::vector<int32_t> start(kMaxRow + 2);
    for (int64_t w0 = 0; w0 < n; w0 += SW) {
      const int64_t w1 = std::min<int64_t>(w0 + SW, n);
      std::fill(start.begin(), start.end(), 0);
      for (int64_t r = w0; r < w1; ++r) ++start[(size_t)(kMaxRow - rlen((int32_t)r)) + 1];
      for (int k = 0; k <= kMaxRow; ++k) start[(size_t)k + 1] += start[(size_t)k];
      for (int64_t r = w0; r < w1; ++r) m[(size_t)(w0 + start[(size_t)(kMaxRow - rlen((int32_t)r))]++)] = (int32_t)r;
    }
    return m;
  };
  auto padded = [&](const std::vector<int32_t> &m) {
    int64_t tot = 0;
    for (int64_t s = 0; s < nslices; ++s) {
      int32_t mx = 0;
      for (int l = 0; l < kSlice; ++l) mx = std::max(mx, rlen(m[(size_t)(s * kSlice + l)]));
      tot += (int64_t)slice_len(mx) * kSlice;
    }
    return tot;
  };
  // Sort windows considered: all of them for the builder's own choice (sort_shift < 0),
  // else only the one asked for (0: none).
  std::vector<int> sshifts;
  if (sort_shift < 0) {
    for (int ss = 8; ss < ws; ss += 2) sshifts.push_back(ss);
    sshifts.push_back(ws);
  } else if (sort_shift > 0) {
    sshifts.push_back(std::max(8, std::min(sort_shift, ws)));
  }
  const int64_t e_ident = padded(ident);
  std::vector<int64_t> e_sorted_at;
  for (int ss : sshifts) e_sorted_at.push_back(padded(sorted_map(ss)));

  // Code width: per-window dictionaries of <= 255 offsets, else 16-bit offsets, else columns.
  int64_t dlo = INT64_MAX, dhi = INT64_MIN;
  for (int32_t r = 0; r < n; ++r)
    for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
      dlo = std::min<int64_t>(dlo, (int64_t)r - ci[e]);
      dhi = std::max<int64_t>(dhi, (int64_t)r - ci[e]);
    }
  const int64_t nwin = (n + W - 1) / W;
  int cb = 32;
  std::vector<int32_t> dcount;
  std::vector<int32_t> lut;  // offset - dlo -> code (valid within the current window)
  if (dhi - dlo < ((int64_t)1 << 26)) {
    std::vector<int32_t> stamp((size_t)(dhi - dlo + 1), -1);
    L.dict.assign((size_t)nwin * 256, 0);
    dcount.assign((size_t)nwin, 0);
    bool ok = true;
    for (int64_t w = 0; w < nwin && ok; ++w) {
      int32_t cnt = 0;
      const int64_t r1 = std::min<int64_t>((w + 1) * W, n);
      for (int64_t r = w * W; r < r1 && ok; ++r)
        for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
          const size_t k = (size_t)(r - ci[e] - dlo);
          if (stamp[k] == (int32_t)w) continue;
          if (cnt == 255) {
            ok = false;
            break;
          }
          stamp[k] = (int32_t)w;
          L.dict[(size_t)w * 256 + cnt++] = (int32_t)(r - ci[e]);
        }
      dcount[(size_t)w] = cnt;
    }
    if (ok) {
      cb = 8;
      lut.assign((size_t)(dhi - dlo + 1), 0);
    } else if (dhi - dlo <= 65534) {
      cb = 16;
    }
  }
  if (cb != 8) L.dict.clear();
  const int cbytes = cb / 8;
  // Layout choice: the bytes of each candidate (padded elements, plus the row permutation
  // when sorted); the smallest sort window (or none) within 3% of the fewest bytes wins.
  // sort_shift >= 0 forces a window (0: unsorted), e.g. for measurements.
  std::vector<double> cost;
  cost.push_back((double)e_ident * (value_bytes + cbytes));
  for (int64_t e : e_sorted_at) cost.push_back((double)e * (value_bytes + cbytes) + 2.0 * nslots);
  const double best = *std::min_element(cost.begin(), cost.end());
  int pick = 0;
  while (cost[(size_t)pick] > 1.03 * best) ++pick;
  if (sort_shift == 0) pick = 0;
  else if (sort_shift > 0) pick = 1;
  const bool sorted = pick > 0;
  const int sshift = sorted ? sshifts[(size_t)pick - 1] : ws;
  // Unsorted with every slice padded to the longest row (uniform ELL) when that costs at
  // most 2% more elements: the kernel then computes slice offsets instead of reading them.
  const int64_t e_uniform = (int64_t)nslices * slice_len(maxlen) * kSlice;
  const bool uniform = !sorted && (double)e_uniform <= 1.02 * (double)e_ident;
  const int64_t elems = sorted ? e_sorted_at[(size_t)pick - 1] : uniform ? e_uniform : e_ident;
  if (elems > (int64_t)INT32_MAX - kSlice || elems > nnz + nnz / 2) {
    L = Layout();
    L.nrows = n;
    return false;
  }
  const std::vector<int32_t> map = sorted ? sorted_map(sshift) : ident;
  L.cb = cb;
  L.nslices = nslices;
  L.wshift = ws;
  L.sshift = sshift;
  L.ulen = uniform ? slice_len(maxlen) : 0;
  L.dmin = cb == 16 ? (int32_t)dlo : 0;
  L.elems = elems;
  L.sorted = sorted;
  if (sorted) {
    L.perm.assign((size_t)nslots, 0xFFFF);
    for (int64_t s = 0; s < nslots; ++s)
      if (map[(size_t)s] >= 0) L.perm[(size_t)s] = (uint16_t)(map[(size_t)s] - ((s >> sshift) << sshift));
  }
  L.sp.assign((size_t)nslices + 1, 0);
  L.val.assign((size_t)elems, 0.0);
  if (cb == 8) L.c8.assign((size_t)elems, 0xFF);
  else if (cb == 16) L.c16.assign((size_t)elems, 0xFFFF);
  else L.c32.assign((size_t)elems, -1);
  int64_t lut_w = -1, pos = 0;
  for (int64_t s = 0; s < nslices; ++s) {
    L.sp[(size_t)s] = (int32_t)pos;
    const int64_t w = (s * kSlice) >> ws;
    if (cb == 8 && w != lut_w) {
      for (int c = 0; c < dcount[(size_t)w]; ++c) lut[(size_t)(L.dict[(size_t)w * 256 + c] - dlo)] = c;
      lut_w = w;
    }
    int32_t mx = 0;
    for (int l = 0; l < kSlice; ++l) mx = std::max(mx, rlen(map[(size_t)(s * kSlice + l)]));
    mx = slice_len(uniform ? maxlen : mx);
    for (int l = 0; l < kSlice; ++l) {
      const int32_t r = map[(size_t)(s * kSlice + l)];
      if (r < 0) continue;
      for (int32_t k = 0; k < rlen(r); ++k) {
        const int32_t e = rp[r] + k;
        const size_t at = (size_t)(pos + elem_at(k, l));
        L.val[at] = v[e];
        const int64_t d = (int64_t)r - ci[e];
        if (cb == 8) L.c8[at] = (uint8_t)lut[(size_t)(d - dlo)];
        else if (cb == 16) L.c16[at] = (uint16_t)(d - dlo);
        else L.c32[at] = ci[e];
      }
    }
    pos += (int64_t)mx * kSlice;
    if (mx > 8) L.unroll = 16;
  }
  L.sp[(size_t)nslices] = (int32_t)pos;
  return true;
}

}  // namespace chrsell

// Host runtime and C ABI of the MI355X AMG solve phase (include/chronos_amg.h).
//
// The hierarchy is uploaded once.  One V-cycle (AMG_Vcycle.m:1-54, zero initial guess on
// every level) is the kernel sequence, per non-coarsest level l:
//   y  = G b              x = omega G^T y          r = b - A x
//   bc = P^T r            (level l+1)              x = x + P xc
//   r  = b - A x          y = G r                  x = x + omega G^T y
// and x = Ainv b on the coarsest level (FP64 dense inverse).  G^T and P^T are explicit
// transposed CSR copies made on upload, so every product is a row-parallel gather with a
// fixed summation order (no atomics).  In CHR_FP32_CYCLE mode the level matrices and the
// cycle vectors are FP32; the level-0 input is read in FP64 and the output written in FP64
// by the kernels that touch them, so no conversion pass is needed.
#include "chronos_amg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "runtime.hpp"
#include "dia_layout.hpp"
#include "pack_layout.hpp"
#include "sell_layout.hpp"

namespace chrrt {

thread_local std::string g_err;

// CHRONOS_UPLOAD_TRACE=1: per-matrix upload and autotuning times on stderr.
double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
bool upload_trace() {
  static const bool on = [] {
    const char *e = std::getenv("CHRONOS_UPLOAD_TRACE");
    return e && std::atoi(e) != 0;
  }();
  return on;
}

uint64_t next_uid() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}

int DevSell::build(int32_t n, int32_t nc, const int32_t *hrp, const int32_t *hci,
                   const double *hv, bool fp32, int sort_shift) {
  chrsell::Layout L;
  chrsell::build(n, nc, hrp, hci, hv, fp32 ? 4 : 8, L, sort_shift);
  return adopt(L, fp32);
}

int DevSell::adopt(const chrsell::Layout &L, bool fp32) {
  release();
  elems = 0;
  nrows = L.nrows;
  if (!L.cb) return CHR_OK;
  const size_t vb = fp32 ? 4 : 8, cbytes = (size_t)L.cb / 8;
  int rc;
  if ((rc = sp.alloc(sizeof(int32_t) * L.sp.size())) || (rc = val.alloc(vb * (size_t)L.elems)) ||
      (rc = code.alloc(cbytes * (size_t)L.elems)))
    return rc;
  HIPC(hipMemcpy(sp.p, L.sp.data(), sizeof(int32_t) * L.sp.size(), hipMemcpyHostToDevice));
  if (fp32) {
    std::vector<float> f(L.val.begin(), L.val.end());
    HIPC(hipMemcpy(val.p, f.data(), vb * (size_t)L.elems, hipMemcpyHostToDevice));
  } else {
    HIPC(hipMemcpy(val.p, L.val.data(), vb * (size_t)L.elems, hipMemcpyHostToDevice));
  }
  const void *cp = L.cb == 8    ? (const void *)L.c8.data()
                   : L.cb == 16 ? (const void *)L.c16.data()
                                : (const void *)L.c32.data();
  HIPC(hipMemcpy(code.p, cp, cbytes * (size_t)L.elems, hipMemcpyHostToDevice));
  if (L.sorted) {
    if ((rc = perm.alloc(sizeof(uint16_t) * L.perm.size()))) return rc;
    HIPC(hipMemcpy(perm.p, L.perm.data(), sizeof(uint16_t) * L.perm.size(), hipMemcpyHostToDevice));
  }
  if (L.cb == 8) {
    if ((rc = dict.alloc(sizeof(int32_t) * L.dict.size()))) return rc;
    HIPC(hipMemcpy(dict.p, L.dict.data(), sizeof(int32_t) * L.dict.size(), hipMemcpyHostToDevice));
  }
  cb = L.cb;
  nslices = L.nslices;
  wshift = L.wshift;
  sshift = L.sshift;
  unroll = L.unroll;
  ulen = L.ulen;
  dmin = L.dmin;
  elems = L.elems;
  sorted = L.sorted;
  return CHR_OK;
}

int DevDia::build(int32_t n_, int32_t ncols, const int32_t *hrp, const int32_t *hci, const double *hv,
                  bool fp32) {
  release();
  n = n_;
  chrdia::Layout L;
  if (!chrdia::build(n_, ncols, hrp, hci, hv, L)) return CHR_OK;
  const size_t cnt = (size_t)L.m * n_;
  int rc;
  if ((rc = val.alloc((fp32 ? 4 : 8) * cnt))) return rc;
  if (fp32) {
    std::vector<float> f(L.val.begin(), L.val.end());
    HIPC(hipMemcpy(val.p, f.data(), 4 * cnt, hipMemcpyHostToDevice));
  } else {
    HIPC(hipMemcpy(val.p, L.val.data(), 8 * cnt, hipMemcpyHostToDevice));
  }
  m = L.m;
  offs = chr::DiaOffs{};
  offs.m = L.m;
  for (int k = 0; k < L.m; ++k) offs.off[k] = L.off[k];
  return CHR_OK;
}

int DevPack::build(int32_t n, int32_t ncols, const int32_t *hrp, const int32_t *hci, int per_) {
  release();
  chrpack::Layout L;
  if (!chrpack::build(n, ncols, hrp, hci, per_, L)) return CHR_OK;
  int rc;
  if ((rc = blk.alloc(sizeof(int32_t) * L.blk.size())) ||
      (rc = code.alloc(sizeof(uint16_t) * (size_t)quad_len((int64_t)L.code.size()))) ||
      (rc = dict.alloc(sizeof(int32_t) * L.dict.size())))
    return rc;
  HIPC(hipMemcpy(blk.p, L.blk.data(), sizeof(int32_t) * L.blk.size(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(code.p, L.code.data(), sizeof(uint16_t) * L.code.size(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dict.p, L.dict.data(), sizeof(int32_t) * L.dict.size(), hipMemcpyHostToDevice));
  per = per_;
  nblk = L.nblk;
  built = true;
  return CHR_OK;
}

// Autotuner for single-vector products.  Candidates:
//  1. the CSR-stream layout with PER in {4, 8, 16} nonzeros per thread (tune_per = false
//     keeps PER 8: the outer operator of PCG, whose fused dot products are summed over the
//     stream row blocks, so that the solver's rounding never depends on a timing);
//  2. the packed-code stream layout (16-bit codes: row offset in the block and a window
//     dictionary index), with the PER of step 1, when the matrix admits it and its blocks
//     span too many columns for 16-bit offsets;
//  3. the sliced-ELL layout, when the matrix admits it: unsorted, then (unless unsorted
//     pads by at most 12%) sorted in 256-row windows and with the builder's own choice of
//     sort window; the sorted ones are skipped when the unsorted layout pads past the
//     builder's limit or when its time, scaled down by its padding, does not beat the best
//     candidate so far (CHRONOS_SELL_PRUNE=0 tries them anyway).
// Each candidate is timed with HIP events, 2 warm-up and 5 timed launches; a matrix of 32 MB
// or more is timed cache-cold (a 512 MiB device copy before every timed launch, as the
// other products of a V-cycle evict it; the median launch counts), smaller ones back to
// back (their total).  The fastest candidate
// wins, but PER 8 is kept unless another is more than 3% faster, so that near-ties do not
// flip from run to run.  Every layout sums a row in the same order, so the choice never
// changes a product.
// Environment: CHRONOS_STREAM_PER=4|8|16 fixes PER (tune_per only); CHRONOS_SELL=0
// excludes and CHRONOS_SELL=1 forces the sliced-ELL layout where the matrix admits it;
// CHRONOS_SELL_SORT=0|8|10|12|14 fixes its sort window (0: unsorted); CHRONOS_PACK=0
// excludes and CHRONOS_PACK=1 forces the packed-code layout (when sliced ELL is not forced).
static int tune_layout_timed(DevCsr &M, const int32_t *hrp, const int32_t *hci, const double *hv, hipStream_t s,
                             const int *zero, bool tune_per, const DevBuf *flush, int tune_k, bool stream_only) {
  M.use_sell = false;
  M.sell_arg = -2;
  M.sell.release();
  M.use_pack = false;
  M.pack.release();
  M.use_dia = false;
  M.dia_pcg = false;
  M.dia.release();
  M.vec4 = false;
  if (M.nnz == 0) return CHR_OK;
  const char *env_pack = std::getenv("CHRONOS_PACK");
  const int pack_mode = env_pack ? std::atoi(env_pack) : -1;  // -1 timed, 0 off, 1 forced
  const char *env_sell = std::getenv("CHRONOS_SELL");
  const int sell_mode = env_sell ? std::atoi(env_sell) : -1;  // -1 timed, 0 off, 1 forced
  const char *env_sort = std::getenv("CHRONOS_SELL_SORT");
  int fixed_per = tune_per ? 0 : 8;
  if (const char *env = std::getenv("CHRONOS_STREAM_PER"); env && tune_per)
    for (int p : kPers)
      if (p == std::atoi(env)) fixed_per = p;
  const size_t sz = M.fp32 ? 4 : 8;
  const bool cold = flush && flush->p && M.csr_bytes() >= 32e6;
  DevBuf xb, yb;
  int rc;
  const int kt = std::max(1, std::min(tune_k, kMaxBlockCols));
  if ((rc = xb.alloc(sz * kt * std::max(M.ncols, 1))) || (rc = yb.alloc(sz * kt * std::max(M.nrows, 1)))) return rc;
  HIPC(hipMemsetAsync(xb.p, 0, xb.bytes, s));
  hipEvent_t ev[12];
  for (auto &e : ev) HIPC(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t *e;
    ~EvGuard() {
      for (int i = 0; i < 12; ++i) (void)hipEventDestroy(e[i]);
    }
  } evg{ev};
  auto time_it = [&](float &ms) -> int {
    auto run = [&]() -> hipError_t {
      if (M.fp32) return launch_op(M, xb.as<float>(), kt, chr::EpiStore<float>{yb.as<float>()}, zero, s);
      return launch_op(M, xb.as<double>(), kt, chr::EpiStore<double>{yb.as<double>()}, zero, s);
    };
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && !e; ++i) e = run();
    if (cold) {
      const size_t half = flush->bytes / 2;
      for (int i = 0; i < 5 && !e; ++i) {
        e = hipMemcpyAsync(flush->as<char>() + half, flush->p, half, hipMemcpyDeviceToDevice, s);
        if (!e) e = hipEventRecord(ev[2 * i], s);
        if (!e) e = run();
        if (!e) e = hipEventRecord(ev[2 * i + 1], s);
      }
      if (!e) e = hipEventSynchronize(ev[9]);
      // The median launch, times 5 (the scale of the warm sum below): one slow launch (a
      // clock change, another process) does not decide the layout.
      float t5[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      for (int i = 0; i < 5 && !e; ++i) e = hipEventElapsedTime(&t5[i], ev[2 * i], ev[2 * i + 1]);
      std::sort(t5, t5 + 5);
      ms = 5.f * t5[2];
    } else {
      if (!e) e = hipEventRecord(ev[10], s);
      for (int i = 0; i < 5 && !e; ++i) e = run();
      if (!e) e = hipEventRecord(ev[11], s);
      if (!e) e = hipEventSynchronize(ev[11]);
      if (!e) e = hipEventElapsedTime(&ms, ev[10], ev[11]);
    }
    if (e) return fail(CHR_EHIP, std::string("autotune: ") + hipGetErrorString(e));
    if (upload_trace()) {
      if (M.use_dia)
        std::fprintf(stderr, "[upload]   %d rows: dia %d diagonals: %.4f ms\n", M.nrows, M.dia.m, ms / 5);
      else if (M.use_sell)
        std::fprintf(stderr, "[upload]   %d rows: sell cb %d sort %d (shift %d) elems/nnz %.3f: %.4f ms\n", M.nrows,
                     M.sell.cb, (int)M.sell.sorted, M.sell.sshift, (double)M.sell.elems / (double)M.nnz, ms / 5);
      else
        std::fprintf(stderr, "[upload]   %d rows: %s PER %d%s: %.4f ms\n", M.nrows, M.use_pack ? "pack" : "stream",
                     M.per, M.vec4 ? "" : " (scalar loads)", ms / 5);
    }
    return CHR_OK;
  };
  const double t_tune0 = now_seconds();
  // Step 1: stream layout.
  float stream_ms = -1.f;
  if (fixed_per) {
    if (fixed_per != M.per) rc = M.layout(fixed_per, hrp, hci);
  } else {
    float best = 1e30f, def_ms = 1e30f;
    int best_per = 8;
    for (int p : kPers) {
      if (p != M.per && (rc = M.layout(p, hrp, hci))) break;
      float ms = 0.f;
      if ((rc = time_it(ms))) break;
      if (p == 8) def_ms = ms;
      if (ms < best) {
        best = ms;
        best_per = p;
      }
    }
    if (!rc) {
      if (best > 0.97f * def_ms) best_per = 8;
      stream_ms = best_per == 8 ? def_ms : best;
      if (best_per != M.per) rc = M.layout(best_per, hrp, hci);
    }
  }
  const double t_tune1 = now_seconds();
  if (kt > 1 || stream_only) return rc;  // block products use the stream blocks only
  // Step 2: packed codes (the stream kernel's blocks with 2-byte codes).
  // (Only where 16-bit block offsets do not apply: otherwise the stream layout's index is
  // already 2 bytes.)
  if (!rc && pack_mode != 0 && sell_mode != 1 && (!M.idx16 || pack_mode == 1) &&
      !(rc = M.pack.build(M.nrows, M.ncols, hrp, hci, M.per)) && M.pack.built) {
    if (pack_mode == 1) {
      M.use_pack = true;
    } else {
      if (stream_ms < 0.f) rc = time_it(stream_ms);
      float ms = 0.f;
      if (!rc) {
        M.use_pack = true;
        rc = time_it(ms);
        M.use_pack = false;
      }
      if (!rc && ms < 0.97f * stream_ms) {
        M.use_pack = true;
        stream_ms = ms;
      } else {
        M.pack.release();
      }
    }
  }
  // Step 3: sliced ELL.  Matrices timed warm (below the cache-cold size) are candidates only
  // when every row fits one 16-element chunk: longer rows chain dependent chunk loads, which
  // time well from a warm L2 but stall inside the cycle, where the operands come from
  // further away (a 5,889-row level of about 70 entries per row took 11-14 us per product
  // in sliced ELL there against 5 us in the stream layout).
  int32_t maxrow = 0;
  for (int32_t i = 0; i < M.nrows; ++i) maxrow = std::max(maxrow, hrp[i + 1] - hrp[i]);
  const bool sell_ok = cold || maxrow <= 16 || sell_mode == 1;
  if (!rc && sell_mode != 0 && sell_ok && !(pack_mode == 1 && M.use_pack)) {
    // Unsorted first; when its padding is within 12% of the nonzeros the sorted variants
    // (fewer padded bytes, scattered rows) are not tried.  The builder's choice last.
    const char *env_prune = std::getenv("CHRONOS_SELL_PRUNE");
    const bool prune = !env_sort && !(env_prune && std::atoi(env_prune) == 0);
    std::vector<int> sorts;
    if (env_sort) sorts.push_back(std::atoi(env_sort));
    else sorts = {0, 8, -1};
    if (sell_mode == 1 && !env_sort) sorts = {-1};
    if (stream_ms < 0.f && sell_mode != 1) rc = time_it(stream_ms);
    float best_ms = sell_mode == 1 ? 1e30f : 0.97f * stream_ms;
    int best_sort = -2, last_sort = -2;  // -2: stream layout
    std::vector<std::pair<int, int>> seen;  // (sorted, sshift) of the layouts built so far
    // Host layouts: the first candidate alone (an unsorted layout that pads little ends the
    // search), the others built concurrently once they are needed.
    auto host_build = [&](int so) {
      chrsell::Layout Ls;
      chrsell::build(M.nrows, M.ncols, hrp, hci, hv, M.fp32 ? 4 : 8, Ls, so);
      return Ls;
    };
    std::vector<chrsell::Layout> layouts(sorts.size());
    for (size_t c = 0; c < sorts.size() && !rc; ++c) {
      if (c == 1) {
        std::vector<std::future<chrsell::Layout>> rest;
        for (size_t q = 1; q < sorts.size(); ++q) rest.push_back(std::async(std::launch::async, host_build, sorts[q]));
        for (size_t q = 1; q < sorts.size(); ++q) layouts[q] = rest[q - 1].get();
      } else if (c == 0) {
        layouts[0] = host_build(sorts[0]);
      }
      if ((rc = M.sell.adopt(layouts[c], M.fp32))) break;
      if (!M.sell.cb) {
        // Not a candidate with this sort window.  An unsorted layout padded past the limit
        // means rows of very uneven length inside a slice; sorting them apart scatters the
        // slice's gathers, which measured slower than the packed layout on every such matrix
        // (level-0 G^T at 128^3: 88-149 us sorted against 69 us packed), so the sorted
        // layouts are not tried either.
        if (sorts[c] == 0 && prune) break;
        continue;
      }
      last_sort = sorts[c];
      const bool tight = sorts[c] == 0 && (double)M.sell.elems <= 1.12 * (double)M.nnz;
      const std::pair<int, int> key{M.sell.sorted ? 1 : 0, M.sell.sorted ? M.sell.sshift : 0};
      if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
      seen.push_back(key);
      float ms = 0.f;
      M.use_sell = true;
      rc = time_it(ms);
      M.use_sell = false;
      if (!rc && (ms < best_ms || sell_mode == 1)) {
        best_ms = ms;
        best_sort = sorts[c];
        if (sell_mode == 1) break;
      }
      if (tight) break;
      // Sorting removes at most the unsorted layout's padding; when the unsorted time scaled
      // down by that padding still does not beat the best so far, no sort window can.
      if (!rc && prune && sorts[c] == 0 && ms * (double)M.nnz / (double)M.sell.elems >= best_ms) break;
    }
    if (!rc && best_sort != -2) {
      if (best_sort != last_sort)
        for (size_t c = 0; c < sorts.size(); ++c)
          if (sorts[c] == best_sort) {
            rc = M.sell.adopt(layouts[c], M.fp32);
            break;
          }
      M.use_sell = !rc && M.sell.cb != 0;
      if (M.use_sell) {
        M.sell_arg = best_sort;
        M.use_pack = false;
        M.pack.release();
      }
    } else {
      M.sell.release();
    }
  }
  // Step 4: the symmetric diagonal layout, when the matrix admits it (a few diagonals,
  // values symmetric bit for bit).  For the outer operator of PCG (tune_per = false) it also
  // carries the solver's p-product whenever it is built, a fixed rule, so that the solver's
  // rounding (its dot products are summed per row block of the product's layout) never
  // depends on a timing or on the single-vector layout.  CHRONOS_DIA=0 excludes it,
  // CHRONOS_DIA=1 forces it for single-vector products; a forced sliced-ELL or packed
  // layout keeps the single-vector products off it unless CHRONOS_DIA=1.
  const char *env_dia = std::getenv("CHRONOS_DIA");
  const int dia_mode = env_dia ? std::atoi(env_dia) : -1;
  if (!rc && dia_mode != 0 && !(rc = M.dia.build(M.nrows, M.ncols, hrp, hci, hv, M.fp32)) && M.dia.m) {
    M.dia_pcg = !tune_per;
    if (dia_mode == 1) {
      M.use_dia = true;
    } else if (sell_mode != 1 && pack_mode != 1) {
      float cur_ms = 0.f, ms = 0.f;
      rc = time_it(cur_ms);
      if (!rc) {
        M.use_dia = true;
        rc = time_it(ms);
        M.use_dia = !rc && ms < 0.97f * cur_ms;
      }
    }
    if (M.use_dia) {
      M.use_sell = false;
      M.sell.release();
      M.use_pack = false;
      M.pack.release();
    } else if (!M.dia_pcg) {
      M.dia.release();
    }
  }
  // Step 5 (FP32 stream and packed layouts): four-element vector loads or one element per
  // load, the former unless the latter is more than 3% faster.  Timed per matrix: on the
  // 128^3 diffusion hierarchy the FP32 solve takes 11.98 ms against 12.21 ms with one element
  // per load, while vector loads on every FP32 matrix slowed the 64^3 elasticity solve (long
  // rows, 16 nonzeros per lane) from 14.8 to 15.6 ms.  FP64 products load one element per
  // lane (with vector loads the FP64 diffusion solve took 16.56 against 16.38 ms).
  // CHRONOS_VEC4=0 turns vector loads off, CHRONOS_VEC4=1 forces them (FP32).
  if (!rc && M.fp32 && !M.use_sell && !M.use_dia) {
    if (const char *env_v4 = std::getenv("CHRONOS_VEC4")) {
      M.vec4 = std::atoi(env_v4) != 0;
    } else {
      float ms4 = 0.f, ms1 = 0.f;
      M.vec4 = true;
      if (!(rc = time_it(ms4))) {
        M.vec4 = false;
        rc = time_it(ms1);
        M.vec4 = !rc && !(ms1 < 0.97f * ms4);
      }
    }
  }
  if (rc) {
    M.use_sell = false;
    M.sell.release();
    M.use_pack = false;
    M.pack.release();
    M.use_dia = false;
    M.dia_pcg = false;
    M.dia.release();
  }
  if (upload_trace())
    std::fprintf(stderr, "[upload] tune %d x %d nnz %lld: stream %.3f s, pack+sell+dia %.3f s -> %s%s PER %d%s%s\n",
                 M.nrows, M.ncols, (long long)M.nnz, t_tune1 - t_tune0, now_seconds() - t_tune1,
                 M.use_dia ? "dia" : M.use_sell ? "sell" : M.use_pack ? "pack" : "stream",
                 M.use_sell && M.sell.sorted ? " sorted" : "", M.per,
                 M.use_sell || M.use_dia || M.vec4 ? "" : " scalar loads", M.dia_pcg ? " (PCG product: dia)" : "");
  return rc;
}


// Layout cache (CHRONOS_LAYOUT_CACHE=path): the tuner's decision for a matrix, keyed by its
// shape, precision, block width, role and a hash of its sparsity pattern, is appended to the
// file; a later upload of a matrix with the same key (in this or another process) rebuilds
// that layout without timing anything.  Uploads become repeatable (a profiler, which slows
// the timed launches, then sees the layouts of an untraced run) and faster.  Every layout
// gives the same products, so a cached decision never changes a product; the PCG operator's
// diagonal copy follows the tuner's fixed rule, so PCG runs keep their bits too.  The cache
// serves default tuning only: with any tuner override in the environment the tuner times
// and records nothing.
struct LayoutChoice {
  int per = 8, pack = 0, sell_arg = -2, dia = 0, vec4 = 1;  // sell_arg: -2 none, else the builder's sort argument
};

static std::mutex g_cache_mu;

static bool tuner_overridden() {
  for (const char *v : {"CHRONOS_STREAM_PER", "CHRONOS_SELL", "CHRONOS_SELL_SORT", "CHRONOS_SELL_PRUNE",
                        "CHRONOS_PACK", "CHRONOS_DIA", "CHRONOS_VEC4"})
    if (std::getenv(v)) return true;
  return false;
}

// The decisions recorded in `path` (read when the path first appears or changes).
static std::map<std::string, LayoutChoice> &layout_cache(const std::string &path) {
  static std::map<std::string, LayoutChoice> m;
  static std::string loaded;
  if (path != loaded) {
    loaded = path;
    m.clear();
    if (FILE *f = std::fopen(path.c_str(), "r")) {
      char key[128];
      LayoutChoice c;
      while (std::fscanf(f, "%127s %d %d %d %d %d", key, &c.per, &c.pack, &c.sell_arg, &c.dia, &c.vec4) == 6)
        m[key] = c;
      std::fclose(f);
    }
  }
  return m;
}

static std::string layout_key(const DevCsr &M, const int32_t *hrp, const int32_t *hci, bool tune_per, int tune_k,
                              bool stream_only) {
  uint64_t hsh = 1469598103934665603ull;  // FNV-1a over the row pointers and columns
  auto mix = [&](uint32_t v) {
    for (int b = 0; b < 4; ++b) {
      hsh ^= (v >> (8 * b)) & 255u;
      hsh *= 1099511628211ull;
    }
  };
  for (int32_t i = 0; i <= M.nrows; ++i) mix((uint32_t)hrp[i]);
  for (int64_t e = 0; e < M.nnz; ++e) mix((uint32_t)hci[e]);
  char buf[128];
  std::snprintf(buf, sizeof buf, "%dx%d:%lld:%c:%d:%d%d:%016llx", M.nrows, M.ncols, (long long)M.nnz,
                M.fp32 ? 's' : 'd', tune_k, (int)tune_per, (int)stream_only, (unsigned long long)hsh);
  return buf;
}

int tune_layout(DevCsr &M, const int32_t *hrp, const int32_t *hci, const double *hv, hipStream_t s,
                const int *zero, bool tune_per, const DevBuf *flush, int tune_k, bool stream_only) {
  const char *path = std::getenv("CHRONOS_LAYOUT_CACHE");
  if (!path || !*path || M.nnz == 0 || tuner_overridden())
    return tune_layout_timed(M, hrp, hci, hv, s, zero, tune_per, flush, tune_k, stream_only);
  const std::string key = layout_key(M, hrp, hci, tune_per, tune_k, stream_only);
  std::unique_lock<std::mutex> lk(g_cache_mu);
  auto &cache = layout_cache(path);
  auto it = cache.find(key);
  if (it != cache.end()) {
    const LayoutChoice c = it->second;
    lk.unlock();
    M.use_sell = M.use_pack = M.use_dia = M.dia_pcg = false;
    M.vec4 = c.vec4 != 0;
    M.sell_arg = -2;
    M.sell.release();
    M.pack.release();
    M.dia.release();
    // The tuner's steps after the stream layout run for single-vector tuning only.
    const bool single = std::max(1, std::min(tune_k, kMaxBlockCols)) == 1 && !stream_only;
    int rc = CHR_OK;
    if (c.per != M.per) rc = M.layout(c.per, hrp, hci);
    if (!rc && single && c.pack && !(rc = M.pack.build(M.nrows, M.ncols, hrp, hci, M.per))) M.use_pack = M.pack.built;
    if (!rc && single && c.sell_arg != -2 &&
        !(rc = M.sell.build(M.nrows, M.ncols, hrp, hci, hv, M.fp32, c.sell_arg))) {
      M.use_sell = M.sell.cb != 0;
      if (M.use_sell) M.sell_arg = c.sell_arg;
    }
    if (!rc && single && (c.dia || !tune_per) && !(rc = M.dia.build(M.nrows, M.ncols, hrp, hci, hv, M.fp32)) &&
        M.dia.m) {
      M.use_dia = c.dia != 0;
      M.dia_pcg = !tune_per;
      if (!M.use_dia && !M.dia_pcg) M.dia.release();
    }
    if (upload_trace())
      std::fprintf(stderr, "[upload] %d x %d nnz %lld: cached layout %s PER %d%s\n", M.nrows, M.ncols,
                   (long long)M.nnz, M.use_dia ? "dia" : M.use_sell ? "sell" : M.use_pack ? "pack" : "stream", M.per,
                   M.dia_pcg ? " (PCG product: dia)" : "");
    return rc;
  }
  lk.unlock();
  int rc = tune_layout_timed(M, hrp, hci, hv, s, zero, tune_per, flush, tune_k, stream_only);
  if (rc) return rc;
  LayoutChoice c;
  c.per = M.per;
  c.pack = M.use_pack ? 1 : 0;
  c.sell_arg = M.use_sell ? M.sell_arg : -2;
  c.dia = M.use_dia ? 1 : 0;
  c.vec4 = M.vec4 ? 1 : 0;
  lk.lock();
  layout_cache(path)[key] = c;
  if (FILE *f = std::fopen(path, "a")) {
    std::fprintf(f, "%s %d %d %d %d %d\n", key.c_str(), c.per, c.pack, c.sell_arg, c.dia, c.vec4);
    std::fclose(f);
  }
  return CHR_OK;
}

// Dense SPD inverse via Cholesky (FP64, host).  Returns false if not positive definite.
bool spd_inverse(int32_t n, std::vector<double> &a) {
  // a: row-major n x n; lower Cholesky in place.
  for (int32_t j = 0; j < n; ++j) {
    double d = a[(size_t)j * n + j];
    for (int32_t k = 0; k < j; ++k) d -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    a[(size_t)j * n + j] = d;
    for (int32_t i = j + 1; i < n; ++i) {
      double s = a[(size_t)i * n + j];
      for (int32_t k = 0; k < j; ++k) s -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
      a[(size_t)i * n + j] = s / d;
    }
  }
  std::vector<double> inv((size_t)n * n, 0.0), col(n);
  for (int32_t c = 0; c < n; ++c) {
    std::fill(col.begin(), col.end(), 0.0);
    col[c] = 1.0;
    for (int32_t i = 0; i < n; ++i) {
      double s = col[i];
      for (int32_t k = 0; k < i; ++k) s -= a[(size_t)i * n + k] * col[k];
      col[i] = s / a[(size_t)i * n + i];
    }
    for (int32_t i = n - 1; i >= 0; --i) {
      double s = col[i];
      for (int32_t k = i + 1; k < n; ++k) s -= a[(size_t)k * n + i] * col[k];
      col[i] = s / a[(size_t)i * n + i];
    }
    for (int32_t i = 0; i < n; ++i) inv[(size_t)i * n + c] = col[i];
  }
  a.swap(inv);
  return true;
}

hipError_t enqueue_vcycle(chr_hier &h, const double *b, double *out, const int *stop, int k,
                          const float *b0f) {
  if (h.fp32()) {
    Cycle<float> c{h, stop, h.stream, k};
    c.b0f = b0f;
    return c.level<double, double>(0, b, out);
  }
  Cycle<double> c{h, stop, h.stream, k};
  return c.level<double, double>(0, b, out);
}

int ensure_block_capacity(chr_hier &h, int k) {
  if (k <= h.kcap) return CHR_OK;
  HIPC(hipStreamSynchronize(h.stream));
  if (h.pcg) h.pcg->drop_graphs();  // they point at the buffers about to be replaced
  ++h.buf_gen;
  const size_t s = h.fp32() ? 4 : 8;
  int rc;
  for (size_t l = 0; l < h.lev.size(); ++l) {
    auto &L = h.lev[l];
    if ((rc = L.x.alloc(s * L.n * k)) || (rc = L.r.alloc(s * L.n * k)) || (rc = L.y.alloc(s * L.n * k)))
      return rc;
    if (l > 0 && (rc = L.b.alloc(s * L.n * k))) return rc;
  }
  const int32_t n0 = h.lev[0].n;
  if ((rc = h.io_b.alloc(8 * (size_t)n0 * k)) || (rc = h.io_x.alloc(8 * (size_t)n0 * k))) return rc;
  h.kcap = k;
  return CHR_OK;
}

void fill_stats(chr_hier &h) {
  chr_hier_stats &st = h.stats;
  std::memset(&st, 0, sizeof(st));
  st.nlev = (int32_t)h.lev.size();
  st.precision = h.precision;
  const double s = h.fp32() ? 4.0 : 8.0;
  double total = 0.0, total_csr = 0.0, dev = 0.0;
  for (size_t l = 0; l < h.lev.size(); ++l) {
    auto &L = h.lev[l];
    if (l < 16) st.n[l] = L.n;
    // Direct solves: the FP64 inverses once + input and output vectors.
    const double nd_direct = (double)(L.n - L.nd);
    double b = 8.0 * (double)L.coarse_m2 + 2.0 * nd_direct * 8.0, bc = b;
    dev += 8.0 * (double)L.coarse_m2;
    if (L.nd > 0) {
      // Per level: 2 A, 2 G, 2 G^T, P, P^T, 18 n_l + 2 n_{l+1} vector words (AMG_Vcycle.m).
      const double nc = h.lev[l + 1].n;
      const double vec = 18.0 * L.nd * s + 2.0 * nc * s;
      // Stored bytes of the products the cycle actually runs (composites where fused).
      const double pre = L.fuse_s ? L.S.stream_bytes() : L.G.stream_bytes() + L.GT.stream_bytes();
      const double down = L.fuse_r ? L.R.stream_bytes() : L.A.stream_bytes() + L.PT.stream_bytes();
      const double up = !L.fuse_up ? L.P.stream_bytes() + L.A.stream_bytes() + L.G.stream_bytes() + L.GT.stream_bytes()
                        : L.U.stream_bytes() + L.W.stream_bytes() + (L.fuse_s ? 0.0 : L.G.stream_bytes());
      b += pre + down + up + vec;
      bc += 2 * L.A.csr_bytes() + 2 * L.G.csr_bytes() + 2 * L.GT.csr_bytes() + L.P.csr_bytes() +
            L.PT.csr_bytes() + vec;
      dev += L.A.device_bytes() + L.G.device_bytes() + L.GT.device_bytes() + L.P.device_bytes() +
             L.PT.device_bytes() + L.R.device_bytes() + L.U.device_bytes() + L.W.device_bytes() +
             L.S.device_bytes() + L.b.bytes + L.x.bytes + L.r.bytes + L.y.bytes;
    }
    if (l < 16) st.level_bytes[l] = bc;
    total += b;
    total_csr += bc;
  }
  st.vcycle_bytes = total;
  st.vcycle_csr_bytes = total_csr;
  dev += h.outer.A.device_bytes();
  st.device_bytes = dev;
  const auto &A0 = h.outer.A;
  st.pcg_iter_bytes = total_csr + A0.csr_bytes() + 18.0 * A0.nrows * 8.0;
}

int sync_stream(hipStream_t s) {
  HIPC(hipStreamSynchronize(s));
  return CHR_OK;
}

}  // namespace chrrt

using namespace chrrt;

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" const char *chr_last_error(void) { return g_err.c_str(); }

namespace chrrt {
int scratch_setup(void **out);
}

extern "C" int chr_init(int32_t device) {
  HIPC(hipSetDevice(device));
  void *p = nullptr;
  return scratch_setup(&p);
}

namespace chrrt {

// Owned host CSR (FP64).
struct HostMat {
  int32_t nrows = 0, ncols = 0;
  std::vector<int32_t> rp{0}, ci;
  std::vector<double> v;
  chr_csr view() const { return {nrows, ncols, rp.data(), ci.data(), v.data()}; }
};

// Stored zeros are dropped before a matrix is uploaded.  A product gains nothing from them:
// every kernel starts a row sum at +0 and adds rounded products, a sum that starts at +0 is
// never -0, and adding a +-0 product leaves any other sum unchanged, so the products keep
// their bits for finite operands while the kernels stream fewer bytes (the 64^3 elasticity
// operator stores 13.4 million zeros among its 63.7 million entries).  Rows longer than a
// stream block are the exception: their strided partial sums regroup when entries drop out,
// which changes the rounding of those rows only.  CHRONOS_KEEP_ZEROS=1 uploads matrices as
// given.  Returns false (and leaves out untouched) when M stores no zero.
bool drop_zeros(const chr_csr &M, HostMat &out) {
  if (const char *e = std::getenv("CHRONOS_KEEP_ZEROS"); e && std::atoi(e) != 0) return false;
  const int64_t nnz = M.rowptr[M.nrows];
  int64_t zeros = 0;
  for (int64_t e = 0; e < nnz; ++e) zeros += M.val[e] == 0.0;
  if (zeros == 0) return false;
  out.nrows = M.nrows;
  out.ncols = M.ncols;
  out.rp.assign((size_t)M.nrows + 1, 0);
  out.ci.clear();
  out.v.clear();
  out.ci.reserve((size_t)(nnz - zeros));
  out.v.reserve((size_t)(nnz - zeros));
  for (int32_t i = 0; i < M.nrows; ++i) {
    for (int32_t e = M.rowptr[i]; e < M.rowptr[i + 1]; ++e)
      if (M.val[e] != 0.0) {
        out.ci.push_back(M.col[e]);
        out.v.push_back(M.val[e]);
      }
    out.rp[(size_t)i + 1] = (int32_t)out.ci.size();
  }
  return true;
}

// C = L R, row by row (Gustavson): row i of C accumulates s * (row j of R) for the entries
// (i, j, s) of L in L's column order; columns sorted.  Exact zeros are kept.
HostMat spgemm_host(const chr_csr &L, const chr_csr &R) {
  HostMat C;
  C.nrows = L.nrows;
  C.ncols = R.ncols;
  C.rp.assign((size_t)L.nrows + 1, 0);
  std::vector<int32_t> mark((size_t)R.ncols, -1), cols;
  std::vector<double> acc((size_t)R.ncols, 0.0);
  for (int32_t i = 0; i < L.nrows; ++i) {
    cols.clear();
    for (int32_t e = L.rowptr[i]; e < L.rowptr[i + 1]; ++e) {
      const int32_t j = L.col[e];
      const double a = L.val[e];
      for (int32_t f = R.rowptr[j]; f < R.rowptr[j + 1]; ++f) {
        const int32_t c = R.col[f];
        if (mark[c] != i) {
          mark[c] = i;
          acc[c] = 0.0;
          cols.push_back(c);
        }
        acc[c] += a * R.val[f];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int32_t c : cols) {
      C.ci.push_back(c);
      C.v.push_back(acc[c]);
    }
    C.rp[i + 1] = (int32_t)C.ci.size();
  }
  return C;
}

// [s1 M1 | s2 M2] (same row count) with interleaved columns: M1's column j becomes 2j and
// M2's column j becomes 2j + 1 (chr::XTwo); each row's entries in ascending column order.
HostMat hcat_host(const chr_csr &M1, double s1, const chr_csr &M2, double s2) {
  HostMat C;
  C.nrows = M1.nrows;
  C.ncols = 2 * std::max(M1.ncols, M2.ncols);
  C.rp.assign((size_t)M1.nrows + 1, 0);
  const int64_t nnz = (int64_t)M1.rowptr[M1.nrows] + M2.rowptr[M2.nrows];
  C.ci.reserve(nnz);
  C.v.reserve(nnz);
  for (int32_t i = 0; i < M1.nrows; ++i) {
    int32_t e1 = M1.rowptr[i], e2 = M2.rowptr[i];
    const int32_t z1 = M1.rowptr[i + 1], z2 = M2.rowptr[i + 1];
    while (e1 < z1 || e2 < z2) {  // merge by interleaved column
      if (e2 >= z2 || (e1 < z1 && 2 * M1.col[e1] < 2 * M2.col[e2] + 1)) {
        C.ci.push_back(2 * M1.col[e1]);
        C.v.push_back(s1 * M1.val[e1++]);
      } else {
        C.ci.push_back(2 * M2.col[e2] + 1);
        C.v.push_back(s2 * M2.val[e2++]);
      }
    }
    C.rp[i + 1] = (int32_t)C.ci.size();
  }
  return C;
}

// Whether R = [-P^T A | P^T] (interleaved columns) can keep 16-bit column offsets, judged
// before forming it: row c of R reaches the columns of A's rows i in P^T's row c, so its
// interleaved columns lie in [2 lo_c, 2 hi_c + 1] with lo_c, hi_c the extreme columns of
// those rows; rows are cut into blocks as the stream layout's smallest candidate does (at
// most 1024 entries and 256 rows, with row lengths bounded by the sum of those A rows), and
// every block must span fewer than 65536 columns.
bool restriction_fits_16bit(const chr_csr &PT, const chr_csr &A) {
  std::vector<int32_t> amin((size_t)A.nrows), amax((size_t)A.nrows);
  for (int32_t i = 0; i < A.nrows; ++i) {
    int32_t lo = i, hi = i;
    for (int32_t e = A.rowptr[i]; e < A.rowptr[i + 1]; ++e) {
      lo = std::min(lo, A.col[e]);
      hi = std::max(hi, A.col[e]);
    }
    amin[(size_t)i] = lo;
    amax[(size_t)i] = hi;
  }
  int64_t blo = INT64_MAX, bhi = INT64_MIN, cnt = 0;
  int32_t rows = 0;
  for (int32_t c = 0; c < PT.nrows; ++c) {
    int64_t lo = INT64_MAX, hi = INT64_MIN, len = 0;
    for (int32_t e = PT.rowptr[c]; e < PT.rowptr[c + 1]; ++e) {
      const int32_t i = PT.col[e];
      lo = std::min<int64_t>(lo, amin[(size_t)i]);
      hi = std::max<int64_t>(hi, amax[(size_t)i]);
      len += 1 + A.rowptr[i + 1] - A.rowptr[i];
    }
    if (rows > 0 && (rows == chr::kBlock || cnt + len > 4 * chr::kBlock)) {
      blo = INT64_MAX;
      bhi = INT64_MIN;
      rows = 0;
      cnt = 0;
    }
    ++rows;
    cnt += len;
    if (len > 0) {
      blo = std::min(blo, 2 * lo);
      bhi = std::max(bhi, 2 * hi + 1);
      if (bhi - blo > 65535) return false;
    }
  }
  return true;
}

// Composite operators of level L (chr_hier::Level).  The residual-and-restriction operator R
// is formed on every level >= 1 where P^T A has no more nonzeros than A and R keeps 16-bit
// column offsets (it then streams fewer bytes than A and P^T, and saves a kernel and the
// residual vector's round trip), and on every small level.  On small levels (A with fewer than CHRONOS_FUSE_NNZ nonzeros,
// default 4,000,000: their products are bound by latency, not bytes) the up path and the
// pre-smoothing are fused too (U, W, S).  CHRONOS_FUSE=0 disables all of it.
int build_composites(chr_hier::Level &L, const LevelIn &in, const std::vector<int32_t> &ptrp,
                     const std::vector<int32_t> &ptci, const std::vector<double> &ptv,
                     const std::vector<int32_t> &gtrp, const std::vector<int32_t> &gtci,
                     const std::vector<double> &gtv, bool f32, hipStream_t s, const int *zf) {
  if (const char *e = std::getenv("CHRONOS_FUSE"); e && std::atoi(e) == 0) return CHR_OK;
  const char *env_nnz = std::getenv("CHRONOS_FUSE_NNZ");
  const int64_t small_nnz = env_nnz ? std::atoll(env_nnz) : 4000000;
  const int64_t nnzA = in.A.rowptr[in.A.nrows];
  const bool small = nnzA < small_nnz;
  const chr_csr PT{in.P.ncols, in.P.nrows, ptrp.data(), ptci.data(), ptv.data()};
  if (!small && !restriction_fits_16bit(PT, in.A)) return CHR_OK;
  HostMat PtA = spgemm_host(PT, in.A);
  if (!small && (int64_t)PtA.ci.size() > nnzA) return CHR_OK;
  int rc;
  auto upload = [&](DevCsr &M, const HostMat &H) -> int {
    int r;
    if ((r = M.upload(H.nrows, H.ncols, H.rp.data(), H.ci.data(), H.v.data(), f32))) return r;
    return tune_layout(M, H.rp.data(), H.ci.data(), H.v.data(), s, zf, true, nullptr, 1, true);
  };
  HostMat R = hcat_host(PtA.view(), -1.0, PT, 1.0);
  PtA = HostMat();
  if ((rc = upload(L.R, R))) return rc;
  if (!small && !L.R.idx16) {  // 4-byte columns: more bytes than A and P^T in their layouts
    L.R = DevCsr();
    return CHR_OK;
  }
  L.fuse_r = true;
  if (!small) return CHR_OK;
  const chr_csr GT{in.G.ncols, in.G.nrows, gtrp.data(), gtci.data(), gtv.data()};
  // CHRONOS_FUSE_S=0: the up path fused with W = [P | omega G^T] and the smoothing sweeps as
  // two products (for comparison).
  const char *env_s = std::getenv("CHRONOS_FUSE_S");
  const bool with_s = !env_s || std::atoi(env_s) != 0;
  HostMat AP = spgemm_host(in.A, in.P);
  HostMat U = hcat_host(AP.view(), 1.0, in.A, 1.0);
  AP = HostMat();
  if ((rc = upload(L.U, U))) return rc;
  if (with_s) {
    HostMat S = spgemm_host(GT, in.G);
    for (double &v : S.v) v *= in.omega;
    HostMat W = hcat_host(in.P, 1.0, S.view(), 1.0);
    if ((rc = upload(L.S, S)) || (rc = upload(L.W, W))) return rc;
  } else {
    HostMat W = hcat_host(in.P, 1.0, GT, in.omega);
    if ((rc = upload(L.W, W))) return rc;
  }
  L.fuse_s = with_s;
  L.fuse_up = true;
  return CHR_OK;
}

// Builds a hierarchy on the device from per-level host descriptions (see chr_hier::Level).
int build_hier(std::vector<LevelIn> &lv, int32_t precision, const chr_csr &outer_in, chr_hier **out, int tune_k) {
  // Matrices without their stored zeros (drop_zeros); the level views point into these.
  std::vector<HostMat> nz;
  nz.reserve(3 * lv.size() + 1);
  auto nonzero = [&](chr_csr &m) {
    HostMat hm;
    if (drop_zeros(m, hm)) {
      nz.push_back(std::move(hm));
      m = nz.back().view();
    }
  };
  for (LevelIn &in : lv)
    if (in.nd > 0) {
      nonzero(in.A);
      nonzero(in.G);
      nonzero(in.P);
    }
  chr_csr outer = outer_in;
  nonzero(outer);
  auto h = std::make_unique<chr_hier>();
  h->precision = precision;
  HIPC(hipGetDevice(&h->device));
  HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  int rc;
  if ((rc = h->zero.alloc(sizeof(int)))) return rc;
  HIPC(hipMemset(h->zero.p, 0, sizeof(int)));
  const int *zf = h->zero.as<int>();
  const bool f32 = precision == CHR_FP32_CYCLE;
  const size_t s = f32 ? 4 : 8;
  const int32_t nlev = (int32_t)lv.size();
  // Cache-cold autotuning of the large level matrices needs a 1 GiB flush buffer for the
  // duration of the upload.
  DevBuf flush;
  if (lv[0].nd > 0 && (double)lv[0].A.rowptr[lv[0].nd] * (s + 4) >= 32e6) {
    if ((rc = flush.alloc(size_t(1) << 30))) return rc;
    HIPC(hipMemsetAsync(flush.p, 0, flush.bytes, h->stream));
  }
  const DevBuf *fl = flush.p ? &flush : nullptr;
  h->lev.resize(nlev);
  std::vector<double> inv_all;  // packed inverses of every direct block
  const double t_all = now_seconds();
  for (int32_t l = 0; l < nlev; ++l) {
    const LevelIn &in = lv[l];
    auto &L = h->lev[l];
    const double t_lev = now_seconds();
    L.n = in.n;
    L.nd = in.nd;
    L.omega = in.omega;
    L.nupre = in.nupre;
    L.nupost = in.nupost;
    if (in.nd > 0) {
      if ((rc = L.A.upload(in.A, f32)) || (rc = tune_layout(L.A, in.A.rowptr, in.A.col, in.A.val, h->stream, zf, true, fl, tune_k)))
        return rc;
      if ((rc = L.G.upload(in.G, f32)) || (rc = tune_layout(L.G, in.G.rowptr, in.G.col, in.G.val, h->stream, zf, true, fl, tune_k)))
        return rc;
      if ((rc = L.P.upload(in.P, f32)) || (rc = tune_layout(L.P, in.P.rowptr, in.P.col, in.P.val, h->stream, zf, true, fl, tune_k)))
        return rc;
      std::vector<int32_t> trp, tci, prp, pci;
      std::vector<double> tv, pv;
      transpose_host(in.G, trp, tci, tv);
      if ((rc = L.GT.upload(in.G.ncols, in.G.nrows, trp.data(), tci.data(), tv.data(), f32)) ||
          (rc = tune_layout(L.GT, trp.data(), tci.data(), tv.data(), h->stream, zf, true, fl, tune_k)))
        return rc;
      transpose_host(in.P, prp, pci, pv);
      if ((rc = L.PT.upload(in.P.ncols, in.P.nrows, prp.data(), pci.data(), pv.data(), f32)) ||
          (rc = tune_layout(L.PT, prp.data(), pci.data(), pv.data(), h->stream, zf, true, fl, tune_k)))
        return rc;
      // Composite operators: single systems (every row smoothed), levels >= 1, V(1,1).
      const double tc = now_seconds();
      if (l >= 1 && in.nd == in.n && tune_k == 1 && in.nupre == 1 && in.nupost == 1 &&
          (rc = build_composites(L, in, prp, pci, pv, trp, tci, tv, f32, h->stream, zf)))
        return rc;
      if (upload_trace())
        std::fprintf(stderr, "[upload] level %d: composites %.3f s (r %d u %d s %d)\n", l, now_seconds() - tc,
                     (int)L.fuse_r, (int)L.fuse_up, (int)L.fuse_s);
      if (!in.omega_rows.empty()) {
        if ((rc = L.omega_rows.alloc(sizeof(double) * in.nd))) return rc;
        HIPC(hipMemcpy(L.omega_rows.p, in.omega_rows.data(), sizeof(double) * in.nd, hipMemcpyHostToDevice));
      }
    }
    // Direct blocks: FP64 dense inverses packed back to back, one cmap entry per row.
    if (!in.direct.empty()) {
      std::vector<int4> cmap((size_t)(in.n - in.nd));
      for (const auto &d : in.direct) {
        const int32_t m = d.A.nrows;
        if ((int64_t)m * m > (int64_t)1 << 26)
          return fail(CHR_EARG, "coarsest level too large for a dense inverse");
        if (d.start < in.nd || d.start + m > in.n) return fail(CHR_EARG, "direct block outside the level");
        std::vector<double> a((size_t)m * m, 0.0);
        for (int32_t i = 0; i < m; ++i)
          for (int32_t e = d.A.rowptr[i]; e < d.A.rowptr[i + 1]; ++e) a[(size_t)i * m + d.A.col[e]] = d.A.val[e];
        if (!spd_inverse(m, a)) return fail(CHR_ENOTSPD, "coarsest operator is not positive definite");
        const int64_t off = (int64_t)inv_all.size();
        if (off > INT32_MAX) return fail(CHR_EARG, "direct blocks too large");
        inv_all.insert(inv_all.end(), a.begin(), a.end());
        for (int32_t i = 0; i < m; ++i) cmap[d.start + i - in.nd] = make_int4(d.start, m, (int32_t)off, 0);
        L.ncoarse_blocks++;
        L.coarse_m2 += (int64_t)m * m;
        h->nc += m;
      }
      if ((rc = L.cmap.alloc(sizeof(int4) * cmap.size()))) return rc;
      HIPC(hipMemcpy(L.cmap.p, cmap.data(), sizeof(int4) * cmap.size(), hipMemcpyHostToDevice));
    }
    if ((rc = L.x.alloc(s * L.n)) || (rc = L.r.alloc(s * L.n)) || (rc = L.y.alloc(s * L.n))) return rc;
    if (l > 0 && (rc = L.b.alloc(s * L.n))) return rc;
    if (upload_trace()) std::fprintf(stderr, "[upload] level %d (%d rows): %.3f s\n", l, L.n, now_seconds() - t_lev);
  }
  if (!inv_all.empty()) {
    if ((rc = h->Ainv.alloc(sizeof(double) * inv_all.size()))) return rc;
    HIPC(hipMemcpy(h->Ainv.p, inv_all.data(), sizeof(double) * inv_all.size(), hipMemcpyHostToDevice));
  }
  // Level-0 operator in FP64 for the outer Krylov loop.
  if ((rc = h->outer.A.upload(outer, false)) ||
      (rc = tune_layout(h->outer.A, outer.rowptr, outer.col, outer.val, h->stream, zf, false, fl)))
    return rc;
  if ((rc = h->outer.zero.alloc(sizeof(int)))) return rc;
  HIPC(hipMemset(h->outer.zero.p, 0, sizeof(int)));
  const int32_t n0 = h->lev[0].n;
  if ((rc = h->io_b.alloc(8 * (size_t)n0))) return rc;
  if ((rc = h->io_x.alloc(8 * (size_t)n0))) return rc;
  fill_stats(*h);
  if (upload_trace()) std::fprintf(stderr, "[upload] hierarchy: %.3f s\n", now_seconds() - t_all);
  *out = h.release();
  return CHR_OK;
}

}  // namespace chrrt

namespace chrrt {

// Tail collapse.  The V-cycle from a small level c down (zero initial guess, smoothing,
// transfers, the coarsest solve) is a fixed linear map b_c -> x_c.  When n_c is at most
// CHRONOS_TAIL_ROWS (default 2048; 0 disables), that map is formed once by running the
// cycle's own kernels from level c on the unit vectors, and level c becomes the last level
// with that matrix applied as its dense FP64 direct solve.  The dozen small kernels of the
// levels below c (each bound by dependent memory latency) become one GEMV; the result equals
// the uncollapsed cycle's up to the rounding of one matrix-vector product.  Single-system
// hierarchies only (batched hierarchies keep every level).
//
// The unit vectors go through the block V-cycle, kTailCols columns per pass (each column of
// a block cycle equals the single-vector cycle of that column bit for bit), with the stream
// synchronised and the pass's columns copied out before the next pass is enqueued, so at
// most one pass (a few dozen kernels) is ever queued.
constexpr int kTailCols = 16;

// B = the unit vectors e_{j0}, ..., e_{j0+k-1} as a row-major m x k block.
template <typename T>
__global__ __launch_bounds__(chr::kBlock) void unit_block_kernel(int m, int k, int j0, T *__restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * chr::kBlock + threadIdx.x;
  if (t >= (int64_t)m * k) return;
  const int i = (int)(t / k), c = (int)(t % k);
  b[t] = i == j0 + c ? T(1) : T(0);
}

template <typename T>
int tail_columns(chr_hier &h, int32_t c, std::vector<double> &M) {
  const int32_t m = h.lev[c].n;
  const int *stop = h.zero.as<int>();
  DevBuf out;
  int rc;
  if ((rc = out.alloc(sizeof(T) * (size_t)m * kTailCols))) return rc;
  std::vector<T> host((size_t)m * kTailCols);
  for (int32_t j0 = 0; j0 < m; j0 += kTailCols) {
    const int k = std::min<int32_t>(kTailCols, m - j0);
    const unsigned grid = (unsigned)(((int64_t)m * k + chr::kBlock - 1) / chr::kBlock);
    hipLaunchKernelGGL(unit_block_kernel<T>, dim3(grid), dim3(chr::kBlock), 0, h.stream, m, k, j0,
                       h.lev[c].b.as<T>());
    HIPC(hipGetLastError());
    Cycle<T> cy{h, stop, h.stream, k};
    HIPC((cy.template level<T, T>(c, h.lev[c].b.as<T>(), out.as<T>())));
    HIPC(hipMemcpyAsync(host.data(), out.p, sizeof(T) * (size_t)m * k, hipMemcpyDeviceToHost, h.stream));
    HIPC(hipStreamSynchronize(h.stream));
    // Row i of the block is (M e_{j0}, ..., M e_{j0+k-1})_i; M is stored row-major.
    for (int32_t i = 0; i < m; ++i)
      for (int q = 0; q < k; ++q) M[(size_t)i * m + j0 + q] = (double)host[(size_t)i * k + q];
  }
  return CHR_OK;
}

int collapse_tail(chr_hier &h) {
  const char *env = std::getenv("CHRONOS_TAIL_ROWS");
  const int64_t limit = env ? std::atoll(env) : 2048;
  const int32_t nlev = (int32_t)h.lev.size();
  int32_t c = -1;
  for (int32_t l = 1; l + 1 < nlev; ++l)
    if (h.lev[l].nd == h.lev[l].n && h.lev[l].n <= limit) {
      c = l;
      break;
    }
  if (c < 0 || limit <= 0) return CHR_OK;
  auto &L = h.lev[c];
  const int32_t m = L.n;
  const size_t sz = h.fp32() ? 4 : 8;
  int rc;
  // Cycle vectors of levels c and below widened to kTailCols columns for the passes (the
  // levels below c are dropped afterwards; level c keeps the wider buffers).
  HIPC(hipStreamSynchronize(h.stream));
  if (h.kcap < kTailCols)
    for (int32_t l = c; l < nlev; ++l) {
      auto &Ll = h.lev[l];
      const size_t vb = sz * (size_t)Ll.n * kTailCols;
      if ((rc = Ll.x.alloc(vb)) || (rc = Ll.r.alloc(vb)) || (rc = Ll.y.alloc(vb)) || (rc = Ll.b.alloc(vb)))
        return rc;
    }
  std::vector<double> M((size_t)m * m);
  if ((rc = h.fp32() ? tail_columns<float>(h, c, M) : tail_columns<double>(h, c, M))) return rc;
  // Level c becomes the last level: all its rows are direct, with M as their operator.
  std::vector<int4> cmap((size_t)m, make_int4(0, m, 0, 0));
  if ((rc = L.cmap.alloc(sizeof(int4) * m)) || (rc = h.Ainv.alloc(sizeof(double) * M.size()))) return rc;
  HIPC(hipMemcpy(L.cmap.p, cmap.data(), sizeof(int4) * m, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(h.Ainv.p, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice));
  L.nd = 0;
  L.ncoarse_blocks = 1;
  L.coarse_m2 = (int64_t)m * m;
  L.G = DevCsr();
  L.GT = DevCsr();
  L.P = DevCsr();
  L.PT = DevCsr();
  L.R = DevCsr();
  L.U = DevCsr();
  L.W = DevCsr();
  L.S = DevCsr();
  L.fuse_r = L.fuse_up = L.fuse_s = false;
  h.lev.resize((size_t)c + 1);
  h.nc = m;
  fill_stats(h);
  return CHR_OK;
}

}  // namespace chrrt

extern "C" int chr_hier_create(int32_t nlev, const chr_level *levels, int32_t precision,
                               chr_hier **out) {
  if (nlev < 1 || nlev > 16 || !levels || !out)
    return fail(CHR_EARG, "chr_hier_create: need 1..16 levels and non-null pointers");
  if (precision != CHR_FP64 && precision != CHR_FP32_CYCLE)
    return fail(CHR_EARG, "chr_hier_create: unknown precision");
  // Shape checks before any device work.
  std::vector<LevelIn> lv(nlev);
  for (int32_t l = 0; l < nlev; ++l) {
    int rc;
    const chr_level &L = levels[l];
    if ((rc = check_csr(L.A, "A"))) return rc;
    if (L.A.nrows != L.A.ncols) return fail(CHR_EARG, "level operator must be square");
    lv[l].n = L.A.nrows;
    if (l + 1 < nlev) {
      if ((rc = check_csr(L.P, "P"))) return rc;
      if ((rc = check_csr(L.G, "G"))) return rc;
      if (L.P.nrows != L.A.nrows || L.P.ncols != levels[l + 1].A.nrows)
        return fail(CHR_EARG, "P shape does not match consecutive level sizes");
      if (L.G.nrows != L.A.nrows || L.G.ncols != L.A.nrows)
        return fail(CHR_EARG, "G shape does not match the level size");
      if (L.nupre < 1 || L.nupost < 0) return fail(CHR_EARG, "nupre must be >= 1, nupost >= 0");
      lv[l].nd = L.A.nrows;
      lv[l].A = L.A;
      lv[l].G = L.G;
      lv[l].P = L.P;
      lv[l].omega = L.omega;
      lv[l].nupre = L.nupre;
      lv[l].nupost = L.nupost;
    } else {
      lv[l].nd = 0;
      lv[l].direct.push_back({0, L.A});
    }
  }
  int rc = build_hier(lv, precision, levels[0].A, out);
  if (rc) return rc;
  const double t_tail = now_seconds();
  rc = collapse_tail(**out);
  if (upload_trace()) std::fprintf(stderr, "[upload] tail collapse: %.3f s\n", now_seconds() - t_tail);
  if (rc) {
    chr_hier_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

extern "C" int chr_hier_destroy(chr_hier *h) {
  if (h) {
    (void)hipStreamSynchronize(h->stream);
    delete h;
  }
  return CHR_OK;
}

extern "C" int chr_hier_get_stats(const chr_hier *h, chr_hier_stats *st) {
  if (!h || !st) return fail(CHR_EARG, "null");
  *st = h->stats;
  return CHR_OK;
}

extern "C" int chr_hier_layout(const chr_hier *h, int32_t level, char which, int32_t *per,
                               int32_t *code_bits, int64_t *nnz, double *stored_bytes) {
  if (!h || level < 0 || level >= (int32_t)h->lev.size() || h->lev[level].nd == 0)
    return fail(CHR_EARG, "chr_hier_layout: level without smoother/transfers");
  const auto &L = h->lev[level];
  const DevCsr *M = which == 'A' ? &L.A : which == 'G' ? &L.G : which == 'T' ? &L.GT
                  : which == 'P' ? &L.P : which == 'R' ? &L.PT
                  : which == 'r' && L.fuse_r ? &L.R : which == 'u' && L.fuse_up ? &L.U
                  : which == 'w' && L.fuse_up ? &L.W : which == 's' && L.fuse_s ? &L.S : nullptr;
  if (!M) return fail(CHR_EARG, "chr_hier_layout: which must be A, G, T, P or R, or a composite (r, u, w, s) the level has");
  if (per) *per = (M->use_sell || M->use_dia) ? 0 : M->use_pack ? -M->pack.per : M->per;
  if (code_bits) *code_bits = M->use_dia ? 0 : M->use_sell ? M->sell.cb : (M->use_pack || M->idx16) ? 16 : 32;
  if (nnz) *nnz = M->nnz;
  if (stored_bytes) *stored_bytes = M->stream_bytes();
  return CHR_OK;
}

extern "C" int chr_vcycle_apply(chr_hier *h, int32_t k, const double *b, int64_t ldb, double *x,
                                int64_t ldx, int32_t ptr_kind) {
  if (!h || k < 0 || k > kMaxBlockCols || (k > 0 && (!b || !x)))
    return fail(CHR_EARG, "chr_vcycle_apply: need 0 <= k <= 16 and non-null blocks");
  const int32_t n = h->lev[0].n;
  if (k > 0 && (ldb < n || ldx < n)) return fail(CHR_EARG, "chr_vcycle_apply: leading dimension < n");
  if (k == 0) return CHR_OK;
  int rc;
  if ((rc = ensure_block_capacity(*h, k))) return rc;
  const int *zf = h->zero.as<int>();
  if (k == 1) {
    if (ptr_kind == CHR_PTR_HOST) {
      HIPC(hipMemcpyAsync(h->io_b.p, b, 8 * (size_t)n, hipMemcpyHostToDevice, h->stream));
      HIPC(enqueue_vcycle(*h, h->io_b.as<double>(), h->io_x.as<double>(), zf));
      HIPC(hipMemcpyAsync(x, h->io_x.p, 8 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    } else {
      HIPC(enqueue_vcycle(*h, b, x, zf));
    }
    return sync_stream(h->stream);
  }
  // Blocks: column-major (ld) at the boundary, row-major n x k inside the cycle.
  DevBuf cm;
  const double *bsrc = b;
  int64_t lds = ldb;
  if (ptr_kind == CHR_PTR_HOST) {
    if ((rc = cm.alloc(8 * (size_t)n * k))) return rc;
    HIPC(copy_cols(cm.p, 8 * (size_t)n, b, 8 * (size_t)ldb, 8 * (size_t)n, k, hipMemcpyHostToDevice,
                   h->stream));
    bsrc = cm.as<double>();
    lds = n;
  }
  const unsigned grid = (unsigned)(((int64_t)n * k + chr::kBlock - 1) / chr::kBlock);
  hipLaunchKernelGGL(chr::cm_to_rm_kernel, dim3(grid), dim3(chr::kBlock), 0, h->stream, n, k, bsrc, lds,
                     h->io_b.as<double>());
  HIPC(hipGetLastError());
  HIPC(enqueue_vcycle(*h, h->io_b.as<double>(), h->io_x.as<double>(), zf, k));
  if (ptr_kind == CHR_PTR_HOST) {
    hipLaunchKernelGGL(chr::rm_to_cm_kernel, dim3(grid), dim3(chr::kBlock), 0, h->stream, n, k,
                       (const double *)h->io_x.as<double>(), cm.as<double>(), (int64_t)n);
    HIPC(hipGetLastError());
    HIPC(copy_cols(x, 8 * (size_t)ldx, cm.p, 8 * (size_t)n, 8 * (size_t)n, k, hipMemcpyDeviceToHost,
                   h->stream));
  } else {
    hipLaunchKernelGGL(chr::rm_to_cm_kernel, dim3(grid), dim3(chr::kBlock), 0, h->stream, n, k,
                       (const double *)h->io_x.as<double>(), x, ldx);
    HIPC(hipGetLastError());
  }
  return sync_stream(h->stream);
}

extern "C" int chr_mat_create(const chr_csr *A, chr_mat **out) {
  if (!A || !out) return fail(CHR_EARG, "null");
  int rc = check_csr(*A, "A");
  if (rc) return rc;
  auto m = std::make_unique<chr_mat>();
  if ((rc = m->zero.alloc(sizeof(int)))) return rc;
  HIPC(hipMemset(m->zero.p, 0, sizeof(int)));
  HostMat nzA;
  const chr_csr Au = drop_zeros(*A, nzA) ? nzA.view() : *A;
  if ((rc = m->A.upload(Au, false)) ||
      (rc = tune_layout(m->A, Au.rowptr, Au.col, Au.val, nullptr, m->zero.as<int>(), false)))
    return rc;
  HIPC(hipDeviceSynchronize());
  *out = m.release();
  return CHR_OK;
}

extern "C" int chr_mat_destroy(chr_mat *m) {
  delete m;
  return CHR_OK;
}

extern "C" int chr_spmv(chr_mat *A, const double *x, double *y, int32_t ptr_kind) {
  if (!A || !x || !y) return fail(CHR_EARG, "null");
  const DevCsr &M = A->A;
  DevBuf dx, dy;
  const double *px = x;
  double *py = y;
  int rc;
  if (ptr_kind == CHR_PTR_HOST) {
    if ((rc = dx.alloc(8 * (size_t)std::max(M.ncols, 1)))) return rc;
    if ((rc = dy.alloc(8 * (size_t)std::max(M.nrows, 1)))) return rc;
    HIPC(hipMemcpy(dx.p, x, 8 * (size_t)M.ncols, hipMemcpyHostToDevice));
    px = dx.as<double>();
    py = dy.as<double>();
  }
  HIPC(launch_spmv(M, px, chr::EpiStore<double>{py}, A->zero.as<int>(), nullptr));
  if (ptr_kind == CHR_PTR_HOST) HIPC(hipMemcpy(y, py, 8 * (size_t)M.nrows, hipMemcpyDeviceToHost));
  HIPC(hipDeviceSynchronize());
  return CHR_OK;
}

// ---------------------------------------------------------------------------------------
// PCG_CJ
// ---------------------------------------------------------------------------------------
namespace {

hipError_t pcg_finalize(int kind, int count, double *part, chr::PcgState *st, double *resvec,
                        hipStream_t s, chr::PcgState *host_out = nullptr) {
  hipLaunchKernelGGL(chr::pcg_finalize_kernel, dim3(1), dim3(chr::kFinBlock), 0, s, kind, count,
                     (const double *)part, st, resvec, host_out);
  return hipGetLastError();
}

template <int PER, typename TI, bool FORMED>
hipError_t launch_pcg_p_f(const DevCsr &A, const PcgPtrs &w, const double *pold, double *pnew,
                          const chr::PcgState *st, hipStream_t s) {
  hipLaunchKernelGGL((chr::pcg_p_stream_kernel<PER, TI, FORMED>), dim3(A.nblk), dim3(chr::kBlock), 0, s,
                     A.nblk, A.blk.as<int4>(), A.rp.as<int>(), A.ci.as<TI>(), A.val.as<double>(),
                     (const double *)w.z, pold, pnew, w.ap, (const double *)w.r, w.part_p, st);
  return hipGetLastError();
}

template <int PER, typename TI>
hipError_t launch_pcg_p_t(const DevCsr &A, const PcgPtrs &w, const double *pold, double *pnew,
                          const chr::PcgState *st, bool formed, hipStream_t s) {
  return formed ? launch_pcg_p_f<PER, TI, true>(A, w, pold, pnew, st, s)
                : launch_pcg_p_f<PER, TI, false>(A, w, pold, pnew, st, s);
}

template <int MAXM>
hipError_t launch_pcg_p_dia(const DevDia &D, const PcgPtrs &w, const double *pold, double *pnew,
                            const chr::PcgState *st, bool formed, hipStream_t s) {
  if (formed)
    hipLaunchKernelGGL((chr::pcg_p_dia_kernel<MAXM, true>), dim3(D.grid()), dim3(chr::kBlock), 0, s, D.n, D.offs,
                       D.val.as<double>(), (const double *)w.z, pold, pnew, w.ap, (const double *)w.r, w.part_p, st);
  else
    hipLaunchKernelGGL((chr::pcg_p_dia_kernel<MAXM, false>), dim3(D.grid()), dim3(chr::kBlock), 0, s, D.n, D.offs,
                       D.val.as<double>(), (const double *)w.z, pold, pnew, w.ap, (const double *)w.r, w.part_p, st);
  return hipGetLastError();
}

// Row blocks of the PCG p-product, each with two partial sums.
int pcg_p_parts(const DevCsr &A) { return A.dia_pcg ? (int)A.dia.grid() : std::max(A.nblk, 1); }

hipError_t launch_pcg_p(const DevCsr &A, const PcgPtrs &w, const double *pold, double *pnew,
                        const chr::PcgState *st, bool formed, hipStream_t s) {
  if (A.dia_pcg) {
    if (A.dia.m <= 4) return launch_pcg_p_dia<4>(A.dia, w, pold, pnew, st, formed, s);
    if (A.dia.m <= 8) return launch_pcg_p_dia<8>(A.dia, w, pold, pnew, st, formed, s);
    return launch_pcg_p_dia<16>(A.dia, w, pold, pnew, st, formed, s);
  }
  switch (A.per * 2 + (A.idx16 ? 1 : 0)) {
    case 8: return launch_pcg_p_t<4, int>(A, w, pold, pnew, st, formed, s);
    case 9: return launch_pcg_p_t<4, uint16_t>(A, w, pold, pnew, st, formed, s);
    case 32: return launch_pcg_p_t<16, int>(A, w, pold, pnew, st, formed, s);
    case 33: return launch_pcg_p_t<16, uint16_t>(A, w, pold, pnew, st, formed, s);
    case 17: return launch_pcg_p_t<8, uint16_t>(A, w, pold, pnew, st, formed, s);
    default: return launch_pcg_p_t<8, int>(A, w, pold, pnew, st, formed, s);
  }
}

// In the stream layout p = z + beta p_old is formed in a pass of its own, so that the
// product gathers one operand per nonzero instead of two (same bits either way; measured
// faster on the 7-point diffusion rows and on the 81-entry elasticity rows alike).  The
// diagonal layout reads its operands as contiguous runs, so it forms p inside the product
// and saves the pass.  CHRONOS_PCG_FORMED=0|1 selects the single-pass or two-pass form.
bool pcg_formed(const DevCsr &A) {
  if (const char *e = std::getenv("CHRONOS_PCG_FORMED")) return std::atoi(e) != 0;
  return !A.dia_pcg;
}

// One PCG_CJ iteration (PCG_CJ.m:88-120), p read from p0 (parity 0) or p1 and written to
// the other.  state_dst (may be null): the device address of host-mapped pinned memory that
// the closing finalisation writes the state to.
hipError_t enqueue_pcg_iter(chr_hier &h, const DevCsr &A, const PcgPtrs &w, int parity,
                            chr::PcgState *state_dst) {
  chr::PcgState *st = w.st;
  const int *stop = &st->stop;
  double *pold = parity ? w.p1 : w.p0;
  double *pnew = parity ? w.p0 : w.p1;
  hipError_t e;
  if ((e = enqueue_vcycle(h, w.r, w.z, stop, 1, w.r32))) return e;
  hipLaunchKernelGGL(chr::pcg_zap_kernel, dim3(w.nblk), dim3(chr::kBlock), 0, h.stream, w.n,
                     (const double *)w.z, (const double *)w.ap, w.part, (const chr::PcgState *)st);
  if ((e = hipGetLastError())) return e;
  if ((e = pcg_finalize(chr::kFinBeta, (int)w.nblk, w.part, st, w.resvec, h.stream))) return e;
  const bool formed = pcg_formed(A);
  if (formed) {
    hipLaunchKernelGGL(chr::pcg_pnew_kernel, dim3(w.nblk), dim3(chr::kBlock), 0, h.stream, w.n,
                       (const double *)w.z, (const double *)pold, pnew, (const chr::PcgState *)st);
    if ((e = hipGetLastError())) return e;
  }
  if ((e = launch_pcg_p(A, w, pold, pnew, st, formed, h.stream))) return e;
  if ((e = pcg_finalize(chr::kFinAlpha, pcg_p_parts(A), w.part_p, st, w.resvec, h.stream))) return e;
  hipLaunchKernelGGL(chr::pcg_update_kernel, dim3(w.nblk), dim3(chr::kBlock), 0, h.stream, w.n,
                     (const double *)pnew, (const double *)w.ap, w.x, w.r, w.r32, w.part,
                     (const chr::PcgState *)st);
  if ((e = hipGetLastError())) return e;
  return pcg_finalize(chr::kFinUpdate, (int)w.nblk, w.part, st, w.resvec, h.stream, state_dst);
}

// r = b - A x, the FP32 copy of r, the norms and the initial scalars (PCG_CJ.m:5-30); the
// state must hold tol and itmax, and p0, p1, Ap must be zero.
hipError_t enqueue_pcg_init(chr_hier &h, const DevCsr &A, const PcgPtrs &w, const double *b) {
  hipError_t e;
  if ((e = launch_spmv(A, (const double *)w.x, chr::EpiResidual<double, double>{b, w.r}, h.zero.as<int>(),
                       h.stream)))
    return e;
  if (w.r32) {
    hipLaunchKernelGGL((chr::convert_kernel<double, float>), dim3((w.n + chr::kBlock - 1) / chr::kBlock),
                       dim3(chr::kBlock), 0, h.stream, w.n, (const double *)w.r, w.r32);
    if ((e = hipGetLastError())) return e;
  }
  hipLaunchKernelGGL(chr::pcg_norms_kernel, dim3(w.nblk), dim3(chr::kBlock), 0, h.stream, w.n, b,
                     (const double *)w.r, w.part);
  if ((e = hipGetLastError())) return e;
  return pcg_finalize(chr::kFinInit, (int)w.nblk, w.part, w.st, w.resvec, h.stream);
}

// Blocks of the PCG vector kernels: enough to fill 256 CUs several times, few enough for the
// single-block finalisation of their partial sums.
unsigned pcg_nblk(int32_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, ((int64_t)n + 1023) / 1024));
}

int pcg_prepare(chr_hier &h, const DevCsr &A, int32_t itmax) {
  if (!h.pcg) h.pcg = std::make_unique<PcgWork>();
  PcgWork &w = *h.pcg;
  const int32_t n = A.nrows;
  int rc;
  if (w.n != n) {
    w.drop_graphs();
    w.n = n;
    const size_t vb = 8 * (size_t)n;
    if ((rc = w.x.alloc(vb)) || (rc = w.r.alloc(vb)) || (rc = w.z.alloc(vb)) ||
        (rc = w.p0.alloc(vb)) || (rc = w.p1.alloc(vb)) || (rc = w.ap.alloc(vb)) ||
        (rc = w.r32.alloc(h.fp32() ? 4 * (size_t)n : 0)))
      return rc;
    w.nblk = pcg_nblk(n);
    if ((rc = w.part.alloc(sizeof(double) * 2 * w.nblk))) return rc;
    if ((rc = w.state.alloc(sizeof(chr::PcgState)))) return rc;
    if (!w.pinned) {
      HIPC(hipHostMalloc((void **)&w.pinned, 2 * sizeof(chr::PcgState), hipHostMallocDefault));
      HIPC(hipHostGetDevicePointer((void **)&w.pinned_dev, w.pinned, 0));
    }
    for (auto &e : w.ev)
      if (!e) HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  // p-SpMV partials: two per row block of the outer operator (its layout may differ from
  // one operator to the next even at equal n).
  const size_t need = 2 * (size_t)pcg_p_parts(A);
  if (w.part_p_cap < need) {
    w.drop_graphs();
    if ((rc = w.part_p.alloc(sizeof(double) * need))) return rc;
    w.part_p_cap = need;
  }
  if (w.itmax_cap < itmax + 1) {
    w.drop_graphs();
    w.itmax_cap = itmax + 1;
    if ((rc = w.resvec.alloc(sizeof(double) * w.itmax_cap))) return rc;
  }
  return CHR_OK;
}

int pcg_graphs(chr_hier &h, const chr_mat *Akey, const DevCsr &A) {
  PcgWork &w = *h.pcg;
  if (w.exec[0] && w.exec[1] && w.graph_uid == Akey->uid) return CHR_OK;
  w.drop_graphs();
  for (int parity = 0; parity < 2; ++parity) {
    if (w.exec[parity]) {
      (void)hipGraphExecDestroy(w.exec[parity]);
      w.exec[parity] = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_pcg_iter(h, A, w.ptrs(h.fp32()), parity, w.pinned_dev + parity);
    hipError_t e2 = hipStreamEndCapture(h.stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(CHR_EHIP, std::string("PCG graph capture: ") +
                                hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&w.exec[parity], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(CHR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  w.graph_uid = Akey->uid;
  return CHR_OK;
}

}  // namespace

extern "C" int chr_pcg_solve(chr_hier *h, chr_mat *Aop, const double *b, double *x,
                             int32_t itmax, double tol, int32_t ptr_kind, int32_t *iter,
                             double *resvec, int32_t *flag) {
  if (!h || !b || !x) return fail(CHR_EARG, "chr_pcg_solve: null argument");
  if (itmax < 0) return fail(CHR_EARG, "chr_pcg_solve: itmax < 0");
  const chr_mat *Akey = Aop ? Aop : &h->outer;
  const DevCsr &A = Akey->A;
  if (A.nrows != h->lev[0].n || A.ncols != A.nrows)
    return fail(CHR_EARG, "chr_pcg_solve: operator size differs from the hierarchy");
  int rc;
  if ((rc = pcg_prepare(*h, A, itmax))) return rc;
  if ((rc = pcg_graphs(*h, Akey, A))) return rc;
  PcgWork &w = *h->pcg;
  const int32_t n = A.nrows;
  const size_t vb = 8 * (size_t)n;
  const hipMemcpyKind in_kind = ptr_kind == CHR_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  // Inputs; p and Ap start at zero so that beta * p_old is defined on the first iteration.
  const double *db = b;
  if (ptr_kind == CHR_PTR_HOST) {
    HIPC(hipMemcpyAsync(h->io_b.p, b, vb, hipMemcpyHostToDevice, h->stream));
    db = h->io_b.as<double>();
  }
  HIPC(hipMemcpyAsync(w.x.p, x, vb, in_kind, h->stream));
  HIPC(hipMemsetAsync(w.p0.p, 0, vb, h->stream));
  HIPC(hipMemsetAsync(w.p1.p, 0, vb, h->stream));
  HIPC(hipMemsetAsync(w.ap.p, 0, vb, h->stream));
  chr::PcgState st0{};
  st0.tol = tol;
  st0.itmax = itmax;
  HIPC(hipMemcpyAsync(w.state.p, &st0, sizeof(st0), hipMemcpyHostToDevice, h->stream));
  HIPC(enqueue_pcg_init(*h, A, w.ptrs(h->fp32()), db));
  HIPC(hipMemcpyAsync(&w.pinned[1], w.state.p, sizeof(chr::PcgState), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  chr::PcgState fin = w.pinned[1];
  if (!fin.stop) {
    // Iterations run back to back; the host waits for iteration it-1 while iteration it is
    // queued, and stops enqueueing once the device reports convergence.  Kernels of an
    // iteration that starts after convergence exit at once.
    for (int32_t it = 1; it <= itmax; ++it) {
      const int parity = (it - 1) & 1;
      HIPC(hipGraphLaunch(w.exec[parity], h->stream));
      HIPC(hipEventRecord(w.ev[parity], h->stream));
      if (it >= 2) {
        HIPC(hipEventSynchronize(w.ev[parity ^ 1]));
        if (w.pinned[parity ^ 1].stop) break;
      }
    }
    HIPC(hipStreamSynchronize(h->stream));
    fin = w.pinned[0].iter >= w.pinned[1].iter ? w.pinned[0] : w.pinned[1];
  }
  if (fin.bnorm == 0.0) {
    HIPC(hipMemsetAsync(w.x.p, 0, vb, h->stream));
  }
  const hipMemcpyKind out_kind = ptr_kind == CHR_PTR_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIPC(hipMemcpyAsync(x, w.x.p, vb, out_kind, h->stream));
  if (resvec) HIPC(hipMemcpyAsync(resvec, w.resvec.p, sizeof(double) * (fin.iter + 1), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  if (iter) *iter = fin.iter;
  if (flag) *flag = fin.resiter > tol ? 1 : 0;
  if (fin.breakdown) return fail(CHR_EBREAKDOWN, "PCG breakdown: p^T A p is zero or not finite");
  return CHR_OK;
}

// ---------------------------------------------------------------------------------------
// PCG_CJ without allocation (chr_pcg_solve_noalloc)
// ---------------------------------------------------------------------------------------
// The solver's working vectors (r, z, p0, p1, Ap, and the FP32 copy of r), its partial sums,
// residual history and state live in one fixed device array per GPU instead of buffers the
// library allocates.  In host code `scratch` names a host-side stand-in: its device address
// is looked up once per device with hipGetSymbolAddress (chr_init, or the first call on a
// device) and only that address is ever passed to kernels.
namespace chrrt {

constexpr size_t kScratchBytes = CHR_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch array.
int scratch_setup(void **out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CHR_EARG, "scratch: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CHR_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return CHR_OK;
}

template <typename T>
__global__ __launch_bounds__(chr::kBlock) void fill_kernel(int64_t n, T v, T *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * chr::kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * chr::kBlock)
    out[i] = v;
}

__global__ void pcg_state_init_kernel(chr::PcgState *st, double tol, int itmax) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    chr::PcgState z{};
    z.tol = tol;
    z.itmax = itmax;
    *st = z;
  }
}

template <typename T>
hipError_t fill(T *p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>(4096, (n + chr::kBlock - 1) / chr::kBlock);
  hipLaunchKernelGGL(fill_kernel<T>, dim3(grid), dim3(chr::kBlock), 0, s, n, v, p);
  return hipGetLastError();
}

}  // namespace chrrt

extern "C" int chr_pcg_solve_noalloc(chr_hier *h, chr_mat *Aop, const double *b, double *x, int32_t itmax,
                                     double tol, int32_t *iter, double *resvec, int32_t *flag) {
  if (!h || !b || !x) return fail(CHR_EARG, "chr_pcg_solve_noalloc: null argument");
  if (itmax < 0) return fail(CHR_EARG, "chr_pcg_solve_noalloc: itmax < 0");
  const DevCsr &A = (Aop ? Aop : &h->outer)->A;
  if (A.nrows != h->lev[0].n || A.ncols != A.nrows)
    return fail(CHR_EARG, "chr_pcg_solve_noalloc: operator size differs from the hierarchy");
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev != h->device) return fail(CHR_EARG, "chr_pcg_solve_noalloc: call it on the hierarchy's device");
  void *base = nullptr;
  int rc;
  if ((rc = scratch_setup(&base))) return rc;
  // Carve the scratch array: 256-byte aligned pieces, every size checked before any work.
  const int32_t n = A.nrows;
  const unsigned nblk = pcg_nblk(n);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) / 256 * 256;
    return at;
  };
  const size_t vb = 8 * (size_t)n;
  const size_t o_r = take(vb), o_z = take(vb), o_p0 = take(vb), o_p1 = take(vb), o_ap = take(vb);
  const size_t o_r32 = take(h->fp32() ? 4 * (size_t)n : 0);
  const size_t o_part = take(sizeof(double) * 2 * nblk);
  const size_t o_partp = take(sizeof(double) * 2 * (size_t)pcg_p_parts(A));
  const size_t o_res = take(sizeof(double) * ((size_t)itmax + 1));
  const size_t o_st = take(sizeof(chr::PcgState));
  if (off > kScratchBytes)
    return fail(CHR_EARG, "chr_pcg_solve_noalloc: the solve needs " + std::to_string(off) +
                              " bytes of scratch, more than its " + std::to_string(kScratchBytes) +
                              " (5 n doubles, 4 n bytes in the FP32 cycle, partial sums, itmax + 1 doubles)");
  char *sb = static_cast<char *>(base);
  PcgPtrs w;
  w.n = n;
  w.nblk = nblk;
  w.x = x;  // iterated in place
  w.r = reinterpret_cast<double *>(sb + o_r);
  w.z = reinterpret_cast<double *>(sb + o_z);
  w.p0 = reinterpret_cast<double *>(sb + o_p0);
  w.p1 = reinterpret_cast<double *>(sb + o_p1);
  w.ap = reinterpret_cast<double *>(sb + o_ap);
  w.r32 = h->fp32() ? reinterpret_cast<float *>(sb + o_r32) : nullptr;
  w.part = reinterpret_cast<double *>(sb + o_part);
  w.part_p = reinterpret_cast<double *>(sb + o_partp);
  w.resvec = reinterpret_cast<double *>(sb + o_res);
  w.st = reinterpret_cast<chr::PcgState *>(sb + o_st);
  hipStream_t s = h->stream;
  // The array keeps the previous call's contents: p0, p1, Ap and the state start from zero.
  HIPC(fill(w.p0, n, 0.0, s));
  HIPC(fill(w.p1, n, 0.0, s));
  HIPC(fill(w.ap, n, 0.0, s));
  hipLaunchKernelGGL(pcg_state_init_kernel, dim3(1), dim3(64), 0, s, w.st, tol, itmax);
  HIPC(hipGetLastError());
  HIPC(enqueue_pcg_init(*h, A, w, b));
  chr::PcgState fin{};
  HIPC(hipMemcpyAsync(&fin, w.st, sizeof(fin), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  // One iteration at a time, the state read back after each (kernels of an iteration that
  // starts after convergence would exit at once; none is enqueued).
  for (int32_t it = 1; it <= itmax && !fin.stop; ++it) {
    HIPC(enqueue_pcg_iter(*h, A, w, (it - 1) & 1, nullptr));
    HIPC(hipMemcpyAsync(&fin, w.st, sizeof(fin), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
  }
  if (fin.bnorm == 0.0) HIPC(fill(x, n, 0.0, s));
  if (resvec) HIPC(hipMemcpyAsync(resvec, w.resvec, sizeof(double) * (fin.iter + 1), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (iter) *iter = fin.iter;
  if (flag) *flag = fin.resiter > tol ? 1 : 0;
  if (fin.breakdown) return fail(CHR_EBREAKDOWN, "PCG breakdown: p^T A p is zero or not finite");
  return CHR_OK;
}

// ---------------------------------------------------------------------------------------
// Adjoint gradient (finite-volume diffusion)
// ---------------------------------------------------------------------------------------
extern "C" int chr_fv_grad(chr_hier *h, chr_mat *Aop, const double *k, const double *x,
                           const double *lam, double *g, int32_t ptr_kind) {
  if ((!h && !Aop) || !k || !x || !lam || !g) return fail(CHR_EARG, "chr_fv_grad: null argument");
  const DevCsr &A = Aop ? Aop->A : h->outer.A;
  if (A.fp32) return fail(CHR_EARG, "chr_fv_grad: needs the FP64 operator");
  const int32_t n = A.nrows;
  hipStream_t s = h ? h->stream : nullptr;
  DevBuf dk, dx, dl, dg;
  const double *pk = k, *px = x, *pl = lam;
  double *pg = g;
  int rc;
  if (ptr_kind == CHR_PTR_HOST) {
    const size_t vb = 8 * (size_t)n;
    if ((rc = dk.alloc(vb)) || (rc = dx.alloc(vb)) || (rc = dl.alloc(vb)) || (rc = dg.alloc(vb))) return rc;
    HIPC(hipMemcpyAsync(dk.p, k, vb, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx.p, x, vb, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dl.p, lam, vb, hipMemcpyHostToDevice, s));
    pk = dk.as<double>();
    px = dx.as<double>();
    pl = dl.as<double>();
    pg = dg.as<double>();
  }
  if (A.idx16)
    hipLaunchKernelGGL(chr::fv_grad_kernel<uint16_t>, dim3(A.nblk), dim3(chr::kBlock), 0, s, A.nblk,
                       A.blk.as<int4>(), A.rp.as<int>(), A.ci.as<uint16_t>(), A.val.as<double>(), pk, px, pl, pg);
  else
    hipLaunchKernelGGL(chr::fv_grad_kernel<int>, dim3(A.nblk), dim3(chr::kBlock), 0, s, A.nblk,
                       A.blk.as<int4>(), A.rp.as<int>(), A.ci.as<int>(), A.val.as<double>(), pk, px, pl, pg);
  HIPC(hipGetLastError());
  if (ptr_kind == CHR_PTR_HOST) HIPC(hipMemcpyAsync(g, pg, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  return CHR_OK;
}

// ---------------------------------------------------------------------------------------
// Timing helpers
// ---------------------------------------------------------------------------------------
extern "C" int chr_time_vcycle(chr_hier *h, int32_t reps, double *ms) {
  if (!h || reps < 1 || !ms) return fail(CHR_EARG, "chr_time_vcycle: arguments");
  const int32_t n = h->lev[0].n;
  std::vector<double> ones(n, 1.0);
  HIPC(hipMemcpy(h->io_b.p, ones.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
  // The FP32 cycle reads an FP32 copy of its input in the first smoother product, as the
  // solver does with its residual.
  DevBuf b32;
  int rc;
  if (h->fp32()) {
    std::vector<float> f(n, 1.0f);
    if ((rc = b32.alloc(4 * (size_t)n))) return rc;
    HIPC(hipMemcpy(b32.p, f.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
  }
  // Up to kPerGraph cycles back to back in one graph (as the solver's iteration graph holds
  // its cycle), the graph launched until reps cycles have run.
  constexpr int kPerGraph = 10;
  const int per = std::min(reps, kPerGraph), launches = (reps + per - 1) / per;
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int c = 0; c < per && !e; ++c)
    e = enqueue_vcycle(*h, h->io_b.as<double>(), h->io_x.as<double>(), h->zero.as<int>(), 1, b32.as<float>());
  hipError_t e2 = hipStreamEndCapture(h->stream, &g);
  if (e || e2) {
    if (g) (void)hipGraphDestroy(g);
    return fail(CHR_EHIP, "V-cycle capture failed");
  }
  HIPC(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  (void)hipGraphDestroy(g);
  hipEvent_t a, b;
  HIPC(hipEventCreate(&a));
  HIPC(hipEventCreate(&b));
  HIPC(hipGraphLaunch(ex, h->stream));
  HIPC(hipEventRecord(a, h->stream));
  for (int i = 0; i < launches; ++i) HIPC(hipGraphLaunch(ex, h->stream));
  HIPC(hipEventRecord(b, h->stream));
  HIPC(hipEventSynchronize(b));
  float t = 0.f;
  HIPC(hipEventElapsedTime(&t, a, b));
  *ms = t / ((double)launches * per);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipGraphExecDestroy(ex);
  return CHR_OK;
}

extern "C" int chr_time_spmv(chr_hier *h, int32_t level, char which, int32_t reps, double *ms,
                             double *bytes) {
  if (!h || level < 0 || level >= (int32_t)h->lev.size() || h->lev[level].nd == 0 || reps < 1 || !ms)
    return fail(CHR_EARG, "chr_time_spmv: arguments");
  auto &L = h->lev[level];
  const DevCsr *M = nullptr;
  switch (which) {
    case 'A': M = &L.A; break;
    case 'G': M = &L.G; break;
    case 'T': M = &L.GT; break;
    case 'P': M = &L.P; break;
    case 'R': M = &L.PT; break;
    default: return fail(CHR_EARG, "chr_time_spmv: which must be A, G, T, P or R");
  }
  // x: a level vector long enough for M's columns; y: another one.
  const bool f32 = h->fp32();
  const double s = f32 ? 4.0 : 8.0;
  DevBuf xb, yb;
  int rc;
  if ((rc = xb.alloc((size_t)s * std::max(M->ncols, 1))) || (rc = yb.alloc((size_t)s * std::max(M->nrows, 1))))
    return rc;
  HIPC(hipMemset(xb.p, 0, xb.bytes));
  hipEvent_t a, b;
  HIPC(hipEventCreate(&a));
  HIPC(hipEventCreate(&b));
  auto probe = [&](auto tag) -> hipError_t {
    constexpr char W = decltype(tag)::value;
    if (f32)
      return launch_spmv(*M, xb.as<float>(), chr::EpiProbe<float, W>{{yb.as<float>()}}, h->zero.as<int>(),
                         h->stream);
    return launch_spmv(*M, xb.as<double>(), chr::EpiProbe<double, W>{{yb.as<double>()}}, h->zero.as<int>(),
                       h->stream);
  };
  auto run = [&]() -> hipError_t {
    switch (which) {
      case 'A': return probe(std::integral_constant<char, 'A'>{});
      case 'G': return probe(std::integral_constant<char, 'G'>{});
      case 'T': return probe(std::integral_constant<char, 'T'>{});
      case 'P': return probe(std::integral_constant<char, 'P'>{});
      default: return probe(std::integral_constant<char, 'R'>{});
    }
  };
  // Every timed launch starts from a cold Infinity Cache: a 512 MiB device copy (1 GiB of
  // traffic, four times the 256 MB cache) runs before it, as the other products of a V-cycle
  // do in a solve.  Without it a matrix of < 256 MB would be re-read from the cache.
  DevBuf flush;
  constexpr size_t kFlushHalf = size_t(512) << 20;
  if ((rc = flush.alloc(2 * kFlushHalf))) return rc;
  HIPC(hipMemsetAsync(flush.p, 0, flush.bytes, h->stream));
  HIPC(run());
  double total = 0.0;
  for (int i = 0; i < reps; ++i) {
    HIPC(hipMemcpyAsync(flush.as<char>() + kFlushHalf, flush.p, kFlushHalf, hipMemcpyDeviceToDevice,
                        h->stream));
    HIPC(hipEventRecord(a, h->stream));
    HIPC(run());
    HIPC(hipEventRecord(b, h->stream));
    HIPC(hipEventSynchronize(b));
    float t = 0.f;
    HIPC(hipEventElapsedTime(&t, a, b));
    total += t;
  }
  *ms = total / reps;
  if (bytes) *bytes = M->csr_bytes() + s * (M->ncols + M->nrows);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return CHR_OK;
}

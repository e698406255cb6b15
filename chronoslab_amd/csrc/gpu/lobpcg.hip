// LOBPCG (sources/EigenSolver/lobpcg.m:21-467 with B_orthonormalize.m:1-25, prodB = [] and
// an optional constraint block Y, as EigDriver.m:202-226 calls it) on the GPU.
//
// The n x k blocks live in HBM as row-major FP64 blocks.  Products with A are block SpMMs,
// the preconditioner is the block V-cycle, and the Gram matrices are tall-skinny products
// reduced in a fixed order.  The Rayleigh-Ritz steps (order <= 3*neig) run in FP64 on the
// host: Cholesky, the generalized symmetric-definite eigenproblem and condition numbers.
#include <cstdlib>
#include <cstring>
#include <utility>

#include "chronos_amg.h"
#include "dense.hpp"
#include "runtime.hpp"

using namespace chrrt;

namespace {

struct Lob {
  chr_hier *h;       // preconditioner (may be null)
  const DevCsr *A;   // prodA
  const int *zero;   // stop flag that is never set
  hipStream_t s;
  int n;
  int wmax = 0;  // columns every n-row work block holds (the block size)
  DevBuf part, gout, stage_dev;
  int nparts = 1, rows_per_wave = 4;
  // Small host operands (Rayleigh-Ritz coefficients, eigenvalues, column lists) go through a
  // pinned staging area, each piece copied asynchronously to its own slot of stage_dev.
  // Slots are reused only after a stream synchronisation (fetch), so no pending copy ever
  // reads a slot the host is rewriting.  Gram matrices are reduced into slots of gout and
  // brought back together by one copy and one synchronisation per batch.
  static constexpr size_t kStage = 16384;  // doubles
  static constexpr int kSlots = 32, kSlot = 256;
  double *hstage = nullptr, *hout = nullptr;
  size_t staged = 0;
  int nslot = 0;

  ~Lob() {
    if (hstage) (void)hipHostFree(hstage);
    if (hout) (void)hipHostFree(hout);
  }
  int init() {
    // Gram partials: nparts workgroups of four waves, each wave a run of rows_per_wave rows
    // (a multiple of 4, the rows of one matrix instruction).
    nparts = std::max(1, std::min(1024, (n + 255) / 256));
    const int waves = nparts * (chr::kBlock / chr::kWave);
    rows_per_wave = ((n + waves - 1) / waves + 3) / 4 * 4;
    int rc;
    if ((rc = part.alloc(sizeof(double) * (size_t)nparts * chr::kGramMaxOut))) return rc;
    if ((rc = gout.alloc(sizeof(double) * kSlots * kSlot))) return rc;
    if ((rc = stage_dev.alloc(sizeof(double) * kStage))) return rc;
    HIPC(hipHostMalloc((void **)&hstage, sizeof(double) * kStage, hipHostMallocDefault));
    HIPC(hipHostMalloc((void **)&hout, sizeof(double) * kSlots * kSlot, hipHostMallocDefault));
    return CHR_OK;
  }
  static unsigned grid(int64_t cnt) { return (unsigned)((cnt + chr::kBlock - 1) / chr::kBlock); }
  // Workgroups of the 16-row-tile kernels: one tile per wave, at most 2048 workgroups (the
  // waves then stride over the tiles).
  unsigned tile_grid() const {
    const int64_t tiles = ((int64_t)n + 15) / 16, per = chr::kBlock / chr::kWave;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (tiles + per - 1) / per));
  }

  // Copies bytes from src to a fresh staging slot and returns its device address (null on error).
  const void *stage(const void *src, size_t bytes) {
    const size_t nd = (bytes + 31) / 32 * 4;  // whole 32-byte units
    if (nd > kStage) return nullptr;
    if (staged + nd > kStage) {
      if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
      staged = 0;
    }
    std::memcpy(hstage + staged, src, bytes);
    void *dst = stage_dev.as<double>() + staged;
    if (hipMemcpyAsync(dst, hstage + staged, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return nullptr;
    staged += nd;
    return dst;
  }
  // Enqueues X^T Y (row-major n x a and n x b blocks) into the next Gram slot; returns the
  // slot, or -1 on error.
  int gram_launch(const double *X, int a, const double *Y, int b) {
    if (a < 1 || b < 1 || a > 16 || b > 16 || nslot >= kSlots) {
      fail(CHR_EARG, "lobpcg: Gram block wider than 16 or too many pending Gram matrices");
      return -1;
    }
    chr::GramSet gs{};
    gs.blk[0] = X;
    gs.blk[1] = Y;
    gs.w[0] = a;
    gs.w[1] = b;
    gs.po[0] = 0;
    gs.po[1] = a * b;
    hipLaunchKernelGGL(chr::gram_mfma_kernel<0>, dim3(nparts), dim3(chr::kBlock), 0, s, n, gs, rows_per_wave,
                       part.as<double>());
    const unsigned rgrid = (unsigned)((a * b * chr::kWave + chr::kBlock - 1) / chr::kBlock);
    hipLaunchKernelGGL(chr::gram_reduce_kernel, dim3(rgrid), dim3(chr::kBlock), 0, s, nparts, a * b,
                       (const double *)part.as<double>(), gout.as<double>() + (size_t)nslot * kSlot);
    if (hipGetLastError() != hipSuccess) {
      fail(CHR_EHIP, "lobpcg: Gram launch");
      return -1;
    }
    return nslot++;
  }
  // Enqueues the Gram products X_k^T Y_k of pairs (X_k, a_k, Y_k, b_k) in one pass over the
  // distinct blocks (gram_mfma_kernel; the pairs must form one of its Rayleigh-Ritz tables),
  // each entry equal to the one gram_launch would give.  Returns the first slot, or -1 on
  // error; product k starts offs[k] entries after that slot's start, row-major a_k x b_k.
  struct GramPair {
    const double *X;
    int a;
    const double *Y;
    int b;
  };
  int gram_multi_launch(const std::vector<GramPair> &pairs, std::vector<int> &offs) {
    chr::GramSet gs{};
    int nblk = 0;
    auto id = [&](const double *p, int w) -> int {
      for (int b = 0; b < nblk; ++b)
        if (gs.blk[b] == p) return gs.w[b] == w ? b : -1;
      if (nblk == chr::kGramMaxBlocks || w < 1 || w > 16) return -1;
      gs.blk[nblk] = p;
      gs.w[nblk] = w;
      return nblk++;
    };
    if (pairs.empty() || (int)pairs.size() > chr::kGramMaxPairs) {
      fail(CHR_EARG, "lobpcg: Gram batch size");
      return -1;
    }
    offs.clear();
    const int np = (int)pairs.size();
    const int table = np == chr::gram_npairs(1) ? 1 : np == chr::gram_npairs(2) ? 2 : -1;
    int total = 0;
    for (int k = 0; k < np; ++k) {
      const int bx = id(pairs[k].X, pairs[k].a), by = id(pairs[k].Y, pairs[k].b);
      if (bx < 0 || by < 0 || table < 0 || bx != chr::gram_pi(table, k) || by != chr::gram_pj(table, k)) {
        fail(CHR_EARG, "lobpcg: Gram batch is not a Rayleigh-Ritz table of blocks of width <= 16");
        return -1;
      }
      gs.po[k] = total;
      offs.push_back(total);
      total += pairs[k].a * pairs[k].b;
    }
    gs.po[np] = total;
    const int slots = (total + kSlot - 1) / kSlot;
    if (total > chr::kGramMaxOut || nslot + slots > kSlots) {
      fail(CHR_EARG, "lobpcg: Gram batch too large");
      return -1;
    }
    if (table == 1)
      hipLaunchKernelGGL(chr::gram_mfma_kernel<1>, dim3(nparts), dim3(chr::kBlock), 0, s, n, gs, rows_per_wave,
                         part.as<double>());
    else
      hipLaunchKernelGGL(chr::gram_mfma_kernel<2>, dim3(nparts), dim3(chr::kBlock), 0, s, n, gs, rows_per_wave,
                         part.as<double>());
    const unsigned rgrid = (unsigned)(((int64_t)total * chr::kWave + chr::kBlock - 1) / chr::kBlock);
    hipLaunchKernelGGL(chr::gram_reduce_kernel, dim3(rgrid), dim3(chr::kBlock), 0, s, nparts, total,
                       (const double *)part.as<double>(), gout.as<double>() + (size_t)nslot * kSlot);
    if (hipGetLastError() != hipSuccess) {
      fail(CHR_EHIP, "lobpcg: Gram batch launch");
      return -1;
    }
    const int base = nslot;
    nslot += slots;
    return base;
  }
  // The Rayleigh-Ritz Gram products with the narrow blocks packed side by side (tables 4
  // and 5 of gram_mfma_kernel; na <= 8): operands X and AX (k columns) and narrow = {AR,
  // AAR} after a restart, {AR, AAR, APa, AAP} otherwise.  Records where each operand sits
  // (block and first column) for get_packed.  Returns the first slot, or -1 on error.
  struct Packed {
    int table = 0, np = 0;
    const double *ptr[6] = {};
    int blk[6] = {}, off[6] = {}, nptr = 0;
    int w[4] = {}, po[8] = {};
  };
  int gram_packed_launch(const double *X, const double *AX, const std::vector<const double *> &narrow, int k,
                         int na, Packed &pk) {
    const int nn = (int)narrow.size();
    if (k < 1 || k > 16 || na < 1 || na > 8 || (nn != 2 && nn != 4)) {
      fail(CHR_EARG, "lobpcg: packed Gram operands");
      return -1;
    }
    chr::GramSet gs{};
    pk = Packed();
    gs.blk[0] = X;
    gs.blk[1] = AX;
    gs.w[0] = gs.w[1] = k;
    gs.na = na;
    gs.sw = na <= 4 ? 4 : 8;
    pk.ptr[0] = X;
    pk.ptr[1] = AX;
    pk.blk[1] = 1;
    pk.nptr = 2;
    const int per_blk = 16 / gs.sw;  // segments a packed block holds
    const int npk = (nn + per_blk - 1) / per_blk;
    pk.table = npk == 1 ? 4 : 5;
    for (int s = 0; s < nn; ++s) {
      const int b = s / per_blk, sg = s % per_blk;
      gs.seg[b][sg] = narrow[s];
      gs.nseg[b] = sg + 1;
      pk.ptr[pk.nptr] = narrow[s];
      pk.blk[pk.nptr] = 2 + b;
      pk.off[pk.nptr++] = sg * gs.sw;
    }
    for (int b = 0; b < npk; ++b) {
      gs.blk[2 + b] = narrow[(size_t)b * per_blk];  // not read: packed blocks load through seg
      gs.w[2 + b] = gs.nseg[b] * gs.sw;
    }
    pk.np = chr::gram_npairs(pk.table);
    int total = 0;
    for (int q = 0; q < pk.np; ++q) {
      gs.po[q] = pk.po[q] = total;
      total += gs.w[chr::gram_pi(pk.table, q)] * gs.w[chr::gram_pj(pk.table, q)];
    }
    gs.po[pk.np] = pk.po[pk.np] = total;
    for (int b = 0; b < 4; ++b) pk.w[b] = gs.w[b];
    const int slots = (total + kSlot - 1) / kSlot;
    if (total > chr::kGramMaxOut || nslot + slots > kSlots) {
      fail(CHR_EARG, "lobpcg: Gram batch too large");
      return -1;
    }
    if (pk.table == 4)
      hipLaunchKernelGGL(chr::gram_mfma_kernel<4>, dim3(nparts), dim3(chr::kBlock), 0, s, n, gs, rows_per_wave,
                         part.as<double>());
    else
      hipLaunchKernelGGL(chr::gram_mfma_kernel<5>, dim3(nparts), dim3(chr::kBlock), 0, s, n, gs, rows_per_wave,
                         part.as<double>());
    const unsigned rgrid = (unsigned)(((int64_t)total * chr::kWave + chr::kBlock - 1) / chr::kBlock);
    hipLaunchKernelGGL(chr::gram_reduce_kernel, dim3(rgrid), dim3(chr::kBlock), 0, s, nparts, total,
                       (const double *)part.as<double>(), gout.as<double>() + (size_t)nslot * kSlot);
    if (hipGetLastError() != hipSuccess) {
      fail(CHR_EHIP, "lobpcg: packed Gram launch");
      return -1;
    }
    const int base = nslot;
    nslot += slots;
    return base;
  }
  // out (column-major a x b) = Xp^T Yp from the fetched packed products at slot base.
  bool get_packed(int base, const Packed &pk, const double *Xp, int a, const double *Yp, int b,
                  std::vector<double> &out) const {
    int ix = -1, iy = -1;
    for (int i = 0; i < pk.nptr; ++i) {
      if (pk.ptr[i] == Xp) ix = i;
      if (pk.ptr[i] == Yp) iy = i;
    }
    if (ix < 0 || iy < 0) return false;
    const int u = pk.blk[ix], v = pk.blk[iy];
    for (int q = 0; q < pk.np; ++q)
      if (chr::gram_pi(pk.table, q) == u && chr::gram_pj(pk.table, q) == v) {
        const double *g = hout + (size_t)base * kSlot + pk.po[q];
        out.assign((size_t)a * b, 0.0);
        for (int i = 0; i < a; ++i)
          for (int j = 0; j < b; ++j) out[i + (size_t)j * a] = g[(size_t)(pk.off[ix] + i) * pk.w[v] + pk.off[iy] + j];
        return true;
      }
    return false;
  }
  // out (column-major a x b) = the fetched product at offset off of slot base.
  void get_at(int base, int off, int a, int b, std::vector<double> &out) const {
    const double *g = hout + (size_t)base * kSlot + off;
    out.assign((size_t)a * b, 0.0);
    for (int p = 0; p < a; ++p)
      for (int q = 0; q < b; ++q) out[p + (size_t)q * a] = g[(size_t)p * b + q];
  }
  // Brings the pending Gram slots to the host (one copy, one synchronisation).
  int fetch() {
    if (nslot > 0)
      HIPC(hipMemcpyAsync(hout, gout.p, sizeof(double) * (size_t)nslot * kSlot, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    staged = 0;
    nslot = 0;
    return CHR_OK;
  }
  // out (column-major a x b) = a fetched slot.
  void get(int slot, int a, int b, std::vector<double> &out) const {
    const double *g = hout + (size_t)slot * kSlot;
    out.assign((size_t)a * b, 0.0);
    for (int p = 0; p < a; ++p)
      for (int q = 0; q < b; ++q) out[p + (size_t)q * a] = g[(size_t)p * b + q];
  }
  // out (column-major a x b, host) = X^T Y.
  int gram(const double *X, int a, const double *Y, int b, std::vector<double> &out) {
    const int slot = gram_launch(X, a, Y, b);
    if (slot < 0) return CHR_EARG;
    if (int rc = fetch()) return rc;
    get(slot, a, b, out);
    return CHR_OK;
  }
  // Z (n x b) = beta Z + X (n x a) Cm (column-major a x b, host).
  int gemm(const double *X, int a, const double *Cm, int b, double beta, double *Z) {
    if (a < 1 || b < 1 || a > 16 || b > 16) return fail(CHR_EARG, "lobpcg: block product wider than 16");
    const double *cdev = static_cast<const double *>(stage(Cm, sizeof(double) * a * b));
    if (!cdev) return fail(CHR_EHIP, "lobpcg: staging a coefficient block");
    hipLaunchKernelGGL((chr::tsgemm_mfma_kernel<4, 1>), dim3(tile_grid()), dim3(chr::kBlock), 0, s, n, a, X, b,
                       cdev, 1, a, 1.0, beta, Z);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  // Z (n x b) = beta Z + X (n x a) (cscale G), G the a x b Gram product in the device slot
  // (row-major, as gram_launch leaves it): no host round trip.
  int gemm_slot(const double *X, int a, int slot, int b, double cscale, double beta, double *Z) {
    if (a < 1 || b < 1 || a > 16 || b > 16) return fail(CHR_EARG, "lobpcg: block product wider than 16");
    hipLaunchKernelGGL((chr::tsgemm_mfma_kernel<4, 1>), dim3(tile_grid()), dim3(chr::kBlock), 0, s, n, a, X, b,
                       (const double *)(gout.as<double>() + (size_t)slot * kSlot), b, 1, cscale, beta, Z);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  int applyA(const double *X, int k, double *Y) {
    HIPC(launch_op(*A, X, k, chr::EpiStore<double>{Y}, zero, s));
    return CHR_OK;
  }
  // X <- M X (block V-cycle) into the scratch block tmp, whose buffer then becomes X's
  // (the two pointers swap instead of a copy back).  A block of 9-11 or 13-15 columns is
  // widened with zero columns to the next multiple of 4 for the cycle (into tmp, the cycle's
  // output in the hierarchy's staging block), where the block products take four columns per
  // lane with 16-byte loads: 96^3 Poisson, FP32 cycle, 1.43 -> 1.22 ms at 9 columns, 1.59 ->
  // 1.45 ms at 13.  Each column of a block cycle is that column's own cycle bit for bit, so
  // the padding changes no result.
  int applyM(double *&X, int k, double *&tmp) {
    if (!h) return CHR_OK;
    const int kp = (k + 3) / 4 * 4;
    if (k > 8 && kp != k && kp <= wmax && kp <= h->kcap) {
      const unsigned gp = (unsigned)(((int64_t)n * kp + chr::kBlock - 1) / chr::kBlock);
      const unsigned gk = (unsigned)(((int64_t)n * k + chr::kBlock - 1) / chr::kBlock);
      double *wide_out = h->io_x.as<double>();
      hipLaunchKernelGGL(chr::block_pad_kernel, dim3(gp), dim3(chr::kBlock), 0, s, n, k, kp, (const double *)X, tmp);
      HIPC(hipGetLastError());
      HIPC(enqueue_vcycle(*h, tmp, wide_out, h->zero.as<int>(), kp));
      hipLaunchKernelGGL(chr::block_unpad_kernel, dim3(gk), dim3(chr::kBlock), 0, s, n, k, kp,
                         (const double *)wide_out, tmp);
      HIPC(hipGetLastError());
    } else {
      HIPC(enqueue_vcycle(*h, X, tmp, h->zero.as<int>(), k));
    }
    std::swap(X, tmp);
    return CHR_OK;
  }
  // Pout = R ER (+ Pa EP), X <- Pout + X EX in one pass (lob_update_mfma_kernel); Pa null on a
  // restart.  ER, EP column-major na x k, EX k x k (host).
  // The update coefficients [ER | EP | EX] (column-major na x k, na x k, k x k) staged as one
  // block for both sides of the update; null on error.
  const double *stage_update(const double *ER, const double *EP, const double *EX, int na, int k) {
    coef.resize((size_t)(2 * na + k) * k);
    std::copy(ER, ER + (size_t)na * k, coef.begin());
    std::copy(EP, EP + (size_t)na * k, coef.begin() + (size_t)na * k);
    std::copy(EX, EX + (size_t)k * k, coef.begin() + (size_t)2 * na * k);
    return static_cast<const double *>(stage(coef.data(), sizeof(double) * coef.size()));
  }
  std::vector<double> coef;
  // Pout = R ER (+ Pa EP), Xout = Pout + X EX in one pass (lob_update_mfma_kernel; Xout may
  // be X); Pa null on a restart.  cdev: the coefficients from stage_update.
  int update(const double *R, const double *Pa, int na, const double *X, int k, const double *cdev, double *Pout,
             double *Xout) {
    if (na < 1 || k < 1 || na > 16 || k > 16) return fail(CHR_EARG, "lobpcg: update block wider than 16");
    if (!cdev) return fail(CHR_EHIP, "lobpcg: staging the update coefficients");
    const double *er = cdev, *ep = cdev + (size_t)na * k, *ex = cdev + (size_t)2 * na * k;
    hipLaunchKernelGGL(chr::lob_update_mfma_kernel, dim3(tile_grid()), dim3(chr::kBlock), 0, s, n, na, k, R, Pa,
                       X, er, Pa ? ep : nullptr, ex, Pout, Xout);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  // Enqueues the Gram product R^T R of the residual block R = AX - X diag(lam) (n x k, never
  // stored) into the next slot; returns the slot, or -1 on error.
  int resid_norms(const double *AX, const double *X, const double *lam_host, int k) {
    const double *ldev = static_cast<const double *>(stage(lam_host, sizeof(double) * k));
    if (!ldev || k < 1 || k > 16 || nslot >= kSlots) {
      fail(CHR_EARG, "lobpcg: residual norms");
      return -1;
    }
    chr::GramSet gs{};
    gs.blk[0] = AX;
    gs.blk[1] = X;
    gs.w[0] = gs.w[1] = k;
    gs.po[0] = 0;
    gs.po[1] = k * k;
    gs.lam = ldev;
    hipLaunchKernelGGL(chr::gram_mfma_kernel<3>, dim3(nparts), dim3(chr::kBlock), 0, s, n, gs, rows_per_wave,
                       part.as<double>());
    const unsigned rgrid = (unsigned)((k * k * chr::kWave + chr::kBlock - 1) / chr::kBlock);
    hipLaunchKernelGGL(chr::gram_reduce_kernel, dim3(rgrid), dim3(chr::kBlock), 0, s, nparts, k * k,
                       (const double *)part.as<double>(), gout.as<double>() + (size_t)nslot * kSlot);
    if (hipGetLastError() != hipSuccess) {
      fail(CHR_EHIP, "lobpcg: residual norms launch");
      return -1;
    }
    return nslot++;
  }
  // Y (n x m) = the columns cols of the residual block AX - X diag(lam) (n x k).
  int resid_gather(const double *AX, const double *X, const double *lam_host, int k, const std::vector<int> &cols,
                   double *Y) {
    const int m = (int)cols.size();
    std::vector<double> st((size_t)k + ((size_t)m + 1) / 2);
    std::copy(lam_host, lam_host + k, st.begin());
    std::memcpy(st.data() + k, cols.data(), sizeof(int) * m);
    const double *ldev = static_cast<const double *>(stage(st.data(), sizeof(double) * st.size()));
    if (!ldev) return fail(CHR_EHIP, "lobpcg: staging eigenvalues and a column list");
    hipLaunchKernelGGL(chr::block_resid_gather_kernel, dim3(grid((int64_t)n * m)), dim3(chr::kBlock), 0, s, n, k,
                       AX, X, ldev, m, reinterpret_cast<const int *>(ldev + k), Y);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  // Y (n x m) = the columns idev[0..m) of X (n x k); idev staged by stage().
  int gather(const double *X, int k, const int *idev, int m, double *Y) {
    if (!idev) return fail(CHR_EHIP, "lobpcg: staging a column list");
    hipLaunchKernelGGL(chr::block_gather_kernel, dim3(grid((int64_t)n * m)), dim3(chr::kBlock), 0, s, n, k,
                       X, m, idev, Y);
    HIPC(hipGetLastError());
    return CHR_OK;
  }
  int copy(double *dst, const double *src, int k) {
    HIPC(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n * k, hipMemcpyDeviceToDevice, s));
    return CHR_OK;
  }
  // Constraint block Y (lobpcg.m:64-75 with prodB = []): row-major chunks of at most 16
  // columns, and Uy (column-major upper) with Uy^T Uy = Y^T Y.
  std::vector<DevBuf> ych;
  std::vector<int> ycw;
  int ym = 0;
  std::vector<double> Uy;
  // Y: device, column-major n x m with leading dimension ld.  Returns CHR_EARG when Y^T Y is
  // not positive definite (MATLAB's chol fails there).
  int set_constraint(const double *Y, int64_t ld, int m) {
    ym = m;
    ych.clear();
    ycw.clear();
    for (int c0 = 0; c0 < m; c0 += 16) {
      const int w = std::min(16, m - c0);
      ych.emplace_back();
      int rc;
      if ((rc = ych.back().alloc(sizeof(double) * (size_t)n * w))) return rc;
      hipLaunchKernelGGL(chr::cm_to_rm_kernel, dim3(grid((int64_t)n * w)), dim3(chr::kBlock), 0, s, n, w,
                         Y + (size_t)c0 * ld, ld, ych.back().as<double>());
      HIPC(hipGetLastError());
      ycw.push_back(w);
    }
    const int nc = (int)ych.size();
    std::vector<int> slot((size_t)nc * nc, -1);
    for (int a = 0; a < nc; ++a)
      for (int b = a; b < nc; ++b)
        if ((slot[(size_t)a * nc + b] = gram_launch(ych[a].as<double>(), ycw[a], ych[b].as<double>(), ycw[b])) < 0)
          return CHR_EARG;
    int rc;
    if ((rc = fetch())) return rc;
    std::vector<double> M((size_t)m * m), g;
    for (int a = 0, ra = 0; a < nc; ra += ycw[a++])
      for (int b = a, rb = ra; b < nc; rb += ycw[b++]) {
        get(slot[(size_t)a * nc + b], ycw[a], ycw[b], g);
        for (int q = 0; q < ycw[b]; ++q)
          for (int p = 0; p < ycw[a]; ++p) {
            M[(size_t)(ra + p) + (size_t)(rb + q) * m] = g[p + (size_t)q * ycw[a]];
            M[(size_t)(rb + q) + (size_t)(ra + p) * m] = g[p + (size_t)q * ycw[a]];
          }
      }
    if (!chrdense::chol_lower(m, M.data())) return fail(CHR_EARG, "chr_lobpcg: the constraint block Y is rank deficient");
    Uy.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i <= j; ++i) Uy[i + (size_t)j * m] = M[j + (size_t)i * m];
    return CHR_OK;
  }
  // X <- X - Y Uy^-1 Uy^-T (Y^T X) for an n x k block (orthog_Y of lobpcg.m:74).
  int orthog(double *X, int k) {
    if (ym == 0) return CHR_OK;
    const int nc = (int)ych.size();
    std::vector<int> slot(nc);
    for (int c = 0; c < nc; ++c)
      if ((slot[c] = gram_launch(ych[c].as<double>(), ycw[c], X, k)) < 0) return CHR_EARG;
    int rc;
    if ((rc = fetch())) return rc;
    std::vector<double> C((size_t)ym * k), g;
    for (int c = 0, r0 = 0; c < nc; r0 += ycw[c++]) {
      get(slot[c], ycw[c], k, g);
      for (int q = 0; q < k; ++q)
        for (int p = 0; p < ycw[c]; ++p) C[(size_t)(r0 + p) + (size_t)q * ym] = g[p + (size_t)q * ycw[c]];
    }
    for (int q = 0; q < k; ++q) {
      double *cq = C.data() + (size_t)q * ym;
      for (int i = 0; i < ym; ++i) {  // Uy^T z = c
        double t = cq[i];
        for (int j = 0; j < i; ++j) t -= Uy[j + (size_t)i * ym] * cq[j];
        cq[i] = t / Uy[i + (size_t)i * ym];
      }
      for (int i = ym - 1; i >= 0; --i) {  // Uy w = z
        double t = cq[i];
        for (int j = i + 1; j < ym; ++j) t -= Uy[i + (size_t)j * ym] * cq[j];
        cq[i] = t / Uy[i + (size_t)i * ym];
      }
    }
    std::vector<double> Wc;
    for (int c = 0, r0 = 0; c < nc; r0 += ycw[c++]) {
      Wc.assign((size_t)ycw[c] * k, 0.0);
      for (int q = 0; q < k; ++q)
        for (int p = 0; p < ycw[c]; ++p) Wc[p + (size_t)q * ycw[c]] = -C[(size_t)(r0 + p) + (size_t)q * ym];
      if ((rc = gemm(ych[c].as<double>(), ycw[c], Wc.data(), k, 1.0, X))) return rc;
    }
    return CHR_OK;
  }
  // B_orthonormalize.m with prodB = []: X <- X U^-1 with U^T U = X^T X (formed in tmp,
  // then the pointers swap); returns 1 when cond(X^T X) > 1/(k eps) (the reference's
  // failure), U (column-major upper) otherwise.
  int borth(double *&X, int k, std::vector<double> &U, double *&tmp, int &failed) {
    std::vector<double> M;
    int rc;
    if ((rc = gram(X, k, X, k, M))) return rc;
    failed = borth_factor(M, k, U) ? 0 : 1;
    return failed ? CHR_OK : borth_apply(X, k, U, tmp);
  }
  // The host half of borth: U from the Gram matrix M = X^T X (destroyed); false on the
  // reference's failure (cond(M) > 1/(k eps), or M not positive definite).
  static bool borth_factor(std::vector<double> &M, int k, std::vector<double> &U) {
    if (chrdense::cond_sym(k, M.data()) > 1.0 / ((double)k * 2.220446049250313e-16) ||
        !chrdense::chol_lower(k, M.data()))
      return false;
    U.assign((size_t)k * k, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i <= j; ++i) U[i + (size_t)j * k] = M[j + (size_t)i * k];
    return true;
  }
  // The device half: X <- X U^-1 (formed in tmp, then the pointers swap).
  int borth_apply(double *&X, int k, const std::vector<double> &U, double *&tmp) {
    std::vector<double> Ui = chrdense::inv_upper(k, U.data());
    if (int rc = gemm(X, k, Ui.data(), k, 0.0, tmp)) return rc;
    std::swap(X, tmp);
    return CHR_OK;
  }
};

}  // namespace

extern "C" int chr_lobpcg(chr_hier *h, chr_mat *Aop, int32_t n, int32_t neig, double *X, int64_t ldx,
                          const double *Y, int32_t m_y, int64_t ldy, int32_t largest, int32_t reslambda_check,
                          int32_t itmax, double tol, int32_t restart_control, int32_t ptr_kind,
                          int32_t *niter_out, double *lambda_out, double *resnorm_hist, double *lambda_hist,
                          int32_t *ierr_out) {
  if (!X || neig < 1 || neig > kMaxBlockCols || itmax < 2 || ldx < n || (int64_t)n < 3 * (int64_t)neig)
    return fail(CHR_EARG, "chr_lobpcg: need 1 <= neig <= 16, itmax >= 2, n >= 3*neig, ldx >= n");
  if (m_y < 0 || m_y > 48 || (m_y > 0 && (!Y || ldy < n)))
    return fail(CHR_EARG, "chr_lobpcg: the constraint block needs 0 <= m <= 48 columns and ldy >= n");
  if (!h && !Aop) return fail(CHR_EARG, "chr_lobpcg: need a hierarchy or an operator");
  const DevCsr *A = Aop ? &Aop->A : &h->outer.A;
  if (A->nrows != n || A->ncols != n || (h && h->lev[0].n != n))
    return fail(CHR_EARG, "chr_lobpcg: operator / hierarchy size differs from n");
  int rc;
  hipStream_t s = h ? h->stream : nullptr;
  if (h && (rc = ensure_block_capacity(*h, neig))) return rc;
  DevBuf zero;
  if ((rc = zero.alloc(sizeof(int)))) return rc;
  HIPC(hipMemsetAsync(zero.p, 0, sizeof(int), s));
  Lob L{h, A, zero.as<int>(), s, n};
  L.wmax = neig;
  if ((rc = L.init())) return rc;
  const size_t blk = sizeof(double) * (size_t)n * neig;
  DevBuf bX, bAX, bP, bAP, bAR, bAAR, bAPa, bAAP, bBest, bT;
  for (DevBuf *b : {&bX, &bAX, &bP, &bAP, &bAR, &bAAR, &bAPa, &bAAP, &bBest, &bT})
    if ((rc = b->alloc(blk))) return rc;
  double *dX = bX.as<double>(), *dAX = bAX.as<double>(), *dP = bP.as<double>(), *dAP = bAP.as<double>();
  double *dAR = bAR.as<double>(), *dAAR = bAAR.as<double>();
  double *dAPa = bAPa.as<double>(), *dAAP = bAAP.as<double>(), *dBest = bBest.as<double>();
  double *dT = bT.as<double>();
  HIPC(hipMemsetAsync(dP, 0, blk, s));
  HIPC(hipMemsetAsync(dAP, 0, blk, s));
  // X0 (column-major) -> row-major.
  const double *xsrc = X;
  int64_t lds = ldx;
  if (ptr_kind == CHR_PTR_HOST) {
    HIPC(copy_cols(dT, 8 * (size_t)n, X, 8 * (size_t)ldx, 8 * (size_t)n, neig, hipMemcpyHostToDevice, s));
    xsrc = dT;
    lds = n;
  }
  hipLaunchKernelGGL(chr::cm_to_rm_kernel, dim3(Lob::grid((int64_t)n * neig)), dim3(chr::kBlock), 0, s, n,
                     neig, xsrc, lds, dX);
  HIPC(hipGetLastError());
  const int64_t H = (int64_t)itmax + 1;
  auto hist = [&](double *hh, int it, int j, double v) {
    if (hh) hh[(int64_t)j * H + it] = v;
  };
  std::vector<double> lam(neig), lam_old(neig, 0.0), relDelta(neig, INFINITY), resn(neig), U, gA, gB, E, w;
  std::vector<int> act;
  int32_t ierr = 0, iter = 0, n_active = neig, failed = 0;
  double smallest = 1e308;
  int restart = 1, forced = 0;
  auto rayleigh_ritz_X = [&]() -> int {  // X, AX <- X E, AX E with E from (X'AX, X'X)
    std::vector<double> xax, xx;
    int r;
    const int s0 = L.gram_launch(dX, neig, dAX, neig), s1 = L.gram_launch(dX, neig, dX, neig);
    if (s0 < 0 || s1 < 0) return CHR_EARG;
    if ((r = L.fetch())) return r;
    L.get(s0, neig, neig, xax);
    L.get(s1, neig, neig, xx);
    chrdense::symmetrize(neig, xax.data());
    chrdense::symmetrize(neig, xx.data());
    w.resize(neig);
    E.resize((size_t)neig * neig);
    if (!chrdense::sygv(neig, xax.data(), xx.data(), w.data(), E.data())) return 1;
    if ((r = L.gemm(dX, neig, E.data(), neig, 0.0, dT))) return r;
    std::swap(dX, dT);
    if ((r = L.gemm(dAX, neig, E.data(), neig, 0.0, dT))) return r;
    std::swap(dAX, dT);
    for (int j = 0; j < neig; ++j) lam[j] = w[j];
    return 0;
  };
  if ((rc = L.copy(dBest, dX, neig))) return rc;
  if (m_y > 0) {  // the constraint block, then X0 projected off it (lobpcg.m:64-75)
    DevBuf ydev;
    const double *ysrc = Y;
    int64_t ly = ldy;
    if (ptr_kind == CHR_PTR_HOST) {
      if ((rc = ydev.alloc(sizeof(double) * (size_t)n * m_y))) return rc;
      HIPC(copy_cols(ydev.p, 8 * (size_t)n, Y, 8 * (size_t)ldy, 8 * (size_t)n, m_y, hipMemcpyHostToDevice, s));
      ysrc = ydev.as<double>();
      ly = n;
    }
    if ((rc = L.set_constraint(ysrc, ly, m_y)) || (rc = L.orthog(dX, neig))) return rc;
  }
  if ((rc = L.borth(dX, neig, U, dT, failed))) return rc;
  if (failed) {
    ierr = 1;
  } else {
    if ((rc = L.applyA(dX, neig, dAX))) return rc;
    int r = rayleigh_ritz_X();
    if (r < 0) return r;
    if (r == 1) ierr = 1;
    for (int j = 0; j < neig; ++j) hist(lambda_hist, 0, j, lam[j]);
  }
  // X buffers: the current block, the best one (best_blockX of lobpcg.m, which aliases the
  // current one until the next update writes a fresh buffer) and a spare; no block copies.
  DevBuf bSpare;
  if ((rc = bSpare.alloc(blk))) return rc;
  double *const xpool[3] = {dX, dBest, bSpare.as<double>()};
  auto xfree = [&]() {  // a pool buffer that is neither the current nor the best block
    for (double *p : xpool)
      if (p != dX && p != dBest) return p;
    return (double *)nullptr;
  };
  // R^T R of R = AX - X diag(lam), fetched (R itself is never stored).
  auto resid_norms = [&](std::vector<double> &rr) -> int {
    const int sl = L.resid_norms(dAX, dX, lam.data(), neig);
    if (sl < 0) return CHR_EARG;
    if (int r = L.fetch()) return r;
    L.get(sl, neig, neig, rr);
    return CHR_OK;
  };
  while (!ierr && iter < itmax - 1) {
    ++iter;
    std::vector<double> rr;
    if ((rc = resid_norms(rr))) return rc;
    double avg = 0.0;
    for (int j = 0; j < neig; ++j) {
      resn[j] = std::sqrt(rr[j + (size_t)j * neig]) / lam[j];
      hist(resnorm_hist, iter - 1, j, resn[j]);
      avg += resn[j];
      relDelta[j] = std::fabs(lam_old[j] - lam[j]) / lam[j];
      lam_old[j] = lam[j];
    }
    avg /= neig;
    bool recompute_ax = false;
    if (avg < smallest) {
      smallest = avg;
      dBest = dX;
    } else if (avg > std::ldexp(1.0, restart_control) * smallest) {
      forced = 1;
      recompute_ax = true;
    }
    act.clear();
    for (int j = 0; j < neig; ++j) {
      const bool conv = reslambda_check ? (relDelta[j] < tol) : (resn[j] < tol);
      if (!conv) act.push_back(j);
    }
    n_active = (int)act.size();
    // The active residual columns come from the residual of this step's AX (lobpcg.m:145-149,
    // 200), formed before a forced restart recomputes AX (lobpcg.m:166-172).
    if (n_active > 0 && (rc = L.resid_gather(dAX, dX, lam.data(), neig, act, dAR))) return rc;
    if (recompute_ax && (rc = L.applyA(dX, neig, dAX))) return rc;
    if (n_active == 0) break;
    const int na = n_active;
    const int *act_dev = static_cast<const int *>(L.stage(act.data(), sizeof(int) * na));
    if (iter > 1) {
      if ((rc = L.gather(dP, neig, act_dev, na, dAPa)) || (rc = L.gather(dAP, neig, act_dev, na, dAAP))) return rc;
    }
    if ((rc = L.applyM(dAR, na, dT))) return rc;
    if ((rc = L.orthog(dAR, na))) return rc;  // lobpcg.m:228-230
    // AR <- AR - X (X^T AR), the coefficients read on the device
    const int sxar = L.gram_launch(dX, neig, dAR, na);
    if (sxar < 0) return CHR_EARG;
    if ((rc = L.gemm_slot(dX, neig, sxar, na, -1.0, 1.0, dAR))) return rc;
    // B-orthonormalization of R, then of P (lobpcg.m:238, :256).  P does not depend on R, so
    // both Gram matrices are formed and fetched together (one host round trip); the
    // factorizations and products follow in the reference's order.
    const int srr = L.gram_launch(dAR, na, dAR, na);
    const int spp = iter > 1 ? L.gram_launch(dAPa, na, dAPa, na) : 0;
    if (srr < 0 || spp < 0) return CHR_EARG;
    if ((rc = L.fetch())) return rc;
    std::vector<double> grr, gpp;
    L.get(srr, na, na, grr);
    if (iter > 1) L.get(spp, na, na, gpp);
    if (!Lob::borth_factor(grr, na, U)) {
      ierr = 2;
      break;
    }
    if ((rc = L.borth_apply(dAR, na, U, dT))) return rc;
    if ((rc = L.applyA(dAR, na, dAAR))) return rc;
    if (iter > 1) {
      if (!Lob::borth_factor(gpp, na, U)) {
        restart = 1;
      } else {
        // P <- P U^-1 and AP <- AP U^-1 (lobpcg.m:256-265)
        std::vector<double> Ui = chrdense::inv_upper(na, U.data());
        if ((rc = L.gemm(dAPa, na, Ui.data(), na, 0.0, dT))) return rc;
        std::swap(dAPa, dT);
        if ((rc = L.gemm(dAAP, na, Ui.data(), na, 0.0, dT))) return rc;
        std::swap(dAAP, dT);
        restart = forced;
      }
    }
    const int m = restart ? neig + na : neig + 2 * na;
    gA.assign((size_t)m * m, 0.0);
    gB.assign((size_t)m * m, 0.0);
    auto put = [&](std::vector<double> &M, int r0, int c0, int rows, int cols, const std::vector<double> &src,
                   bool sym) {
      for (int jj = 0; jj < cols; ++jj)
        for (int ii = 0; ii < rows; ++ii) {
          const double v = src[ii + (size_t)jj * rows];
          M[(r0 + ii) + (size_t)(c0 + jj) * m] = v;
          if (sym) M[(c0 + jj) + (size_t)(r0 + ii) * m] = v;
        }
    };
    // All Gram blocks of the step are reduced on the device, then fetched together.
    struct GB {
      const double *X;
      int a;
      const double *Y;
      int b;
      bool toA;
      int r0, c0;
      bool diag;
      int slot;
    };
    const int c2 = neig + na;
    std::vector<GB> blocks = {{dX, neig, dAX, neig, true, 0, 0, true, -1},
                              {dX, neig, dAAR, na, true, 0, neig, false, -1},
                              {dAR, na, dAAR, na, true, neig, neig, true, -1},
                              {dX, neig, dX, neig, false, 0, 0, true, -1},
                              {dX, neig, dAR, na, false, 0, neig, false, -1},
                              {dAR, na, dAR, na, false, neig, neig, true, -1}};
    if (!restart) {
      blocks.push_back({dX, neig, dAAP, na, true, 0, c2, false, -1});
      blocks.push_back({dAR, na, dAAP, na, true, neig, c2, false, -1});
      blocks.push_back({dAPa, na, dAAP, na, true, c2, c2, true, -1});
      blocks.push_back({dX, neig, dAPa, na, false, 0, c2, false, -1});
      blocks.push_back({dAR, na, dAPa, na, false, neig, c2, false, -1});
      blocks.push_back({dAPa, na, dAPa, na, false, c2, c2, true, -1});
    }
    std::vector<Lob::GramPair> pairs;
    for (const auto &bk : blocks) pairs.push_back({bk.X, bk.a, bk.Y, bk.b});
    std::vector<int> offs;
    int base;
    const char *env_multi = std::getenv("CHRONOS_GRAM_MULTI");
    const bool one_pass = !(env_multi && std::atoi(env_multi) == 0);
    Lob::Packed pk;
    const bool packed = one_pass && na <= 8;
    if (packed) {
      std::vector<const double *> narrow = {dAR, dAAR};
      if (!restart) {
        narrow.push_back(dAPa);
        narrow.push_back(dAAP);
      }
      if ((base = L.gram_packed_launch(dX, dAX, narrow, neig, na, pk)) < 0) return CHR_EARG;
    } else if (!one_pass) {
      // one pass per product (same entries)
      base = -1;
      offs.clear();
      for (const auto &pr : pairs) {
        const int sl = L.gram_launch(pr.X, pr.a, pr.Y, pr.b);
        if (sl < 0) return CHR_EARG;
        if (base < 0) base = sl;
        offs.push_back((sl - base) * Lob::kSlot);
      }
    } else {
      base = L.gram_multi_launch(pairs, offs);
      if (base < 0) return CHR_EARG;
    }
    if ((rc = L.fetch())) return rc;
    std::vector<double> g;
    for (size_t k = 0; k < blocks.size(); ++k) {
      const auto &bk = blocks[k];
      if (packed) {
        if (!L.get_packed(base, pk, bk.X, bk.a, bk.Y, bk.b, g)) return fail(CHR_EARG, "lobpcg: packed Gram lookup");
      } else {
        L.get_at(base, offs[k], bk.a, bk.b, g);
      }
      if (bk.diag) {
        chrdense::symmetrize(bk.a, g.data());
        put(bk.toA ? gA : gB, bk.r0, bk.c0, bk.a, bk.b, g, false);
      } else {
        put(bk.toA ? gA : gB, bk.r0, bk.c0, bk.a, bk.b, g, true);
      }
    }
    w.resize(m);
    E.resize((size_t)m * m);
    if (!chrdense::sygv(m, gA.data(), gB.data(), w.data(), E.data())) {
      ierr = 2;
      break;
    }
    std::vector<double> EX((size_t)neig * neig), ER((size_t)na * neig), EP((size_t)na * neig);
    for (int j = 0; j < neig; ++j) {
      const int src = largest ? m - 1 - j : j;
      lam[j] = w[src];
      hist(lambda_hist, iter, j, lam[j]);
      const double *e = E.data() + (size_t)src * m;
      for (int i = 0; i < neig; ++i) EX[i + (size_t)j * neig] = e[i];
      for (int i = 0; i < na; ++i) ER[i + (size_t)j * na] = e[neig + i];
      if (!restart)
        for (int i = 0; i < na; ++i) EP[i + (size_t)j * na] = e[neig + na + i];
    }
    // P = AR ER + APa EP, X = P + X EX; AP and AX alike (lobpcg.m:389-400).
    const double *cdev = L.stage_update(ER.data(), EP.data(), EX.data(), na, neig);
    double *xout = dX == dBest ? xfree() : dX;  // keep the best block intact
    if ((rc = L.update(dAR, restart ? nullptr : dAPa, na, dX, neig, cdev, dP, xout)) ||
        (rc = L.update(dAAR, restart ? nullptr : dAAP, na, dAX, neig, cdev, dAP, dAX)))
      return rc;
    dX = xout;
  }
  if (ierr != 1) {
    if (n_active > 0) {
      ++iter;
      std::vector<double> rr;
      if ((rc = resid_norms(rr))) return rc;
      for (int j = 0; j < neig; ++j) {
        resn[j] = std::sqrt(rr[j + (size_t)j * neig]) / lam[j];
        hist(resnorm_hist, iter - 1, j, resn[j]);
        hist(lambda_hist, iter - 1, j, lam[j]);
      }
    }
    double mx = 0.0;
    for (int j = 0; j < neig; ++j) mx = std::max(mx, reslambda_check ? relDelta[j] : resn[j]);
    if (mx > tol) {
      ierr = -1;
      dX = dBest;
    }
    // Final Rayleigh-Ritz without preconditioning (lobpcg.m:420-459).
    ++iter;
    if ((rc = L.orthog(dX, neig))) return rc;  // lobpcg.m:384-387
    if ((rc = L.applyA(dX, neig, dAX))) return rc;
    int r = rayleigh_ritz_X();
    if (r < 0) return r;
    if (r == 1) {
      ierr = 1;
    } else {
      std::vector<double> rr;
      if ((rc = resid_norms(rr))) return rc;
      for (int j = 0; j < neig; ++j) {
        resn[j] = std::sqrt(rr[j + (size_t)j * neig]) / lam[j];
        hist(resnorm_hist, iter - 1, j, resn[j]);
        hist(lambda_hist, iter - 1, j, lam[j]);
      }
    }
  }
  // Row-major X -> caller's column-major block.
  double *xdst = ptr_kind == CHR_PTR_HOST ? dT : X;
  const int64_t ldd = ptr_kind == CHR_PTR_HOST ? n : ldx;
  hipLaunchKernelGGL(chr::rm_to_cm_kernel, dim3(Lob::grid((int64_t)n * neig)), dim3(chr::kBlock), 0, s, n, neig,
                     (const double *)dX, xdst, ldd);
  HIPC(hipGetLastError());
  if (ptr_kind == CHR_PTR_HOST)
    HIPC(copy_cols(X, 8 * (size_t)ldx, dT, 8 * (size_t)n, 8 * (size_t)n, neig, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (lambda_out) std::copy(lam.begin(), lam.end(), lambda_out);
  if (niter_out) *niter_out = iter - 1;
  if (ierr_out) *ierr_out = ierr;
  return CHR_OK;
}

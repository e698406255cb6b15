// Batched AMG-PCG: many independent systems of uneven size, each with a block of k right-
// hand sides, advanced together by one captured graph of several iterations.
//
// The systems' hierarchies are merged level by level into one block-diagonal hierarchy
// (systems sorted by depth, deepest first).  On level l the systems that still descend
// form a prefix of the rows and go through smoothing and the transfers; the systems whose
// coarsest level is l follow and are solved with their dense inverses (chr_hier::Level).
// Every row keeps its own nonzeros and summation order, so the merged V-cycle equals each
// system's own V-cycle bit for bit.  The outer PCG_CJ (sources/LinearSolver/PCG_CJ.m:1-73)
// runs one independent recurrence per (system, column): its scalars are per solve, the
// reductions are per system (row chunks that never straddle two systems, added in order by
// one thread per solve), and a solve that reaches its tolerance is frozen (alpha = beta = 0)
// while the others continue.
#include "chronos_amg.h"
#include "runtime.hpp"

using namespace chrrt;

namespace {

constexpr int kChunkRows = 4096;
constexpr int kMaxSolves = 1024;

struct BState {
  int git;      // iterations done by the batch
  int itmax;
  int nactive;  // solves not yet converged
  int stop;     // 1 once nactive == 0 or git == itmax (kernels then exit)
  double tol;
};

struct BArrays {  // per solve q = system * k + column
  double *bnorm, *resiter, *ptap, *pr, *alpha, *beta;
  int *active, *iters, *breakdown;
};

// Row -> system (internal order) by binary search over the level-0 offsets (nsys + 1).
__device__ __forceinline__ int sys_of(int r, const int *__restrict__ off, int nsys) {
  int lo = 0, hi = nsys;  // off[lo] <= r < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= r) lo = mid; else hi = mid;
  }
  return lo;
}

// Per chunk and column: sums of u.v (and of w.w when w is given) over the chunk's rows.
// chunks[i] = {first row, end row, system}.
__global__ __launch_bounds__(chr::kBlock) void bpcg_dots_kernel(const int4 *__restrict__ chunks, int k,
                                                                const double *__restrict__ u,
                                                                const double *__restrict__ v,
                                                                const double *__restrict__ w,
                                                                double *__restrict__ part,
                                                                const BState *st) {
  if (st->stop) return;
  __shared__ double lds[chr::kBlock];
  const int4 ch = chunks[blockIdx.x];
  const int c = threadIdx.x % k, g = threadIdx.x / k, ng = chr::kBlock / k;
  double s1 = 0.0, s2 = 0.0;
  if (g < ng && threadIdx.x < ng * k) {
    for (int r = ch.x + g; r < ch.y; r += ng) {
      const size_t i = (size_t)r * k + c;
      s1 += u[i] * v[i];
      if (w) s2 += w[i] * w[i];
    }
  }
  // Reduce over the ng groups for each column (fixed order).
  for (int pass = 0; pass < (w ? 2 : 1); ++pass) {
    lds[threadIdx.x] = pass ? s2 : s1;
    __syncthreads();
    if (threadIdx.x < k) {
      double t = 0.0;
      for (int q = 0; q < ng; ++q) t += lds[q * k + threadIdx.x];
      part[((size_t)blockIdx.x * 2 + pass) * k + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// x += alpha p ; r -= alpha Ap (alpha per solve), partials of r.r per chunk and column.
__global__ __launch_bounds__(chr::kBlock) void bpcg_update_kernel(const int4 *__restrict__ chunks, int k,
                                                                  const double *__restrict__ p,
                                                                  const double *__restrict__ ap,
                                                                  double *__restrict__ x,
                                                                  double *__restrict__ r,
                                                                  const double *__restrict__ alpha,
                                                                  double *__restrict__ part,
                                                                  const BState *st) {
  if (st->stop) return;
  __shared__ double lds[chr::kBlock];
  const int4 ch = chunks[blockIdx.x];
  const int c = threadIdx.x % k, g = threadIdx.x / k, ng = chr::kBlock / k;
  double s = 0.0;
  if (threadIdx.x < ng * k) {
    const double a = alpha[ch.z * k + c];
    for (int row = ch.x + g; row < ch.y; row += ng) {
      const size_t i = (size_t)row * k + c;
      x[i] += a * p[i];
      const double ri = r[i] - a * ap[i];
      r[i] = ri;
      s += ri * ri;
    }
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int q = 0; q < ng; ++q) t += lds[q * k + threadIdx.x];
    part[((size_t)blockIdx.x * 2) * k + threadIdx.x] = t;
  }
}

// p_new = z + beta p_old (row-major n x k, beta per solve: the chunk's system and the
// column), formed before the product so that the product gathers one operand per nonzero.
__global__ __launch_bounds__(chr::kBlock) void bpcg_pnew_kernel(const int4 *__restrict__ chunks, int k,
                                                                const double *__restrict__ beta,
                                                                const double *__restrict__ z,
                                                                const double *__restrict__ p_old,
                                                                double *__restrict__ p_new,
                                                                const BState *st) {
  if (st->stop) return;
  const int4 ch = chunks[blockIdx.x];
  const int c = threadIdx.x % k, g = threadIdx.x / k, ng = chr::kBlock / k;
  if (threadIdx.x >= ng * k) return;
  const double bq = beta[ch.z * k + c];
  for (int row = ch.x + g; row < ch.y; row += ng) {
    const size_t i = (size_t)row * k + c;
    p_new[i] = z[i] + bq * p_old[i];
  }
}

// Ap = A p_new (row-major n x k).  FORMED = false: p_new = z + beta p_old formed at every
// gathered column and written for the block's rows; FORMED = true: p_new formed by
// bpcg_pnew_kernel (the same expression, so the same bits).  As in the block product, the
// stream block's values and columns are staged in LDS (cap = PER * kBlock nonzeros) and
// each (row, column) pair is a lane that issues the gathers of 8 staged nonzeros before
// adding their products in nonzero order; a block holding one longer row is summed
// straight from memory.
template <int KL, typename TI, bool FORMED>
__global__ __launch_bounds__(chr::kBlock) void bpcg_p_kernel(int nblk, const int4 *__restrict__ blk,
                                                             const int *__restrict__ rp,
                                                             const TI *__restrict__ ci,
                                                             const double *__restrict__ val, int k, int cap,
                                                             const int *__restrict__ off, int nsys,
                                                             const double *__restrict__ beta,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ p_old,
                                                             double *__restrict__ p_new,
                                                             double *__restrict__ ap,
                                                             const BState *st) {
  extern __shared__ __align__(16) char smem[];
  const int b = chr::xcd_block(blockIdx.x, nblk);
  const int4 lo = blk[b], hi = blk[b + 1];
  if (st->stop) return;
  const int base = lo.z, e0 = lo.y, e1 = hi.y, tid = threadIdx.x;
  if (e1 - e0 > cap) {  // one long row
    const int c = tid % KL, r = lo.x;
    if (tid >= KL || c >= k) return;  // no cross-lane operations below
    const double bq = FORMED ? 0.0 : beta[sys_of(r, off, nsys) * k + c];
    double s = 0.0;
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
      const size_t j = (size_t)(base + (int)ci[e]) * k + c;
      s = s + chr::mul_rn(val[e], FORMED ? p_new[j] : z[j] + bq * p_old[j]);
    }
    const size_t i = (size_t)r * k + c;
    if constexpr (!FORMED) p_new[i] = z[i] + bq * p_old[i];
    ap[i] = s;
    return;
  }
  double *sv = reinterpret_cast<double *>(smem);
  TI *sc = reinterpret_cast<TI *>(smem + (size_t)cap * sizeof(double));
  for (int t = tid; t < e1 - e0; t += chr::kBlock) {
    sv[t] = val[e0 + t];
    sc[t] = ci[e0 + t];
  }
  __syncthreads();
  const int npair = (hi.x - lo.x) * KL;
  for (int p = tid; p < npair; p += chr::kBlock) {
    const int r = lo.x + p / KL, c = p % KL;
    if (c >= k) continue;
    const double bq = FORMED ? 0.0 : beta[sys_of(r, off, nsys) * k + c];
    const int ra = rp[r] - e0, rz = rp[r + 1] - e0;
    double s = 0.0;
    for (int q = ra; q < rz; q += 8) {
      const int n = min(8, rz - q);
      double vv[8], pc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int qq = q + min(u, n - 1);
        const size_t j = (size_t)(base + (int)sc[qq]) * k + c;
        vv[u] = sv[qq];
        if constexpr (FORMED) {
          pc[u] = p_new[j];
        } else {
          pc[u] = z[j] + bq * p_old[j];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < n) s = s + chr::mul_rn(vv[u], pc[u]);
    }
    const size_t i = (size_t)r * k + c;
    if constexpr (!FORMED) p_new[i] = z[i] + bq * p_old[i];
    ap[i] = s;
  }
}

enum { kBInit = 0, kBBeta = 1, kBAlpha = 2, kBUpdate = 3 };

// One thread per solve: adds its system's chunk partials in order and updates the scalars.
// cs[s], cs[s+1]: the chunk range of system s.
__global__ void bpcg_finalize_kernel(int kind, int nsys, int k, const int *__restrict__ cs,
                                     const double *__restrict__ part, BArrays a, BState *st,
                                     double *__restrict__ resvec) {
  __shared__ int nact;
  const int Q = nsys * k;
  const int q = threadIdx.x;
  if (kind != kBInit && st->stop) return;
  if (threadIdx.x == 0) nact = 0;
  __syncthreads();
  if (q < Q) {
    const int s = q / k, c = q % k;
    double s0 = 0.0, s1 = 0.0;
    for (int i = cs[s]; i < cs[s + 1]; ++i) {
      s0 += part[((size_t)i * 2) * k + c];
      s1 += part[((size_t)i * 2 + 1) * k + c];
    }
    const int git = st->git;
    switch (kind) {
      case kBInit: {  // s0 = b.b, s1 = r.r
        a.bnorm[q] = sqrt(s0);
        a.breakdown[q] = 0;
        a.iters[q] = 0;
        a.beta[q] = 0.0;
        a.alpha[q] = 0.0;
        const double res = s0 == 0.0 ? 0.0 : sqrt(s1) / a.bnorm[q];
        a.resiter[q] = res;
        a.active[q] = (s0 != 0.0 && st->itmax > 0) ? 1 : 0;
        resvec[q] = res;
        break;
      }
      case kBBeta:  // s0 = z.Ap_old
        a.beta[q] = (a.active[q] && git > 0) ? -s0 / a.ptap[q] : 0.0;
        break;
      case kBAlpha:  // s0 = p.Ap, s1 = p.r
        if (a.active[q]) {
          a.ptap[q] = s0;
          a.pr[q] = s1;
          if (!(s0 != 0.0) || !isfinite(s0)) {
            a.breakdown[q] = 1;
            a.active[q] = 0;
            a.alpha[q] = 0.0;
          } else {
            a.alpha[q] = s1 / s0;
          }
        } else {
          a.alpha[q] = 0.0;
        }
        break;
      default:  // s0 = r.r after the update
        if (a.active[q]) {
          const int it = git + 1;
          a.iters[q] = it;
          a.resiter[q] = sqrt(s0) / a.bnorm[q];
          resvec[(size_t)it * Q + q] = a.resiter[q];
          if (a.resiter[q] < st->tol || it >= st->itmax) a.active[q] = 0;
        }
    }
    if (a.active[q]) atomicAdd(&nact, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (kind == kBUpdate) st->git = st->git + 1;
    st->nactive = nact;
    if (nact == 0 || (kind == kBUpdate && st->git >= st->itmax)) st->stop = 1;
    if (kind == kBInit) st->stop = nact == 0 ? 1 : 0;
  }
}

// Column-major (user order) <-> row-major (internal order) for one system's rows.
__global__ __launch_bounds__(chr::kBlock) void bperm_in_kernel(int n, int k, const double *__restrict__ src,
                                                               int64_t ld, int64_t src_row0,
                                                               double *__restrict__ dst, int64_t dst_row0) {
  const int64_t t = (int64_t)blockIdx.x * chr::kBlock + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  const int64_t i = t / k;
  const int c = (int)(t % k);
  dst[(dst_row0 + i) * k + c] = src[c * ld + src_row0 + i];
}
__global__ __launch_bounds__(chr::kBlock) void bperm_out_kernel(int n, int k, const double *__restrict__ src,
                                                                int64_t src_row0, double *__restrict__ dst,
                                                                int64_t ld, int64_t dst_row0) {
  const int64_t t = (int64_t)blockIdx.x * chr::kBlock + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  const int64_t i = t % n;
  const int c = (int)(t / n);
  dst[c * ld + dst_row0 + i] = src[(src_row0 + i) * k + c];
}

struct HostCsr {
  int32_t nrows = 0, ncols = 0;
  std::vector<int32_t> rp{0}, ci;
  std::vector<double> v;
  chr_csr view() const { return {nrows, ncols, rp.data(), ci.data(), v.data()}; }
  void append(const chr_csr &m, int32_t col_off) {
    const int32_t e0 = rp.back();
    for (int32_t i = 0; i < m.nrows; ++i) rp.push_back(e0 + m.rowptr[i + 1]);
    for (int32_t e = 0; e < m.rowptr[m.nrows]; ++e) {
      ci.push_back(m.col[e] + col_off);
      v.push_back(m.val[e]);
    }
    nrows += m.nrows;
  }
};

}  // namespace

struct chr_batch {
  chr_hier *h = nullptr;  // merged hierarchy
  int32_t nsys = 0;
  std::vector<int32_t> order;  // internal position -> user system
  std::vector<int64_t> user_off, int_off;  // level-0 row offsets (user order / internal order)
  std::vector<int32_t> n0;     // level-0 rows per user system
  DevBuf off_dev, chunks, cstart;
  int32_t nchunks = 0;
  // PCG work (sized for kcap columns)
  int32_t kcap = 0, itcap = 0;
  DevBuf x, r, z, p0, p1, ap, b, part, state, arr, resv;
  BArrays a{};
  BState *pinned = nullptr;
  // The solve's graph: graph_iters consecutive batch iterations (an even number, so that the
  // p buffers are back in place for a further launch).
  hipGraphExec_t exec = nullptr;
  int32_t graph_k = 0, graph_iters = 0;
  uint64_t graph_gen = 0;  // h->buf_gen the graph was captured with
  void drop_graphs() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    graph_k = 0;
    graph_iters = 0;
  }
  ~chr_batch() {
    drop_graphs();
    if (pinned) (void)hipHostFree(pinned);
    delete h;
  }
};

extern "C" int chr_batch_create(int32_t nsys, const int32_t *nlev, const chr_level *const *levels,
                                int32_t precision, chr_batch **out) {
  if (nsys < 1 || !nlev || !levels || !out) return fail(CHR_EARG, "chr_batch_create: arguments");
  int32_t Lmax = 0;
  const chr_level *ref = nullptr;  // sweeps must agree across systems
  for (int32_t s = 0; s < nsys; ++s)
    if (nlev && levels && levels[s] && nlev[s] > 1 && !ref) ref = &levels[s][0];
  chr_level dummy{};
  dummy.nupre = 1;
  dummy.nupost = 1;
  if (!ref) ref = &dummy;
  for (int32_t s = 0; s < nsys; ++s) {
    if (nlev[s] < 1 || nlev[s] > 16 || !levels[s]) return fail(CHR_EARG, "chr_batch_create: 1..16 levels each");
    Lmax = std::max(Lmax, nlev[s]);
    for (int32_t l = 0; l < nlev[s]; ++l) {
      const chr_level &L = levels[s][l];
      int rc;
      if ((rc = check_csr(L.A, "A"))) return rc;
      if (L.A.nrows != L.A.ncols) return fail(CHR_EARG, "level operator must be square");
      if (l + 1 < nlev[s]) {
        if ((rc = check_csr(L.P, "P")) || (rc = check_csr(L.G, "G"))) return rc;
        if (L.P.nrows != L.A.nrows || L.P.ncols != levels[s][l + 1].A.nrows || L.G.nrows != L.A.nrows ||
            L.G.ncols != L.A.nrows)
          return fail(CHR_EARG, "chr_batch_create: inconsistent level shapes");
        if (L.nupre != ref->nupre || L.nupost != ref->nupost || L.nupre < 1 || L.nupost < 0)
          return fail(CHR_EARG, "chr_batch_create: every system needs the same nupre / nupost");
      }
    }
  }
  auto bt = std::make_unique<chr_batch>();
  bt->nsys = nsys;
  bt->order.resize(nsys);
  for (int32_t s = 0; s < nsys; ++s) bt->order[s] = s;
  std::stable_sort(bt->order.begin(), bt->order.end(), [&](int a, int b) { return nlev[a] > nlev[b]; });
  // Merged levels.
  std::vector<HostCsr> mA(Lmax), mG(Lmax), mP(Lmax), mA0(1);
  std::vector<LevelIn> lv(Lmax);
  std::vector<std::vector<int32_t>> off(Lmax + 1, std::vector<int32_t>(nsys, 0));
  for (int32_t l = 0; l < Lmax; ++l) {
    int32_t pos = 0;
    for (int32_t t = 0; t < nsys; ++t) {  // internal order
      const int32_t s = bt->order[t];
      if (nlev[s] > l) {
        off[l][t] = pos;
        pos += levels[s][l].A.nrows;
      }
    }
    lv[l].n = pos;
  }
  for (int32_t l = 0; l < Lmax; ++l) {
    LevelIn &in = lv[l];
    in.nupre = ref->nupre;
    in.nupost = ref->nupost;
    int32_t nd = 0;
    bool same_omega = true;
    double om0 = 0.0;
    for (int32_t t = 0; t < nsys; ++t) {
      const int32_t s = bt->order[t];
      if (nlev[s] > l + 1) {  // descends
        const chr_level &L = levels[s][l];
        mA[l].append(L.A, off[l][t]);
        mG[l].append(L.G, off[l][t]);
        mP[l].append(L.P, off[l + 1][t]);
        for (int32_t i = 0; i < L.A.nrows; ++i) in.omega_rows.push_back(L.omega);
        if (nd == 0) om0 = L.omega;
        same_omega &= L.omega == om0;
        nd += L.A.nrows;
      }
    }
    for (int32_t t = 0; t < nsys; ++t) {
      const int32_t s = bt->order[t];
      if (nlev[s] == l + 1) in.direct.push_back({off[l][t], levels[s][l].A});
    }
    in.nd = nd;
    in.omega = om0;
    if (same_omega) in.omega_rows.clear();
    if (nd > 0) {
      mA[l].ncols = nd;
      mG[l].ncols = nd;
      mP[l].ncols = lv[l + 1].n;
      in.A = mA[l].view();
      in.G = mG[l].view();
      in.P = mP[l].view();
    }
  }
  for (int32_t t = 0; t < nsys; ++t) mA0[0].append(levels[bt->order[t]][0].A, off[0][t]);
  mA0[0].ncols = lv[0].n;
  // The batch's level products are block products (8 right-hand sides in its configuration),
  // so the layouts are tuned with 8-column products.
  int rc = build_hier(lv, precision, mA0[0].view(), &bt->h, 8);
  if (rc) return rc;
  // Offsets in both orders, row chunks that never straddle systems.
  bt->n0.resize(nsys);
  bt->user_off.assign(nsys + 1, 0);
  bt->int_off.assign(nsys + 1, 0);
  for (int32_t s = 0; s < nsys; ++s) {
    bt->n0[s] = levels[s][0].A.nrows;
    bt->user_off[s + 1] = bt->user_off[s] + bt->n0[s];
  }
  std::vector<int32_t> offi(nsys + 1, 0), cs(nsys + 1, 0);
  std::vector<int4> ch;
  for (int32_t t = 0; t < nsys; ++t) {
    const int32_t s = bt->order[t];
    offi[t + 1] = offi[t] + bt->n0[s];
    bt->int_off[t + 1] = offi[t + 1];
    cs[t] = (int32_t)ch.size();
    for (int32_t r = offi[t]; r < offi[t + 1]; r += kChunkRows)
      ch.push_back(make_int4(r, std::min(offi[t + 1], r + kChunkRows), t, 0));
  }
  cs[nsys] = (int32_t)ch.size();
  bt->nchunks = (int32_t)ch.size();
  if ((rc = bt->off_dev.alloc(sizeof(int32_t) * (nsys + 1))) || (rc = bt->chunks.alloc(sizeof(int4) * ch.size())) ||
      (rc = bt->cstart.alloc(sizeof(int32_t) * (nsys + 1))))
    return rc;
  HIPC(hipMemcpy(bt->off_dev.p, offi.data(), sizeof(int32_t) * (nsys + 1), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(bt->chunks.p, ch.data(), sizeof(int4) * ch.size(), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(bt->cstart.p, cs.data(), sizeof(int32_t) * (nsys + 1), hipMemcpyHostToDevice));
  *out = bt.release();
  return CHR_OK;
}

extern "C" int chr_batch_destroy(chr_batch *bt) {
  if (bt) {
    if (bt->h) (void)hipStreamSynchronize(bt->h->stream);
    delete bt;
  }
  return CHR_OK;
}

extern "C" chr_hier *chr_batch_hier(chr_batch *bt) { return bt ? bt->h : nullptr; }

namespace {

unsigned grid_of(int64_t cnt) { return (unsigned)((cnt + chr::kBlock - 1) / chr::kBlock); }

// User column-major block (rows in user system order, leading dimension ld) -> internal
// row-major block, and back.
int batch_in(chr_batch &bt, int k, const double *src, int64_t ld, double *dst, hipStream_t s) {
  for (int32_t t = 0; t < bt.nsys; ++t) {
    const int32_t u = bt.order[t];
    hipLaunchKernelGGL(bperm_in_kernel, dim3(grid_of((int64_t)bt.n0[u] * k)), dim3(chr::kBlock), 0, s, bt.n0[u],
                       k, src, ld, bt.user_off[u], dst, bt.int_off[t]);
  }
  HIPC(hipGetLastError());
  return CHR_OK;
}
int batch_out(chr_batch &bt, int k, const double *src, double *dst, int64_t ld, hipStream_t s) {
  for (int32_t t = 0; t < bt.nsys; ++t) {
    const int32_t u = bt.order[t];
    hipLaunchKernelGGL(bperm_out_kernel, dim3(grid_of((int64_t)bt.n0[u] * k)), dim3(chr::kBlock), 0, s, bt.n0[u],
                       k, src, bt.int_off[t], dst, ld, bt.user_off[u]);
  }
  HIPC(hipGetLastError());
  return CHR_OK;
}

int batch_prepare(chr_batch &bt, int k, int itmax) {
  chr_hier &h = *bt.h;
  int rc;
  if ((rc = ensure_block_capacity(h, k))) return rc;
  const int64_t n = h.lev[0].n;
  if (bt.kcap < k) {
    bt.drop_graphs();
    const size_t vb = 8 * (size_t)n * k;
    for (DevBuf *b : {&bt.x, &bt.r, &bt.z, &bt.p0, &bt.p1, &bt.ap, &bt.b})
      if ((rc = b->alloc(vb))) return rc;
    if ((rc = bt.part.alloc(sizeof(double) * 2 * (size_t)bt.nchunks * k))) return rc;
    if ((rc = bt.state.alloc(sizeof(BState)))) return rc;
    const int Q = bt.nsys * k;
    if ((rc = bt.arr.alloc((sizeof(double) * 6 + sizeof(int) * 3) * (size_t)Q))) return rc;
    double *d = bt.arr.as<double>();
    int *ip = reinterpret_cast<int *>(d + 6 * (size_t)Q);
    bt.a = BArrays{d, d + Q, d + 2 * Q, d + 3 * Q, d + 4 * Q, d + 5 * Q, ip, ip + Q, ip + 2 * Q};
    if (!bt.pinned) HIPC(hipHostMalloc((void **)&bt.pinned, 2 * sizeof(BState), hipHostMallocDefault));
    bt.kcap = k;
    bt.itcap = 0;  // the residual histories hold (itmax + 1) * nsys * kcap entries: regrow them
  }
  if (bt.itcap < itmax + 1) {
    bt.drop_graphs();
    if ((rc = bt.resv.alloc(sizeof(double) * (size_t)(itmax + 1) * bt.nsys * bt.kcap))) return rc;
    bt.itcap = itmax + 1;
  }
  return CHR_OK;
}

hipError_t launch_bp(chr_batch &bt, int k, const double *pold, double *pnew, hipStream_t s) {
  const DevCsr &A = bt.h->outer.A;
  auto *st = bt.state.as<BState>();
  const int cap = A.per * chr::kBlock;
  // p = z + beta p_old in a pass of its own (CHRONOS_PCG_FORMED=0: formed inside the product).
  const char *env = std::getenv("CHRONOS_PCG_FORMED");
  const bool formed = !env || std::atoi(env) != 0;
  if (formed)
    hipLaunchKernelGGL(bpcg_pnew_kernel, dim3(bt.nchunks), dim3(chr::kBlock), 0, s, bt.chunks.as<int4>(), k,
                       (const double *)bt.a.beta, (const double *)bt.z.as<double>(), pold, pnew,
                       (const BState *)st);
#define BP(KL, TI)                                                                                        \
  do {                                                                                                    \
  if (formed) hipLaunchKernelGGL((bpcg_p_kernel<KL, TI, true>), dim3(A.nblk), dim3(chr::kBlock),         \
                     (size_t)cap * (sizeof(double) + sizeof(TI)), s, A.nblk, A.blk.as<int4>(),            \
                     A.rp.as<int>(), A.ci.as<TI>(), A.val.as<double>(), k, cap, bt.off_dev.as<int>(), bt.nsys, \
                     (const double *)bt.a.beta, (const double *)bt.z.as<double>(), pold, pnew,            \
                     bt.ap.as<double>(), (const BState *)st);                                             \
  else hipLaunchKernelGGL((bpcg_p_kernel<KL, TI, false>), dim3(A.nblk), dim3(chr::kBlock),               \
                     (size_t)cap * (sizeof(double) + sizeof(TI)), s, A.nblk, A.blk.as<int4>(),            \
                     A.rp.as<int>(), A.ci.as<TI>(), A.val.as<double>(), k, cap, bt.off_dev.as<int>(), bt.nsys, \
                     (const double *)bt.a.beta, (const double *)bt.z.as<double>(), pold, pnew,            \
                     bt.ap.as<double>(), (const BState *)st);                                             \
  } while (0)
  if (A.idx16) {
    if (k <= 2) BP(2, uint16_t); else if (k <= 4) BP(4, uint16_t); else if (k <= 8) BP(8, uint16_t); else BP(16, uint16_t);
  } else {
    if (k <= 2) BP(2, int); else if (k <= 4) BP(4, int); else if (k <= 8) BP(8, int); else BP(16, int);
  }
#undef BP
  return hipGetLastError();
}

// One batch iteration, p read from p0 (parity 0) or p1; with copy_state the device state is
// copied to pinned host memory at the end.
hipError_t batch_iter(chr_batch &bt, int k, int parity, bool copy_state) {
  chr_hier &h = *bt.h;
  hipStream_t s = h.stream;
  auto *st = bt.state.as<BState>();
  const int *stop = &st->stop;
  double *pold = parity ? bt.p1.as<double>() : bt.p0.as<double>();
  double *pnew = parity ? bt.p0.as<double>() : bt.p1.as<double>();
  double *part = bt.part.as<double>();
  const int4 *chunks = bt.chunks.as<int4>();
  double *resv = bt.resv.as<double>();
  const int Q = bt.nsys * k;
  hipError_t e;
  if ((e = enqueue_vcycle(h, bt.r.as<double>(), bt.z.as<double>(), stop, k))) return e;
  hipLaunchKernelGGL(bpcg_dots_kernel, dim3(bt.nchunks), dim3(chr::kBlock), 0, s, chunks, k,
                     (const double *)bt.z.as<double>(), (const double *)bt.ap.as<double>(), (const double *)nullptr,
                     part, (const BState *)st);
  hipLaunchKernelGGL(bpcg_finalize_kernel, dim3(1), dim3(kMaxSolves), 0, s, (int)kBBeta, bt.nsys, k,
                     (const int *)bt.cstart.as<int>(), (const double *)part, bt.a, st, resv);
  if ((e = launch_bp(bt, k, pold, pnew, s))) return e;
  // p.Ap in slot 0 and p.r in slot 1 of each chunk (two dot launches share the layout).
  hipLaunchKernelGGL(bpcg_dots_kernel, dim3(bt.nchunks), dim3(chr::kBlock), 0, s, chunks, k, (const double *)pnew,
                     (const double *)bt.ap.as<double>(), (const double *)nullptr, part, (const BState *)st);
  hipLaunchKernelGGL(bpcg_dots_kernel, dim3(bt.nchunks), dim3(chr::kBlock), 0, s, chunks, k, (const double *)pnew,
                     (const double *)bt.r.as<double>(), (const double *)nullptr, part + (size_t)k, (const BState *)st);
  hipLaunchKernelGGL(bpcg_finalize_kernel, dim3(1), dim3(kMaxSolves), 0, s, (int)kBAlpha, bt.nsys, k,
                     (const int *)bt.cstart.as<int>(), (const double *)part, bt.a, st, resv);
  hipLaunchKernelGGL(bpcg_update_kernel, dim3(bt.nchunks), dim3(chr::kBlock), 0, s, chunks, k, (const double *)pnew,
                     (const double *)bt.ap.as<double>(), bt.x.as<double>(), bt.r.as<double>(),
                     (const double *)bt.a.alpha, part, (const BState *)st);
  hipLaunchKernelGGL(bpcg_finalize_kernel, dim3(1), dim3(kMaxSolves), 0, s, (int)kBUpdate, bt.nsys, k,
                     (const int *)bt.cstart.as<int>(), (const double *)part, bt.a, st, resv);
  if ((e = hipGetLastError())) return e;
  (void)Q;
  if (!copy_state) return hipSuccess;
  return hipMemcpyAsync(&bt.pinned[0], st, sizeof(BState), hipMemcpyDeviceToHost, s);
}

}  // namespace

extern "C" int chr_batch_vcycle(chr_batch *bt, int32_t k, const double *b, double *x, int32_t ptr_kind) {
  if (!bt || k < 1 || k > kMaxBlockCols || !b || !x) return fail(CHR_EARG, "chr_batch_vcycle: arguments");
  chr_hier &h = *bt->h;
  int rc;
  if ((rc = ensure_block_capacity(h, k))) return rc;
  const int64_t n = h.lev[0].n;
  hipStream_t s = h.stream;
  DevBuf tmp;
  const double *src = b;
  if (ptr_kind == CHR_PTR_HOST) {
    if ((rc = tmp.alloc(8 * (size_t)n * k))) return rc;
    HIPC(hipMemcpyAsync(tmp.p, b, 8 * (size_t)n * k, hipMemcpyHostToDevice, s));
    src = tmp.as<double>();
  }
  if ((rc = batch_in(*bt, k, src, n, h.io_b.as<double>(), s))) return rc;
  HIPC(enqueue_vcycle(h, h.io_b.as<double>(), h.io_x.as<double>(), h.zero.as<int>(), k));
  double *dst = ptr_kind == CHR_PTR_HOST ? tmp.as<double>() : x;
  if ((rc = batch_out(*bt, k, h.io_x.as<double>(), dst, n, s))) return rc;
  if (ptr_kind == CHR_PTR_HOST) HIPC(hipMemcpyAsync(x, tmp.p, 8 * (size_t)n * k, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  return CHR_OK;
}

extern "C" int chr_batch_pcg(chr_batch *bt, int32_t k, const double *b, double *x, int32_t itmax, double tol,
                             int32_t ptr_kind, int32_t *iters, int32_t *flags, double *resvec) {
  if (!bt || k < 1 || k > kMaxBlockCols || !b || !x || itmax < 0)
    return fail(CHR_EARG, "chr_batch_pcg: need 1 <= k <= 16, blocks and itmax >= 0");
  if (bt->nsys * k > kMaxSolves) return fail(CHR_EARG, "chr_batch_pcg: at most 1024 solves (systems x columns)");
  chr_hier &h = *bt->h;
  int rc;
  if ((rc = batch_prepare(*bt, k, itmax))) return rc;
  hipStream_t s = h.stream;
  const int64_t n = h.lev[0].n;
  const int Q = bt->nsys * k;
  // Inputs.
  DevBuf tmp;
  const double *bs = b, *xs = x;
  if (ptr_kind == CHR_PTR_HOST) {
    if ((rc = tmp.alloc(16 * (size_t)n * k))) return rc;
    HIPC(hipMemcpyAsync(tmp.p, b, 8 * (size_t)n * k, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(tmp.as<double>() + n * k, x, 8 * (size_t)n * k, hipMemcpyHostToDevice, s));
    bs = tmp.as<double>();
    xs = tmp.as<double>() + n * k;
  }
  if ((rc = batch_in(*bt, k, bs, n, bt->b.as<double>(), s)) || (rc = batch_in(*bt, k, xs, n, bt->x.as<double>(), s)))
    return rc;
  const size_t vb = 8 * (size_t)n * k;
  HIPC(hipMemsetAsync(bt->p0.p, 0, vb, s));
  HIPC(hipMemsetAsync(bt->p1.p, 0, vb, s));
  HIPC(hipMemsetAsync(bt->ap.p, 0, vb, s));
  BState st0{};
  st0.itmax = itmax;
  st0.tol = tol;
  HIPC(hipMemcpyAsync(bt->state.p, &st0, sizeof(st0), hipMemcpyHostToDevice, s));
  // r = b - A x, norms, scalars.
  HIPC(launch_op(h.outer.A, bt->x.as<double>(), k,
                 chr::EpiResidual<double, double>{bt->b.as<double>(), bt->r.as<double>()}, h.zero.as<int>(), s));
  hipLaunchKernelGGL(bpcg_dots_kernel, dim3(bt->nchunks), dim3(chr::kBlock), 0, s, (const int4 *)bt->chunks.as<int4>(),
                     k, (const double *)bt->b.as<double>(), (const double *)bt->b.as<double>(),
                     (const double *)bt->r.as<double>(), bt->part.as<double>(), (const BState *)bt->state.as<BState>());
  hipLaunchKernelGGL(bpcg_finalize_kernel, dim3(1), dim3(kMaxSolves), 0, s, (int)kBInit, bt->nsys, k,
                     (const int *)bt->cstart.as<int>(), (const double *)bt->part.as<double>(), bt->a,
                     bt->state.as<BState>(), bt->resv.as<double>());
  HIPC(hipGetLastError());
  if (bt->resv.bytes < sizeof(double) * (size_t)(itmax + 1) * Q || bt->itcap < itmax + 1 || bt->kcap < k)
    return fail(CHR_EHIP, "chr_batch_pcg: work buffers smaller than the solve (internal error)");
  // One graph holds graph_iters batch iterations (p buffers alternating) and ends with the
  // state copy; it embeds the cycle vectors' addresses, so a reallocation of those (a wider
  // chr_batch_vcycle call) invalidates it.  A solve is one launch of it when it converges
  // within graph_iters iterations; kernels of iterations after every solve has stopped exit
  // at once.  graph_iters: CHRONOS_BATCH_GRAPH_ITERS (default 16), rounded up to an even
  // number, at most itmax rounded up likewise.
  const char *env_gi = std::getenv("CHRONOS_BATCH_GRAPH_ITERS");
  int gi = env_gi ? std::max(1, std::atoi(env_gi)) : 16;
  gi = std::min(gi, std::max(itmax, 1));
  gi += gi & 1;
  if (bt->graph_k != k || bt->graph_iters != gi || bt->graph_gen != h.buf_gen || !bt->exec) {
    bt->drop_graphs();
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < gi && e == hipSuccess; ++i) e = batch_iter(*bt, k, i & 1, i == gi - 1);
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(CHR_EHIP, std::string("batch graph capture: ") + hipGetErrorString(e ? e : e2));
    }
    e = hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(CHR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    bt->graph_k = k;
    bt->graph_iters = gi;
    bt->graph_gen = h.buf_gen;
  }
  HIPC(hipMemcpyAsync(&bt->pinned[0], bt->state.p, sizeof(BState), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  for (int32_t done = 0; !bt->pinned[0].stop && done < itmax; done += gi) {
    HIPC(hipGraphLaunch(bt->exec, s));
    HIPC(hipStreamSynchronize(s));
  }
  // Outputs.
  double *xd = ptr_kind == CHR_PTR_HOST ? tmp.as<double>() : x;
  if ((rc = batch_out(*bt, k, bt->x.as<double>(), xd, n, s))) return rc;
  if (ptr_kind == CHR_PTR_HOST) HIPC(hipMemcpyAsync(x, tmp.p, 8 * (size_t)n * k, hipMemcpyDeviceToHost, s));
  std::vector<int> it_int(Q), br(Q);
  std::vector<double> res(Q), bn(Q);
  HIPC(hipMemcpyAsync(it_int.data(), bt->a.iters, sizeof(int) * Q, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(br.data(), bt->a.breakdown, sizeof(int) * Q, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(res.data(), bt->a.resiter, sizeof(double) * Q, hipMemcpyDeviceToHost, s));
  std::vector<double> rv;
  if (resvec) {
    rv.resize((size_t)(itmax + 1) * Q);
    HIPC(hipMemcpyAsync(rv.data(), bt->resv.p, sizeof(double) * rv.size(), hipMemcpyDeviceToHost, s));
  }
  HIPC(hipStreamSynchronize(s));
  // Internal solve index (internal system t, column c) -> user (system order[t], c).
  for (int32_t t = 0; t < bt->nsys; ++t) {
    const int32_t u = bt->order[t];
    for (int c = 0; c < k; ++c) {
      const int qi = t * k + c, qu = u * k + c;
      if (iters) iters[qu] = it_int[qi];
      if (flags) flags[qu] = br[qi] ? -1 : (res[qi] > tol ? 1 : 0);
      if (resvec)
        for (int i = 0; i <= itmax; ++i) resvec[(size_t)qu * (itmax + 1) + i] = i <= it_int[qi] ? rv[(size_t)i * Q + qi] : 0.0;
    }
  }
  return CHR_OK;
}

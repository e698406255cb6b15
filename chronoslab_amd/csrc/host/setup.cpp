// Host-side producer of the AMG hierarchy (CPU, C++17 + OpenMP).
//
// This is the classical + adaptive-FSAI setup that ChronosLab runs before its solve phase
// (sources/Preconditioner/AMG/aspAMG/cpt_aspAMG.m:68-351), restated from the reference's
// behaviour.  The GPU solve phase consumes its output through chronos_amg.h.
#include "bamg.hpp"
#include "emin.hpp"
#include "chronos_setup.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include <omp.h>

#include "rng.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int nthreads_or_default(int nt) { return nt > 0 ? nt : omp_get_max_threads(); }

}  // namespace

struct chs_mat {
  int32_t nrows = 0, ncols = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
  int64_t nnz() const { return (int64_t)ci.size(); }
};

struct chs_hier {
  struct Level {
    chs_mat A, P, G;
    double omega = 1.0, lambda = 0.0;
  };
  std::vector<Level> lev;
  double t[5] = {0, 0, 0, 0, 0};
};

// ----------------------------------------------------------------------------------------
// RNG (rng.hpp): SplitMix64's output function over a counter-addressed Weyl sequence.
// ----------------------------------------------------------------------------------------
extern "C" uint64_t chs_rng_u64(uint64_t seed, uint64_t stream, uint64_t k) {
  return chrng::u64(seed, stream, k);
}

static inline double u01(uint64_t seed, uint64_t stream, uint64_t k) { return chrng::u01(seed, stream, k); }

extern "C" void chs_rng_uniform(uint64_t seed, uint64_t stream, int64_t offset, int64_t n,
                                double lo, double hi, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * u01(seed, stream, offset + i);
}

extern "C" void chs_rng_normal(uint64_t seed, uint64_t stream, int64_t offset, int64_t n,
                               double *out) {
  const double two_pi = 6.283185307179586476925286766559;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int64_t g = offset + i;
    int64_t pair = g >> 1;
    double a = u01(seed, stream, 2 * pair), b = u01(seed, stream, 2 * pair + 1);
    double rad = std::sqrt(-2.0 * std::log1p(-a));  // 1-a in (0,1]
    out[i] = (g & 1) ? rad * std::sin(two_pi * b) : rad * std::cos(two_pi * b);
  }
}

extern "C" void chs_params_default(chs_params *p) {
  p->nlev_max = 10;
  p->max_coarse = 1000;
  p->soc_tau = 0.25;
  p->nstep = 15;
  p->step_size = 1;
  p->fsai_eps = 1e-3;
  p->nupre = 1;
  p->nupost = 1;
  p->smoother = 0;
  p->prolong = 0;
  p->nthreads = 0;
  p->seed = 20240607ULL;
  p->verbose = 0;
  p->afsai_fn = nullptr;
  p->afsai_ctx = nullptr;
  p->omega_fn = nullptr;
  p->omega_ctx = nullptr;
  p->soc_type = 0;
  p->ntv = 6;
  p->ntv0 = 0;
  p->tv = nullptr;
  p->bamg.itmax_vol = 100;
  p->bamg.tol_vol = 0.01;
  p->bamg.dist_min = 1;
  p->bamg.dist_max = 10;
  p->bamg.maxcond = 1e13;
  p->bamg.maxrownrm = 5.0;
  p->bamg.eps_prol = 1e-10;
  p->prol_smooth = 0;
  p->galerkin_fn = nullptr;
  p->galerkin_ctx = nullptr;
  p->nl_agg = 0;
  p->emin.patt_pow = 2;
  p->emin.patt_tau = 0.01;
  p->emin.nnzr_max = 30;
  p->emin.itmax = 10;
  p->emin.energ_tol = 1e-3;
  p->emin.condmax = 1e10;
  p->emin_kp.open = nullptr;
  p->emin_kp.apply = nullptr;
  p->emin_kp.close = nullptr;
  p->emin_kp.ctx = nullptr;
}

extern "C" const char *chs_last_error(void) { return g_err.c_str(); }

// ----------------------------------------------------------------------------------------
// Opaque matrix
// ----------------------------------------------------------------------------------------
static int check_csr(int32_t nr, int32_t nc, const int32_t *rp, const int32_t *ci) {
  if (nr < 0 || nc < 0 || !rp) return fail(CHS_EARG, "csr: bad dimensions or null rowptr");
  if (rp[0] != 0) return fail(CHS_EARG, "csr: rowptr[0] != 0");
  for (int32_t i = 0; i < nr; ++i) {
    if (rp[i + 1] < rp[i]) return fail(CHS_EARG, "csr: rowptr not monotone");
    for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
      if (ci[e] < 0 || ci[e] >= nc) return fail(CHS_EARG, "csr: column index out of range");
      if (e > rp[i] && ci[e] <= ci[e - 1])
        return fail(CHS_EARG, "csr: columns not strictly increasing within a row");
    }
  }
  return CHS_OK;
}

extern "C" int chs_mat_from_csr(int32_t nrows, int32_t ncols, const int32_t *rowptr,
                                const int32_t *col, const double *val, chs_mat **out) {
  if (!out) return fail(CHS_EARG, "null out");
  int rc = check_csr(nrows, ncols, rowptr, col);
  if (rc) return rc;
  try {
    auto *m = new chs_mat;
    m->nrows = nrows;
    m->ncols = ncols;
    m->rp.assign(rowptr, rowptr + nrows + 1);
    m->ci.assign(col, col + rowptr[nrows]);
    m->v.assign(val, val + rowptr[nrows]);
    *out = m;
  } catch (std::bad_alloc &) {
    return fail(CHS_EALLOC, "allocation failed");
  }
  return CHS_OK;
}

extern "C" int chs_mat_dims(const chs_mat *m, int32_t *nrows, int32_t *ncols, int64_t *nnz) {
  if (!m) return fail(CHS_EARG, "null matrix");
  if (nrows) *nrows = m->nrows;
  if (ncols) *ncols = m->ncols;
  if (nnz) *nnz = m->nnz();
  return CHS_OK;
}

extern "C" int chs_mat_copy(const chs_mat *m, int32_t *rowptr, int32_t *col, double *val) {
  if (!m) return fail(CHS_EARG, "null matrix");
  if (rowptr) std::memcpy(rowptr, m->rp.data(), sizeof(int32_t) * (m->nrows + 1));
  if (col) std::memcpy(col, m->ci.data(), sizeof(int32_t) * m->nnz());
  if (val) std::memcpy(val, m->v.data(), sizeof(double) * m->nnz());
  return CHS_OK;
}

extern "C" void chs_mat_free(chs_mat *m) { delete m; }

// ----------------------------------------------------------------------------------------
// Generators
// ----------------------------------------------------------------------------------------
extern "C" int chs_gen_poisson7(int32_t N, chs_mat **out) {
  if (N <= 0 || (int64_t)N * N * N > (1LL << 30)) return fail(CHS_EARG, "bad N");
  auto *m = new (std::nothrow) chs_mat;
  if (!m) return fail(CHS_EALLOC, "alloc");
  const int64_t n = (int64_t)N * N * N;
  m->nrows = m->ncols = (int32_t)n;
  m->rp.resize(n + 1);
  // Row lengths: 1 + number of in-grid neighbours.
  m->rp[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int x = i % N, y = (i / N) % N, z = (int)(i / ((int64_t)N * N));
    int deg = (x > 0) + (x < N - 1) + (y > 0) + (y < N - 1) + (z > 0) + (z < N - 1);
    m->rp[i + 1] = m->rp[i] + 1 + deg;
  }
  m->ci.resize(m->rp[n]);
  m->v.resize(m->rp[n]);
  const int64_t NN = (int64_t)N * N;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int x = i % N, y = (i / N) % N, z = (int)(i / NN);
    int32_t e = m->rp[i];
    auto put = [&](int64_t j, double a) { m->ci[e] = (int32_t)j; m->v[e] = a; ++e; };
    if (z > 0) put(i - NN, -1.0);
    if (y > 0) put(i - N, -1.0);
    if (x > 0) put(i - 1, -1.0);
    put(i, 6.0);
    if (x < N - 1) put(i + 1, -1.0);
    if (y < N - 1) put(i + N, -1.0);
    if (z < N - 1) put(i + NN, -1.0);
  }
  *out = m;
  return CHS_OK;
}

extern "C" void chs_gen_lnk(int32_t N, double sigma, uint64_t seed, double *lnk) {
  const int64_t n = (int64_t)N * N * N;
  chs_rng_normal(seed, 1, 0, n, lnk);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) lnk[i] *= sigma;
}

extern "C" int chs_gen_diffusion(int32_t N, const double *k, chs_mat **out) {
  if (N <= 0 || !k || (int64_t)N * N * N > (1LL << 30)) return fail(CHS_EARG, "bad N or k");
  auto *m = new (std::nothrow) chs_mat;
  if (!m) return fail(CHS_EALLOC, "alloc");
  const int64_t n = (int64_t)N * N * N, NN = (int64_t)N * N;
  m->nrows = m->ncols = (int32_t)n;
  m->rp.resize(n + 1);
  m->rp[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int x = i % N, y = (i / N) % N, z = (int)(i / NN);
    int deg = (x > 0) + (x < N - 1) + (y > 0) + (y < N - 1) + (z > 0) + (z < N - 1);
    m->rp[i + 1] = m->rp[i] + 1 + deg;
  }
  m->ci.resize(m->rp[n]);
  m->v.resize(m->rp[n]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int x = i % N, y = (i / N) % N, z = (int)(i / NN);
    const double ki = k[i];
    auto T = [&](int64_t j) { return 2.0 * ki * k[j] / (ki + k[j]); };
    double diag = 0.0;
    int32_t e = m->rp[i], ediag = -1;
    auto put = [&](int64_t j) {
      double t = T(j);
      m->ci[e] = (int32_t)j;
      m->v[e] = -t;
      diag += t;
      ++e;
    };
    if (z > 0) put(i - NN);
    if (y > 0) put(i - N);
    if (x > 0) put(i - 1);
    ediag = e++;
    if (x < N - 1) put(i + 1);
    if (y < N - 1) put(i + N);
    if (z < N - 1) put(i + NN);
    if (x == 0) diag += 2.0 * ki;      // Dirichlet half cell on the x = 0 face
    if (x == N - 1) diag += 2.0 * ki;  // Dirichlet half cell on the x = 1 face
    m->ci[ediag] = (int32_t)i;
    m->v[ediag] = diag;
  }
  *out = m;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Linear elasticity on trilinear hexahedra (Q1), unit cube of ne^3 elements, (ne+1)^3
// nodes with 3 displacement dofs each (dof 3*node + component, node = ix + m*(iy + m*iz),
// m = ne + 1).  Isotropic material (E, nu): D in Voigt order (xx, yy, zz, yz, xz, xy) with
// engineering shears; the element matrix is the 2x2x2 Gauss rule of B^T D B, the same for
// every element, and rows are assembled node by node over the (up to) 8 elements around the
// node, so the pattern holds all 27 neighbour nodes x 3 components (81 per interior row).
// A uniform body force -1 in z gives the load, each element contributing a eighth of its
// volume to the z dof of each of its nodes.  clamp != 0: the dofs of the z = 0 face become
// unit rows and columns (zero load), as the reference's elasticity tests keep them.
// ----------------------------------------------------------------------------------------
namespace {

void q1_element_matrix(double h, double E, double nu, double Ke[24][24]) {
  const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double D[6][6] = {};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i][j] = c * (i == j ? 1.0 - nu : nu);
  for (int i = 3; i < 6; ++i) D[i][i] = c * (1.0 - 2.0 * nu) / 2.0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) Ke[i][j] = 0.0;
  const double g = 1.0 / std::sqrt(3.0);
  const double detJ = (h / 2.0) * (h / 2.0) * (h / 2.0);
  for (int gp = 0; gp < 8; ++gp) {
    const double xi = (gp & 1) ? g : -g, eta = (gp & 2) ? g : -g, zeta = (gp & 4) ? g : -g;
    double dN[8][3];
    for (int a = 0; a < 8; ++a) {
      const double sa = (a & 1) ? 1.0 : -1.0, ta = (a & 2) ? 1.0 : -1.0, ua = (a & 4) ? 1.0 : -1.0;
      // dN/dx = (2/h) dN/dxi
      dN[a][0] = (2.0 / h) * 0.125 * sa * (1.0 + eta * ta) * (1.0 + zeta * ua);
      dN[a][1] = (2.0 / h) * 0.125 * ta * (1.0 + xi * sa) * (1.0 + zeta * ua);
      dN[a][2] = (2.0 / h) * 0.125 * ua * (1.0 + xi * sa) * (1.0 + eta * ta);
    }
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
      const double bx = dN[a][0], by = dN[a][1], bz = dN[a][2];
      B[0][3 * a] = bx;
      B[1][3 * a + 1] = by;
      B[2][3 * a + 2] = bz;
      B[3][3 * a + 1] = bz;
      B[3][3 * a + 2] = by;
      B[4][3 * a] = bz;
      B[4][3 * a + 2] = bx;
      B[5][3 * a] = by;
      B[5][3 * a + 1] = bx;
    }
    double DB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 24; ++j) {
        double t = 0.0;
        for (int k = 0; k < 6; ++k) t += D[i][k] * B[k][j];
        DB[i][j] = t;
      }
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double t = 0.0;
        for (int k = 0; k < 6; ++k) t += B[k][i] * DB[k][j];
        Ke[i][j] += t * detJ;
      }
  }
}

}  // namespace

extern "C" int chs_gen_elasticity(int32_t ne, double E, double nu, int32_t clamp, chs_mat **out,
                                  double *rhs) {
  if (ne <= 0 || !out || E <= 0.0 || nu <= -1.0 || nu >= 0.5) return fail(CHS_EARG, "elasticity: arguments");
  const int64_t m = (int64_t)ne + 1, nnode = m * m * m, n = 3 * nnode;
  if (n > INT32_MAX / 81) return fail(CHS_EARG, "elasticity: too large");
  const double h = 1.0 / ne;
  double Ke[24][24];
  q1_element_matrix(h, E, nu, Ke);
  auto *mat = new (std::nothrow) chs_mat;
  if (!mat) return fail(CHS_EALLOC, "alloc");
  mat->nrows = mat->ncols = (int32_t)n;
  auto clamped = [&](int64_t node) { return clamp && node / (m * m) == 0; };
  // Row lengths: 3 per existing neighbour node (clamped neighbours dropped), 1 for clamped rows.
  mat->rp.assign(n + 1, 0);
  for (int64_t node = 0; node < nnode; ++node) {
    const int64_t ix = node % m, iy = (node / m) % m, iz = node / (m * m);
    int32_t cnt = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
          if (jx < 0 || jy < 0 || jz < 0 || jx >= m || jy >= m || jz >= m) continue;
          if (!clamped(jx + m * (jy + m * jz))) cnt += 3;
        }
    for (int c = 0; c < 3; ++c) mat->rp[3 * node + c + 1] = clamped(node) ? 1 : cnt;
  }
  for (int64_t i = 0; i < n; ++i) mat->rp[i + 1] += mat->rp[i];
  mat->ci.resize(mat->rp[n]);
  mat->v.resize(mat->rp[n]);
#pragma omp parallel for schedule(static)
  for (int64_t node = 0; node < nnode; ++node) {
    const int64_t ix = node % m, iy = (node / m) % m, iz = node / (m * m);
    if (clamped(node)) {
      for (int c = 0; c < 3; ++c) {
        mat->ci[mat->rp[3 * node + c]] = (int32_t)(3 * node + c);
        mat->v[mat->rp[3 * node + c]] = 1.0;
      }
      continue;
    }
    double blk[27][3][3] = {};
    bool has[27] = {};
    // Elements around the node: corner (ix - i, iy - j, iz - k), the node is local corner a.
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          const int64_t ex = ix - i, ey = iy - j, ez = iz - k;
          if (ex < 0 || ey < 0 || ez < 0 || ex >= ne || ey >= ne || ez >= ne) continue;
          const int a = i + 2 * j + 4 * k;
          for (int b = 0; b < 8; ++b) {
            const int bi = b & 1, bj = (b >> 1) & 1, bk = (b >> 2) & 1;
            const int slot = (bi - i + 1) + 3 * (bj - j + 1) + 9 * (bk - k + 1);
            has[slot] = true;
            for (int c = 0; c < 3; ++c)
              for (int d = 0; d < 3; ++d) blk[slot][c][d] += Ke[3 * a + c][3 * b + d];
          }
        }
    for (int c = 0; c < 3; ++c) {
      int64_t e = mat->rp[3 * node + c];
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
            if (jx < 0 || jy < 0 || jz < 0 || jx >= m || jy >= m || jz >= m) continue;
            const int64_t jn = jx + m * (jy + m * jz);
            if (clamped(jn)) continue;
            const int slot = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
            for (int d = 0; d < 3; ++d) {
              mat->ci[e] = (int32_t)(3 * jn + d);
              mat->v[e] = has[slot] ? blk[slot][c][d] : 0.0;
              ++e;
            }
          }
    }
  }
  if (rhs) {
    const double f = -h * h * h / 8.0;  // body force -1 in z, an eighth of each element volume
#pragma omp parallel for schedule(static)
    for (int64_t node = 0; node < nnode; ++node) {
      const int64_t ix = node % m, iy = (node / m) % m, iz = node / (m * m);
      const int64_t nel = (int64_t)(1 + (ix > 0 && ix < ne)) * (1 + (iy > 0 && iy < ne)) * (1 + (iz > 0 && iz < ne));
      rhs[3 * node] = 0.0;
      rhs[3 * node + 1] = 0.0;
      rhs[3 * node + 2] = clamped(node) ? 0.0 : f * (double)nel;
    }
  }
  *out = mat;
  return CHS_OK;
}

// Rigid-body modes of the (ne+1)^3 nodes (sources/utils/mk_rbm_3d.m:1-27): three translations and three
// rotations about the centroid, column-major n x 6.
extern "C" void chs_gen_rbm(int32_t ne, double *rbm) {
  const int64_t m = (int64_t)ne + 1, nnode = m * m * m, n = 3 * nnode;
  const double h = 1.0 / ne, xb = 0.5;  // the centroid of the unit cube's nodes
  for (int64_t q = 0; q < 6 * n; ++q) rbm[q] = 0.0;
  for (int64_t node = 0; node < nnode; ++node) {
    const double x = (node % m) * h - xb, y = ((node / m) % m) * h - xb, z = (node / (m * m)) * h - xb;
    const int64_t d = 3 * node;
    rbm[d + 0 * n] = 1.0;
    rbm[d + 1 + 1 * n] = 1.0;
    rbm[d + 2 + 2 * n] = 1.0;
    rbm[d + 1 + 3 * n] = -z;
    rbm[d + 2 + 3 * n] = y;
    rbm[d + 0 + 4 * n] = z;
    rbm[d + 2 + 4 * n] = -x;
    rbm[d + 0 + 5 * n] = -y;
    rbm[d + 1 + 5 * n] = x;
  }
}

// ----------------------------------------------------------------------------------------
// Sparse helpers
// ----------------------------------------------------------------------------------------
static void transpose_into(const chs_mat &M, chs_mat &T) {
  T.nrows = M.ncols;
  T.ncols = M.nrows;
  T.rp.assign((size_t)T.nrows + 1, 0);
  for (int32_t c : M.ci) T.rp[c + 1]++;
  for (int32_t i = 0; i < T.nrows; ++i) T.rp[i + 1] += T.rp[i];
  T.ci.resize(M.ci.size());
  T.v.resize(M.v.size());
  std::vector<int32_t> pos(T.rp.begin(), T.rp.end() - 1);
  // Row order of M is ascending, so columns of T come out sorted.
  for (int32_t i = 0; i < M.nrows; ++i)
    for (int32_t e = M.rp[i]; e < M.rp[i + 1]; ++e) {
      int32_t p = pos[M.ci[e]]++;
      T.ci[p] = i;
      T.v[p] = M.v[e];
    }
}

static void spmv(const chs_mat &M, const double *x, double *y, int nthreads = 0) {
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthreads))
  for (int32_t i = 0; i < M.nrows; ++i) {
    double s = 0.0;
    for (int32_t e = M.rp[i]; e < M.rp[i + 1]; ++e) s += M.v[e] * x[M.ci[e]];
    y[i] = s;
  }
}

// Gustavson SpGEMM C = A*B, rows in parallel, exact zeros dropped, columns sorted.
static void spgemm(const chs_mat &A, const chs_mat &B, chs_mat &C, int nthreads) {
  C.nrows = A.nrows;
  C.ncols = B.ncols;
  C.rp.assign((size_t)A.nrows + 1, 0);
  std::vector<std::vector<int32_t>> rc(A.nrows);
  std::vector<std::vector<double>> rv(A.nrows);
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<int32_t> mark(B.ncols, -1);
    std::vector<double> acc(B.ncols, 0.0);
    std::vector<int32_t> cols;
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < A.nrows; ++i) {
      cols.clear();
      for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
        const int32_t k = A.ci[e];
        const double a = A.v[e];
        for (int32_t f = B.rp[k]; f < B.rp[k + 1]; ++f) {
          const int32_t j = B.ci[f];
          if (mark[j] != i) {
            mark[j] = i;
            acc[j] = a * B.v[f];
            cols.push_back(j);
          } else {
            acc[j] += a * B.v[f];
          }
        }
      }
      std::sort(cols.begin(), cols.end());
      auto &oc = rc[i];
      auto &ov = rv[i];
      oc.reserve(cols.size());
      ov.reserve(cols.size());
      for (int32_t j : cols)
        if (acc[j] != 0.0) {
          oc.push_back(j);
          ov.push_back(acc[j]);
        }
      C.rp[i + 1] = (int32_t)oc.size();
    }
  }
  for (int32_t i = 0; i < A.nrows; ++i) C.rp[i + 1] += C.rp[i];
  C.ci.resize(C.rp[A.nrows]);
  C.v.resize(C.rp[A.nrows]);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int32_t i = 0; i < A.nrows; ++i) {
    std::copy(rc[i].begin(), rc[i].end(), C.ci.begin() + C.rp[i]);
    std::copy(rv[i].begin(), rv[i].end(), C.v.begin() + C.rp[i]);
  }
}

extern "C" int chs_transpose(const chs_mat *M, chs_mat **MT) {
  if (!M || !MT) return fail(CHS_EARG, "null");
  auto *t = new (std::nothrow) chs_mat;
  if (!t) return fail(CHS_EALLOC, "alloc");
  transpose_into(*M, *t);
  *MT = t;
  return CHS_OK;
}

extern "C" int chs_galerkin(const chs_mat *A, const chs_mat *P, int nthreads, chs_mat **Ac) {
  if (!A || !P || !Ac || A->ncols != P->nrows) return fail(CHS_EARG, "galerkin: shapes");
  nthreads = nthreads_or_default(nthreads);
  chs_mat AP, PT;
  spgemm(*A, *P, AP, nthreads);
  transpose_into(*P, PT);
  auto *c = new (std::nothrow) chs_mat;
  if (!c) return fail(CHS_EALLOC, "alloc");
  spgemm(PT, AP, *c, nthreads);
  *Ac = c;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Classical strength of connection (cpt_SoC_cla.m:100-152).
// The measure of entry (i,j) is a_ij / max(m_i, m_j), with m_c the minimum of column c of
// the off-diagonal part including its implicit zeros, replaced by 1 where it is zero; if no
// column has a negative off-diagonal entry the absolute-value variant is used instead.
// ----------------------------------------------------------------------------------------
extern "C" int chs_soc_cla(const chs_mat *Ap, double tau, int8_t *strong) {
  if (!Ap || !strong) return fail(CHS_EARG, "null");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  std::vector<double> mincol(A.ncols, 0.0), maxabs(A.ncols, 0.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
      const int32_t j = A.ci[e];
      if (j == i) continue;
      mincol[j] = std::min(mincol[j], A.v[e]);
      maxabs[j] = std::max(maxabs[j], std::fabs(A.v[e]));
    }
  // The off-diagonal part always holds an implicit zero on the diagonal position, so the
  // column minimum (and maximum magnitude) includes 0: hence the 0 start values above.
  bool any_neg = false;
  for (int32_t j = 0; j < A.ncols; ++j) any_neg |= (mincol[j] < 0.0);
  if (any_neg) {
    for (auto &m : mincol)
      if (m == 0.0) m = 1.0;
  } else {
    for (auto &m : maxabs)
      if (m == 0.0) m = 1.0;
  }
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
      const int32_t j = A.ci[e];
      double s = any_neg ? A.v[e] / std::max(mincol[i], mincol[j])
                         : std::fabs(A.v[e]) / std::max(maxabs[i], maxabs[j]);
      strong[e] = (j == i) ? 1 : (s > tau ? 1 : -1);
    }
  return CHS_OK;
}

// Strength by diagonal dominance (cpt_SoC_dom.m:26-52, tau >= 0): s_ij = |a_ij| /
// sqrt(a_ii a_jj); entries with s_ij <= tau are weak.  The diagonal (s_ii = 1) is kept.
extern "C" int chs_soc_dom(const chs_mat *Ap, double tau, int8_t *strong) {
  if (!Ap || !strong) return fail(CHS_EARG, "null");
  if (tau < 0.0) return fail(CHS_EARG, "soc_dom: tau < 0 (kept-entry count) is not supported");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  std::vector<double> d(n, 0.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.ci[e] == i) d[i] = A.v[e];
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
      const int32_t j = A.ci[e];
      const double sij = std::fabs(A.v[e]) / std::sqrt(d[i] * d[j]);
      strong[e] = (j == i) ? 1 : (sij > tau ? 1 : -1);
    }
  return CHS_OK;
}

extern "C" int chs_prol_bamg(const chs_mat *Ap, const int8_t *strong, const int32_t *fcnode,
                             int32_t nc, int32_t ntv, const double *tv, const chs_bamg_params *bp,
                             int nthreads, chs_mat **Pout, int32_t *c_mark, int32_t *nmarked) {
  if (!Ap || !strong || !fcnode || !tv || !bp || !Pout) return fail(CHS_EARG, "null");
  if (ntv < 1 || bp->dist_min < 0 || bp->dist_max < bp->dist_min)
    return fail(CHS_EARG, "bamg: arguments");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  nthreads = nthreads_or_default(nthreads);
  // Strength pattern with the diagonal, in A's (ascending) column order.
  std::vector<int32_t> srp(n + 1, 0), sci;
  sci.reserve(A.nnz());
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (strong[e] > 0) sci.push_back(A.ci[e]);
    srp[i + 1] = (int32_t)sci.size();
  }
  std::vector<double> tvr((size_t)n * ntv);  // row-major copy (cpt_Prolongation_BAMG.cpp:43-54)
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < ntv; ++j) tvr[(size_t)i * ntv + j] = tv[(size_t)j * n + i];
  chsb::Params prm{bp->itmax_vol, bp->tol_vol, bp->dist_min, bp->dist_max, ntv,
                   bp->maxcond, bp->maxrownrm, bp->eps_prol};
  std::vector<int32_t> mark(n, 0);
  auto P = std::make_unique<chs_mat>();
  P->nrows = n;
  P->ncols = nc;
  std::string err;
  int rc = chsb::prolong(n, srp.data(), sci.data(), fcnode, ntv, tvr.data(), prm, nthreads, P->rp,
                         P->ci, P->v, mark.data(), err);
  if (rc) return fail(CHS_ENUM, err);
  int32_t nm = 0;
  for (int32_t i = 0; i < n; ++i) nm += mark[i];
  if (c_mark) std::copy(mark.begin(), mark.end(), c_mark);
  if (nmarked) *nmarked = nm;
  *Pout = P.release();
  return CHS_OK;
}

namespace {
chse::Params emin_params(const chs_emin_params &e) {
  return chse::Params{e.patt_pow, e.patt_tau, e.nnzr_max, e.itmax, e.energ_tol, e.condmax};
}
}  // namespace

extern "C" int chs_emin_pattern(const chs_mat *Ap, const chs_mat *Pp, const int32_t *fcnode, int32_t ntv,
                                const chs_emin_params *prm, int nthreads, chs_mat **patt) {
  if (!Ap || !Pp || !fcnode || !prm || !patt) return fail(CHS_EARG, "null");
  if (Pp->nrows != Ap->nrows || Ap->nrows != Ap->ncols || ntv < 1)
    return fail(CHS_EARG, "emin pattern: dimensions");
  auto T = std::make_unique<chs_mat>();
  T->nrows = Pp->nrows;
  T->ncols = Pp->ncols;
  chse::pattern(Ap->nrows, Pp->ncols, Ap->rp.data(), Ap->ci.data(), Ap->v.data(), Pp->rp.data(),
                Pp->ci.data(), Pp->v.data(), fcnode, ntv, emin_params(*prm),
                nthreads_or_default(nthreads), T->rp, T->ci);
  T->v.assign(T->ci.size(), 1.0);
  *patt = T.release();
  return CHS_OK;
}

namespace {
// EMIN on (A, P, pattern), with the CG operator from ops when ops->open accepts the level.
int emin_improve(const chs_mat *Ap, const chs_mat *Pp, const chs_mat *patt, const int32_t *fcnode,
                 int32_t ntv, const double *tv, const chs_emin_params *prm, const chs_emin_kp_ops *ops,
                 int nthreads, chs_mat **Pout, int32_t *iter) {
  if (!Ap || !Pp || !patt || !fcnode || !tv || !prm || !Pout) return fail(CHS_EARG, "null");
  if (Pp->nrows != Ap->nrows || Ap->nrows != Ap->ncols || patt->nrows != Pp->nrows ||
      patt->ncols != Pp->ncols || ntv < 1)
    return fail(CHS_EARG, "emin: dimensions");
  void *h = nullptr;
  if (ops && ops->open && ops->apply &&
      ops->open(ops->ctx, Ap->nrows, Pp->ncols, Ap->rp.data(), Ap->ci.data(), Ap->v.data(), patt->rp.data(),
                patt->ci.data(), &h) != 0)
    h = nullptr;  // declined: the host product
  std::function<int(const double *, double *)> ext;
  if (h) ext = [&](const double *x, double *y) { return ops->apply(h, x, y); };
  auto P = std::make_unique<chs_mat>();
  P->nrows = Pp->nrows;
  P->ncols = Pp->ncols;
  int32_t it = 0;
  std::string err;
  const int rc = chse::improve(Ap->nrows, Pp->ncols, Ap->rp.data(), Ap->ci.data(), Ap->v.data(),
                               Pp->rp.data(), Pp->ci.data(), Pp->v.data(), patt->rp.data(),
                               patt->ci.data(), fcnode, ntv, tv, emin_params(*prm),
                               nthreads_or_default(nthreads), P->rp, P->ci, P->v, it, err, ext);
  if (h && ops->close) ops->close(h);
  if (rc == 1) return fail(CHS_EARG, err);
  if (rc) return fail(CHS_ENUM, err);
  if (iter) *iter = it;
  *Pout = P.release();
  return CHS_OK;
}
}  // namespace

extern "C" int chs_emin_improve(const chs_mat *Ap, const chs_mat *Pp, const chs_mat *patt,
                                const int32_t *fcnode, int32_t ntv, const double *tv,
                                const chs_emin_params *prm, int nthreads, chs_mat **Pout,
                                int32_t *iter) {
  return emin_improve(Ap, Pp, patt, fcnode, ntv, tv, prm, nullptr, nthreads, Pout, iter);
}

extern "C" int chs_emin_improve_ops(const chs_mat *Ap, const chs_mat *Pp, const chs_mat *patt,
                                    const int32_t *fcnode, int32_t ntv, const double *tv,
                                    const chs_emin_params *prm, const chs_emin_kp_ops *ops,
                                    int nthreads, chs_mat **Pout, int32_t *iter) {
  return emin_improve(Ap, Pp, patt, fcnode, ntv, tv, prm, ops, nthreads, Pout, iter);
}

extern "C" int chs_orthonormalize(int64_t n, int32_t k, double *V) {
  if (!V || n < 0 || k < 0) return fail(CHS_EARG, "orthonormalize: arguments");
  return chsb::orthonormalize(n, k, V);
}

// ----------------------------------------------------------------------------------------
// Dirichlet marking (mark_isolated_nodes.m) + sequential PMIS (cpt_PMIS.m:174-266).
// S is the strong off-diagonal pattern.  A node's weight is the number of nodes it
// strongly influences (column count of S) plus its tie-break draw.  Free nodes are visited
// in list order; a node becomes coarse when no free strong neighbour (in S + S^T) has a
// larger weight, and then the nodes it strongly influences become fine.  Removal swaps the
// last list entry into the visited slot, which is then visited next.
// ----------------------------------------------------------------------------------------
static int pmis_core(const chs_mat *Ap, const int8_t *strong, const double *rand01, bool dirichlet,
                     int32_t *fcnode, int32_t *nc_out);

extern "C" int chs_pmis(const chs_mat *Ap, const int8_t *strong, const double *rand01,
                        int32_t *fcnode, int32_t *nc_out) {
  return pmis_core(Ap, strong, rand01, true, fcnode, nc_out);
}

// dirichlet: mark the nodes whose column of A holds one entry as fine first
// (mark_isolated_nodes.m); the second pass of aggressive coarsening starts from all zeros.
static int pmis_core(const chs_mat *Ap, const int8_t *strong, const double *rand01, bool dirichlet,
                     int32_t *fcnode, int32_t *nc_out) {
  if (!Ap || !strong || !rand01 || !fcnode) return fail(CHS_EARG, "null");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  if (A.ncols != n) return fail(CHS_EARG, "pmis: square matrix required");
  // Column counts of A (Dirichlet marking) and of S (influence weights).
  std::vector<int32_t> colcnt(n, 0), wcnt(n, 0);
  chs_mat S;  // strong off-diagonal pattern, rows
  S.nrows = S.ncols = n;
  S.rp.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
      colcnt[A.ci[e]]++;
      if (A.ci[e] != i && strong[e] > 0) S.rp[i + 1]++;
    }
  for (int32_t i = 0; i < n; ++i) S.rp[i + 1] += S.rp[i];
  S.ci.resize(S.rp[n]);
  S.v.assign(S.rp[n], 1.0);
  for (int32_t i = 0, p = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.ci[e] != i && strong[e] > 0) {
        S.ci[p++] = A.ci[e];
        wcnt[A.ci[e]]++;
      }
  chs_mat ST;
  transpose_into(S, ST);  // row j of ST = nodes i with S(i,j) != 0 (column j of S)
  // Symmetrised neighbourhood S + S^T, per node, merged from the sorted rows.
  std::vector<int32_t> symrp(n + 1, 0);
  std::vector<int32_t> symci;
  symci.reserve(2 * S.ci.size());
  for (int32_t i = 0; i < n; ++i) {
    int32_t a = S.rp[i], ae = S.rp[i + 1], b = ST.rp[i], be = ST.rp[i + 1];
    while (a < ae || b < be) {
      int32_t ca = a < ae ? S.ci[a] : INT32_MAX, cb = b < be ? ST.ci[b] : INT32_MAX;
      int32_t c = std::min(ca, cb);
      symci.push_back(c);
      if (ca == c) ++a;
      if (cb == c) ++b;
    }
    symrp[i + 1] = (int32_t)symci.size();
  }
  std::vector<double> w(n), wfree(n);
  for (int32_t i = 0; i < n; ++i) {
    fcnode[i] = (dirichlet && colcnt[i] == 1) ? -1 : 0;  // isolated (Dirichlet) nodes are fine
    w[i] = (double)wcnt[i] + rand01[i];
    wfree[i] = fcnode[i] == -1 ? -1.0 : w[i];
  }
  std::vector<int32_t> list;
  list.reserve(n);
  for (int32_t i = 0; i < n; ++i)
    if (fcnode[i] == 0) list.push_back(i);
  int64_t len = (int64_t)list.size();
  int64_t ind = 0;
  while (len > 0) {
    const int32_t inod = list[ind];
    if (fcnode[inod] != 0) {
      list[ind] = list[len - 1];
      --len;
    } else {
      bool is_coarse = true;
      const double wi = w[inod];
      for (int32_t e = symrp[inod]; e < symrp[inod + 1]; ++e)
        if (wfree[symci[e]] > wi) {
          is_coarse = false;
          break;
        }
      if (!is_coarse) {
        ++ind;
      } else {
        for (int32_t e = ST.rp[inod]; e < ST.rp[inod + 1]; ++e) {
          fcnode[ST.ci[e]] = -1;
          wfree[ST.ci[e]] = -1.0;
        }
        fcnode[inod] = 1;
        wfree[inod] = -1.0;
        list[ind] = list[len - 1];
        --len;
      }
    }
    if (ind >= len) ind = 0;
  }
  int32_t nc = 0;
  for (int32_t i = 0; i < n; ++i) fcnode[i] = fcnode[i] > 0 ? nc++ : -1;
  if (nc_out) *nc_out = nc;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Classical interpolation (ProlStripe_Classical.cpp:44-160).  Coarse rows are e_c.  A fine
// row i interpolates from its strong coarse neighbours C_i with
//   w_ij = -(a_ij + sum_{k in FS_i, s_k != 0} a_ik * min(0,a_kj) / s_k) / d_i,
//   s_k  = sum_{l in C_i} min(0, a_kl),
//   d_i  = a_ii + sum_{weak j} a_ij + sum_{k in FS_i, |s_k| <= eps*a_ii} a_ik,
// FS_i being the strong fine neighbours.  Accumulation orders follow the row orders of A
// so that the result is reproducible bit for bit.
// ----------------------------------------------------------------------------------------
extern "C" int chs_prol_classical(const chs_mat *Ap, const int8_t *strong,
                                  const int32_t *fcnode, int32_t nc, int nthreads,
                                  chs_mat **Pout) {
  if (!Ap || !strong || !fcnode || !Pout) return fail(CHS_EARG, "null");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  nthreads = nthreads_or_default(nthreads);
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<std::vector<int32_t>> rc(n);
  std::vector<std::vector<double>> rv(n);
  int bad = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : bad)
  {
    std::vector<int32_t> slot(n, 0);  // 1-based position of a strong C neighbour, 0 if none
    std::vector<int32_t> cc, fs, pos;
    std::vector<double> cv, fv, akj, w;
#pragma omp for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) {
      if (fcnode[i] >= 0) {
        rc[i].push_back(fcnode[i]);
        rv[i].push_back(1.0);
        continue;
      }
      cc.clear(); cv.clear(); fs.clear(); fv.clear();
      double aii = 0.0, denom = 0.0;
      bool has_diag = false;
      for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
        const int32_t j = A.ci[e];
        if (j == i) {
          aii = A.v[e];
          denom += aii;
          has_diag = true;
        } else if (strong[e] > 0) {
          if (fcnode[j] >= 0) {
            cc.push_back(j);
            cv.push_back(A.v[e]);
            slot[j] = (int32_t)cc.size();
          } else {
            fs.push_back(j);
            fv.push_back(A.v[e]);
          }
        } else {
          denom += A.v[e];
        }
      }
      if (!has_diag) bad++;
      w.assign(cv.begin(), cv.end());
      for (size_t k = 0; k < fs.size(); ++k) {
        const int32_t kn = fs[k];
        const double aik = fv[k];
        double ext = 0.0;
        akj.clear();
        pos.clear();
        for (int32_t f = A.rp[kn]; f < A.rp[kn + 1]; ++f) {
          const int32_t l = A.ci[f];
          if (slot[l] > 0) {
            const double a = std::min(0.0, A.v[f]);
            ext += a;
            akj.push_back(a);
            pos.push_back(slot[l] - 1);
          }
        }
        if (std::fabs(ext) > eps * aii) {
          for (size_t q = 0; q < akj.size(); ++q) w[pos[q]] += aik * akj[q] / ext;
        } else {
          denom += aik;
        }
      }
      // Map to coarse numbering; coarse indices increase with fine indices and cc is in
      // ascending fine order, so the row is already sorted by coarse column.
      auto &oc = rc[i];
      auto &ov = rv[i];
      oc.resize(cc.size());
      ov.resize(cc.size());
      for (size_t q = 0; q < cc.size(); ++q) {
        slot[cc[q]] = 0;
        oc[q] = fcnode[cc[q]];
        ov[q] = -w[q] / denom;
      }
    }
  }
  if (bad) return fail(CHS_EARG, "classical prolongation: fine row without a diagonal entry");
  auto *P = new (std::nothrow) chs_mat;
  if (!P) return fail(CHS_EALLOC, "alloc");
  P->nrows = n;
  P->ncols = nc;
  P->rp.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) P->rp[i + 1] = P->rp[i] + (int32_t)rc[i].size();
  P->ci.resize(P->rp[n]);
  P->v.resize(P->rp[n]);
  for (int32_t i = 0; i < n; ++i) {
    std::copy(rc[i].begin(), rc[i].end(), P->ci.begin() + P->rp[i]);
    std::copy(rv[i].begin(), rv[i].end(), P->v.begin() + P->rp[i]);
  }
  *Pout = P;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Extended+i interpolation (De Sterck et al., NLAA 15 (2008) 115-139), as the reference's
// EXTI core computes it (MEX_EXTI_Prol/ONELP_ProlStripe_EXTI.cpp:27-293, driven by
// EXTI_prolongation.cpp:52-199).  Coarse rows are e_c.  A fine row i starts its
// interpolatory set with its strong coarse neighbours (weights a_ij) and its denominator
// with a_ii plus the weak coarse and weak fine entries; every strong fine neighbour k then
// extends the set with k's strong coarse neighbours (distance two; a weak coarse neighbour
// of i that is reached this way joins with its a_ij, which leaves the denominator) and
// distributes a_ik over them in proportion to min(0, a_kj), i's own min(0, a_ki) share
// going to the denominator; when that extended sum vanishes a_ik is spread evenly.  The
// weights are -w/denom, sorted by coarse column.  Accumulations follow the row orders of A,
// so the result is reproducible bit for bit.
// ----------------------------------------------------------------------------------------
extern "C" int chs_prol_exti(const chs_mat *Ap, const int8_t *strong, const int32_t *fcnode,
                             int32_t nc, int nthreads, chs_mat **Pout) {
  if (!Ap || !strong || !fcnode || !Pout) return fail(CHS_EARG, "null");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  nthreads = nthreads_or_default(nthreads);
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<std::vector<int32_t>> rc(n);
  std::vector<std::vector<double>> rv(n);
#pragma omp parallel num_threads(nthreads)
  {
    // mark[j] > 0: j is interpolatory for the current row, at set position mark[j] - 1;
    // mark[j] < 0: j is a weak coarse neighbour whose entry is A.v[-mark[j] - 1]; 0: neither.
    std::vector<int32_t> mark(n, 0);
    std::vector<int32_t> set, weak, fs;
    std::vector<double> w, fsv;
#pragma omp for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) {
      if (fcnode[i] >= 0) {
        rc[i].push_back(fcnode[i]);
        rv[i].push_back(1.0);
        continue;
      }
      set.clear();
      w.clear();
      weak.clear();
      double aii = 0.0, denom = 0.0;
      for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
        const int32_t j = A.ci[e];
        if (j == i) {
          aii = A.v[e];
          denom += aii;
        } else if (fcnode[j] >= 0) {
          if (strong[e] > 0) {
            set.push_back(j);
            w.push_back(A.v[e]);
            mark[j] = (int32_t)set.size();
          } else {
            denom += A.v[e];
            mark[j] = -(e + 1);
            weak.push_back(j);
          }
        } else if (strong[e] < 0) {
          denom += A.v[e];
        }
      }
      for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
        const int32_t k = A.ci[e];
        if (strong[e] <= 0 || k == i || fcnode[k] >= 0) continue;
        const double aik = A.v[e];
        double aki = 0.0, ext = 0.0;
        fs.clear();
        fsv.clear();
        for (int32_t f = A.rp[k]; f < A.rp[k + 1]; ++f) {
          if (strong[f] <= 0) continue;
          const int32_t l = A.ci[f];
          if (l == i) {
            aki = std::min(0.0, A.v[f]);
            ext += aki;
          } else if (fcnode[l] >= 0) {
            const double akl = std::min(0.0, A.v[f]);
            fs.push_back(l);
            fsv.push_back(akl);
            ext += akl;
            if (mark[l] <= 0) {
              double init = 0.0;
              if (mark[l] < 0) {  // a weak coarse neighbour becomes interpolatory
                init = A.v[-mark[l] - 1];
                denom -= init;
              }
              set.push_back(l);
              w.push_back(init);
              mark[l] = (int32_t)set.size();
            }
          }
        }
        if (std::fabs(ext) > eps * std::fabs(aii)) {
          for (size_t q = 0; q < fs.size(); ++q) w[mark[fs[q]] - 1] += (aik * fsv[q]) / ext;
          denom += (aik * aki) / ext;
        } else {
          const double share = 1.0 / (double)(fs.size() + 1);
          denom += aik * share;
          for (size_t q = 0; q < fs.size(); ++q) w[mark[fs[q]] - 1] += aik * share;
        }
      }
      if (std::fabs(denom) < eps) denom = std::fabs(aii) > eps ? aii : 1.0;
      // Coarse numbering, weights, and the row sorted by coarse column.
      std::vector<std::pair<int32_t, double>> row(set.size());
      for (size_t q = 0; q < set.size(); ++q) {
        mark[set[q]] = 0;
        row[q] = {fcnode[set[q]], -w[q] / denom};
      }
      for (int32_t j : weak) mark[j] = 0;
      std::sort(row.begin(), row.end(),
                [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) { return a.first < b.first; });
      auto &oc = rc[i];
      auto &ov = rv[i];
      oc.resize(row.size());
      ov.resize(row.size());
      for (size_t q = 0; q < row.size(); ++q) {
        oc[q] = row[q].first;
        ov[q] = row[q].second;
      }
    }
  }
  auto *P = new (std::nothrow) chs_mat;
  if (!P) return fail(CHS_EALLOC, "alloc");
  P->nrows = n;
  P->ncols = nc;
  P->rp.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) P->rp[i + 1] = P->rp[i] + (int32_t)rc[i].size();
  P->ci.resize(P->rp[n]);
  P->v.resize(P->rp[n]);
  for (int32_t i = 0; i < n; ++i) {
    std::copy(rc[i].begin(), rc[i].end(), P->ci.begin() + P->rp[i]);
    std::copy(rv[i].begin(), rv[i].end(), P->v.begin() + P->rp[i]);
  }
  *Pout = P;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Adaptive symmetric FSAI (cpt_afsai_coef.cpp:316-554, compute_local_fsai.cpp:6-116).
//
// Row i grows a pattern P_i of columns j < i.  At every step the candidates are the
// columns j < i, j not in P_i, of row i of A and of the rows of A indexed by P_i; the
// candidate gradient is g_j = a_ij + sum_{k in P_i} x_k a_kj (x = current solution).  The
// step_size candidates of largest |g_j| join P_i (selection by the first-element-pivot
// partial partition below, which fixes how ties are broken), then A_PP x = -a_Pi is solved
// by Cholesky.  Growth stops after nstep steps, when the pattern does not grow, or when the
// Kaporin quantity a_Pi^T A_PP^-1 a_Pi changes by less than eps relative.  The row is
// s*[x; 1] with s = 1/sqrt(a_ii - a_Pi^T A_PP^-1 a_Pi), diagonal last.
// ----------------------------------------------------------------------------------------
namespace {

// Partial partition: on exit the first ncut entries of (val, idx) hold the ncut largest
// |val|.  Lomuto partition around the first element of the active range with a strict '>'
// comparison (cpt_afsai_coef.cpp:40-89 is the procedure whose tie behaviour this keeps).
void select_largest(int n, int ncut, double *val, int32_t *idx) {
  int first = 0, last = n - 1;
  if (ncut < 1 || ncut > n) return;
  const int target = ncut - 1;
  for (;;) {
    int mid = first;
    const double pivot = std::fabs(val[first]);
    for (int j = first + 1; j <= last; ++j)
      if (std::fabs(val[j]) > pivot) {
        ++mid;
        std::swap(val[mid], val[j]);
        std::swap(idx[mid], idx[j]);
      }
    std::swap(val[mid], val[first]);
    std::swap(idx[mid], idx[first]);
    if (mid == target) return;
    if (mid > target)
      last = mid - 1;
    else
      first = mid + 1;
  }
}

// In-place lower Cholesky of the leading m x m block of K (column-major, leading dim ld).
bool chol_lower(int m, double *K, int ld) {
  for (int j = 0; j < m; ++j) {
    double d = K[j + j * ld];
    for (int k = 0; k < j; ++k) d -= K[j + k * ld] * K[j + k * ld];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    K[j + j * ld] = d;
    for (int i = j + 1; i < m; ++i) {
      double s = K[i + j * ld];
      for (int k = 0; k < j; ++k) s -= K[i + k * ld] * K[j + k * ld];
      K[i + j * ld] = s / d;
    }
  }
  return true;
}

void chol_solve(int m, const double *L, int ld, double *b) {
  for (int i = 0; i < m; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[i + k * ld] * b[k];
    b[i] = s / L[i + i * ld];
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < m; ++k) s -= L[k + i * ld] * b[k];
    b[i] = s / L[i + i * ld];
  }
}

struct FsaiRowWork {
  std::vector<int32_t> mark;  // 0 untouched, -1 in pattern, >0 candidate slot (1-based)
  std::vector<int32_t> pat, cand;
  std::vector<double> grad, K, w, wsav;
  FsaiRowWork(int32_t n, int mmax)
      : mark(n, 0), pat(mmax), cand(), grad(), K((size_t)mmax * mmax), w(mmax + 1, 0.0),
        wsav(mmax + 1, 0.0) {}
};

// Gathers K = A_PP (lower part, column-major) and w = -a_iP.  Returns false when row i has
// no entry left of the diagonal (the right-hand side is null).
bool gather_system(const chs_mat &A, int32_t i, int m, const int32_t *pat, double *K, int ld,
                   double *w) {
  for (int c = 0; c < m; ++c) {
    const int32_t r = pat[c];
    int32_t e = A.rp[r], ee = A.rp[r + 1];
    for (int q = c; q < m; ++q) {
      const int32_t want = pat[q];
      while (e < ee && A.ci[e] < want) ++e;
      K[q + c * ld] = (e < ee && A.ci[e] == want) ? A.v[e] : 0.0;
    }
  }
  int32_t e = A.rp[i], ee = A.rp[i + 1];
  if (e >= ee || A.ci[e] >= i) return false;
  for (int q = 0; q < m; ++q) {
    const int32_t want = pat[q];
    while (e < ee && A.ci[e] < want) ++e;
    w[q] = (e < ee && A.ci[e] == want) ? -A.v[e] : 0.0;
  }
  return true;
}

}  // namespace

extern "C" int chs_afsai(const chs_mat *Ap, int32_t nstep, int32_t step_size, double eps,
                         int nthreads, chs_mat **Gout, int32_t *is_posdef) {
  if (!Ap || !Gout || nstep < 0 || step_size < 1) return fail(CHS_EARG, "afsai: arguments");
  const chs_mat &A = *Ap;
  const int32_t n = A.nrows;
  nthreads = nthreads_or_default(nthreads);
  const int mmax = std::max(1, nstep * step_size);
  const int kmax = mmax + 1;
  std::vector<int32_t> rowlen(n);
  std::vector<int32_t> gcol((size_t)n * kmax);
  std::vector<double> gval((size_t)n * kmax);
  int posdef_all = 1;
  const int chunk = std::max(1, n / (20 * nthreads));
#pragma omp parallel num_threads(nthreads) reduction(& : posdef_all)
  {
    FsaiRowWork W(n, mmax);
    int posdef = 1;
#pragma omp for schedule(dynamic, chunk)
    for (int32_t i = 0; i < n; ++i) {
      int m = 0;
      double dk_old = 0.0;
      bool refine = nstep >= 1;
      int istep = 0;
      while (refine) {
        ++istep;
        const int m_old = m;
        // Candidate gradient.
        W.cand.clear();
        W.grad.clear();
        for (int32_t e = A.rp[i]; e < A.rp[i + 1] && A.ci[e] < i; ++e) {
          const int32_t j = A.ci[e];
          if (W.mark[j] >= 0) {
            W.cand.push_back(j);
            W.grad.push_back(A.v[e]);
            W.mark[j] = (int32_t)W.cand.size();
          }
        }
        for (int k = 0; k < m; ++k) {
          const int32_t r = W.pat[k];
          const double xk = W.w[k];
          for (int32_t e = A.rp[r]; e < A.rp[r + 1]; ++e) {
            const int32_t j = A.ci[e];
            if (j >= i) continue;
            const int32_t s = W.mark[j];
            if (s == 0) {
              W.cand.push_back(j);
              W.grad.push_back(xk * A.v[e]);
              W.mark[j] = (int32_t)W.cand.size();
            } else if (s > 0) {
              W.grad[s - 1] += xk * A.v[e];
            }
          }
        }
        const int ncand = (int)W.cand.size();
        const int ncut = std::min<int>(step_size, ncand);
        select_largest(ncand, ncut, W.grad.data(), W.cand.data());
        for (int q = 0; q < ncut; ++q) W.mark[W.cand[q]] = -1;
        for (int q = ncut; q < ncand; ++q) W.mark[W.cand[q]] = 0;
        for (int q = 0; q < ncut; ++q) W.pat[m + q] = W.cand[q];
        m += ncut;
        std::sort(W.pat.begin(), W.pat.begin() + m);
        if (m > m_old) {
          const bool has_rhs = gather_system(A, i, m, W.pat.data(), W.K.data(), mmax, W.w.data());
          if (has_rhs && m > 0) {
            if (!chol_lower(m, W.K.data(), mmax)) {
              posdef = 0;
            } else {
              std::copy(W.w.begin(), W.w.end(), W.wsav.begin());
              chol_solve(m, W.K.data(), mmax, W.w.data());
            }
          }
          double dk_new = 0.0;
          for (int q = 0; q < m; ++q) dk_new += W.w[q] * W.wsav[q];
          if (istep == nstep) {
            refine = false;
          } else {
            refine = (std::fabs(dk_new - dk_old) >= eps * dk_old) && (dk_new != 0.0);
            dk_old = dk_new;
          }
        } else {
          refine = false;
        }
      }
      // Scaling factor from the diagonal entry.
      double aii = 0.0;
      bool found = false;
      for (int32_t e = A.rp[i]; e < A.rp[i + 1] && A.ci[e] <= i; ++e)
        if (A.ci[e] == i) {
          aii = A.v[e];
          found = true;
          break;
        }
      double sc = 0.0;
      if (!found) {
        posdef = 0;
      } else {
        double d = 0.0;
        for (int q = 0; q < m; ++q) d += W.wsav[q] * W.w[q];
        sc = aii - d;
      }
      if (sc < 0.0) {
        posdef = 0;
        if (m > 0 && found) {
          gather_system(A, i, m, W.pat.data(), W.K.data(), mmax, W.w.data());
          if (chol_lower(m, W.K.data(), mmax)) {
            std::copy(W.w.begin(), W.w.end(), W.wsav.begin());
            chol_solve(m, W.K.data(), mmax, W.w.data());
            double d = 0.0;
            for (int q = 0; q < m; ++q) d += W.wsav[q] * W.w[q];
            sc = aii - d;
          }
        }
      }
      if (sc <= 0.0) {
        posdef = 0;
        sc = 1.0;
      }
      sc = 1.0 / std::sqrt(sc);
      int32_t *gc = &gcol[(size_t)i * kmax];
      double *gv = &gval[(size_t)i * kmax];
      for (int q = 0; q < m; ++q) {
        gc[q] = W.pat[q];
        gv[q] = sc * W.w[q];
        W.mark[W.pat[q]] = 0;
      }
      gc[m] = i;
      gv[m] = sc;
      rowlen[i] = m + 1;
    }
    posdef_all &= posdef;
  }
  auto *G = new (std::nothrow) chs_mat;
  if (!G) return fail(CHS_EALLOC, "alloc");
  G->nrows = G->ncols = n;
  G->rp.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) G->rp[i + 1] = G->rp[i] + rowlen[i];
  G->ci.resize(G->rp[n]);
  G->v.resize(G->rp[n]);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int32_t i = 0; i < n; ++i) {
    std::copy(&gcol[(size_t)i * kmax], &gcol[(size_t)i * kmax] + rowlen[i], G->ci.begin() + G->rp[i]);
    std::copy(&gval[(size_t)i * kmax], &gval[(size_t)i * kmax] + rowlen[i], G->v.begin() + G->rp[i]);
  }
  if (is_posdef) *is_posdef = posdef_all;
  *Gout = G;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Largest eigenvalue by Lanczos with full reorthogonalisation.  The tridiagonal's largest
// eigenvalue is found by Sturm-sequence bisection.
// ----------------------------------------------------------------------------------------
namespace {

int sturm_count(int k, const double *a, const double *b, double x) {
  int c = 0;
  double q = a[0] - x;
  if (q < 0) ++c;
  for (int i = 1; i < k; ++i) {
    double qq = q == 0.0 ? 1e-300 : q;
    q = a[i] - x - b[i - 1] * b[i - 1] / qq;
    if (q < 0) ++c;
  }
  return c;  // number of eigenvalues < x
}

double tridiag_max_eig(int k, const double *a, const double *b) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < k; ++i) {
    double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i < k - 1 ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (sturm_count(k, a, b, mid) >= k) hi = mid; else lo = mid;
  }
  return hi;
}

// x . y with partial sums over fixed 4096-entry blocks added in block order: the same bits
// for any thread count and schedule (an OpenMP reduction clause combines its thread
// partials in an unspecified order).
double det_dot(int32_t n, const double *x, const double *y, std::vector<double> &part) {
  constexpr int32_t kBlk = 4096;
  const int32_t nb = (n + kBlk - 1) / kBlk;
  part.resize(nb);
#pragma omp parallel for schedule(static)
  for (int32_t q = 0; q < nb; ++q) {
    double t = 0.0;
    const int32_t e = std::min(n, (q + 1) * kBlk);
    for (int32_t i = q * kBlk; i < e; ++i) t += x[i] * y[i];
    part[q] = t;
  }
  double s = 0.0;
  for (int32_t q = 0; q < nb; ++q) s += part[q];
  return s;
}

template <class Op>
double lanczos_max(int32_t n, Op op, uint64_t seed, int maxit, double tol, int nthreads) {
  nthreads = nthreads_or_default(nthreads);
  maxit = std::max(2, std::min<int>(maxit, n));
  std::vector<std::vector<double>> V;
  std::vector<double> a, b, v(n), w(n), part;
  chs_rng_uniform(seed, 7, 0, n, -1.0, 1.0, v.data());
  double nv = 0.0;
  for (double t : v) nv += t * t;
  nv = std::sqrt(nv);
  for (double &t : v) t /= nv;
  double lam = 0.0, lam_old = 0.0;
  for (int j = 0; j < maxit; ++j) {
    V.push_back(v);
    op(v.data(), w.data());
    const double alpha = det_dot(n, w.data(), v.data(), part);
    a.push_back(alpha);
    // Full reorthogonalisation (twice is enough).
    for (int pass = 0; pass < 2; ++pass)
      for (auto &u : V) {
        const double d = det_dot(n, w.data(), u.data(), part);
#pragma omp parallel for num_threads(nthreads)
        for (int32_t i = 0; i < n; ++i) w[i] -= d * u[i];
      }
    const double beta = std::sqrt(det_dot(n, w.data(), w.data(), part));
    lam = tridiag_max_eig((int)a.size(), a.data(), b.data());
    if (j > 3 && std::fabs(lam - lam_old) <= tol * std::fabs(lam)) break;
    lam_old = lam;
    if (beta <= 1e-14 * std::fabs(lam)) break;
    b.push_back(beta);
#pragma omp parallel for num_threads(nthreads)
    for (int32_t i = 0; i < n; ++i) v[i] = w[i] / beta;
  }
  return lam;
}

}  // namespace

extern "C" int chs_fsai_omega(const chs_mat *Ap, const chs_mat *Gp, uint64_t seed,
                              int nthreads, double *omega, double *lambda) {
  if (!Ap || !Gp) return fail(CHS_EARG, "null");
  nthreads = nthreads_or_default(nthreads);
  chs_mat GT;
  transpose_into(*Gp, GT);
  const int32_t n = Ap->nrows;
  std::vector<double> t1(n), t2(n);
  auto op = [&](const double *x, double *y) {
    spmv(GT, x, t1.data(), nthreads);
    spmv(*Ap, t1.data(), t2.data(), nthreads);
    spmv(*Gp, t2.data(), y, nthreads);
  };
  double lam = lanczos_max(n, op, seed, 40, 1e-7, nthreads);
  if (lambda) *lambda = lam;
  if (omega) *omega = std::min(1.0, 1.9 / lam);
  return CHS_OK;
}

// Dirichlet rows of the input (rows assembled as unit rows) are rescaled to the spectrum of
// A: lambda_max(A) by Lanczos (the reference calls eigs(A,1,'lm') with tolerance 1e-3), then
//   mode CHS_BC_DIAG: diagonal entries equal to 1 become lambda_max/10 (computeAMG.m:5-10);
//   mode CHS_BC_ALL:  every stored entry equal to 1 becomes lambda_max/10
//                     (sources/Drivers/driver_AMG/aspAMG_main.m:89-90, A(A==1) = lmax/10).
// The Lanczos dot products use fixed blocks, so the result is the same bits for any thread
// count.
extern "C" int chs_bc_fix(const chs_mat *Ap, uint64_t seed, int32_t mode, int nthreads, chs_mat **Afix,
                          int32_t *nfixed) {
  if (!Ap || !Afix) return fail(CHS_EARG, "null");
  if (mode != CHS_BC_DIAG && mode != CHS_BC_ALL) return fail(CHS_EARG, "chs_bc_fix: unknown mode");
  auto *m = new (std::nothrow) chs_mat(*Ap);
  if (!m) return fail(CHS_EALLOC, "alloc");
  int32_t cnt = 0;
  std::vector<int32_t> where;
  for (int32_t i = 0; i < m->nrows; ++i)
    for (int32_t e = m->rp[i]; e < m->rp[i + 1]; ++e)
      if ((mode == CHS_BC_ALL || m->ci[e] == i) && m->v[e] == 1.0) where.push_back(e);
  if (!where.empty()) {
    auto op = [&](const double *x, double *y) { spmv(*Ap, x, y, nthreads); };
    double lmax = lanczos_max(m->nrows, op, seed, 40, 1e-6, nthreads);
    for (int32_t e : where) m->v[e] = lmax / 10.0;
    cnt = (int32_t)where.size();
  }
  if (nfixed) *nfixed = cnt;
  *Afix = m;
  return CHS_OK;
}

// The aFSAI factor's repair for very ill-conditioned rows (smoother.m:59-71): every row whose
// diagonal entry is NaN becomes the single entry 1/sqrt(a_ii) on the diagonal.
extern "C" int chs_fsai_nan_repair(const chs_mat *Ap, chs_mat *Gp, int32_t *nfixed) {
  if (!Ap || !Gp || Ap->nrows != Gp->nrows) return fail(CHS_EARG, "chs_fsai_nan_repair: arguments");
  const chs_mat &A = *Ap;
  chs_mat &G = *Gp;
  std::vector<int32_t> bad;
  for (int32_t i = 0; i < G.nrows; ++i)
    for (int32_t e = G.rp[i]; e < G.rp[i + 1]; ++e)
      if (G.ci[e] == i && std::isnan(G.v[e])) {
        bad.push_back(i);
        break;
      }
  if (nfixed) *nfixed = (int32_t)bad.size();
  if (bad.empty()) return CHS_OK;
  std::vector<int32_t> rp((size_t)G.nrows + 1, 0), ci;
  std::vector<double> v;
  ci.reserve(G.ci.size());
  v.reserve(G.v.size());
  size_t q = 0;
  for (int32_t i = 0; i < G.nrows; ++i) {
    if (q < bad.size() && bad[q] == i) {
      double aii = 0.0;
      for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e)
        if (A.ci[e] == i) aii = A.v[e];
      ci.push_back(i);
      v.push_back(1.0 / std::sqrt(aii));
      ++q;
    } else {
      for (int32_t e = G.rp[i]; e < G.rp[i + 1]; ++e) {
        ci.push_back(G.ci[e]);
        v.push_back(G.v[e]);
      }
    }
    rp[i + 1] = (int32_t)ci.size();
  }
  G.rp.swap(rp);
  G.ci.swap(ci);
  G.v.swap(v);
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// Jacobi smoother factor: G = diag(1/sqrt(a_ii)) (smoother.m:112-132).
// ----------------------------------------------------------------------------------------
static int jacobi_factor(const chs_mat &A, chs_mat &G) {
  G.nrows = G.ncols = A.nrows;
  G.rp.resize(A.nrows + 1);
  G.ci.resize(A.nrows);
  G.v.resize(A.nrows);
  for (int32_t i = 0; i <= A.nrows; ++i) G.rp[i] = i;
  for (int32_t i = 0; i < A.nrows; ++i) {
    double d = 0.0;
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.ci[e] == i) d = A.v[e];
    if (!(d > 0.0)) return fail(CHS_ENOTSPD, "jacobi: non-positive diagonal");
    G.ci[i] = i;
    G.v[i] = 1.0 / std::sqrt(d);
  }
  return CHS_OK;
}

// Aggressive coarsening (agg_coarsen.m:1-100): PMIS on the strength pattern S splits the
// nodes into coarse candidates and fine nodes Ff; a second PMIS on the candidates, over the
// distance-two pattern S_Fc = S_sub^T S_sub with S_sub the candidate columns of |S| + |S^T|
// (mk_SoC_FcC.m; its values are positive, so its pattern is the structural product's), keeps
// the coarse nodes C and leaves the other candidates as Fc.  kind[i] = 2 (C), 1 (Fc), 0 (Ff).
static int agg_coarsen(const chs_mat &A, const int8_t *strong, uint64_t seed, int level,
                       std::vector<int8_t> &kind) {
  const int32_t n = A.nrows;
  std::vector<double> r01(n);
  chs_rng_uniform(seed, 100 + level, 0, n, 0.0, 1.0, r01.data());
  std::vector<int32_t> fc1(n);
  int32_t ncand = 0;
  int rc = pmis_core(&A, strong, r01.data(), true, fc1.data(), &ncand);
  if (rc) return rc;
  // S_sym pattern (strong off-diagonal entries of A and their transposes), rows.
  std::vector<std::vector<int32_t>> sym(n);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e) {
      const int32_t j = A.ci[e];
      if (j != i && strong[e] > 0) {
        sym[i].push_back(j);
        sym[j].push_back(i);
      }
    }
  for (auto &r : sym) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  // S_Fc over the candidates (numbered fc1): a ~ b when some node m has both as S_sym
  // neighbours, i.e. m in the S_sym rows of both (S_sym is symmetric).
  chs_mat S;
  S.nrows = S.ncols = ncand;
  S.rp.assign((size_t)ncand + 1, 0);
  std::vector<std::vector<int32_t>> rows(ncand);
  std::vector<int32_t> cand;
  cand.reserve(ncand);
  for (int32_t i = 0; i < n; ++i)
    if (fc1[i] >= 0) cand.push_back(i);
#pragma omp parallel
  {
    std::vector<int32_t> mark(ncand, -1);
#pragma omp for schedule(dynamic, 256)
    for (int32_t a = 0; a < ncand; ++a) {
      auto &out = rows[a];
      for (int32_t m : sym[cand[a]])
        for (int32_t b : sym[m])
          if (fc1[b] >= 0 && mark[fc1[b]] != a) {
            mark[fc1[b]] = a;
            out.push_back(fc1[b]);
          }
      std::sort(out.begin(), out.end());
    }
  }
  for (int32_t a = 0; a < ncand; ++a) S.rp[a + 1] = S.rp[a] + (int32_t)rows[a].size();
  S.ci.resize(S.rp[ncand]);
  S.v.assign(S.rp[ncand], 1.0);
  for (int32_t a = 0; a < ncand; ++a) std::copy(rows[a].begin(), rows[a].end(), S.ci.begin() + S.rp[a]);
  std::vector<int8_t> sstrong(S.ci.size(), 1);
  std::vector<double> r2(ncand);
  chs_rng_uniform(seed, 300 + level, 0, ncand, 0.0, 1.0, r2.data());
  std::vector<int32_t> fc2(ncand);
  int32_t nc = 0;
  if ((rc = pmis_core(&S, sstrong.data(), r2.data(), false, fc2.data(), &nc))) return rc;
  kind.assign(n, 0);
  for (int32_t i = 0; i < n; ++i)
    if (fc1[i] >= 0) kind[i] = fc2[fc1[i]] >= 0 ? 2 : 1;
  return CHS_OK;
}

extern "C" int chs_agg_coarsen(const chs_mat *Ap, const int8_t *strong, uint64_t seed, int32_t level,
                               int8_t *kind) {
  if (!Ap || !strong || !kind) return fail(CHS_EARG, "null");
  std::vector<int8_t> k;
  int rc = agg_coarsen(*Ap, strong, seed, level, k);
  if (rc) return rc;
  std::copy(k.begin(), k.end(), kind);
  return CHS_OK;
}

// Two-stage extended+i prolongation (cpt_TwoStageProlMEX_ExtI.m:1-60): P1 interpolates the
// Ff nodes from the candidates C + Fc, P2 the Fc nodes from C (the Ff nodes taking part only
// as fine neighbours, without rows), and P = P1 P2.
static int two_stage_exti(const chs_mat &A, const int8_t *strong, const std::vector<int8_t> &kind,
                          int nt, chs_mat **Pout) {
  const int32_t n = A.nrows;
  std::vector<int32_t> f1(n), f2(n);
  int32_t n1 = 0, n2 = 0;
  for (int32_t i = 0; i < n; ++i) {
    f1[i] = kind[i] > 0 ? n1++ : -1;
    f2[i] = kind[i] == 2 ? n2++ : (kind[i] == 1 ? -1 : -2);
  }
  chs_mat *P1 = nullptr, *P2f = nullptr;
  int rc = chs_prol_exti(&A, strong, f1.data(), n1, nt, &P1);
  if (rc) return rc;
  std::unique_ptr<chs_mat> g1(P1);
  if ((rc = chs_prol_exti(&A, strong, f2.data(), n2, nt, &P2f))) return rc;
  std::unique_ptr<chs_mat> g2(P2f);
  // P2: the rows of the candidates, in node order (the Ff rows are not part of it).
  chs_mat P2;
  P2.nrows = n1;
  P2.ncols = n2;
  P2.rp.assign((size_t)n1 + 1, 0);
  for (int32_t i = 0, r = 0; i < n; ++i)
    if (kind[i] > 0) {
      P2.rp[r + 1] = P2.rp[r] + (P2f->rp[i + 1] - P2f->rp[i]);
      P2.ci.insert(P2.ci.end(), P2f->ci.begin() + P2f->rp[i], P2f->ci.begin() + P2f->rp[i + 1]);
      P2.v.insert(P2.v.end(), P2f->v.begin() + P2f->rp[i], P2f->v.begin() + P2f->rp[i + 1]);
      ++r;
    }
  auto *P = new (std::nothrow) chs_mat;
  if (!P) return fail(CHS_EALLOC, "alloc");
  spgemm(*P1, P2, *P, nt);
  *Pout = P;
  return CHS_OK;
}

// Prolongation smoothing (prolong.m:180-197): P <- P - omega D^-1 A P with omega =
// 1.9 / lambda_max(D^-1 A); lambda_max by the same Lanczos as the smoother damping, on
// D^-1/2 A D^-1/2 (the same spectrum).  The product is formed in MATLAB's order: the scaled
// operator omega (a_ij / a_ii) first, then its product with P, then the difference, exact
// zeros dropped.
static int smooth_prolongation(const chs_mat &A, chs_mat &P, uint64_t seed, int nt, double *omega_out) {
  chs_mat Dh;
  int rc = jacobi_factor(A, Dh);
  if (rc) return rc;
  double lam = 0.0;
  rc = chs_fsai_omega(&A, &Dh, seed, nt, nullptr, &lam);
  if (rc) return rc;
  const double omega = 1.9 / lam;
  chs_mat W = A;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int32_t i = 0; i < A.nrows; ++i) {
    double d = 0.0;
    for (int32_t e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.ci[e] == i) d = A.v[e];
    for (int32_t e = W.rp[i]; e < W.rp[i + 1]; ++e) W.v[e] = omega * (A.v[e] / d);
  }
  chs_mat WP;
  spgemm(W, P, WP, nt);
  chs_mat R;
  R.nrows = P.nrows;
  R.ncols = P.ncols;
  R.rp.assign((size_t)P.nrows + 1, 0);
  std::vector<std::vector<std::pair<int32_t, double>>> rows(P.nrows);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int32_t i = 0; i < P.nrows; ++i) {
    auto &r = rows[i];
    int32_t a = P.rp[i], ae = P.rp[i + 1], b = WP.rp[i], be = WP.rp[i + 1];
    while (a < ae || b < be) {
      const int32_t ca = a < ae ? P.ci[a] : INT32_MAX, cb = b < be ? WP.ci[b] : INT32_MAX;
      double v;
      int32_t c;
      if (ca == cb) { c = ca; v = P.v[a++] - WP.v[b++]; }
      else if (ca < cb) { c = ca; v = P.v[a++]; }
      else { c = cb; v = -WP.v[b++]; }
      if (v != 0.0) r.emplace_back(c, v);
    }
  }
  for (int32_t i = 0; i < P.nrows; ++i) R.rp[i + 1] = R.rp[i] + (int32_t)rows[i].size();
  R.ci.resize(R.rp[P.nrows]);
  R.v.resize(R.rp[P.nrows]);
  for (int32_t i = 0; i < P.nrows; ++i)
    for (size_t q = 0; q < rows[i].size(); ++q) {
      R.ci[R.rp[i] + q] = rows[i][q].first;
      R.v[R.rp[i] + q] = rows[i][q].second;
    }
  P = std::move(R);
  if (omega_out) *omega_out = omega;
  return CHS_OK;
}

// ----------------------------------------------------------------------------------------
// The recursion of cpt_aspAMG.m:68-351 for the classical + afsai_sym / jacobi path.
// ----------------------------------------------------------------------------------------
extern "C" int chs_build(const chs_mat *A0, const chs_params *pp, chs_hier **out) {
  if (!A0 || !out) return fail(CHS_EARG, "null");
  chs_params p;
  if (pp) p = *pp; else chs_params_default(&p);
  const int nt = nthreads_or_default(p.nthreads);
  auto h = std::make_unique<chs_hier>();
  const double t_start = now_s();
  chs_mat A = *A0;
  // Test space (cpt_aspAMG.m:29-46): the caller's columns, padded with uniform draws, then
  // orthonormalised at every level (tspace method NONE) and injected to the coarse nodes.
  std::vector<double> tv;
  if (p.prolong == 2) {
    if (p.ntv < 1) return fail(CHS_EARG, "bamg: ntv < 1");
    const int64_t n0 = A.nrows;
    tv.assign((size_t)n0 * p.ntv, 0.0);
    const int32_t ncopy = p.tv ? std::min(p.ntv0, p.ntv) : 0;
    for (int32_t j = 0; j < ncopy; ++j)
      std::copy(p.tv + (size_t)j * n0, p.tv + (size_t)(j + 1) * n0, tv.begin() + (size_t)j * n0);
    if (ncopy < p.ntv)
      chs_rng_uniform(p.seed, 7, 0, (int64_t)(p.ntv - ncopy) * n0, 0.0, 1.0, tv.data() + (size_t)ncopy * n0);
  }
  for (int level = 1;; ++level) {
    chs_hier::Level L;
    const int32_t n = A.nrows;
    if (n <= p.max_coarse || level >= p.nlev_max) {
      L.A = std::move(A);
      h->lev.push_back(std::move(L));
      break;
    }
    double t0 = now_s();
    if (p.smoother == 1) {
      int rc = jacobi_factor(A, L.G);
      if (rc) return rc;
    } else {
      bool host = true;
      if (p.afsai_fn) {
        if (p.nstep < 0 || p.step_size < 1) return fail(CHS_EARG, "afsai: arguments");
        const int64_t kmax = (int64_t)std::max(1, p.nstep * p.step_size) + 1;
        int32_t pd = 1;
        int64_t nnz_g = 0;
        L.G.nrows = L.G.ncols = n;
        L.G.rp.assign((size_t)n + 1, 0);
        L.G.ci.resize((size_t)n * kmax);
        L.G.v.resize((size_t)n * kmax);
        int rc = p.afsai_fn(p.afsai_ctx, n, A.rp.data(), A.ci.data(), A.v.data(), p.nstep, p.step_size,
                            p.fsai_eps, L.G.rp.data(), L.G.ci.data(), L.G.v.data(), &nnz_g, &pd);
        if (rc < 0) return fail(rc, std::string("afsai (external): error ") + std::to_string(rc));
        if (rc == 0) {
          host = false;
          L.G.ci.resize((size_t)nnz_g);
          L.G.v.resize((size_t)nnz_g);
          L.G.ci.shrink_to_fit();
          L.G.v.shrink_to_fit();
          if (!pd) return fail(CHS_ENOTSPD, "afsai_sym: matrix deemed indefinite at level " +
                                                std::to_string(level));
        }
        // rc == 1: the replacement declined this level's row lengths; the host factors it.
      }
      if (host) {
        chs_mat *G = nullptr;
        int32_t pd = 1;
        int rc = chs_afsai(&A, p.nstep, p.step_size, p.fsai_eps, nt, &G, &pd);
        if (rc) return rc;
        L.G = std::move(*G);
        delete G;
        if (!pd) return fail(CHS_ENOTSPD, "afsai_sym: matrix deemed indefinite at level " +
                                              std::to_string(level));
      }
      int32_t nfix = 0;
      int rc = chs_fsai_nan_repair(&A, &L.G, &nfix);
      if (rc) return rc;
    }
    if (p.omega_fn) {
      int rc = p.omega_fn(p.omega_ctx, n, A.rp.data(), A.ci.data(), A.v.data(), L.G.rp.data(),
                          L.G.ci.data(), L.G.v.data(), p.seed + level, &L.omega, &L.lambda);
      if (rc) return fail(rc, std::string("omega (external): error ") + std::to_string(rc));
    } else {
      chs_fsai_omega(&A, &L.G, p.seed + level, nt, &L.omega, &L.lambda);
    }
    double t1 = now_s();
    std::vector<int8_t> strong(A.nnz());
    int rcs = p.soc_type == 1 ? chs_soc_dom(&A, p.soc_tau, strong.data())
                              : chs_soc_cla(&A, p.soc_tau, strong.data());
    if (rcs) return rcs;
    std::vector<int32_t> fc(n);
    int32_t nc = 0;
    const bool aggressive = level <= p.nl_agg;  // cpt_aspAMG.m:183
    std::vector<int8_t> kind;
    if (aggressive) {
      if (p.prolong == 0)
        return fail(CHS_EARG, "aggressive coarsening needs the EXTI or BAMG prolongation (prolong.m:101-165)");
      if ((rcs = agg_coarsen(A, strong.data(), p.seed, level, kind))) return rcs;
      for (int32_t i = 0; i < n; ++i) fc[i] = kind[i] == 2 ? nc++ : -1;
    } else {
      std::vector<double> r01(n);
      chs_rng_uniform(p.seed, 100 + level, 0, n, 0.0, 1.0, r01.data());
      chs_pmis(&A, strong.data(), r01.data(), fc.data(), &nc);
    }
    double t2 = now_s();
    if (nc <= 0 || nc >= n) return fail(CHS_ECONSIST, "coarsening produced no reduction at level " +
                                                         std::to_string(level));
    chs_mat *P = nullptr;
    int32_t nmarked = 0;
    int rc;
    std::vector<double> tv_f;  // this level's test space, kept for EMIN
    if (p.prolong == 2) {
      chs_orthonormalize(n, p.ntv, tv.data());
      rc = chs_prol_bamg(&A, strong.data(), fc.data(), nc, p.ntv, tv.data(), &p.bamg, nt, &P,
                         nullptr, &nmarked);
      if (!rc && p.prol_smooth == 2) tv_f = tv;
      if (!rc) {  // next level's test space: the coarse nodes' rows (cpt_aspAMG.m:242-243)
        std::vector<double> tvc((size_t)nc * p.ntv);
        for (int32_t j = 0; j < p.ntv; ++j)
          for (int32_t i = 0; i < n; ++i)
            if (fc[i] >= 0) tvc[(size_t)j * nc + fc[i]] = tv[(size_t)j * n + i];
        tv.swap(tvc);
      }
    } else if (aggressive) {
      rc = two_stage_exti(A, strong.data(), kind, nt, &P);
    } else {
      rc = p.prolong == 1 ? chs_prol_exti(&A, strong.data(), fc.data(), nc, nt, &P)
                          : chs_prol_classical(&A, strong.data(), fc.data(), nc, nt, &P);
    }
    if (rc) return rc;
    if (p.prol_smooth == 1) {
      rc = smooth_prolongation(A, *P, p.seed + 1000 + level, nt, nullptr);
      if (rc) { delete P; return rc; }
    } else if (p.prol_smooth == 2) {
      // prolong.m:207-232: pattern, then energy minimisation with the level's test space.
      if (tv_f.empty()) {
        delete P;
        return fail(CHS_EARG, "EMIN needs the BAMG prolongation's test space (prolong = 2)");
      }
      chs_mat *T = nullptr, *Pe = nullptr;
      int32_t it = 0;
      rc = chs_emin_pattern(&A, P, fc.data(), p.ntv, &p.emin, nt, &T);
      if (!rc) rc = emin_improve(&A, P, T, fc.data(), p.ntv, tv_f.data(), &p.emin, &p.emin_kp, nt, &Pe, &it);
      delete T;
      delete P;
      if (rc) return rc;
      P = Pe;
    } else if (p.prol_smooth != 0) {
      delete P;
      return fail(CHS_EARG, "prol_smooth: 0 (NONE), 1 (SMOOTH) or 2 (EMIN)");
    }
    double t3 = now_s();
    chs_mat *Ac = nullptr;
    if (p.galerkin_fn) {
      auto C = std::make_unique<chs_mat>();
      C->nrows = C->ncols = nc;
      C->rp.assign((size_t)nc + 1, 0);
      auto reserve = [](void *sink, int64_t nnz, int32_t **col, double **val) -> int {
        auto *m = static_cast<chs_mat *>(sink);
        if (nnz < 0 || nnz > INT32_MAX) return -1;
        m->ci.resize((size_t)nnz);
        m->v.resize((size_t)nnz);
        *col = m->ci.data();
        *val = m->v.data();
        return 0;
      };
      rc = p.galerkin_fn(p.galerkin_ctx, n, A.rp.data(), A.ci.data(), A.v.data(), nc, P->rp.data(),
                         P->ci.data(), P->v.data(), C->rp.data(), C.get(), reserve);
      if (rc < 0) {
        delete P;
        return fail(rc, std::string("galerkin (external): error ") + std::to_string(rc));
      }
      if (rc == 0) Ac = C.release();  // rc == 1: declined, formed on the host below
    }
    if (!Ac) {
      rc = chs_galerkin(&A, P, nt, &Ac);
      if (rc) { delete P; return rc; }
    }
    double t4 = now_s();
    h->t[0] += t1 - t0;
    h->t[1] += t2 - t1;
    h->t[2] += t3 - t2;
    h->t[3] += t4 - t3;
    if (p.verbose)
      std::fprintf(stderr, "level %d: n=%d nnz=%lld nnzG=%lld nc=%d nnzP=%lld omega=%.6f "
                   "inexact=%d (smoother %.2fs coarsen %.2fs prol %.2fs galerkin %.2fs)\n",
                   level, n, (long long)A.nnz(), (long long)L.G.nnz(), nc,
                   (long long)P->nnz(), L.omega, nmarked, t1 - t0, t2 - t1, t3 - t2, t4 - t3);
    L.A = std::move(A);
    L.P = std::move(*P);
    delete P;
    A = std::move(*Ac);
    delete Ac;
    h->lev.push_back(std::move(L));
  }
  h->t[4] = now_s() - t_start;
  *out = h.release();
  return CHS_OK;
}

extern "C" int chs_hier_nlev(const chs_hier *h) { return h ? (int)h->lev.size() : 0; }

extern "C" const chs_mat *chs_hier_mat(const chs_hier *h, int32_t level, char which) {
  if (!h || level < 0 || level >= (int32_t)h->lev.size()) return nullptr;
  const auto &L = h->lev[level];
  switch (which) {
    case 'A': return &L.A;
    case 'P': return level + 1 < (int32_t)h->lev.size() ? &L.P : nullptr;
    case 'G': return level + 1 < (int32_t)h->lev.size() ? &L.G : nullptr;
    default: return nullptr;
  }
}

extern "C" double chs_hier_omega(const chs_hier *h, int32_t level) {
  if (!h || level < 0 || level >= (int32_t)h->lev.size()) return 0.0;
  return h->lev[level].omega;
}

extern "C" double chs_hier_setup_seconds(const chs_hier *h, int32_t phase) {
  if (!h || phase < 0 || phase > 4) return 0.0;
  return h->t[phase];
}

extern "C" void chs_hier_free(chs_hier *h) { delete h; }

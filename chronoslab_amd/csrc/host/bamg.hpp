// Bootstrap-AMG least-squares interpolation (host setup library, internal interface).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace chsb {

struct Params {
  int32_t itmax_vol;  // maximum volume iterations
  double tol_vol;     // stop when every |Z| <= 1 + tol_vol
  int32_t dist_min;   // neighbourhood belts before the first fit
  int32_t dist_max;   // belts at most
  int32_t mmax;       // interpolation points per row at most
  double maxcond;     // bound on the selected rows' conditioning estimate
  double maxrownrm;   // bound on a row's weight 2-norm
  double eps;         // relative least-squares residual accepted as exact
};

// Interpolation rows of an n-node level.  S: strength pattern (rp, ci; the diagonal may be
// present), fc: coarse index of coarse nodes, < 0 for fine nodes, tv: test space row-major
// (n x ntv).  Fills the CSR arrays of P (n x nc) and mark[i] = 1 for the fine rows whose
// test-vector fit stays inexact.  Returns 0, or 2 with err set (a singular fit).
int prolong(int32_t n, const int32_t *rp, const int32_t *ci, const int32_t *fc, int32_t ntv,
            const double *tv, const Params &prm, int nthreads, std::vector<int32_t> &prp,
            std::vector<int32_t> &pci, std::vector<double> &pv, int32_t *mark, std::string &err);

// Orthonormal basis of the column span of V (n x k column-major, in place): the Q factor of
// a thin QR with a positive R diagonal, by modified Gram-Schmidt with one full
// reorthogonalisation; dot products are summed over fixed 4096-row blocks, so the result does
// not depend on the thread count.  Returns 1 if a column is dependent on the previous ones
// (that column is left zero), else 0.
int orthonormalize(int64_t n, int32_t k, double *V);

}  // namespace chsb

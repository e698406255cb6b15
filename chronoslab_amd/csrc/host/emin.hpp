// Energy-minimising prolongation (host setup library, internal interface).
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace chse {

struct Params {
  int32_t patt_pow;  // powers of the strength pattern grown onto P's pattern
  double patt_tau;   // DOM strength threshold of that pattern
  int32_t nnzr_max;  // entries per fine row at most (raised to the number of test vectors)
  int32_t itmax;     // CG iterations at most
  double en_tol;     // stop once an iteration lowers the energy by less than en_tol times
                     // the first iteration's decrease
  double condmax;    // conditioning bound of a row's constraint block before the SVD fallback
};

// Interpolation pattern of the fine rows (coarse rows empty): P's pattern grown by the
// strongest entries of S^(k-1) P, S the DOM strength matrix of A with threshold patt_tau.
// fc: coarse index of coarse nodes, < 0 for fine nodes.  Returns CSR (trp, tci), n rows,
// columns ascending.
void pattern(int32_t n, int32_t nc, const int32_t *arp, const int32_t *aci, const double *av,
             const int32_t *prp, const int32_t *pci, const double *pv, const int32_t *fc,
             int32_t ntv, const Params &prm, int nthreads, std::vector<int32_t> &trp,
             std::vector<int32_t> &tci);

// Energy minimisation of P (n x nc, coarse rows identity) over the fine-row pattern
// (trp, tci) under the test-space constraint of every fine row.  tv: n x ntv column-major.
// Fills the CSR arrays of the improved P (coarse rows e_c, exact zeros dropped) and the
// CG iteration count.  kp_ext, when given, replaces the host's masked product y = (A X) on
// the pattern (it must sum in the same order).  Returns 0, or 4 (a constraint factorisation
// failed) / 1 (bad input) / 5 (kp_ext failed) with err set.
int improve(int32_t n, int32_t nc, const int32_t *arp, const int32_t *aci, const double *av,
            const int32_t *prp, const int32_t *pci, const double *pv, const int32_t *trp,
            const int32_t *tci, const int32_t *fc, int32_t ntv, const double *tv,
            const Params &prm, int nthreads, std::vector<int32_t> &orp, std::vector<int32_t> &oci,
            std::vector<double> &ov, int32_t &iter, std::string &err,
            const std::function<int(const double *, double *)> &kp_ext = nullptr);

}  // namespace chse

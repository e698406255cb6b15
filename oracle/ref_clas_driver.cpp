// Test-infrastructure driver: exposes the reference's classical-prolongation core
// (sources/Preconditioner/AMG/prolong/MEX_CLAS_prol/Classical_prolongation.cpp:40-147),
// compiled from the reference tree by oracle/Makefile into oracle/_ref/, through a C ABI so
// that tests can compare the host setup's classical interpolation with it.  It replaces the
// MATLAB wrapper cpt_ProlMEX_Classical.m:18-67: contiguous row stripes per thread (the
// vecstart split of :31-40) and 0-based int32 inputs.
#include <cstdlib>
#include <cstring>

int Classical_prolongation(const int level, const int nthreads, const int *vecstart,
                           const int nn_A, const int nt_A, const int *const iat_A,
                           const int *const ja_A, const double *const coef_A,
                           const int *const coef_S, const int *const fcnodes, const int nr_I,
                           const int nc_I, int &nt_I, int *&iat_I, int *&ja_I, double *&coef_I);

extern "C" int ref_classical_prol(int nthreads, int n, const int *iat, const int *ja,
                                  const double *coef, const int *strong, const int *fcnodes,
                                  int nc, int *nt_out, int **iat_out, int **ja_out,
                                  double **coef_out) {
  int *vecstart = new int[nthreads + 1];
  int bsize = n / nthreads, rest = n % nthreads;
  vecstart[0] = 0;
  for (int i = 1; i <= nthreads; ++i) vecstart[i] = vecstart[i - 1] + bsize + (i <= rest ? 1 : 0);
  int nt = 0, *ip = nullptr, *jp = nullptr;
  double *cp = nullptr;
  int ierr = Classical_prolongation(0, nthreads, vecstart, n, iat[n], iat, ja, coef, strong,
                                    fcnodes, n, nc, nt, ip, jp, cp);
  delete[] vecstart;
  *nt_out = nt;
  *iat_out = ip;
  *ja_out = jp;
  *coef_out = cp;
  return ierr;
}

extern "C" void ref_free(int *ip, int *jp, double *cp) {
  delete[] ip;
  delete[] jp;
  delete[] cp;
}

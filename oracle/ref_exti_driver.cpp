// Test-infrastructure driver: exposes the reference's extended+i prolongation core
// (sources/Preconditioner/AMG/prolong/MEX_EXTI_Prol/EXTI_prolongation.cpp:52-199), compiled
// from the reference tree by oracle/Makefile into oracle/_ref/, through a C ABI so that
// tests can compare the host setup's EXTI interpolation with it.  It replaces the MATLAB
// wrappers cpt_ProlMEX_EXTI.m:20-38 (0-based int32 inputs, fine nodes marked -1) and, for
// the second stage of cpt_TwoStageProlMEX_ExtI.m:41-56, nodes marked -2, which take part as
// fine neighbours but get no row: the row count is that of the nodes marked >= -1.
int EXTI_prolongation(const int level, const int nthreads, const int nn_A, const int nt_A,
                      const int *const iat_A, const int *const ja_A, const double *const coef_A,
                      const int *const coef_S, const int *const fcnodes, const int nr_I,
                      const int nc_I, int &nt_I, int *&iat_I, int *&ja_I, double *&coef_I);

extern "C" int ref_exti_prol(int nthreads, int n, const int *iat, const int *ja, const double *coef,
                             const int *strong, const int *fcnodes, int nc, int *nt_out, int **iat_out,
                             int **ja_out, double **coef_out) {
  int nt = 0, *ip = nullptr, *jp = nullptr;
  double *cp = nullptr;
  int nr = 0;
  for (int i = 0; i < n; ++i) nr += fcnodes[i] >= -1 ? 1 : 0;
  const int ierr = EXTI_prolongation(0, nthreads, n, iat[n], iat, ja, coef, strong, fcnodes, nr, nc, nt,
                                     ip, jp, cp);
  *nt_out = nt;
  *iat_out = ip;
  *ja_out = jp;
  *coef_out = cp;
  return ierr;
}

extern "C" void ref_exti_free(int *ip, int *jp, double *cp) {
  delete[] ip;
  delete[] jp;
  delete[] cp;
}

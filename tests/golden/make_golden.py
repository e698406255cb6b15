"""Writes the golden fixtures of the classical interpolation from the reference's own core.

Run in a container that holds the reference tree (oracle/_ref/libref_clas.so is built from
it by `make -C oracle ref`):

    python tests/golden/make_golden.py

Each fixture stores an input (A in CSR, the strong-connection signs aligned with A, the C/F
split) and the prolongation the reference's Classical_prolongation core
(sources/Preconditioner/AMG/prolong/MEX_CLAS_prol/Classical_prolongation.cpp:40-147)
returns for it.  Inputs: a 7-point Poisson cube, a heterogeneous-diffusion cube, and the
leading 1200 x 1200 block of the reference test matrix examples/matrices/Flux.mat (the
FlowCLAS_* tests' matrix, read as data with scipy.io.loadmat).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from chronoslab_amd import hierarchy as H  # noqa: E402

REF = "/root/reference"


def inputs():
    yield "poisson7_n9", H.poisson7(9), 0.25
    yield "diffusion_n10_s2", H.heterogeneous_diffusion(10, 2.0, 17)[0], 0.25
    flux = os.path.join(REF, "examples", "matrices", "Flux.mat")
    if os.path.exists(flux):
        import scipy.io

        d = scipy.io.loadmat(flux)
        A = d["Amat"].tocsr()[:1200, :1200]
        yield "flux_block1200", H.Csr.from_scipy(A), 0.25


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref_clas.so is not built (make -C oracle ref)")
    for name, A, tau in inputs():
        strong = H.soc_cla(A, tau)
        r01 = H.rng_uniform(99, 100, A.nrows)
        fc, nc = H.pmis(A, strong, r01)
        iat, ja, co = oracle.ref_classical_prol(A, strong, fc, nc, nthreads=1)
        np.savez_compressed(os.path.join(HERE, f"clas_{name}.npz"), rowptr=A.rowptr, col=A.col, val=A.val,
                            strong=strong, fcnode=fc, nc=nc, P_rowptr=iat, P_col=ja, P_val=co)
        print(name, A.nrows, A.nnz, nc, len(ja))


if __name__ == "__main__":
    main()

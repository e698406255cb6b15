"""Writes the golden fixtures of the classical and extended+i interpolations from the
reference's own cores, and the reference test matrix the known-answer test runs on.

Run in a container that holds the reference tree (oracle/_ref/libref_clas.so and
libref_exti.so are built from it by `make -C oracle ref`):

    python tests/golden/make_golden.py

Each clas_* / exti_* fixture stores an input (A in CSR, the strong-connection signs aligned
with A, the C/F split) and the prolongation the reference's Classical_prolongation core
(sources/Preconditioner/AMG/prolong/MEX_CLAS_prol/Classical_prolongation.cpp:40-147) or
EXTI_prolongation core (MEX_EXTI_Prol/EXTI_prolongation.cpp:52-199) returns for it.
Inputs: a 7-point Poisson cube, a heterogeneous-diffusion cube, and the leading 1200 x 1200
block of the reference test matrix examples/matrices/Flux.mat (the FlowCLAS_* tests'
matrix, read as data with scipy.io.loadmat).  flux.npz holds that whole matrix (Amat, as
CSR) and its right-hand side b: the data of Tests/FlowCLAS_0.m.  cubo4820.npz holds
examples/matrices/Cubo_4820.mat (the upper triangle of the symmetric A, the test space TV0
and rhs): the data of Tests/CubeBAMG_NONE.m and Tests/CubeBAMG_SMOOTH.m.
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


def data_fixtures():
    """The reference test matrices the known-answer tests run on, as plain arrays."""
    import scipy.io
    import scipy.sparse as sp

    cubo = os.path.join(REF, "examples", "matrices", "Cubo_4820.mat")
    if os.path.exists(cubo):
        d = scipy.io.loadmat(cubo, variable_names=["A", "TV0", "rhs"])
        A = d["A"].tocsr()
        assert abs(A - A.T).max() == 0.0  # symmetric: the upper triangle holds it
        U = sp.triu(A).tocsr()
        U.sort_indices()
        np.savez_compressed(os.path.join(HERE, "cubo4820.npz"), n=A.shape[0],
                            U_rowptr=U.indptr.astype(np.int32), U_col=U.indices.astype(np.int32),
                            U_val=U.data.astype(np.float64), TV0=np.asarray(d["TV0"], np.float64),
                            rhs=d["rhs"].ravel().astype(np.float64))
        print("cubo4820", A.shape, A.nnz)


def main():
    data_fixtures()
    if not (oracle.ref_available() and oracle.ref_exti_available()):
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    for name, A, tau in inputs():
        strong = H.soc_cla(A, tau)
        r01 = H.rng_uniform(99, 100, A.nrows)
        fc, nc = H.pmis(A, strong, r01)
        iat, ja, co = oracle.ref_classical_prol(A, strong, fc, nc, nthreads=1)
        np.savez_compressed(os.path.join(HERE, f"clas_{name}.npz"), rowptr=A.rowptr, col=A.col, val=A.val,
                            strong=strong, fcnode=fc, nc=nc, P_rowptr=iat, P_col=ja, P_val=co)
        print("clas", name, A.nrows, A.nnz, nc, len(ja))
        iat, ja, co = oracle.ref_exti_prol(A, strong, fc, nc, nthreads=1)
        np.savez_compressed(os.path.join(HERE, f"exti_{name}.npz"), rowptr=A.rowptr, col=A.col, val=A.val,
                            strong=strong, fcnode=fc, nc=nc, P_rowptr=iat, P_col=ja, P_val=co)
        print("exti", name, A.nrows, A.nnz, nc, len(ja))
    flux = os.path.join(REF, "examples", "matrices", "Flux.mat")
    if os.path.exists(flux):
        import scipy.io

        d = scipy.io.loadmat(flux)
        M = d["Amat"].tocsr()
        M.sort_indices()
        np.savez_compressed(os.path.join(HERE, "flux.npz"), rowptr=M.indptr.astype(np.int32),
                            col=M.indices.astype(np.int32), val=M.data.astype(np.float64),
                            b=d["b"].ravel().astype(np.float64))
        print("flux", M.shape, M.nnz)


if __name__ == "__main__":
    main()

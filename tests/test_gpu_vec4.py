"""Vector-load staging in the FP32 stream and packed-code products: the same bits as one
element per load.

FP32 stream and packed-code products load four consecutive nonzeros per lane from a
4-element-aligned position (elements of the neighbouring row blocks masked out);
CHRONOS_VEC4=0 loads one element per lane.  The products and their summation order are the
same, so V-cycles and PCG runs must agree bit for bit.  The sliced-ELL and diagonal layouts
are excluded here so that every single-vector product runs in the stream or packed layout,
and the stream blocks are fixed at 8 nonzeros per lane (the autotuner's timed choice could
differ between the two uploads); block starts fall at every offset modulo 4 on these
hierarchies."""
import numpy as np
import pytest

from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def _make(monkeypatch, env, make):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    try:
        return make()
    finally:
        for k in env:
            monkeypatch.delenv(k)


def _hiers():
    A, _ = H.heterogeneous_diffusion(24, 2.0, seed=5)
    E, _ = H.elasticity(6)
    E, _ = H.bc_fix(E)
    he = H.build(E, H.SetupParams(prolong="bamg", soc="dom", soc_tau=0.01, prol_emin="emin",
                                  tv=H.rigid_body_modes(6), max_coarse=200))
    return [("diffusion24", A, H.build(A)), ("elasticity6", E, he)]


@pytest.mark.parametrize("pack", ["0", "1"])
@pytest.mark.parametrize("name,A,hier", _hiers())
def test_vcycle_and_pcg_vec4_bit_identical(gpu, monkeypatch, name, A, hier, pack):
    base = {"CHRONOS_SELL": "0", "CHRONOS_DIA": "0", "CHRONOS_PACK": pack, "CHRONOS_STREAM_PER": "8"}
    M4 = _make(monkeypatch, base, lambda: gpu.AMGPreconditioner(hier, "fp32"))
    M1 = _make(monkeypatch, dict(base, CHRONOS_VEC4="0"), lambda: gpu.AMGPreconditioner(hier, "fp32"))
    print(name, "pack", pack, "layouts", M4.layout()[0], flush=True)
    assert M4.layout() == M1.layout()
    b = H.rng_uniform(9, 2, A.nrows, -1, 1)
    np.testing.assert_array_equal(M4(b), M1(b))
    x4, it4, rv4, f4 = M4.pcg(b, itmax=200, tol=1e-8)
    x1, it1, rv1, f1 = M1.pcg(b, itmax=200, tol=1e-8)
    print(name, "iterations", it4, it1, flush=True)
    assert f4 == f1 == 0 and it4 == it1
    np.testing.assert_array_equal(rv4, rv1)
    np.testing.assert_array_equal(x4, x1)

"""The layout cache (CHRONOS_LAYOUT_CACHE): the first upload of a hierarchy times the layout
candidates and records the decisions; a second upload of the same matrices rebuilds the
same layouts without timing, and gives the same bits."""
import numpy as np
import pytest

from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def _hiers():
    A, _ = H.heterogeneous_diffusion(20, 2.0, seed=7)
    E, _ = H.elasticity(6)
    E, _ = H.bc_fix(E)
    he = H.build(E, H.SetupParams(prolong="bamg", soc="dom", soc_tau=0.01, prol_emin="emin",
                                  tv=H.rigid_body_modes(6), max_coarse=200))
    return [pytest.param(A, H.build(A), id="diffusion20"), pytest.param(E, he, id="elasticity6")]


@pytest.mark.parametrize("A,hier", _hiers())
def test_cached_layouts_equal_tuned(gpu, monkeypatch, tmp_path, A, hier):
    path = tmp_path / "layouts.txt"
    b = H.rng_uniform(3, 2, A.nrows, -1, 1)
    monkeypatch.setenv("CHRONOS_LAYOUT_CACHE", str(path))
    for prec in ("fp32", "fp64"):
        first = gpu.AMGPreconditioner(hier, prec)
        n_lines = len(path.read_text().splitlines())
        second = gpu.AMGPreconditioner(hier, prec)
        print(prec, "cache entries", n_lines, "layouts", first.layout()[0], flush=True)
        assert n_lines > 0 and len(path.read_text().splitlines()) == n_lines  # nothing re-timed
        assert first.layout() == second.layout()
        np.testing.assert_array_equal(first(b), second(b))
        if prec == "fp32":
            first_fp32_layout, first_fp32_out = first.layout(), first(b)
        x1, it1, _, _ = first.pcg(b, itmax=200, tol=1e-8)
        x2, it2, _, _ = second.pcg(b, itmax=200, tol=1e-8)
        assert it1 == it2
        np.testing.assert_array_equal(x1, x2)
    # The decisions as another process reads them: a copy of the file under a new path is
    # loaded afresh, and an upload from it times nothing and gives the same bits.
    copy = tmp_path / "layouts_copy.txt"
    copy.write_text(path.read_text())
    monkeypatch.setenv("CHRONOS_LAYOUT_CACHE", str(copy))
    n_copy = len(copy.read_text().splitlines())
    third = gpu.AMGPreconditioner(hier, "fp32")
    assert len(copy.read_text().splitlines()) == n_copy
    assert third.layout() == first_fp32_layout
    np.testing.assert_array_equal(third(b), first_fp32_out)

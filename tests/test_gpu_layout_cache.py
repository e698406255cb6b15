"""The layout cache (CHRONOS_LAYOUT_CACHE): the first upload of a hierarchy times the layout
candidates and records the decisions; a second upload of the same matrices rebuilds the
same layouts without timing, and gives the same bits."""
import os

import numpy as np
import pytest

from chronoslab_amd import hierarchy as H

pytestmark = pytest.mark.gpu


def test_cached_layouts_equal_tuned(gpu, monkeypatch, tmp_path):
    path = tmp_path / "layouts.txt"
    A, _ = H.heterogeneous_diffusion(20, 2.0, seed=7)
    hier = H.build(A)
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
        x1, it1, _, _ = first.pcg(b, itmax=100, tol=1e-8)
        x2, it2, _, _ = second.pcg(b, itmax=100, tol=1e-8)
        assert it1 == it2
        np.testing.assert_array_equal(x1, x2)

"""fsph_reductions_noalloc (partial sums in the library's static device
scratch array) against fsph_reductions (partial sums in the context's own
buffer), through the C ABI: bit-equal on two calls in a row, the second on
a smaller input, so the second call reuses the scratch the first one left
behind."""
import numpy as np
import pytest

from flecsph_amd import BodySystem, _abi, generators

pytestmark = pytest.mark.gpu


def _stepped(parts, prm, steps):
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    for _ in range(steps):
        g.step()
    return g


def test_noalloc_reductions_equal_the_allocating_entry_point():
    # 60k self-gravitating particles: 235 workgroups of partials and a
    # non-zero potential term; then 3k particles: 12 workgroups.
    big, pb = generators.polytrope(60_000)
    small, ps = generators.random_box(3000, seed=7)
    for name, parts, prm, steps in (("polytrope60k", big, pb, 2),
                                    ("random_box3k", small, ps, 2)):
        print(f"noalloc reductions: {name}", flush=True)
        g = _stepped(parts, prm, steps)
        a = g.reductions_noalloc()
        b = g.reductions()
        c = g.reductions_noalloc()
        assert np.array_equal(a, b), (name, a, b)
        assert np.array_equal(c, b), (name, c, b)
        assert a[0] > 0 and np.all(np.isfinite(a))
        if prm.get("gravity"):
            assert a[6] < 0  # 1/2 sum m phi of a bound star
        g.close()

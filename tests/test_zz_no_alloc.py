"""tsp_get_statistics_noalloc, the statistics entry point that allocates nothing (its kernels take
a static per-GPU scratch array), against tsp_get_statistics on the same context: every value
identical. Two calls in a row, the second on a smaller scene, each made before the allocating
variant has run on its context; both through the C ABI.
"""
import ctypes

import pytest

from turbospharticles_amd import _capi, scenes
from turbospharticles_amd.backend import HipBackend

pytestmark = pytest.mark.gpu


def _stats(g, fn):
    out = _capi.TspStatistics()
    rc = getattr(g.lib, fn)(g.h, ctypes.byref(out))
    assert rc == 0, (fn, rc, g.lib.tsp_last_error(g.h).decode())
    return {name: (list(v) if hasattr(v, "__len__") else v)
            for name, v in ((f, getattr(out, f)) for f, _ in _capi.TspStatistics._fields_)}


@pytest.mark.parametrize("n_side,steps", [(40, 2), (12, 1)])
def test_statistics_without_allocation_equal_the_allocating_entry_point(n_side, steps):
    print(f"case: statistics of a {n_side}^3 dam break after {steps} steps", flush=True)
    g = HipBackend(0)
    g.set_params(scenes.make_params(0.01, termination_cond=scenes.TERM_POS))
    g.pull(scenes.build_scene(scenes.dam_break_spec(n_side, 0.01)))
    g.step(steps)
    a = _stats(g, "tsp_get_statistics_noalloc")
    b = _stats(g, "tsp_get_statistics")
    c = _stats(g, "tsp_get_statistics_noalloc")
    g.close()
    assert a["n_fluid"] == n_side ** 3 and a["nbr_max"][0] > 0
    assert a == b == c

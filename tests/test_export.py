"""Export snapshots (Sync::ExportSnapshot) on the CPU side: the oracle's surface classification
pinned by a closed form.

Classification (IsNearSurfaceFluidParticle, core/backend_sim.h:109-117; cpu_backend.cpp:910-981):
on an exact cubic lattice of spacing h (a power of two, so every pair distance is exact and the
pairs at exactly 2h are neighbors) an interior particle has 32 fluid neighbors, one on a face
at most 22, one a layer below the face 31; with the default threshold 27 the near-surface
particles are exactly the outer layer, and the export carries them and their neighbors: the
three outer layers, n^3 - (n - 6)^3 particles.
"""
import numpy as np
import pytest

from turbospharticles_amd import _capi, scenes

import oracle_lib as O

H = 2.0 ** -7


def lattice_state(n, origin=0.5):
    st = _capi.HostState(n ** 3, 0, 0, 0)
    idx = np.arange(n, dtype=np.float32)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    st.fluid["x"][:] = np.float32(origin) + x.ravel() * np.float32(H)
    st.fluid["y"][:] = np.float32(origin) + y.ravel() * np.float32(H)
    st.fluid["z"][:] = np.float32(origin) + z.ravel() * np.float32(H)
    st.fluid["density"][:] = 1000.0
    st.fluid_id[:] = np.arange(n ** 3, dtype=np.uint64)
    st.domain = (np.zeros(3, np.float32), np.ones(3, np.float32))
    return st


def classify(st, threshold=27):
    o = O.Oracle(scenes.make_params(H), threads=4)
    o.pull(st)
    o.neighbor_search()
    flags = o.export_surface(threshold)
    return flags, o.push()


@pytest.mark.parametrize("n", [7, 10])
def test_lattice_export_is_the_three_outer_layers(n):
    flags, s = classify(lattice_state(n))
    layer = np.stack([np.rint((s.fluid[c].astype(np.float64) - 0.5) / H) for c in ("x", "y", "z")], 1)
    depth = np.minimum(layer, n - 1 - layer).min(1)
    assert np.array_equal(flags.astype(bool), depth <= 2)
    assert int(flags.sum()) == n ** 3 - max(n - 6, 0) ** 3


def test_threshold_bounds():
    st = lattice_state(8)
    assert classify(st, threshold=0)[0].sum() == 0            # no particle has fewer than 0 neighbors
    assert classify(st, threshold=33)[0].sum() == 8 ** 3      # every particle has at most 32


def test_particles_touching_walls_are_exported():
    st = scenes.build_scene(scenes.dam_break_spec(12, 0.01))
    o = O.Oracle(scenes.make_params(0.01), threads=4)
    o.pull(st)
    o.neighbor_search()
    flags = o.export_surface(0)   # only the boundary rule can mark a particle
    cfb, _, _ = o.lists(O.FB)
    assert cfb.sum() > 0
    near = cfb > 0
    assert np.all(flags[near] == 1)
    # and the marks spread exactly to their FF neighbors
    cff, iff, _ = o.lists(O.FF)
    off = np.concatenate([[0], np.cumsum(cff.astype(np.int64))])
    expect = near.copy()
    for i in np.nonzero(near)[0]:
        expect[iff[off[i]:off[i + 1]]] = True
    assert np.array_equal(flags.astype(bool), expect)

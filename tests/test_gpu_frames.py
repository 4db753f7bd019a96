"""Cell frames beyond 1024 cells per axis, and 64-bit sort keys on the HIP path.

A 21 m channel (21 x 0.1 x 0.1 m at h = 0.01) with a small fluid block at one end: its walls give
a cell frame of 1070 x 25 x 25 cells, whose Morton ids reach bit 30 (the reference's frames take
up to 2^21 cells per axis with 63-bit ids, core/cell_frame.h:29, neighbor_search.cpp:2032-2070).
A 42 m channel (2,120 cells along x) has Morton ids beyond 2^32, so the context sorts 64-bit keys.
Morton order, every list and three steps must equal the oracle bit for bit; the channels in a
batch with an ordinary dam break (64-bit keys for all three) equal each scene alone.
"""
import numpy as np
import pytest

from turbospharticles_amd import scenes
from turbospharticles_amd.backend import HipBackend, HipBatch

import oracle_lib as O
from helpers import FIELDS_STEP

pytestmark = pytest.mark.gpu

H = 0.01


def channel(length=21.0):
    return scenes.build_scene(scenes.dam_break_spec(0, H, counts=(20, 8, 8), tank=(length, 0.1, 0.1)))


@pytest.mark.parametrize("length,key_bits", [(21.0, 30), (42.0, 33)])
def test_frame_beyond_1024_cells_per_axis_matches_the_oracle(length, key_bits):
    print(f"case: {length} m channel, Morton ids beyond {key_bits} bits", flush=True)
    st = channel(length)
    p = scenes.make_params(H, termination_cond=scenes.TERM_POS)
    g = HipBackend(0)
    g.set_params(p)
    g.pull(st)
    o = O.Oracle(p, chunk_width=1, threads=8)
    o.pull(st)
    g.neighbor_search()
    o.neighbor_search()
    gi, oi = g.info(), o.info()
    assert list(gi.frame_dim) == list(oi.frame_dim) and max(gi.frame_dim) > 1024
    assert list(gi.frame_orig) == list(oi.frame_orig)
    sg, so = g.push(), o.push()
    assert np.array_equal(sg.fluid_id, so.fluid_id) and np.array_equal(sg.boundary_id, so.boundary_id)
    _, kf_g, kb_g = g.sort_state()
    _, kf_o, _, kb_o = o.order()
    assert np.array_equal(kf_g, kf_o)
    # the boundary's cell ids as a multiset, and ids beyond 30 bits (x cells past 1024)
    assert np.array_equal(np.sort(kb_g), np.sort(kb_o))
    assert int(kb_o.max()).bit_length() > key_bits
    for rel in (O.FF, O.FB):
        cg, ig = g.neighbors(rel)
        co, io, _ = o.lists(rel)
        assert np.array_equal(cg, co) and np.array_equal(ig, io), rel
    for s in range(3):
        g.step(1)
        o.step()
        assert g.info().solver_iters == o.info().solver_iters, s
    sg, so = g.push(), o.push()
    assert np.array_equal(sg.fluid_id, so.fluid_id)
    for f in FIELDS_STEP:
        assert np.array_equal(sg.fluid[f], so.fluid[f]), f
    g.close()


def test_large_frame_scene_in_a_batch_matches_each_scene_alone():
    print("case: both channels and a dam break in one context, 4 steps", flush=True)
    p = scenes.make_params(H, termination_cond=scenes.TERM_POS)
    states = [channel(), scenes.build_scene(scenes.dam_break_spec(12, H, tank=(0.4, 0.3, 0.3))), channel(42.0)]
    alone = []
    for st in states:
        g = HipBackend(0)
        g.set_params(p)
        g.pull(st)
        g.step(4)
        alone.append(g.push())
        g.close()
    b = HipBatch(0)
    ids = b.update(add=[(p, st) for st in states])
    b.step(4)
    for sid, ref in zip(ids, alone):
        got = b.push(sid)
        assert np.array_equal(got.fluid_id, ref.fluid_id), sid
        for f in FIELDS_STEP:
            assert np.array_equal(got.fluid[f], ref.fluid[f]), (sid, f)
    b.close()

"""Scenes generated on the device (tsp_pull_spec, tsp_batch_add_scene_spec) against the host
builder (tsp_scene_build + tsp_pull): every pushed array bit-identical -- fluid positions with
their SplitMix64 jitter, the boundary layers with their psi / V kernel sums, the rigid body, the
ids -- and so is a step from either, alone and inside a batch.
"""
import time

import numpy as np
import pytest

from turbospharticles_amd import scenes
from turbospharticles_amd.backend import HipBackend, HipBatch

pytestmark = pytest.mark.gpu


def same_state(a, b):
    assert a.n == b.n
    for k in ("fluid_id", "boundary_id", "rigid_id", "rigid_body", "body_id", "body_count"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for g in ("fluid", "boundary", "rigid", "body"):
        for k in getattr(a, g):
            x, y = getattr(a, g)[k], getattr(b, g)[k]
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (g, k)


SPECS = {
    "dam40": lambda: scenes.dam_break_spec(40, 0.01),
    "dam13_3layers": lambda: scenes.dam_break_spec(13, 0.01, tank=(0.3, 0.2, 0.25), layers=3),
    "cube_drop31": lambda: scenes.scaled_cube_drop_spec(31, 0.01),
    "dam12_nojitter": lambda: scenes.dam_break_spec(12, 0.01, jitter_frac=0.0),
}


@pytest.mark.parametrize("case", sorted(SPECS))
def test_device_builder_equals_host_builder(case):
    print(f"case: device scene builder, {case}", flush=True)
    spec = SPECS[case]()
    p = scenes.make_params(spec.particle_size, termination_cond=scenes.TERM_POS)
    host = scenes.build_scene(spec)
    a = HipBackend(0)
    a.set_params(p)
    a.pull(host)
    b = HipBackend(0)
    b.set_params(p)
    b.pull_spec(spec)
    same_state(a.push(), b.push())
    a.step(2)
    b.step(2)
    same_state(a.push(), b.push())
    a.close()
    b.close()


def test_device_builder_in_a_batch():
    print("case: device-built scenes in a batch", flush=True)
    specs = [SPECS["cube_drop31"](), SPECS["dam40"]()]
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    ref = HipBatch(0)
    rid = [ref.add(p, scenes.build_scene(s)) for s in specs]
    dev = HipBatch(0)
    did = [dev.add_spec(p, s) for s in specs]
    for x, y in zip(rid, did):
        same_state(ref.push(x), dev.push(y))
    ref.step(3)
    dev.step(3)
    for x, y in zip(rid, did):
        same_state(ref.push(x), dev.push(y))
    ref.close()
    dev.close()


def test_one_million_particle_scene_from_the_device():
    print("case: device scene builder, 1M rigid cube scene", flush=True)
    spec = scenes.rigid_cube_spec(100, 0.004, side=0.1, density=500.0, center=(0.7, 0.3, 0.25), rotation_seed=7)
    p = scenes.make_params(0.004, termination_cond=scenes.TERM_POS)
    t0 = time.perf_counter()
    host = scenes.build_scene(spec)
    a = HipBackend(0)
    a.set_params(p)
    a.pull(host)
    a.synchronize()
    t1 = time.perf_counter()
    b = HipBackend(0)
    b.set_params(p)
    b.pull_spec(spec)
    b.synchronize()
    t2 = time.perf_counter()
    print(f"  host build + pull {t1 - t0:.2f} s, device build {t2 - t1:.2f} s", flush=True)
    same_state(a.push(), b.push())
    a.close()
    b.close()

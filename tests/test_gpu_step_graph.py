"""Steps as captured graphs against steps launched one by one: the same bits.

The captured form (tsp_set_step_graph) replays the search, the prediction, the first sweep group
and the tail as graphs with one status read; the device gates call the rest of a step off when
the host has to decide (a frame rebuild, storage growth, a contact solve, more sweeps), and the
host finishes that step launch by launch. A batch of three scenes -- two dam breaks that spread
and rebuild their frames, a cube dropped onto the floor so that the contact solve starts -- is
stepped both ways and compared after every step call.
"""
import numpy as np
import pytest

from turbospharticles_amd import _capi, scenes
from turbospharticles_amd.backend import HipBatch

from helpers import FIELDS_STEP

pytestmark = pytest.mark.gpu


def _scenes():
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    dams = [scenes.build_scene(scenes.dam_break_spec(n, 0.01, tank=(0.5, 0.3, 0.3))) for n in (14, 20)]
    cube = scenes.build_scene(scenes.rigid_cube_spec(12, 0.01, side=0.05, center=(0.22, 0.03, 0.1),
                                                     tank=(0.3, 0.25, 0.2)))
    return [(p, st) for st in dams + [cube]]


def test_captured_steps_equal_launched_steps():
    print("case: batch of two dam breaks and a cube drop, 12 calls of 5 steps, graphs on and off", flush=True)
    runs = []
    for on in (True, False):
        b = HipBatch(0)
        b.set_step_graph(on)
        ids = b.update(add=_scenes())
        infos = []
        for call in range(12):
            b.step(5)
            infos.append([b.info(sid) for sid in ids])
        runs.append((b, ids, infos))
    (bg, ig, fg), (bl, il, fl) = runs
    graph_calls = 0
    for call, (a, c) in enumerate(zip(fg, fl)):
        for sa, sc in zip(a, c):
            assert sa.solver_iters == sc.solver_iters, call
            assert sa.contact_iters == sc.contact_iters, call
            assert (sa.flags & ~_capi.FLAG_STEP_GRAPH) == sc.flags, call
            assert not sc.flags & _capi.FLAG_STEP_GRAPH
        graph_calls += int(any(sa.flags & _capi.FLAG_STEP_GRAPH for sa in a))
    print(f"  calls with captured steps: {graph_calls} of {len(fg)}", flush=True)
    assert graph_calls > 0
    for a, c in zip(ig, il):
        sa, sc = bg.push(a), bl.push(c)
        assert np.array_equal(sa.fluid_id, sc.fluid_id)
        for f in FIELDS_STEP:
            assert np.array_equal(sa.fluid[f], sc.fluid[f]), f
        for f in sa.body:
            assert np.array_equal(sa.body[f], sc.body[f]), f
    bg.close()
    bl.close()

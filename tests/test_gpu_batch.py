"""Batched scenes: many independent scenes in one context give, scene by scene, the results of the
same scene run alone.

The scenes differ in particle size, counts (ragged, not multiples of the 256-particle block),
termination criterion, rigid bodies and frame behaviour (one scene's lone fast particle forces
frame rebuilds that the others must not notice). Every field, id, iteration count, sweep count
and frame-build count must be bit-identical between the scene alone and the scene in the mix,
after every step; so must the scenes that stay resident while others join and leave.
"""
import numpy as np
import pytest

from turbospharticles_amd import _capi, scenes
from turbospharticles_amd.backend import HipBackend, HipBatch

from helpers import FIELDS_STEP

pytestmark = pytest.mark.gpu


def _fast_particle():
    st = _capi.HostState(1, 0, 0, 0)
    st.fluid_id[:] = 0
    st.fluid["x"][:], st.fluid["y"][:], st.fluid["z"][:] = 0.5, 0.5, 0.5
    st.fluid["vx"][:] = 30.0
    return st


def _cases():
    small_tank = (0.4, 0.3, 0.3)
    return {
        "dam12_pos": (scenes.make_params(0.01, termination_cond=scenes.TERM_POS),
                      scenes.build_scene(scenes.dam_break_spec(12, 0.01, tank=small_tank))),
        "ragged_abs": (scenes.make_params(0.01, termination_cond=scenes.TERM_ABS, max_solver_iters=40),
                       scenes.build_scene(scenes.dam_break_spec(0, 0.01, counts=(17, 13, 11), tank=small_tank))),
        "rigid_cube": (scenes.make_params(0.01, termination_cond=scenes.TERM_POS),
                       scenes.build_scene(scenes.rigid_cube_spec(10, 0.01, side=0.06, center=(0.25, 0.15, 0.15),
                                                                 tank=small_tank))),
        "h02_raw": (scenes.make_params(0.02, termination_cond=scenes.TERM_RAW, jacobi_warm_start=0.5),
                    scenes.build_scene(scenes.dam_break_spec(8, 0.02, tank=(0.6, 0.4, 0.4)))),
        "fast_particle": (scenes.make_params(0.01, termination_cond=scenes.TERM_POS), _fast_particle()),
    }


def _alone(p, st, n_steps):
    g = HipBackend(0)
    g.set_params(p)
    g.pull(st)
    trace = []
    for _ in range(n_steps):
        g.step(1)
        i = g.info()
        trace.append((i.solver_iters, i.sweeps_total, i.frame_builds, i.step))
    return g, trace


def _record(b, sid):
    i = b.info(sid)
    return (i.solver_iters, i.sweeps_total, i.frame_builds, i.step)


def _same_state(sa, sb, name):
    assert np.array_equal(sa.fluid_id, sb.fluid_id), name
    for f in FIELDS_STEP:
        assert np.array_equal(sa.fluid[f], sb.fluid[f]), (name, f)
    assert np.array_equal(sa.boundary_id, sb.boundary_id), name
    for f in _capi.B_FIELDS:
        assert np.array_equal(sa.boundary[f], sb.boundary[f]), (name, f)
    for f in _capi.RB_FIELDS:
        assert np.array_equal(sa.body[f], sb.body[f]), (name, f)
    assert np.array_equal(sa.rigid_body, sb.rigid_body), name
    assert (sa.step, sa.time) == (sb.step, sb.time), name


def test_scenes_in_one_batch_match_each_scene_alone():
    n_steps = 8
    cases = _cases()
    print(f"case: batch of {len(cases)} scenes, {n_steps} steps", flush=True)
    b = HipBatch(0)
    ids = b.update(add=[(p, st) for p, st in cases.values()])
    traces = {sid: [] for sid in ids}
    for _ in range(n_steps):
        b.step(1)
        for sid in ids:
            traces[sid].append(_record(b, sid))
    for sid, (name, (p, st)) in zip(ids, cases.items()):
        print(f"  scene {name} alone", flush=True)
        g, trace = _alone(p, st, n_steps)
        assert trace == traces[sid], name
        _same_state(g.push(), b.push(sid), name)
        g.close()
    fast = ids[list(cases).index("fast_particle")]
    assert b.info(fast).frame_builds > 1, "the fast particle must force a frame rebuild"


def test_scenes_joining_and_leaving_leave_the_others_unchanged():
    cases = _cases()
    pa, sa = cases["dam12_pos"]
    pb, sb = cases["ragged_abs"]
    pc, sc = cases["rigid_cube"]
    print("case: admission and removal between steps", flush=True)
    b = HipBatch(0)
    ia, ib = b.update(add=[(pa, sa), (pb, sb)])
    b.step(3)
    (ic,) = b.update(add=[(pc, sc)])
    b.step(3)
    state_a = b.push(ia)
    b.remove(ia)
    assert b.ids() == [ib, ic]
    b.step(2)
    g, _ = _alone(pa, sa, 6)
    _same_state(g.push(), state_a, "dam12_pos (6 steps, then removed)")
    g, _ = _alone(pb, sb, 8)
    _same_state(g.push(), b.push(ib), "ragged_abs (8 steps)")
    g, _ = _alone(pc, sc, 5)
    _same_state(g.push(), b.push(ic), "rigid_cube (joined after 3 steps)")


def test_served_stream_with_a_tight_budget_matches_the_plain_form():
    # five requests arriving at once; the budget holds at most two of them, so the rest wait and
    # join as others finish; every result must equal the plain one-scene-per-context run
    from turbospharticles_amd import serving
    print("case: served stream, 5 requests, budget for 2", flush=True)
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    specs = [(scenes.scaled_dam_break_spec(8, 0.01), 7), (scenes.scaled_cube_drop_spec(10, 0.01), 12),
             (scenes.scaled_dam_break_spec(12, 0.01, jitter_seed=3), 5), (scenes.scaled_dam_break_spec(9, 0.01), 9),
             (scenes.scaled_cube_drop_spec(8, 0.01, rotation_seed=2), 4)]

    def make():
        return [serving.Request(q, 0.0, p, scenes.build_scene(s), n) for q, (s, n) in enumerate(specs)]

    rb = make()
    budget = 2.5 * max(r.footprint for r in rb)
    serving.serve_batched(rb, budget, chunk_steps=3)
    rp = make()
    serving.serve_plain(rp)
    assert max(r.admitted_s for r in rb) > min(r.done_s for r in rb), "some request must wait for another to finish"
    for a, c in zip(rb, rp):
        _same_state(a.result, c.result, f"request {a.rid}")
        assert a.result.step == a.steps


def test_served_stream_of_scene_specs_matches_host_built_states():
    # the same five requests, every other one carried as a scene spec (generated on the device
    # at admission); each result must equal the plain run from host-built states
    from turbospharticles_amd import serving
    print("case: served stream, specs and states mixed, budget for 2", flush=True)
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    specs = [(scenes.scaled_cube_drop_spec(10, 0.01), 12), (scenes.scaled_dam_break_spec(8, 0.01), 7),
             (scenes.scaled_dam_break_spec(12, 0.01, jitter_seed=3), 5), (scenes.scaled_dam_break_spec(9, 0.01), 9),
             (scenes.scaled_cube_drop_spec(8, 0.01, rotation_seed=2), 4)]
    rb = [serving.Request(q, 0.0, p, None, n, spec=s) if q % 2 == 0 else
          serving.Request(q, 0.0, p, scenes.build_scene(s), n) for q, (s, n) in enumerate(specs)]
    budget = 2.5 * max(r.footprint for r in rb)
    serving.serve_batched(rb, budget, chunk_steps=3)
    rp = [serving.Request(q, 0.0, p, scenes.build_scene(s), n) for q, (s, n) in enumerate(specs)]
    serving.serve_plain(rp)
    rs = [serving.Request(q, 0.0, p, None, n, spec=s) for q, (s, n) in enumerate(specs)]
    serving.serve_plain(rs)
    for a, c, e in zip(rb, rp, rs):
        _same_state(a.result, c.result, f"request {a.rid}")
        _same_state(e.result, c.result, f"request {e.rid} (plain, from its spec)")


# cells of three far-pair frames and the small dam breaks, not of four far pairs (the frames of
# _far_pair(0..8) hold 360,639 to 528,471 cells, the dam breaks 16,128 to 31,500)
CELL_LIMIT = 1_500_000


def _far_pair(q):
    """Two fluid particles 19.6 m apart along x: a cell frame of 999 x (19..23) x (19..23) cells
    at h 0.01; a context limited to CELL_LIMIT cells holds three such scenes."""
    st = _capi.HostState(2, 0, 0, 0)
    st.fluid_id[:] = [2 * q, 2 * q + 1]
    st.fluid["x"][:] = [0.0, 19.6]
    st.fluid["y"][:] = [0.5 + 0.01 * q, 0.5]
    st.fluid["z"][:] = [0.5, 0.5 - 0.01 * q]
    return st


def test_served_stream_waits_when_the_cell_tables_are_full():
    """Six requests of scenes with large cell frames arrive together; the context's cell limit
    holds three of them at a time, so admissions are refused (TSP_ERR_CAPACITY) and requeued until
    scenes leave. Dam breaks in between are admitted as usual; every result equals its plain run."""
    from turbospharticles_amd import serving
    print("case: served stream, cell limit for three large-frame scenes", flush=True)
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)

    def make():
        reqs = []
        for q in range(6):
            reqs.append(serving.Request(2 * q, 0.0, p, _far_pair(q), 3 + q))
            reqs.append(serving.Request(2 * q + 1, 0.0, p, None, 2 + q % 3,
                                        spec=scenes.scaled_dam_break_spec(6 + q, 0.01, jitter_seed=q)))
        return reqs

    rb = make()
    serving.serve_batched(rb, 200e9, chunk_steps=2, cell_limit=CELL_LIMIT)
    rp = make()
    serving.serve_plain(rp)
    far = [r for r in rb if r.rid % 2 == 0]
    assert max(r.admitted_s for r in far) >= min(r.done_s for r in far), "a large-frame scene must wait"
    for a, c in zip(rb, rp):
        assert a.result.step == a.steps
        _same_state(a.result, c.result, f"request {a.rid}")


def test_refused_admission_keeps_the_resident_scenes_and_drops_removed_ids():
    """A batch update that removes one scene and adds two that do not fit the cell limit: the
    context refuses the new scenes, the removal takes effect, the remaining scenes step on as if
    nothing happened, and the wrapper's scene table follows the context."""
    from turbospharticles_amd.backend import TspError
    print("case: refused admission with a removal", flush=True)
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    b = HipBatch(0)
    b.set_cell_limit(CELL_LIMIT)
    ids = b.update(add=[(p, _far_pair(q)) for q in range(3)])
    dam = scenes.build_scene(scenes.dam_break_spec(8, 0.01, tank=(0.3, 0.2, 0.2)))
    (kept,) = b.update(add=[(p, dam)])
    b.step(2)
    with pytest.raises(TspError) as e:
        b.update(remove=[ids[0]], add=[(p, _far_pair(7)), (p, _far_pair(8))])
    assert e.value.code == _capi.TSP_ERR_CAPACITY
    assert sorted(b.counts) == sorted(b.ids()) == sorted([ids[1], ids[2], kept])
    b.step(3)
    g, _ = _alone(p, dam, 5)
    _same_state(g.push(), b.push(kept), "dam break kept through the refusal")
    g.close()
    g, _ = _alone(p, _far_pair(2), 5)
    _same_state(g.push(), b.push(ids[2]), "large-frame scene kept through the refusal")
    g.close()
    b.close()

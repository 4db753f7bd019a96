"""The rigid contact solve on the device against the oracle's restatement of the reference CPU
solver (rigid_solver.cpp:23-799).

- RB and RR lists: the same entries in the same order as the oracle's (exact).
- A cube dropped onto the floor and two cubes colliding: per step the contact iteration counts
  and convergence flags equal the oracle's and the contact error average agrees to 1e-5
  relative; after the steps the body states agree to 1e-5 relative (the FP64 sums of the
  contact loads and errors run in another order than the oracle's serial loop, so the last bit
  of a body sum may differ), and every fluid field of the landing case is bit-identical.
- A contact scene inside a batch of mixed scenes gives bit-identical results to the same scene
  alone.
- The contact parameters are checked before a step of a scene with bodies.
"""
import numpy as np
import pytest

from turbospharticles_amd import scenes
from turbospharticles_amd.backend import HipBackend, HipBatch

import oracle_lib as O
from helpers import FIELDS_STEP, cube_scene, dam_break, merge_bodies, rel_err

pytestmark = pytest.mark.gpu

H = 0.01
BODY_FIELDS = ("xcm", "ycm", "zcm", "vx", "vy", "vz", "wx", "wy", "wz", "lx", "ly", "lz",
               "a00", "a01", "a02", "a10", "a11", "a12", "a20", "a21", "a22")


def params(**kw):
    return scenes.make_params(H, termination_cond=scenes.TERM_POS, **kw)


def landing_cube():
    return cube_scene(center=(0.15, 0.035, 0.15), vel=(0.0, -1.0, 0.0), omega=(0.0, 0.0, 2.0))


def colliding_cubes():
    a = cube_scene(center=(0.12, 0.15, 0.15), vel=(1.5, 0.0, 0.0), omega=(0.0, 2.0, 0.0))
    b = cube_scene(center=(0.188, 0.155, 0.152), vel=(-1.5, 0.0, 0.0), omega=(1.0, 0.0, 0.0),
                   quat=(0.99, 0.1, 0.05, 0.0))
    return merge_bodies(a, b)


def pressed_cube():
    return cube_scene(center=(0.15, 0.028, 0.15), vel=(0.0, -1.0, 0.0), omega=(0.0, 0.0, 3.0))


def gpu(st, p):
    g = HipBackend(0)
    g.set_params(p)
    g.pull(st)
    return g


@pytest.mark.parametrize("case", ["pressed_cube", "colliding_cubes"])
def test_contact_lists_match_the_oracle(case):
    print(f"case: contact lists, {case}", flush=True)
    st = pressed_cube() if case == "pressed_cube" else colliding_cubes()
    p = params()
    g = gpu(st, p)
    o = O.Oracle(p, threads=4)
    o.pull(st)
    g.neighbor_search()
    o.neighbor_search()
    total = 0
    for rel in (O.RF, O.RB, O.RR):
        cg, ig = g.neighbors(rel)
        co, io, _ = o.lists(rel)
        assert np.array_equal(cg, co), rel
        assert np.array_equal(ig, io), rel
        if rel != O.RF:
            total += int(co.sum())
    assert total > 0
    g.close()


def _run_pair(st, p, steps):
    g = gpu(st, p)
    o = O.Oracle(p, threads=4)
    o.pull(st)
    solved = 0
    for s in range(steps):
        g.step(1)
        o.step()
        gi, oi = g.info(), o.info()
        assert gi.solver_iters == oi.solver_iters, s
        assert gi.contact_iters == oi.contact_iters, s
        assert gi.contact_converged == oi.contact_converged, s
        assert gi.contact_boundary == oi.contact_boundary, s
        assert gi.contacts == oi.contacts, s
        assert gi.contact_avg_err == pytest.approx(oi.contact_avg_err, rel=1e-5, abs=1e-9), s
        assert bool(gi.flags & 4) == (oi.n_rb + oi.n_rr > 0), s
        solved += gi.contact_iters > 0
    sg, so = g.push(), o.push()
    g.close()
    return sg, so, solved


def test_landing_cube_matches_the_oracle():
    print("case: cube dropped on the floor, 14 steps", flush=True)
    st = landing_cube()
    sg, so, solved = _run_pair(st, params(), 14)
    assert solved >= 2
    for f in BODY_FIELDS:
        assert rel_err(sg.body[f], so.body[f]) <= 1e-5, f
    assert np.array_equal(sg.fluid_id, so.fluid_id)
    for f in FIELDS_STEP:
        assert np.array_equal(sg.fluid[f], so.fluid[f]), f
    # the floor stopped the cube (it fell at 1 m/s)
    assert abs(float(sg.body["vy"][0])) < 0.2


def test_colliding_cubes_match_the_oracle():
    print("case: two cubes colliding, 10 steps", flush=True)
    st = colliding_cubes()
    sg, so, solved = _run_pair(st, params(), 10)
    assert solved >= 1
    for f in BODY_FIELDS:
        assert rel_err(sg.body[f], so.body[f]) <= 1e-5, f
    # the bodies met at 1.5 m/s each and bounced: they now move apart
    assert float(sg.body["vx"][0]) < 0.0 < float(sg.body["vx"][1])


def test_contact_scene_in_a_batch_is_identical_to_alone():
    print("case: landing cube, colliding cubes and a dam break in one batch, 12 steps", flush=True)
    p = params()
    states = [landing_cube(), dam_break(12, H), colliding_cubes()]
    alone = []
    for st in states:
        g = gpu(st, p)
        g.step(12)
        alone.append(g.push())
        g.close()
    b = HipBatch(0)
    ids = [b.add(p, st) for st in states]
    b.step(12)
    for sid, ref in zip(ids, alone):
        got = b.push(sid)
        for f in FIELDS_STEP:
            assert np.array_equal(got.fluid[f], ref.fluid[f]), (sid, f)
        for f in got.body:
            assert np.array_equal(got.body[f], ref.body[f]), (sid, f)
    b.close()


def test_contact_parameters_are_checked_before_a_step():
    print("case: contact iteration limit 0", flush=True)
    g = gpu(landing_cube(), params(rigid_contact_max_solver_iters=0))
    with pytest.raises(RuntimeError, match="contact"):
        g.step(1)
    g.close()

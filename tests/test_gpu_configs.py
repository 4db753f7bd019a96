"""The benchmark workloads at their own sizes, on the HIP path:

  - the rigid-body coupling scene: 1,000,000 fluid (100^3 at h = 0.004, seeded jitter), 320,564
    wall particles and a hollow 0.1 m cube of density 500 (6,364 rigid particles) released at rest
    at height 0.3 with a seeded rotation; three steps against the FP32 oracle (the restatement of
    the reference CPU path, SimBackend_CPU::Step_IISPH, sph_solver.cpp:34-319): ids and every
    fluid field bit-identical, equal solver iteration counts, the body within 1e-6 (its FP64 force
    sums run in a fixed block order on the device, serially in the oracle); and the same
    scene with the cube resting on the floor, so the rigid contact solve runs at 1M
    (rigid_solver.cpp:602-799): equal contact iteration counts, the body within 1e-5;
  - 256 mixed dam breaks (fluid counts log-uniform in [2,000, 200,000], tanks scaled to their
    blocks) advanced in one context: after five steps every scene equals its own plain run, and
    three of them (smallest, median, largest) equal the oracle;
  - a served stream of dam breaks and dropped cubes of 5,000 to 1,000,000 fluid particles (step
    counts cut to 4-16 from [200, 2,000] to fit a test), admitted under an HBM budget that holds
    the largest with little room beside it: every result equals its plain run.
"""
import os

import numpy as np
import pytest

from turbospharticles_amd import _capi, scenes, serving
from turbospharticles_amd.backend import HipBackend, HipBatch

import oracle_lib as O
from helpers import FIELDS_STEP, rel_err

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
BODY_FIELDS = ("xcm", "ycm", "zcm", "vx", "vy", "vz", "wx", "wy", "wz", "lx", "ly", "lz",
               "a00", "a01", "a02", "a10", "a11", "a12", "a20", "a21", "a22")


def cube_1m(resting=False):
    if not resting:
        return scenes.rigid_cube_spec(100, 0.004, side=0.1, density=500.0, center=(0.7, 0.3, 0.25), rotation_seed=7)
    # axis-aligned, its bottom layer half a particle above the floor: contacts from the first step
    spec = scenes.rigid_cube_spec(100, 0.004, side=0.1, density=500.0, center=(0.7, 0.05, 0.25), rotation_seed=7)
    for a, q in enumerate((1.0, 0.0, 0.0, 0.0)):
        spec.rigid_quat[a] = q
    return spec


def _steps_against_the_oracle(spec, steps, body_tol):
    st = scenes.build_scene(spec)
    assert st.n == (1_000_000, 320_564, 6_364, 1)
    p = scenes.make_params(0.004, termination_cond=scenes.TERM_POS)
    g = HipBackend(0)
    g.set_params(p)
    g.set_step_graph(True)
    g.pull(st)
    o = O.Oracle(p, chunk_width=1, threads=THREADS)
    o.pull(st)
    contact = graph = 0
    for s in range(steps):
        g.step(1)
        o.step()
        gi, oi = g.info(), o.info()
        graph += int(gi.flags & _capi.FLAG_STEP_GRAPH != 0)
        print(f"  step {s + 1}: iterations {gi.solver_iters} / {oi.solver_iters}, contact iterations "
              f"{gi.contact_iters} / {oi.contact_iters}, contact boundary particles {gi.contact_boundary} / "
              f"{oi.contact_boundary}", flush=True)
        assert gi.solver_iters == oi.solver_iters, s
        assert gi.contact_iters == oi.contact_iters, s
        assert gi.contact_converged == oi.contact_converged, s
        assert gi.contacts == oi.contacts, s
        assert gi.contact_boundary == oi.contact_boundary, s
        # the contact solve ran (a resting body may converge in its first evaluation)
        assert bool(gi.flags & 4) == (oi.n_rb + oi.n_rr > 0), s
        contact += int(gi.flags & 4 != 0)
    # steps after the first run as one captured graph unless the contact solve runs
    assert graph == (0 if contact else steps - 1)
    sg, so = g.push(), o.push()
    g.close()
    assert np.array_equal(sg.fluid_id, so.fluid_id)
    for f in FIELDS_STEP:
        assert np.array_equal(sg.fluid[f], so.fluid[f]), f
    exact = [f for f in BODY_FIELDS if np.array_equal(sg.body[f], so.body[f])]
    print(f"  body fields bit-identical to the oracle: {len(exact)} of {len(BODY_FIELDS)}", flush=True)
    for f in BODY_FIELDS:
        assert rel_err(sg.body[f], so.body[f]) <= body_tol, f
    return sg, contact


@pytest.mark.timeout(300)
def test_1m_cube_drop_three_steps_match_the_oracle():
    print("case: 1M fluid + 0.1 m cube released at height 0.3, 3 steps", flush=True)
    sg, contact = _steps_against_the_oracle(cube_1m(), 3, 1e-6)
    assert contact == 0
    assert float(sg.body["vy"][0]) < 0.0   # the cube falls


@pytest.mark.timeout(300)
def test_1m_cube_resting_on_the_floor_runs_the_contact_solve():
    print("case: 1M fluid + 0.1 m cube resting on the floor, 3 steps", flush=True)
    _, contact = _steps_against_the_oracle(cube_1m(resting=True), 3, 1e-5)
    assert contact == 3


def _same_fluid(a, b, what, steps=True):
    assert np.array_equal(a.fluid_id, b.fluid_id), what
    for f in FIELDS_STEP:
        assert np.array_equal(a.fluid[f], b.fluid[f]), (what, f)
    if steps:
        assert (a.step, a.time) == (b.step, b.time), what


@pytest.mark.timeout(600)
def test_256_mixed_scenes_in_one_context_match_plain_runs_and_the_oracle():
    steps = 5
    sides = scenes.log_uniform_sides(256, 2000, 200000, seed=2026)
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    specs = [scenes.scaled_dam_break_spec(n, 0.01, jitter_seed=67 + q) for q, n in enumerate(sides)]
    print(f"case: 256 dam breaks, {sum(n ** 3 for n in sides)} fluid, {steps} steps", flush=True)
    b = HipBatch(0)
    ids = b.update_specs(add=[(p, s) for s in specs])
    b.step(steps)
    order = np.argsort(sides, kind="stable")
    sample = {int(order[0]), int(order[len(order) // 2]), int(order[-1])}
    for q, (sid, spec) in enumerate(zip(ids, specs)):
        if q % 64 == 0:
            print(f"  plain run of scene {q}", flush=True)
        got = b.push(sid)
        g = HipBackend(0)
        g.set_params(p)
        g.pull_spec(spec)
        g.step(steps)
        _same_fluid(got, g.push(), f"scene {q} (side {sides[q]})")
        g.close()
        if q in sample:
            print(f"  oracle run of scene {q} (side {sides[q]})", flush=True)
            o = O.Oracle(p, chunk_width=1, threads=THREADS)
            o.pull(scenes.build_scene(spec))
            for _ in range(steps):
                o.step()
            so = o.push()
            _same_fluid(got, so, f"scene {q} against the oracle", steps=False)   # the oracle keeps no step count
            assert b.info(sid).solver_iters == o.info().solver_iters
    b.close()


@pytest.mark.timeout(600)
def test_served_stream_of_5k_to_1m_scenes_matches_plain_runs():
    n = 8
    rng = np.random.default_rng(17)
    sides = scenes.log_uniform_sides(n, 5000, 1000000, seed=18)
    sides[int(np.argmax(sides))] = 100          # the largest request is the 1M scene
    steps = rng.integers(4, 17, n)
    kinds = rng.integers(0, 2, n)
    arrivals = np.cumsum(rng.exponential(1.0 / 20.0, n))
    arrivals -= arrivals[0]
    p = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)

    def make():
        return [serving.Request(q, arrivals[q], p, None, int(steps[q]),
                                spec=(scenes.scaled_cube_drop_spec(sides[q], 0.01, rotation_seed=q) if kinds[q]
                                      else scenes.scaled_dam_break_spec(sides[q], 0.01, jitter_seed=67 + q)))
                for q in range(n)]

    rb = make()
    budget = 1.2 * max(r.footprint for r in rb)
    print(f"case: served stream, {n} requests, fluid {min(s ** 3 for s in sides)}..{max(s ** 3 for s in sides)}, "
          f"steps {int(steps.min())}..{int(steps.max())}", flush=True)
    serving.serve_batched(rb, budget, chunk_steps=4)
    rp = make()
    serving.serve_plain(rp)
    assert max(r.admitted_s for r in rb) > min(r.done_s for r in rb), "some request must wait for another"
    for a, c in zip(rb, rp):
        assert a.result.step == a.steps
        _same_fluid(a.result, c.result, f"request {a.rid} ({a.n[0]} fluid)")
        for f in a.result.body:
            assert np.array_equal(a.result.body[f], c.result.body[f]), (a.rid, f)

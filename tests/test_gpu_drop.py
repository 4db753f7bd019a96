"""The 1M rigid cube drop over its 500 steps (bench.py's workload) on the GPU against the oracle's
record of the same run (tests/golden/cube1m_drop500.json, written by
tests/golden/make_cube1m_drop.py from the FP32 restatement of the reference's TBB path): every
step's solver and contact iterations, mean density error, kinetic energy and body centre equal,
and the final state (every fluid field of all 1,000,000 particles and the body) bit-identical by
its SHA-256. The cube lands on the floor late in the run, so the contact solve is covered too."""
import json
import os

import pytest

from turbospharticles_amd import scenes

from helpers import drop_record, state_digest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cube1m_drop500.json")


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_1m_cube_drop_500_steps_match_the_oracle_record():
    from turbospharticles_amd.backend import HipBackend
    with open(GOLD) as f:
        gold = json.load(f)
    h = 0.004
    spec = scenes.rigid_cube_spec(100, h, side=0.1, density=500.0, center=(0.7, 0.3, 0.25), rotation_seed=7)
    params = scenes.make_params(h, termination_cond=scenes.TERM_POS)
    print("case: 1M fluid + 0.1 m cube released at height 0.3, 500 steps against the oracle record", flush=True)
    g = HipBackend(0)
    try:
        g.set_params(params)
        g.pull(scenes.build_scene(spec))
        contact_steps = 0
        for s in range(gold["steps"]):
            g.step(1)
            inf = g.info()
            assert inf.solver_iters == gold["iters"][s], s
            assert inf.contact_iters == gold["contact_iters"][s], s
            contact_steps += int(inf.contact_iters > 0)
            rec = drop_record(g.push(), params)
            for k, v in rec.items():
                assert v == gold[k][s], (s, k)
            if (s + 1) % 100 == 0:
                print(f"  step {s + 1}: {inf.solver_iters} iterations, identical so far", flush=True)
        assert contact_steps > 0
        assert state_digest(g.push()) == gold["final_digest"]
    finally:
        g.close()

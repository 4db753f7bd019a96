#!/usr/bin/env python3
"""Writes cube1m_drop500.json: the oracle's record of the 1M rigid cube drop (bench.py's workload:
100^3 fluid at h 0.004, a 0.1 m cube of density 500 released at height 0.3 with a seeded
rotation, 0.1% average compression error, 2..100 iterations) over 500 steps from the initial
state -- per step the solver and contact iterations, the mean density error, the kinetic energy
and the body's centre of mass, and at the end the SHA-256 of the whole state (tests/helpers.py
state_digest).

The oracle is the FP32 restatement of the reference's TBB path (oracle/), chunk width 1, the form
the HIP path reproduces bit for bit; tests/test_gpu_drop.py checks the GPU against this record.

  python tests/golden/make_cube1m_drop.py [--threads T] [--steps N]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from turbospharticles_amd import scenes  # noqa: E402
import oracle_lib as O  # noqa: E402
from helpers import drop_record, state_digest  # noqa: E402


def cube1m():
    h = 0.004
    spec = scenes.rigid_cube_spec(100, h, side=0.1, density=500.0, center=(0.7, 0.3, 0.25), rotation_seed=7)
    return scenes.build_scene(spec), scenes.make_params(h, termination_cond=scenes.TERM_POS)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--out", default=os.path.join(HERE, "cube1m_drop500.json"))
    args = ap.parse_args()
    state, params = cube1m()
    o = O.Oracle(params, precision=32, chunk_width=1, threads=args.threads)
    o.pull(state)
    rec = {"iters": [], "contact_iters": [], "mean_density_error": [], "kinetic_energy": [], "xcm": []}
    t0 = time.perf_counter()
    for s in range(args.steps):
        o.step()
        inf = o.info()
        rec["iters"].append(int(inf.solver_iters))
        rec["contact_iters"].append(int(inf.contact_iters))
        for k, v in drop_record(o.push(), params).items():
            rec[k].append(v)
        if (s + 1) % 25 == 0:
            print(f"step {s + 1}/{args.steps}: {rec['iters'][-1]} iterations ({time.perf_counter() - t0:.0f} s)", flush=True)
    out = {"case": "1M rigid cube drop (bench.py's workload), FP32 oracle, chunk width 1",
           "steps": args.steps, **rec, "final_digest": state_digest(o.push())}
    with open(args.out, "w") as f:
        json.dump(out, f)
    print("wrote", args.out, out["final_digest"])


if __name__ == "__main__":
    main()

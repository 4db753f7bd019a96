"""Autotuner of the build-time variants of the step (the TSP_TUNE_* switches in csrc/tsp_device.h).

Every variant gives the same bits; they differ in how the work maps onto the GPU: the SELL slice
height (32 or 64 particles per slice), the radix sort's digit width (8, 6 or 4 bits: 4 to 8 passes
over 28-bit keys), the threads of a solver sweep tile (256, or 128 with two particles each), the
sweeps' LDS reads of four halo records ahead of their sums, the size of a step's first sweep
group against the previous step's sweep count, the whole step as one graph launch or launch by
launch, the FF halo built from page bitmaps or a window bitmap and hash set, prefetched list entries, 24- or 32-byte Jacobi halo records, stored or
recomputed gradient factors, the halo cap (LDS per sweep block), register budgets of the sweep
kernels, sweep groups as HIP graphs, the neighbor query's walk of the stencil (cell by cell, or
rows of three cells) and its candidate loads in flight. The
tuner walks the switches one at a time from the current settings (coordinate descent), keeping a
change when it shortens the measured step.

  python -m turbospharticles_amd.tune build             # CPU: one libtsp_hip.so per candidate
  python -m turbospharticles_amd.tune measure [--steps N] # GPU: check and time every built candidate
  python -m turbospharticles_amd.tune apply              # write csrc/tsp_tuned.h from the results

`build` and `measure` walk one round of the descent (each switch against the current best);
run them again after `apply` for the next round. `measure` first checks each candidate against
the current build: five steps of a dam break with a rigid body must give identical state, else
the candidate is rejected. The timing is bench.py's 1M rigid-cube step (its value window, no
per-launch events). Results go to lib/tune/results.json; `apply` picks the fastest checked
candidate.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys

from . import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TUNE_DIR = os.path.join(B.LIB, "tune")
TUNED_H = os.path.join(HERE, "csrc", "tsp_tuned.h")

# switch -> values tried (the first is the default of tsp_device.h)
SPACE = {
    "TSP_TUNE_SLICE": [32, 64],
    "TSP_TUNE_RADIX_BITS": [8, 6, 4],
    "TSP_TUNE_SWEEP_THREADS": [256, 128],
    "TSP_TUNE_PREFETCH": [1, 0],
    "TSP_TUNE_GATHER4": [0, 1],
    "TSP_TUNE_SWEEP_AHEAD": [-1, 0, 1, 2],
    "TSP_TUNE_STEP_GRAPH": [0, 1],
    "TSP_TUNE_QUERY_WAVES": [0, 6, 8],
    "TSP_TUNE_RANGE_LOADS": [0, 1],
    "TSP_TUNE_BOX_CELLS": [0, 512, 768, 1024],
    "TSP_TUNE_JACOBI_REC24": [1, 0],
    "TSP_TUNE_KG_STORE": [0, 1],
    "TSP_TUNE_HALO_MAX": [1000, 960, 1024, 1064],
    "TSP_TUNE_GRAPHS": [1, 0],
    "TSP_TUNE_PA_WAVES": [0, 6, 8],
    "TSP_TUNE_JACOBI_WAVES": [0, 6],
    "TSP_TUNE_SEARCH_ROW": [1, 0, 2],
    "TSP_TUNE_LANE_SORT": [1, 0, 2],
    "TSP_TUNE_SLOT16": [1, 0],
    "TSP_TUNE_SEARCH_FLAT": [0, 1, 2],
    "TSP_TUNE_SEARCH_BATCH": [4, 8],
    "TSP_TUNE_PERSISTENT": [0, 1],
}


def current():
    """The settings of csrc/tsp_tuned.h, or the defaults."""
    cfg = {k: v[0] for k, v in SPACE.items()}
    if os.path.exists(TUNED_H):
        for line in open(TUNED_H):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "#define" and parts[1] in cfg:
                cfg[parts[1]] = int(parts[2])
    return cfg


def name_of(cfg):
    """The variant's directory name: the settings spelled out, or their hash when that is longer
    than a file name may be (the directory's config.json holds the settings either way)."""
    full = "_".join(f"{k.replace('TSP_TUNE_', '').lower()}{v}" for k, v in sorted(cfg.items()))
    return full if len(full) <= 200 else "cfg_" + hashlib.sha1(full.encode()).hexdigest()[:16]


def candidates():
    """The current settings and every one-switch change of them."""
    base = current()
    out = [dict(base)]
    for k, vals in SPACE.items():
        for v in vals:
            if v != base[k]:
                c = dict(base)
                c[k] = v
                out.append(c)
    return out


def build_variant(cfg):
    out_dir = os.path.join(TUNE_DIR, name_of(cfg))
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libtsp_hip.so")
    B._run(B.hip_command(out + ".tmp", [f"-D{k}={v}" for k, v in cfg.items()]))
    os.replace(out + ".tmp", out)
    with open(os.path.join(out_dir, "config.json"), "w") as f:
        json.dump(cfg, f)
    return out


def cmd_build(args):
    from concurrent.futures import ThreadPoolExecutor

    def one(cfg):
        build_variant(cfg)
        print("built", name_of(cfg), flush=True)

    with ThreadPoolExecutor(max_workers=max(1, args.jobs)) as pool:
        list(pool.map(one, candidates()))


CHECK = r"""
import sys, numpy as np
sys.path.insert(0, {root!r})
from turbospharticles_amd import scenes
from turbospharticles_amd.backend import HipBackend
st = scenes.build_scene(scenes.rigid_cube_spec(20, 0.01, side=0.06, center=(0.25, 0.15, 0.15), tank=(0.4, 0.3, 0.3)))
g = HipBackend(0)
g.set_params(scenes.make_params(0.01, termination_cond=scenes.TERM_POS))
g.pull(st)
g.step(5)
s = g.push()
np.savez({out!r}, id=s.fluid_id, **{{k: v for k, v in s.fluid.items()}}, **{{"b_" + k: v for k, v in s.body.items()}})
"""


def run_py(code, lib):
    env = dict(os.environ)
    if lib:
        env["TSP_HIP_LIB"] = lib
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)


def cmd_measure(args):
    import numpy as np
    results = {}
    ref = os.path.join(TUNE_DIR, "reference_state.npz")
    r = run_py(CHECK.format(root=ROOT, out=ref), None)
    if r.returncode != 0:
        raise RuntimeError("reference run failed:\n" + r.stderr)
    want = np.load(ref)
    for d in sorted(os.listdir(TUNE_DIR)):
        lib = os.path.join(TUNE_DIR, d, "libtsp_hip.so")
        if not os.path.exists(lib):
            continue
        cfg = json.load(open(os.path.join(TUNE_DIR, d, "config.json")))
        out = os.path.join(TUNE_DIR, d, "state.npz")
        r = run_py(CHECK.format(root=ROOT, out=out), lib)
        same = r.returncode == 0 and all(np.array_equal(want[k], np.load(out)[k]) for k in want.files)
        entry = {"config": cfg, "identical": bool(same)}
        if same:
            env = dict(os.environ, TSP_HIP_LIB=lib)
            b = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--no-cpu", "--no-batch", "--no-serve",
                                "--steps", str(args.steps)], env=env, capture_output=True, text=True, timeout=600)
            line = [x for x in b.stdout.splitlines() if x.startswith("{")]
            if b.returncode == 0 and line:
                j = json.loads(line[-1])
                entry["ms_per_step"] = j["ms_per_step"]
                entry["value"] = j["value"]
        results[d] = entry
        print(d, entry.get("ms_per_step"), "identical" if same else "REJECTED (results differ)", flush=True)
    with open(args.out or os.path.join(TUNE_DIR, "results.json"), "w") as f:
        json.dump(results, f, indent=1)


def cmd_apply(args):
    res = json.load(open(args.results or os.path.join(TUNE_DIR, "results.json")))
    ok = {k: v for k, v in res.items() if v.get("identical") and "ms_per_step" in v}
    best = min(ok.values(), key=lambda v: v["ms_per_step"])
    lines = ["// tsp_tuned.h -- build-time variant chosen by turbospharticles_amd/tune.py on MI355X",
             f"// (1M rigid cube drop: {best['ms_per_step']} ms per step; every candidate checked bit-identical)"]
    for k, v in sorted(best["config"].items()):
        lines += [f"#ifndef {k}", f"#define {k} {v}", "#endif"]
    with open(TUNED_H, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build")
    b.add_argument("--jobs", type=int, default=4, help="candidates compiled at once")
    m = sub.add_parser("measure")
    m.add_argument("--steps", type=int, default=30)
    m.add_argument("--out", default=None, help="results file (default lib/tune/results.json)")
    a = sub.add_parser("apply")
    a.add_argument("--results", default=None, help="results file (default lib/tune/results.json)")
    args = ap.parse_args()
    {"build": cmd_build, "measure": cmd_measure, "apply": cmd_apply}[args.cmd](args)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Trajectory comparison of the HIP path against the oracle over many steps.

  python tests/trajectory_report.py run --side gpu|f32|f64 --case dam64k|cube1m|dam8m \
      --steps N --out DIR [--sample K] [--threads T]
  python tests/trajectory_report.py compare DIR_A DIR_B

`run` advances one scene and records after every step: the mean density error (the average of
rho/rho0 - 1 over the fluid, the reference's statistic at cpu_backend.cpp:1509-1514), the kinetic
energy 1/2 m_F sum |v|^2, the solver iterations, and for scenes with a body its centre of mass and
orientation; at the end it keeps every K-th fluid particle by id (density, pressure, velocity,
position). Sides: `gpu` is the HIP path (libtsp_hip.so); `f32` and `f64` are the oracle's FP32
restatement of the reference CPU path and its FP64 build. `compare` prints the differences of
two runs: per step, relative to the quantity's scale over the run, and on the final sample
normwise (max |a - b| / max |b|).

Cases (the benchmark configurations): dam64k -- 40^3 fluid at h 0.01 in the 1.0 x 0.5 x 0.5 tank;
cube1m -- the 1M rigid cube drop of bench.py; dam8m -- 200^3 fluid at h 0.002 in the same tank.
All solve to 0.1% average compression error (the Pos criterion), 2..100 iterations.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from turbospharticles_amd import scenes  # noqa: E402


def make_case(case):
    if case == "dam64k":
        h = 0.01
        spec = scenes.dam_break_spec(40, h)
    elif case == "cube1m":
        h = 0.004
        spec = scenes.rigid_cube_spec(100, h, side=0.1, density=500.0, center=(0.7, 0.3, 0.25), rotation_seed=7)
    elif case == "dam8m":
        h = 0.002
        spec = scenes.dam_break_spec(200, h)
    else:
        raise ValueError(case)
    return scenes.build_scene(spec), scenes.make_params(h, termination_cond=scenes.TERM_POS)


def stats_of(st, params):
    rho0 = np.float64(params.rest_density)
    rho = st.fluid["density"].astype(np.float64)
    v2 = sum(st.fluid[c].astype(np.float64) ** 2 for c in ("vx", "vy", "vz"))
    out = {"mean_density_error": float(np.mean(rho / rho0 - 1.0)) if rho.size else 0.0,
           "kinetic_energy": float(0.5 * np.float64(params.mass) * np.sum(v2))}
    if st.n[3]:
        b = st.body
        out["xcm"] = [float(b[c][0]) for c in ("xcm", "ycm", "zcm")]
        out["A"] = [float(b[f"a{r}{c}"][0]) for r in range(3) for c in range(3)]
    return out


def run(args):
    state, params = make_case(args.case)
    os.makedirs(args.out, exist_ok=True)
    if args.side == "gpu":
        from turbospharticles_amd.backend import HipBackend
        eng = HipBackend(0)
        eng.set_params(params)
        eng.pull(state)

        def step():
            eng.step(1)
            return eng.info().solver_iters
    else:
        import oracle_lib as O
        eng = O.Oracle(params, precision=32 if args.side == "f32" else 64, chunk_width=1,
                       threads=args.threads or os.cpu_count())
        eng.pull(state)

        def step():
            eng.step()
            return eng.info().solver_iters
    rec = {k: [] for k in ("mean_density_error", "kinetic_energy", "iters", "xcm", "A")}
    t0 = time.perf_counter()
    last = t0
    for s in range(args.steps):
        rec["iters"].append(int(step()))
        st = eng.push()
        for k, v in stats_of(st, params).items():
            rec[k].append(v)
        now = time.perf_counter()
        if now - last > 20 or s + 1 == args.steps:
            print(f"[{args.side} {args.case}] step {s + 1}/{args.steps} iters {rec['iters'][-1]} "
                  f"mde {rec['mean_density_error'][-1]:.3e} ke {rec['kinetic_energy'][-1]:.6e} "
                  f"({now - t0:.0f} s)", flush=True)
            last = now
    ids = st.fluid_id
    keep = np.nonzero(ids % np.uint64(args.sample) == 0)[0]
    order = keep[np.argsort(ids[keep])]
    np.savez_compressed(os.path.join(args.out, "final_sample.npz"), id=ids[order],
                        **{f: st.fluid[f][order] for f in ("x", "y", "z", "vx", "vy", "vz", "density", "pressure")})
    with open(os.path.join(args.out, "stats.json"), "w") as f:
        json.dump({"case": args.case, "side": args.side, "steps": args.steps, "sample": args.sample,
                   "wall_s": time.perf_counter() - t0, **{k: v for k, v in rec.items() if v}}, f)


def rel_norm(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def compare(a_dir, b_dir):
    A = json.load(open(os.path.join(a_dir, "stats.json")))
    B = json.load(open(os.path.join(b_dir, "stats.json")))
    n = min(len(A["kinetic_energy"]), len(B["kinetic_energy"]))
    out = {"a": f"{A['side']} {A['case']}", "b": f"{B['side']} {B['case']}", "steps": n}
    mde_a = np.array(A["mean_density_error"][:n])
    mde_b = np.array(B["mean_density_error"][:n])
    ke_a = np.array(A["kinetic_energy"][:n])
    ke_b = np.array(B["kinetic_energy"][:n])
    out["mean_density_error_run_avg"] = [float(mde_a.mean()), float(mde_b.mean())]
    out["mean_density_error_run_avg_rel_diff"] = float(abs(mde_a.mean() - mde_b.mean()) / abs(mde_b.mean()))
    out["mean_density_error_max_step_diff_rel_to_run_scale"] = float(np.max(np.abs(mde_a - mde_b)) / np.max(np.abs(mde_b)))
    out["kinetic_energy_max_step_rel_diff"] = float(np.max(np.abs(ke_a - ke_b) / np.maximum(np.abs(ke_b), 1e-30)))
    out["kinetic_energy_final"] = [float(ke_a[-1]), float(ke_b[-1])]
    out["iters_equal_steps"] = int(np.sum(np.array(A["iters"][:n]) == np.array(B["iters"][:n])))
    if "xcm" in A and "xcm" in B:
        xa = np.array(A["xcm"][:n])
        xb = np.array(B["xcm"][:n])
        disp = np.max(np.linalg.norm(xb - xb[0], axis=1))
        out["rigid_xcm_max_diff_rel_to_displacement"] = float(np.max(np.linalg.norm(xa - xb, axis=1)) / max(disp, 1e-30))
        out["rigid_displacement"] = float(disp)
        Aa = np.array(A["A"][:n]).reshape(n, 3, 3)
        Ab = np.array(B["A"][:n]).reshape(n, 3, 3)
        out["rigid_orientation_max_frobenius_diff"] = float(np.max(np.linalg.norm(Aa - Ab, axis=(1, 2))))
    sa = np.load(os.path.join(a_dir, "final_sample.npz"))
    sb = np.load(os.path.join(b_dir, "final_sample.npz"))
    if np.array_equal(sa["id"], sb["id"]):
        out["final_sample_particles"] = int(sa["id"].size)
        for f in ("density", "pressure"):
            out[f"final_{f}_rel_err"] = rel_norm(sa[f], sb[f])
        va = np.stack([sa[c] for c in ("vx", "vy", "vz")], 1)
        vb = np.stack([sb[c] for c in ("vx", "vy", "vz")], 1)
        out["final_velocity_rel_err"] = rel_norm(va, vb)
        xa = np.stack([sa[c] for c in ("x", "y", "z")], 1)
        xb = np.stack([sb[c] for c in ("x", "y", "z")], 1)
        out["final_position_max_abs_diff"] = float(np.max(np.abs(xa - xb)))
    print(json.dumps(out, indent=1))
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--side", choices=("gpu", "f32", "f64"), required=True)
    r.add_argument("--case", choices=("dam64k", "cube1m", "dam8m"), required=True)
    r.add_argument("--steps", type=int, required=True)
    r.add_argument("--out", required=True)
    r.add_argument("--sample", type=int, default=1)
    r.add_argument("--threads", type=int, default=0)
    c = sub.add_parser("compare")
    c.add_argument("a")
    c.add_argument("b")
    args = ap.parse_args()
    if args.cmd == "run":
        run(args)
    else:
        compare(args.a, args.b)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""IISPH particle-steps/s of the gfx950 step on one MI355X.

Workload: the rigid-body coupling dam break -- 1,000,000 fluid particles (100^3 lattice at
h = 0.004, seeded +-1% jitter) in a 1.0 x 0.5 x 0.5 tank with 320,564 wall particles, and a
hollow 0.1 m cube of density 500 (6,364 rigid particles) released at rest at height 0.3 with a
seeded random rotation. dt 1e-3, omega 0.5, solve to 0.1% average compression error with 2..100
solver iterations. A step is one full IISPH step, neighbor search included; the state stays on
the device, inputs are resident before the timed region starts.

Prints one JSON line: value = N_F * steps / wall time of the timed steps; roofline of the
dominant kernel from HIP events recorded on the context's stream during the same timed steps run
a second time from the same initial state (the events add gaps of their own, so the value is
taken without them); the CPU baseline is the oracle (CPU restatement of the reference TBB path)
on a bounded sample. The batched and served sections compare mixed-size scenes in one context
against one context per scene (see batched_bench, served_bench).
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "IISPH particle-steps/s at 1M; batched-vs-plain throughput; p99 scene latency"
HBM_PEAK_GBS = 8000.0
H = 0.004
N_SIDE = 100


def algorithmic_bytes(info, n_f):
    """Compulsory bytes per fluid particle of each kernel (each SoA array counted once per
    pass; neighbor gathers hit in cache). L: the list bytes per particle at 2 B per entry, the
    three u16 counts and the slice metadata."""
    entries = (info.n_ff + info.n_fb + info.n_fr) / max(n_f, 1)
    L = 2.0 * entries + 6.0 + 48.0 / 32.0
    return {
        "density": 12 + L + 4,
        "nonpressure": 12 + 12 + 4 + L + 12,
        "source_diag": 12 + 12 + 4 + L + 12,
        "pressure_acc": 12 + 4 + L + 12,
        "pressure_acc_final": 12 + 4 + L + 12,
        "jacobi": 12 + 12 + 4 + 4 + 8 + L,
        "integration": 60.0,
        "_L": L,
    }


VALU_CLOCK_HZ = 2.4e9      # MI355X peak engine clock
N_SIMDS = 1024             # 256 CUs x 4 SIMDs; a wave64 VALU instruction takes 4 cycles of one SIMD


def measured_counters(kernel, n_side):
    """The committed rocprofv3 counter summary of `kernel` on this workload
    (profiles/<kernel>_1M_cube_traffic.json): HBM bytes per launch (FETCH_SIZE / WRITE_SIZE scaled
    by the copy calibration) and VALU instructions per wave (SQ_INSTS_VALU / SQ_WAVES), or {}."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_1M_cube_traffic.json")
    if n_side != N_SIDE or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """(model name, fp32 SIMD lanes, oracle build): 16 lanes and the AVX-512 build where the host
    has AVX-512, else 8 lanes and the AVX2 build."""
    model, flags = "unknown", ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                if line.startswith("flags") and not flags:
                    flags = line
    except OSError:
        pass
    return (model, 16, "v4") if " avx512f" in flags else (model, 8, "v3")


def cpu_baseline(state, params, warmup, steps, threads, lanes, variant):
    """The oracle (the TBB path restated, -O3 build for the host's ISA, chunk width = the host's
    fp32 SIMD lanes, as the reference's SIMD chunks) on the host cores: the same warm-up steps
    untimed, then the first `steps` of the GPU's timed window."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    o = O.Oracle(params, chunk_width=lanes, threads=threads, variant=variant)
    o.pull(state)
    for _ in range(warmup):
        o.step()
    t0 = time.perf_counter()
    sweeps = 0
    for _ in range(steps):
        o.step()
        sweeps += max(o.info().solver_iters - 1, 0)
    dt = time.perf_counter() - t0
    return state.n[0] * steps / dt, dt, sweeps / steps


def state_digest(st):
    """SHA-256 over a state's fluid ids and fields (both forms compare their final states)."""
    h = hashlib.sha256(st.fluid_id.tobytes())
    for f in sorted(st.fluid):
        h.update(st.fluid[f].tobytes())
    return h.hexdigest()


def batched_bench(args):
    """Mixed-size scenes in one context against the same scenes one context each: 256 dam breaks,
    fluid counts log-uniform in [2,000, 200,000] (seeded), each tank scaled to its block, h = 0.01.
    Both forms take the same warm-up steps, then the same timed steps."""
    import numpy as np
    from turbospharticles_amd import scenes
    from turbospharticles_amd.backend import HipBackend, HipBatch

    sides = scenes.log_uniform_sides(args.batch_scenes, 2000, 200000, seed=2026)
    params = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)
    states = [scenes.build_scene(scenes.scaled_dam_break_spec(n, 0.01, jitter_seed=67 + q)) for q, n in enumerate(sides)]
    n_total = sum(st.n[0] for st in states)
    progress(f"batched: {len(states)} scenes, {n_total} fluid")
    b = HipBatch(0)
    ids = b.update(add=[(params, st) for st in states])
    b.step(args.batch_warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(args.batch_steps)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    sweeps = sum(b.info(i).sweeps_total for i in ids)
    digests = [state_digest(b.push(i)) for i in ids]
    b.close()
    progress(f"batched: {t_batch:.2f} s for {args.batch_steps} steps; plain form next")
    t_plain = 0.0
    same = 0
    for q, st in enumerate(states):
        if q % 64 == 0:
            progress(f"batched: plain form, scene {q}/{len(states)}")
        g = HipBackend(0)
        g.set_params(params)
        g.pull(st)
        g.step(args.batch_warmup)
        g.synchronize()
        t0 = time.perf_counter()
        g.step(args.batch_steps)
        g.synchronize()
        t_plain += time.perf_counter() - t0
        same += int(state_digest(g.push()) == digests[q])
        g.close()
    v_b = n_total * args.batch_steps / t_batch
    v_p = n_total * args.batch_steps / t_plain
    return {"workload": f"{len(states)} dam-break scenes, fluid counts log-uniform in [2000, 200000] "
                        f"(lattice sides {min(sides)}..{max(sides)}, {n_total} fluid in total), tanks scaled to "
                        f"their blocks, h 0.01, steps {args.batch_warmup + 1}..{args.batch_warmup + args.batch_steps}",
            "value": round(v_b, 1), "unit": "particle-steps/s", "plain_value": round(v_p, 1),
            "speedup_vs_plain": round(v_b / v_p, 2), "ms_per_step": round(t_batch * 1e3 / args.batch_steps, 3),
            "plain_ms_per_step_all_scenes": round(t_plain * 1e3 / args.batch_steps, 3),
            "mean_jacobi_sweeps_per_scene_step": round(sweeps / len(ids) / (args.batch_warmup + args.batch_steps), 2),
            "results_identical_to_plain": f"{same}/{len(states)}"}


def served_bench(args):
    """A Poisson stream of dam-break / dropped-cube scenes (50/50), sizes log-uniform in
    [5,000, 1,000,000] fluid, steps uniform in [200, 2,000], admitted against a 200 GB budget;
    served with continuous batching and, as the plain form, one scene at a time. Each request
    carries its scene spec, generated on the device at admission."""
    import numpy as np
    from turbospharticles_amd import scenes, serving

    rng = np.random.default_rng(args.serve_seed)
    n = args.serve_requests
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, n))
    arrivals -= arrivals[0]
    sides = scenes.log_uniform_sides(n, 5000, 1000000, seed=args.serve_seed + 1)
    steps = rng.integers(200, 2001, n)
    kinds = rng.integers(0, 2, n)
    params = scenes.make_params(0.01, termination_cond=scenes.TERM_POS)

    def make():
        reqs = []
        for q in range(n):
            spec = (scenes.scaled_cube_drop_spec(sides[q], 0.01, rotation_seed=q) if kinds[q]
                    else scenes.scaled_dam_break_spec(sides[q], 0.01, jitter_seed=67 + q))
            reqs.append(serving.Request(q, arrivals[q], params, None, steps[q], spec=spec))
        return reqs

    budget = 200e9
    rb = make()
    progress(f"served: {n} requests, continuous batching")
    wall_b = serving.serve_batched(rb, budget, chunk_steps=args.serve_chunk)
    rp = make()
    progress(f"served: batched {wall_b:.1f} s; plain form next")
    wall_p = serving.serve_plain(rp)
    work = sum(r.particle_steps for r in rb)
    same = 0
    for a, c in zip(rb, rp):
        same += int(all(np.array_equal(a.result.fluid[f], c.result.fluid[f]) for f in a.result.fluid)
                    and np.array_equal(a.result.fluid_id, c.result.fluid_id))
    lb, lp = serving.latency_summary(rb), serving.latency_summary(rp)
    # per size bucket: its share of the work, the GPU time the plain form spent on its scenes
    # (admission to completion, one scene at a time) and when the served form finished its last
    # scene; a bucket whose scenes fill the GPU alone runs no faster batched than plain
    edges = sorted(serving.DEFAULT_BUCKETS)
    by_bucket = []
    for b in range(len(edges) + 1):
        sel = [q for q in range(n) if sum(rb[q].n[0] >= e for e in edges) == b]
        if not sel:
            continue
        lo = 0 if b == 0 else edges[b - 1]
        by_bucket.append({
            "fluid": f">={lo}" + (f",<{edges[b]}" if b < len(edges) else ""),
            "requests": len(sel),
            "work_share": round(sum(rb[q].particle_steps for q in sel) / work, 3),
            "plain_service_s": round(sum(rp[q].done_s - rp[q].admitted_s for q in sel), 2),
            "served_first_admit_s": round(min(rb[q].admitted_s for q in sel), 2),
            "served_last_done_s": round(max(rb[q].done_s for q in sel), 2)})
    return {"requests": n,
            "workload": f"{n} requests, Poisson arrivals at {args.serve_rate}/s (seed {args.serve_seed}), "
                        f"dam break / cube drop 50/50, fluid {min(s ** 3 for s in sides)}..{max(s ** 3 for s in sides)}, "
                        f"steps {int(steps.min())}..{int(steps.max())}, 200 GB admission budget, joins and leaves "
                        f"every <= {args.serve_chunk} steps, scenes generated on the device at admission",
            "value": round(work / wall_b, 1), "unit": "particle-steps/s",
            "p50_latency_s": lb["p50_s"], "p99_latency_s": lb["p99_s"], "latency_samples": n,
            "plain_value": round(work / wall_p, 1), "plain_p50_latency_s": lp["p50_s"],
            "plain_p99_latency_s": lp["p99_s"], "speedup_vs_plain": round(wall_p / wall_b, 2),
            "last_arrival_s": round(float(arrivals[-1]), 2), "served_wall_s": round(wall_b, 2),
            "plain_wall_s": round(wall_p, 2), "by_bucket": by_bucket,
            "results_identical_to_plain": f"{same}/{n}"}


def copy_probe():
    """Three plain device-to-device copies of 1 GiB: the counters' reading for a kernel that
    reads and writes exactly 1 GiB per launch."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": src.numel() * 4, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="oracle steps timed for the CPU baseline (after the same warm-up as the GPU)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the CPU baseline (a GPU box's share of its host is 16 cores)")
    ap.add_argument("--full-steps", type=int, default=500,
                    help="steps of the whole-run figure (the workload's full 500-step drop from its initial state)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n-side", type=int, default=N_SIDE)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain mixed scenes")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=295,
                    help="timed steps of the batched section (with the warm-up, 300 steps per scene)")
    ap.add_argument("--batch-warmup", type=int, default=5)
    ap.add_argument("--no-serve", action="store_true", help="skip the served request stream")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-chunk", type=int, default=10)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 1 GiB device buffer 3 times (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs one scene on one GPU; --gpus {args.gpus} is not supported "
              "(independent scenes would be spread over GPUs by HBM footprint).", file=sys.stderr)
        sys.exit(2)

    from turbospharticles_amd import scenes
    from turbospharticles_amd.backend import HipBackend

    h = H * N_SIDE / args.n_side
    spec = scenes.rigid_cube_spec(args.n_side, h, side=0.1, density=500.0, center=(0.7, 0.3, 0.25),
                                  rotation_seed=7)
    state = scenes.build_scene(spec)
    params = scenes.make_params(h, termination_cond=scenes.TERM_POS)
    n_f = state.n[0]

    def window(profile):
        # the same warm-up and the same timed steps from the same initial state; with profile,
        # HIP events on the context's stream bracket every launch group of the timed steps
        gpu = HipBackend(0)
        gpu.set_params(params)
        gpu.pull(state)
        gpu.step(args.warmup)
        gpu.synchronize()
        sweeps0 = gpu.info().sweeps_total
        gpu.profile_enable(profile)
        gpu.synchronize()
        t0 = time.perf_counter()
        gpu.step(args.steps)
        gpu.synchronize()
        t1 = time.perf_counter()
        prof = gpu.profile_read() if profile else {}
        gpu.profile_enable(False)
        info = gpu.info()
        gpu.close()
        return t1 - t0, info, prof, (info.sweeps_total - sweeps0) / args.steps

    def full_run(steps):
        # the whole drop from the initial state, step by step (the cube lands and rests on the
        # floor late in it, and the contact solve runs): wall time, sweeps and contact steps
        gpu = HipBackend(0)
        gpu.set_params(params)
        gpu.pull(state)
        gpu.synchronize()
        contact = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            gpu.step(1)
            contact += int(gpu.info().flags & 4 != 0)
        gpu.synchronize()
        wall_full = time.perf_counter() - t0
        info = gpu.info()
        gpu.close()
        return {"steps": steps, "ms_per_step": round(wall_full * 1e3 / steps, 3),
                "value": round(n_f * steps / wall_full, 1), "mean_jacobi_sweeps": round(info.sweeps_total / steps, 2),
                "contact_steps": contact, "note": "steps 1..%d from the initial state, one tsp_step call per step" % steps}

    # the value comes from the plain timed steps; the per-kernel times from the same steps run
    # again with the events (which add their own gaps between launches)
    wall, info, _, k_bar = window(False)
    wall_prof, _, prof, _ = window(True)
    value = n_f * args.steps / wall
    full = full_run(args.full_steps) if args.full_steps > 0 else None

    ab = algorithmic_bytes(info, n_f)
    candidates = {k: v for k, v in prof.items() if k in ab}
    dominant = max(candidates, key=lambda k: candidates[k][1]) if candidates else None
    roof = None
    if dominant:
        launches, ms = candidates[dominant]
        avg_s = ms / 1e3 / launches
        bytes_launch = ab[dominant] * n_f
        achieved = bytes_launch / avg_s / 1e9
        counters = measured_counters(dominant, args.n_side)
        roof = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": int(counters["hbm_bytes_per_launch"]) if counters else None,
                "bytes_per_launch": int(bytes_launch), "avg_launch_us": round(avg_s * 1e6, 2)}
        if counters.get("valu_insts_per_wave"):
            # instruction issue: the kernel's VALU instructions (per wave, from the counters) over
            # what the 1024 SIMDs issue in the launch time measured here (4 cycles per wave64
            # instruction at the peak clock); the sweeps' pair arithmetic, not HBM, bounds them
            waves = (n_f + 63) // 64
            valu = counters["valu_insts_per_wave"] * waves * 4 / (N_SIMDS * VALU_CLOCK_HZ * avg_s)
            roof["valu_issue_frac"] = round(valu, 4)
            roof["valu_insts_per_wave"] = counters["valu_insts_per_wave"]
    b_step = 846.0 + 214.0 * k_bar
    step_gbs = n_f * b_step * args.steps / wall / 1e9

    cpu = None
    if not args.no_cpu:
        progress("CPU baseline (oracle)")
        model, lanes, variant = host_cpu()
        v, dt, kc = cpu_baseline(state, params, args.warmup, args.cpu_steps, args.cpu_threads, lanes, variant)
        cpu = {"value": round(v, 1), "unit": "particle-steps/s", "cores": args.cpu_threads, "kind": "port",
               "cpu_model": model, "simd_fp32_lanes": lanes,
               "sample": f"oracle (C restatement of the TBB path, OpenMP, gcc -O3 -march=x86-64-{variant}, "
                         f"chunk width {lanes}) on steps {args.warmup + 1}.."
                         f"{args.warmup + args.cpu_steps} of the same workload (the first {args.cpu_steps} of the GPU's "
                         f"timed window, after the same {args.warmup} warm-up steps): {dt:.1f} s, "
                         f"mean {kc:.1f} Jacobi sweeps/step"}

    progress(f"1M step: {value / 1e6:.1f} M particle-steps/s")
    batched = None if args.no_batch else batched_bench(args)
    served = None if args.no_serve else served_bench(args)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"rigid cube drop, {n_f} fluid ({args.n_side}^3, h={h:g}) + {state.n[1]} boundary "
                               f"+ {state.n[2]} rigid, dt 1e-3, 0.1% avg compression error, 2..100 iterations",
                   "mean_jacobi_sweeps": round(k_bar, 2),
                   "full_run": full,
                   "profiled_ms_per_step": round(wall_prof * 1e3 / args.steps, 3),
                   "step_algorithmic_GBps": round(step_gbs, 1),
                   "kernel_ms": {k: round(v[1], 3) for k, v in prof.items()}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

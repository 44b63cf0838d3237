"""Benchmark of the PCISPH timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--no-cpu-baseline] [--no-learned] [--timing-every T]

A step is one PCI_SPHSystem::update of the whole scene (grid build, neighbour
search, non-pressure forces, the tolerance-driven corrector, integration,
box collision, velocity interpolation) with all state resident in HBM.  The
default workload is the 262,144-particle PCISPH dam break (64^3 block,
s = 5 mm, 1.28 x 0.64 x 0.64 m box, 256 x 128 x 128 cells, 1 % tolerance,
3..50 iterations).  W untimed steps, then K steps timed between two stream
synchronisations; per-kernel durations come from HIP events recorded on the
library's stream around every launch of every T-th timed step.

Then the batched serving workload: 64 independent scenes (log-uniform
4,096..131,072 particles, dam breaks and drops) advanced by one launch per
phase, block sizes autotuned, steps launched on the stream, timed with and
without the learned start, with per-scene step latency p50/p99 (the plain
workload also replayed from HIP graphs, reported as value_graphs).

The learned warm start is trained the way it is meant to be used
(untimed): plain PCISPH records converged pressures on 16 seeded scenes of
the same shape for 200 steps, the MLP is trained on the GPU, and 4 held-out
seeds compare plain, previous-pressure and learned starts.  The same K steps
of the workload are then timed again with the learned start.

Prints one JSON line with the metric, the roofline of the dominant kernel,
the CPU baseline (the oracle on this host, a bounded sample) and the learned
warm start's iterations and throughput.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per particle per launch of each kernel:
# FP32 streams, each array read and written once, neighbour gathers assumed
# cache resident (a neighbour's data is reused by its ~40 neighbours).
ALG_BYTES = {
    "pci_density": 24,        # x* 12 + rho* 4 + p read/write 8
    "pci_force": 68,          # x*,p 16 + x 12 + v_half 12 + a_np 12 + write x* 12 + p 4
    "pci_nonpressure": 40,    # x 12 + v 12 + rho 4 + write a_np 12
    "neighbors_density": 16,  # x 12 + write rho 4
    "density_epilogue": 12,
    "wcsph_forces": 116,      # forces 44 + integrate/box/interpolate 72
    "grid_build": 72,         # hash 20 + counting sort / ranges 4 + reorder 48 (+ 8 n_cells/N)
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="pcisph_dam_break",
                    choices=["pcisph_dam_break", "wcsph_dam_break"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP events")
    ap.add_argument("--timing-every", type=int, default=16,
                    help="per-kernel events on every T-th timed step")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm start")
    ap.add_argument("--train-seeds", type=int, default=16)
    ap.add_argument("--eval-seeds", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=200)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scene leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-warmup", type=int, default=10)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the surface-extraction and checkpoint timings")
    ap.add_argument("--no-dem", action="store_true", help="skip the DEM double dam break")
    ap.add_argument("--dem-steps", type=int, default=1000)
    ap.add_argument("--dem-warmup", type=int, default=5)
    return ap.parse_args()


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc, if a summary exists."""
    f = ROOT / "profiles" / f"traffic_{kernel}_{workload}.txt"
    if not f.exists():
        return None
    m = re.search(r"^hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read_text(), flags=re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(sc, start_state, first_step: int, steps: int):
    """The FP32 oracle on this host's cores, started from the GPU state at the
    middle of the timed region (the trajectories are bit-identical, so the
    CPU sees the corrector iteration counts of those GPU steps)."""
    from oracle import pyoracle
    pyoracle.build()
    cores = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", str(min(16, cores))))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = pyoracle.OracleSystem(sc.params, sc.n)
    o.set_particles(start_state["pos"], start_state["vel"], start_state["vh"])
    t0 = time.perf_counter()
    o.update(steps)
    dt = time.perf_counter() - t0
    st = o.stats()
    o.close()
    return {
        "value": sc.n * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sc.name} {sc.n} particles, steps {first_step + 1}..{first_step + steps} "
                  f"(from the GPU state at the middle of the timed region), FP32 oracle, "
                  f"{threads} OpenMP threads, "
                  f"{st['total_iterations'] / steps:.2f} iters/step",
    }


def timed_run(ctx, fs, steps: int, split: bool):
    """K steps between stream synchronisations; with `split` the state at the
    midpoint is copied out (untimed) for the CPU baseline."""
    half = steps // 2 if split else steps
    total, stats, mid = 0.0, [], None
    for k, nsteps in enumerate((half, steps - half)):
        ctx.sync()
        t0 = time.perf_counter()
        if nsteps:
            fs.update(nsteps)
        ctx.sync()
        total += time.perf_counter() - t0
        if nsteps:
            stats.append(fs.stats())
        if k == 0 and split:
            mid = fs.state()
    out = {"total_iterations": sum(x["total_iterations"] for x in stats),
           "worst_error": max((x["worst_error"] for x in stats), default=0.0),
           "flags": 0}
    for x in stats:
        out["flags"] |= x["flags"]
    return total, out, mid


def learned_leg(ctx, sc, args):
    """Train the warm-start MLP (untimed) and time the workload with it."""
    from psb_project_amd import learned, scenes
    from psb_project_amd._lib import WARM_MLP
    from psb_project_amd.system import FluidSystem
    n_side = sc.meta["n_side"]
    t0 = time.perf_counter()
    model, summary = learned.train_and_evaluate(
        ctx, lambda i: scenes.pcisph_dam_break(n_side=n_side, scene_id=100 + i),
        range(args.train_seeds), range(args.train_seeds, args.train_seeds + args.eval_seeds),
        steps=args.train_steps, epochs=args.train_epochs)
    train_s = time.perf_counter() - t0
    p = sc.params.copy()
    p.warm_start = WARM_MLP
    fs = FluidSystem(ctx, p, sc.n)
    fs.set_mlp(model)
    fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
    if args.warmup:
        fs.update(args.warmup)
    elapsed, st, _ = timed_run(ctx, fs, args.steps, split=False)
    fs.close()
    ev = summary["eval"]
    return model, {
        "value": round(sc.n * args.steps / elapsed, 1),
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "iters_per_step": round(st["total_iterations"] / args.steps, 3),
        "max_density_error": st["worst_error"],
        "flags": st["flags"],
        "heldout": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in ev.items()},
        "train": {"scenes": args.train_seeds, "steps": args.train_steps,
                  "samples": summary["samples"], "epochs": summary["epochs"],
                  "mse_per_epoch": [round(x, 5) for x in summary["mse_per_epoch"]],
                  "seconds": round(train_s, 1)},
    }


def batched_leg(ctx, args, model):
    """The 64-scene serving workload: autotuned batch steps, timed with the
    learned warm start (when a model is given) and without.  Steps are
    launched directly on the stream; the plain workload is also timed with
    each step replayed from HIP graphs (`value_graphs`), which measured
    slower on MI355X."""
    from psb_project_amd import scenes
    from psb_project_amd._lib import WARM_MLP, WARM_NONE
    from psb_project_amd.batch import BatchSystem
    out = {}
    modes = [("plain", WARM_NONE)] + ([("learned", WARM_MLP)] if model is not None else [])
    graph_value = None
    runs = [("plain", WARM_NONE, True)] + [(n, w, False) for n, w in modes]
    for name, warm, graphs in runs:
        scs = scenes.batched_scenes(num_scenes=args.batch_scenes, warm_start=warm)
        b = BatchSystem(ctx, scs)
        if warm == WARM_MLP:
            b.set_mlp(model)
        b.set_all()
        b.update(args.batch_warmup)
        tuned = b.autotune(steps=2)
        b.set_graphs(graphs)
        b.set_latency(True)
        ctx.sync()
        t0 = time.perf_counter()
        b.update(args.batch_steps)
        ctx.sync()
        dt = time.perf_counter() - t0
        if graphs:
            graph_value = round(b.n * args.batch_steps / dt, 1)
            b.close()
            continue
        lat = b.latency()
        its = [b.stats(i)["total_iterations"] / args.batch_steps for i in range(b.num_scenes)]
        worst = max(b.stats(i)["worst_error"] for i in range(b.num_scenes))
        flags = b.stats(0)["flags"]
        out[name] = {
            "value": round(b.n * args.batch_steps / dt, 1),
            "ms_per_step": round(1e3 * dt / args.batch_steps, 4),
            "iters_per_step_mean": round(float(sum(its) / len(its)), 3),
            "iters_per_step_max_scene": round(float(max(its)), 3),
            "latency_p50_ms": round(lat["p50_ms"], 4),
            "latency_p99_ms": round(lat["p99_ms"], 4),
            "latency_samples": lat["samples"],
            "max_density_error": worst,
            "flags": flags,
            "block_sizes": tuned,
        }
        if name == "plain":
            out[name]["value_graphs"] = graph_value
        b.close()
    out["scenes"] = args.batch_scenes
    out["particles"] = sum(sc.n for sc in scenes.batched_scenes(num_scenes=args.batch_scenes))
    out["steps"] = args.batch_steps
    out["graphs"] = False
    return out


def dem_leg(ctx, args, model):
    """Double dam break with 512 rigid DEM spheres (2 x 128 x 128 x 64 fluid
    particles, lattice walls), timed over --dem-steps steps with the zero,
    previous-pressure and (given a model) learned corrector starts."""
    from psb_project_amd import scenes
    from psb_project_amd._lib import WARM_MLP, WARM_NONE, WARM_PREVIOUS
    from psb_project_amd.system import FluidSystem, sphere_particle_count
    out = {}
    starts = [("plain", WARM_NONE), ("previous_pressure", WARM_PREVIOUS)]
    if model is not None:
        starts.append(("learned", WARM_MLP))
    for name, warm in starts:
        sc = scenes.double_dam_break_dem(warm_start=warm)
        m = sc.meta
        per = sphere_particle_count(sc.params, m["sphere_radius"])
        fs = FluidSystem(ctx, sc.params, sc.n + per * len(m["sphere_centers"]))
        if warm == WARM_MLP:
            fs.set_mlp(model)
        fs.set_spheres(m["sphere_centers"], None, m["sphere_radius"], m["sphere_density"],
                       m["spring"], m["damping"])
        fs.set_particles(sc.pos4)
        fs.update(args.dem_warmup)
        elapsed, st, _ = timed_run(ctx, fs, args.dem_steps, split=False)
        sp = fs.spheres()
        out[name] = {
            "value": round(sc.n * args.dem_steps / elapsed, 1),
            "ms_per_step": round(1e3 * elapsed / args.dem_steps, 4),
            "iters_per_step": round(st["total_iterations"] / args.dem_steps, 3),
            "max_density_error": st["worst_error"],
            "flags": st["flags"],
            "sphere_speed_max": round(float(np.linalg.norm(sp["velocities"], axis=1).max()), 4),
        }
        out["boundary_particles"] = fs.n - sc.n
        out["fluid_particles"] = sc.n
        fs.close()
    out["unit"] = "fluid particle-steps/s"
    out["spheres"] = int(len(sc.meta["sphere_centers"]))
    out["steps"] = args.dem_steps
    return out


def state_file_leg(ctx, fs, sc, reps: int = 5):
    """Surface extraction and checkpoint save / load on the timed run's final
    state (the rows downstream of the step)."""
    import tempfile
    from psb_project_amd import surface
    from psb_project_amd.system import FluidSystem
    out = {}
    s = surface.Surface(ctx)
    p = surface.default_params(cell_size=sc.meta["s"])
    nv, nt = s.run(fs, p)                                   # warm-up, buffers sized
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        nv, nt = s.run(fs, p)
    dt = (time.perf_counter() - t0) / reps
    nodes = int(np.prod(s.field()["dims"].astype(np.int64)))
    s.close()
    out["surface"] = {"ms_per_extraction": round(1e3 * dt, 3), "nodes": nodes,
                      "nodes_per_s": round(nodes / dt, 1), "vertices": nv, "triangles": nt,
                      "cell_size": sc.meta["s"], "scale": p.scale, "iso_level": p.iso_level,
                      "particles": sc.n}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "state.ckpt")
        t0 = time.perf_counter()
        fs.save_checkpoint(path)
        t_save = time.perf_counter() - t0
        size = os.path.getsize(path)
        other = FluidSystem(ctx, sc.params, sc.n)
        t0 = time.perf_counter()
        other.load_checkpoint(path)
        t_load = time.perf_counter() - t0
        other.close()
    out["checkpoint"] = {"save_ms": round(1e3 * t_save, 2), "load_ms": round(1e3 * t_load, 2),
                         "bytes": size, "particles": sc.n}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is not built)"}))
        sys.exit(2)
    from psb_project_amd import scenes
    from psb_project_amd.device import Context
    from psb_project_amd.system import FluidSystem

    sc = scenes.pcisph_dam_break() if args.workload == "pcisph_dam_break" else scenes.wcsph_dam_break()
    ctx = Context(0)
    sys_ = FluidSystem(ctx, sc.params, sc.n)
    sys_.set_particles(sc.pos4, sc.vel4, sc.vh4)
    print(f"bench: {sc.name} n={sc.n} warmup={args.warmup}", file=sys.stderr, flush=True)
    if args.warmup:
        sys_.update(args.warmup)
    ctx.sync()
    sys_.reset_timing()
    sys_.set_timing(not args.no_timing, every=args.timing_every)
    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    elapsed, stats, start_state = timed_run(ctx, sys_, args.steps, split=not args.no_cpu_baseline)
    timing = sys_.timing()
    sys_.set_timing(False)
    extras = None
    if not args.no_extras:
        print("bench: surface extraction and checkpoint of the final state", file=sys.stderr,
              flush=True)
        extras = state_file_leg(ctx, sys_, sc)
    sys_.close()

    value = sc.n * args.steps / elapsed
    iters = stats["total_iterations"] / max(1, args.steps)
    kernels = {k: {"avg_us": 1e3 * t / c, "launches": c, "total_ms": t}
               for k, (t, c) in timing.items() if c}
    roofline = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
        per = ALG_BYTES.get(dom)
        if per is not None:
            alg = per * sc.n
            ach = alg / (kernels[dom]["avg_us"] * 1e-6) / 1e9
            roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": read_traffic(dom, sc.name),
                        "alg_bytes_per_launch": alg,
                        "avg_launch_us": round(kernels[dom]["avg_us"], 3),
                        "sampled_launches": kernels[dom]["launches"]}
    learned = None
    model = None
    if not args.no_learned and sc.params.solver == 1:
        print("bench: learned warm start (train untimed, then time)", file=sys.stderr, flush=True)
        model, learned = learned_leg(ctx, sc, args)
    batched = None
    if not args.no_batch and sc.params.solver == 1:
        print("bench: batched scenes", file=sys.stderr, flush=True)
        batched = batched_leg(ctx, args, model)
        if model is not None and "learned" in batched:
            batched["learned_vs_plain_single"] = round(batched["learned"]["value"] / value, 3)
    dem = None
    if not args.no_dem and sc.params.solver == 1:
        print("bench: double dam break with DEM spheres", file=sys.stderr, flush=True)
        dem = dem_leg(ctx, args, model)
    if model is not None:
        model.close()
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sc, start_state, args.warmup + args.steps // 2, args.cpu_steps)
    line = {
        "metric": "particle-steps/s at <=1% density error; PCISPH iters/step, plain vs learned",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "particles": sc.n,
                   "grid": list(sc.params.grid.dims), "tolerance": sc.params.tolerance,
                   "min_iterations": sc.params.min_iterations,
                   "max_iterations": sc.params.max_iterations, "warm_start": "none",
                   "parallelism": "single-gpu"},
        "iters_per_step": round(iters, 3),
        "max_density_error": stats["worst_error"],
        "flags": stats["flags"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned": learned,
        "batched": batched,
        "dem": dem,
        "state_files": extras,
        "kernels": {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"]}
                    for k, v in kernels.items()},
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

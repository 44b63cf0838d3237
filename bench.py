"""Benchmark of the PCISPH timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--no-cpu-baseline] [--no-learned] [--timing-every T]

A step is one PCI_SPHSystem::update of the whole scene (grid build, neighbour
search, non-pressure forces, the tolerance-driven corrector, integration,
box collision, velocity interpolation) with all state resident in HBM.  The
default workload is the 262,144-particle PCISPH dam break (64^3 block,
s = 5 mm, 1.28 x 0.64 x 0.64 m box, 256 x 128 x 128 cells, 1 % tolerance,
3..50 iterations, failed steps redone as dt/2 substeps).  W untimed steps,
then K steps timed between two stream synchronisations (`value`); per-kernel
durations come from HIP events on the library's stream around every launch
of every T-th timed step, and the roofline's from the corrector passes
replayed back to back on the final state of the timed region.

`full_run` times the workload's whole configured horizon (2000 steps from
the initial state), plain and with the learned start, whatever K is, with
the tolerance accounting of each: iterations per step, steps redone as
substeps, the worst final error and the flags.

Then the learned warm start is trained the way it is meant to be used
(untimed): plain PCISPH records converged pressures on 16 seeded scenes of
the same shape, 200 steps each, the windows spread over the 2000-step
horizon (starting at steps 0, 120, ..., 1800), the MLP is trained on the
GPU, and 4 held-out seeds compare plain, previous-pressure and learned
starts over their first 200 steps.  The same K steps
of the workload are then timed again with the learned start.

Then the batched serving workload: 64 independent scenes (log-uniform
4,096..131,072 particles, dam breaks and drops) advanced by one launch per
phase, block sizes autotuned, steps launched on the stream, timed with and
without the learned start, with per-scene step latency p50/p99 (the plain
workload also replayed from HIP graphs, reported as value_graphs); and the
2.1 M-particle double dam break with 512 DEM spheres, whose corrector passes
(a working set beyond the 256 MB Infinity Cache) give a second roofline.

The CPU baseline is the oracle on this host's cores over the same step window
as `value`, from the GPU state at the start of that window (the trajectories
are bit-identical).

Prints one JSON line.
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
# batch steps whose per-scene latency is recorded (events after every force
# pass); the others pace their queue without events
LATENCY_EVERY = 4

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
    # per-particle materials (one thread per particle, material index 1 B)
    "mix_pci_density": 29,    # x* 12 + material 1 + p read/write 8 + rho* 4 + p/rho^2 4
    "mix_pci_force": 80,      # x*,p 16 + rho* 4 + a_np 12 + a_s 12 + x 12 + v_half 12 + write x* 12
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
    ap.add_argument("--timing-every", type=int, default=64,
                    help="per-kernel events on every T-th timed step")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm start")
    ap.add_argument("--train-seeds", type=int, default=16)
    ap.add_argument("--eval-seeds", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=200)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--train-horizon", type=int, default=None,
                    help="spread the training windows over this many steps "
                         "(default: the workload's horizon, 2000; 0: all from step 0)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scene leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-warmup", type=int, default=10)
    ap.add_argument("--no-sequential", action="store_true",
                    help="skip the batched scenes run one by one through the single-system step")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the 2000-step full-horizon block")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the surface-extraction and checkpoint timings")
    ap.add_argument("--no-dem", action="store_true", help="skip the DEM double dam break")
    ap.add_argument("--no-materials", action="store_true",
                    help="skip the per-particle-material (water + mucus) reference-mode leg")
    ap.add_argument("--dem-steps", type=int, default=1000)
    ap.add_argument("--dem-warmup", type=int, default=5)
    ap.add_argument("--dem-all-starts", action="store_true",
                    help="also time the DEM scene with the previous-pressure and learned starts")
    return ap.parse_args()


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc, if a summary exists."""
    f = ROOT / "profiles" / f"traffic_{kernel}_{workload}.txt"
    if not f.exists():
        return None
    m = re.search(r"^hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read_text(), flags=re.M)
    return float(m.group(1)) if m else None


def read_issue_limits(kernel: str, workload: str):
    """VALU issue and texture-data (TD) busy fractions of a kernel's dispatch
    from the committed rocprofv3 counter summary, if there is one: the
    limits that stand beside the HBM fraction for the neighbour passes."""
    f = ROOT / "profiles" / f"counters_corrector_{workload}.txt"
    if not f.exists():
        return None
    text = f.read_text()
    m = re.search(rf"^\S*k_{kernel}<false.*?(?=^\S|\Z)", text, flags=re.M | re.S)
    if not m:
        return None
    block = m.group(0)
    v = re.search(r"VALU issue [0-9.]+ us = (\d+)% of the dispatch", block)
    t = re.search(r"TD busy (\d+)% per CU", block)
    if not (v and t):
        return None
    return {"valu_issue_frac": int(v.group(1)) / 100, "td_busy_frac": int(t.group(1)) / 100,
            "source": f"profiles/{f.name}"}


def host_cpu():
    """CPU model and core counts of this host (lscpu)."""
    import subprocess
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        if kv.get("Core(s) per socket") and kv.get("Socket(s)"):
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        pass
    return info


def cpu_baseline(sc, start_state, first_step: int, steps: int):
    """The FP32 oracle on this host's cores (OpenMP, threads bound close),
    over the GPU's timed window from the GPU state at its start (the
    trajectories are bit-identical, so the CPU runs the same iterations)."""
    # binding must be set before the OpenMP runtime starts (first oracle load)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import pyoracle
    pyoracle.build()
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", str(cpu["physical_cores"] or cpu["logical_cpus"] or 1)))
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
        "cpu_model": cpu["model"],
        "host_physical_cores": cpu["physical_cores"],
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "iters_per_step": round(st["total_iterations"] / steps, 3),
        "sample": f"{sc.name} {sc.n} particles, steps {first_step + 1}..{first_step + steps} "
                  f"(the GPU's timed window, from the GPU state at its start), FP32 oracle, "
                  f"{threads} OpenMP threads bound close, "
                  f"{st['total_iterations'] / steps:.2f} iters/step",
    }


def timed_run(ctx, fs, steps: int, progress: str | None = None, chunk: int = 100):
    """K steps between stream synchronisations, with the tolerance accounting.
    With `progress`, the steps run in chunks of `chunk` (each update call
    ends with the host knowing the last iteration anyway) and a line per
    chunk goes to stderr; the chunks' statistics are combined."""
    out = {"total_iterations": 0, "worst_error": 0.0, "flags": 0, "retried_steps": 0,
           "substeps": 0}
    sizes = [steps] if not progress else [min(chunk, steps - k) for k in range(0, steps, chunk)]
    elapsed = 0.0
    done = 0
    for n in sizes:
        if not n:
            continue
        ctx.sync()
        t0 = time.perf_counter()
        fs.update(n)
        ctx.sync()
        elapsed += time.perf_counter() - t0
        done += n
        st = fs.stats()
        out["total_iterations"] += st["total_iterations"]
        out["worst_error"] = max(out["worst_error"], st["worst_error"])
        out["flags"] |= st["flags"]
        out["retried_steps"] += st["retried_steps"]
        out["substeps"] += st["substeps"]
        if progress:
            print(f"bench: {progress} {done}/{steps} steps, {elapsed:.1f} s, "
                  f"{out['total_iterations'] / done:.1f} iters/step, worst {out['worst_error']:.4f}",
                  file=sys.stderr, flush=True)
    return elapsed, out


def tolerance_report(sc, steps, st):
    return {"iters_per_step": round(st["total_iterations"] / max(1, steps), 3),
            "retried_steps": st["retried_steps"], "substeps": st["substeps"],
            "max_density_error": st["worst_error"], "flags": st["flags"],
            "within_tolerance": bool(st["worst_error"] <= sc.params.tolerance * (1 + 1e-6)
                                     and (st["flags"] & 0x9) == 0)}


def full_run(ctx, sc, model=None):
    """The workload's configured horizon (2000 steps from the initial state)
    as one timed block, plain or with the learned start."""
    from psb_project_amd._lib import WARM_MLP
    from psb_project_amd.system import FluidSystem
    steps = int(sc.meta.get("steps", 2000))
    p = sc.params.copy()
    if model is not None:
        p.warm_start = WARM_MLP
    fs = FluidSystem(ctx, p, sc.n)
    if model is not None:
        fs.set_mlp(model)
    fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
    elapsed, st = timed_run(ctx, fs, steps)
    fs.close()
    out = {"steps": steps, "value": round(sc.n * steps / elapsed, 1),
           "ms_per_step": round(1e3 * elapsed / steps, 4)}
    out.update(tolerance_report(sc, steps, st))
    return out


def trace_average(kernel: str, workload: str):
    """Full-dispatch average (us) of a kernel over the timed run in the
    committed rocprofv3 summary of this workload, if there is one."""
    f = ROOT / "profiles" / f"kernels_{workload}.txt"
    if not f.exists():
        return None
    text = f.read_text()
    i = text.find("full dispatches")
    if i < 0:
        return None
    m = re.search(rf"^\S*k_{kernel}\S*.*?\s+(\d+)\s+([0-9.]+)\s+(\d+)\s+([0-9.]+)\s*$",
                  text[i:], flags=re.M)
    return float(m.group(2)) if m else None


def roofline_of(kernel: str, n: int, us: float, workload: str, trace: bool = True):
    """`trace` adds the committed kernel trace's full-dispatch average: only
    where every launch there covers all n particles (not for the batch, whose
    launches skip the scenes that have converged)."""
    alg = ALG_BYTES[kernel] * n
    ach = alg / (us * 1e-6) / 1e9
    r = {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
         "traffic": read_traffic(kernel, workload), "alg_bytes_per_launch": alg,
         "avg_launch_us": round(us, 3)}
    lim = read_issue_limits(kernel, workload) if trace else None
    if lim:
        r["issue_limits"] = lim
    tr = trace_average(kernel, workload) if trace else None
    if tr:
        r["trace_avg_us"] = tr
        r["frac_trace"] = round(alg / (tr * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    return r


def learned_leg(ctx, sc, args):
    """Train the warm-start MLP (untimed) and time the workload with it."""
    from psb_project_amd import learned, scenes
    from psb_project_amd._lib import WARM_MLP
    from psb_project_amd.system import FluidSystem
    n_side = sc.meta["n_side"]
    horizon = (int(sc.meta.get("steps", 2000)) if args.train_horizon is None
               else args.train_horizon)
    t0 = time.perf_counter()
    model, summary = learned.train_and_evaluate(
        ctx, lambda i: scenes.pcisph_dam_break(n_side=n_side, scene_id=100 + i),
        range(args.train_seeds), range(args.train_seeds, args.train_seeds + args.eval_seeds),
        steps=args.train_steps, epochs=args.train_epochs, horizon=horizon)
    train_s = time.perf_counter() - t0
    p = sc.params.copy()
    p.warm_start = WARM_MLP
    fs = FluidSystem(ctx, p, sc.n)
    fs.set_mlp(model)
    fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
    if args.warmup:
        fs.update(args.warmup)
    elapsed, st = timed_run(ctx, fs, args.steps)
    fs.close()
    ev = summary["eval"]
    out = {
        "value": round(sc.n * args.steps / elapsed, 1),
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
    }
    out.update(tolerance_report(sc, args.steps, st))
    return model, dict(out, **{
        "heldout": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in ev.items()},
        "train": {"scenes": args.train_seeds, "steps": args.train_steps,
                  "window_starts": summary["window_starts"],
                  "samples": summary["samples"], "epochs": summary["epochs"],
                  "mse_per_epoch": [round(x, 5) for x in summary["mse_per_epoch"]],
                  "seconds": round(train_s, 1)},
    })


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
        b.set_latency(True, every=LATENCY_EVERY)
        ctx.sync()
        t0 = time.perf_counter()
        b.update(args.batch_steps)
        ctx.sync()
        dt = time.perf_counter() - t0
        if graphs:
            graph_value = round(b.n * args.batch_steps / dt, 1)
            b.close()
            continue
        if name == "plain":
            # the corrector passes over all 64 scenes at once: a launch that
            # fills the chip, on the final state of the timed run
            t = b.time_corrector(reps=20)
            out["roofline"] = roofline_of("pci_force", b.n, t["pci_force"], "batched_scenes",
                                          trace=False)
            out["roofline"]["source"] = "corrector passes replayed over every scene on the final state"
            out["roofline_density"] = roofline_of("pci_density", b.n, t["pci_density"],
                                                  "batched_scenes", trace=False)
        lat = b.latency()
        its = [b.stats(i)["total_iterations"] / args.batch_steps for i in range(b.num_scenes)]
        worst = max(b.stats(i)["worst_error"] for i in range(b.num_scenes))
        flags = b.stats(0)["flags"]
        over = sum(1 for i in range(b.num_scenes)
                   if b.stats(i)["worst_error"] > scs[i].params.tolerance * (1 + 1e-6))
        out[name] = {
            "value": round(b.n * args.batch_steps / dt, 1),
            "ms_per_step": round(1e3 * dt / args.batch_steps, 4),
            "iters_per_step_mean": round(float(sum(its) / len(its)), 3),
            "iters_per_step_max_scene": round(float(max(its)), 3),
            "latency_p50_ms": round(lat["p50_ms"], 4),
            "latency_p99_ms": round(lat["p99_ms"], 4),
            "latency_samples": lat["samples"],
            "latency_every": LATENCY_EVERY,
            "max_density_error": worst,
            "flags": flags,
            "retried_scene_steps": int(sum(b.stats(i)["retried_steps"] for i in range(b.num_scenes))),
            "substeps": int(sum(b.stats(i)["substeps"] for i in range(b.num_scenes))),
            "scenes_over_tolerance": over,
            "within_tolerance": over == 0 and (flags & 0x9) == 0,
            "block_sizes": tuned,
        }
        if name == "plain":
            out[name]["value_graphs"] = graph_value
        b.close()
    if "learned" in out:
        out["learned_vs_plain_batched"] = round(out["learned"]["value"] / out["plain"]["value"], 3)
    # the same scenes and steps through the plain single-system path, one
    # scene after another: the baseline the batched, warm-started serving
    # path replaces
    if not args.no_sequential:
        seq = plain_sequential(ctx, args)
        out["plain_sequential"] = seq
        for name in ("plain", "learned"):
            if name in out:
                out[f"{name}_batched_vs_plain_sequential"] = round(out[name]["value"] / seq["value"], 3)
    out["scenes"] = args.batch_scenes
    out["particles"] = sum(sc.n for sc in scenes.batched_scenes(num_scenes=args.batch_scenes))
    out["steps"] = args.batch_steps
    out["graphs"] = False
    return out


def plain_sequential(ctx, args):
    """The batched workload's scenes advanced one after another, each by the
    plain single-system step (FluidSystem, zero start), over the same warm-up
    and timed steps: total particle-steps over the summed timed time."""
    from psb_project_amd import scenes
    from psb_project_amd._lib import WARM_NONE
    from psb_project_amd.system import FluidSystem
    scs = scenes.batched_scenes(num_scenes=args.batch_scenes, warm_start=WARM_NONE)
    total_t = 0.0
    total_ps = 0
    its = []
    worst = 0.0
    flags = 0
    for sc in scs:
        fs = FluidSystem(ctx, sc.params, sc.n)
        fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
        fs.update(args.batch_warmup)
        elapsed, st = timed_run(ctx, fs, args.batch_steps)
        fs.close()
        total_t += elapsed
        total_ps += sc.n * args.batch_steps
        its.append(st["total_iterations"] / args.batch_steps)
        worst = max(worst, st["worst_error"])
        flags |= st["flags"]
    return {"value": round(total_ps / total_t, 1), "seconds": round(total_t, 3),
            "iters_per_step_mean": round(float(np.mean(its)), 3), "max_density_error": worst,
            "flags": flags, "within_tolerance": worst <= scs[0].params.tolerance * (1 + 1e-6)
            and (flags & 0x9) == 0}


def dem_leg(ctx, args, model):
    """Double dam break with 512 rigid DEM spheres (2 x 128 x 128 x 64 fluid
    particles, lattice walls), timed over --dem-steps steps with the zero,
    previous-pressure and (given a model) learned corrector starts."""
    from psb_project_amd import scenes
    from psb_project_amd._lib import WARM_MLP, WARM_NONE, WARM_PREVIOUS
    from psb_project_amd.system import FluidSystem, sphere_particle_count
    out = {}
    starts = [("plain", WARM_NONE)]
    if args.dem_all_starts:
        starts.append(("previous_pressure", WARM_PREVIOUS))
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
        elapsed, st = timed_run(ctx, fs, args.dem_steps, progress=f"dem {name}")
        sp = fs.spheres()
        out[name] = {
            "value": round(sc.n * args.dem_steps / elapsed, 1),
            "ms_per_step": round(1e3 * elapsed / args.dem_steps, 4),
            "sphere_speed_max": round(float(np.linalg.norm(sp["velocities"], axis=1).max()), 4),
        }
        out[name].update(tolerance_report(sc, args.dem_steps, st))
        if name == "plain":
            # corrector passes on the final state: a working set beyond the
            # 256 MB Infinity Cache (positions 34 MB, compact list 0.4 GB)
            t = fs.time_corrector(reps=20)
            out["roofline"] = roofline_of("pci_force", fs.n, t["pci_force"], "double_dam_break_dem")
            out["roofline_density"] = roofline_of("pci_density", fs.n, t["pci_density"],
                                                  "double_dam_break_dem")
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


def materials_leg(ctx, n_side: int = 32, steps: int = 100):
    """Reference-mode PCISPH steps with per-particle materials: a water block
    and a mucus block (the reference's Water.xml / Mucus.xml), n_side^3 each,
    each on its own rest spacing, dt 2 ms, 5 untimed steps then `steps` timed.

    PCISPH only: under the reference's WCSPH mixing rule (neighbour j's own
    support radius in the kernel, the list cut to particle i's) a mucus
    particle sums water neighbours beyond water's support with negative
    weight, and the mixture reaches zero density within tens of steps at any
    time step tried (the oracle flags it the same way), so a WCSPH mixture
    makes no throughput figure."""
    from psb_project_amd import scenes
    from psb_project_amd._lib import SOLVER_PCISPH
    from psb_project_amd.system import FluidSystem
    sc = scenes.two_fluid_scene(SOLVER_PCISPH, n_side=n_side, rest_spacing=True, dt=0.002)
    fs = FluidSystem(ctx, sc.params, sc.n)
    fs.set_particles(sc.pos4)
    fs.set_materials(sc.meta["materials"], sc.meta["index"])
    fs.update(5)
    fs.reset_timing()
    fs.set_timing(True, every=4)
    elapsed, st = timed_run(ctx, fs, steps)
    timing = fs.timing()
    fs.set_timing(False)
    kernels = {k: {"avg_us": round(1e3 * t / c, 3), "launches": c}
               for k, (t, c) in timing.items() if c}
    out = {"value": round(sc.n * steps / elapsed, 1),
           "ms_per_step": round(1e3 * elapsed / steps, 4),
           "particles": sc.n, "materials": fs.num_materials, "steps": steps,
           "dt": sc.params.dt, "flags": st["flags"],
           "iters_per_step": st["total_iterations"] / steps, "kernels": kernels}
    if "mix_pci_force" in kernels:
        # HIP events around the launches of every 4th timed step
        out["roofline"] = roofline_of("mix_pci_force", sc.n, kernels["mix_pci_force"]["avg_us"],
                                      sc.name, trace=False)
        if "mix_pci_density" in kernels:
            out["roofline"]["other_kernel"] = roofline_of(
                "mix_pci_density", sc.n, kernels["mix_pci_density"]["avg_us"], sc.name, trace=False)
    fs.close()
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
    pcisph = sc.params.solver == 1
    ctx = Context(0)
    sys_ = FluidSystem(ctx, sc.params, sc.n)
    sys_.set_particles(sc.pos4, sc.vel4, sc.vh4)
    print(f"bench: {sc.name} n={sc.n} warmup={args.warmup}", file=sys.stderr, flush=True)
    if args.warmup:
        sys_.update(args.warmup)
    ctx.sync()
    start_state = None if args.no_cpu_baseline else sys_.state()
    sys_.reset_timing()
    sys_.set_timing(not args.no_timing, every=args.timing_every)
    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    elapsed, stats = timed_run(ctx, sys_, args.steps)
    timing = sys_.timing()
    sys_.set_timing(False)
    kernels = {k: {"avg_us": 1e3 * t / c, "launches": c, "total_ms": t}
               for k, (t, c) in timing.items() if c}
    roofline = None
    if pcisph and args.steps:
        # the dominant kernel of the step by sampled time; its average launch
        # duration over the timed region from the HIP events around its
        # launches (every T-th step), the corrector passes replayed back to
        # back on the final state beside it
        t = sys_.time_corrector(reps=50)
        pair = ("pci_force", "pci_density")
        have = all(k in kernels for k in pair)
        dom = max(pair, key=(lambda k: kernels[k]["total_ms"]) if have else (lambda k: t[k]))
        other = "pci_density" if dom == "pci_force" else "pci_force"
        us = (lambda k: kernels[k]["avg_us"]) if have else (lambda k: t[k])
        roofline = roofline_of(dom, sc.n, us(dom), sc.name)
        roofline["source"] = ("HIP events around every launch of every "
                              f"{args.timing_every}th timed step" if have else
                              "corrector passes replayed on the final state")
        roofline["replay_final_state_us"] = round(t[dom], 3)
        roofline["other_kernel"] = roofline_of(other, sc.n, us(other), sc.name)
        roofline["other_kernel"]["replay_final_state_us"] = round(t[other], 3)
    elif kernels:
        dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
        if dom in ALG_BYTES:
            roofline = roofline_of(dom, sc.n, kernels[dom]["avg_us"], sc.name)
    extras = None
    if not args.no_extras:
        print("bench: surface extraction and checkpoint of the final state", file=sys.stderr,
              flush=True)
        extras = state_file_leg(ctx, sys_, sc)
    sys_.close()

    value = sc.n * args.steps / elapsed
    full = None
    if pcisph and not args.no_full_run:
        print("bench: full 2000-step horizon, plain", file=sys.stderr, flush=True)
        full = {"plain": full_run(ctx, sc)}
    learned = None
    model = None
    if not args.no_learned and pcisph:
        print("bench: learned warm start (train untimed, then time)", file=sys.stderr, flush=True)
        model, learned = learned_leg(ctx, sc, args)
        if full is not None:
            print("bench: full 2000-step horizon, learned", file=sys.stderr, flush=True)
            full["learned"] = full_run(ctx, sc, model)
            pl, le = full["plain"], full["learned"]
            full["learned_vs_plain"] = round(le["value"] / pl["value"], 3)
            full["iteration_reduction"] = round(1 - le["iters_per_step"] / pl["iters_per_step"], 4)
    batched = None
    if not args.no_batch and pcisph:
        print("bench: batched scenes", file=sys.stderr, flush=True)
        batched = batched_leg(ctx, args, model)
        if "learned" in batched and full is not None:
            # side figure: batched learned against the plain single scene's
            # full horizon (different scenes and horizons)
            batched["learned_batched_vs_plain_single_full"] = round(
                batched["learned"]["value"] / full["plain"]["value"], 3)
    dem = None
    if not args.no_dem and pcisph:
        print("bench: double dam break with DEM spheres", file=sys.stderr, flush=True)
        dem = dem_leg(ctx, args, model)
    if model is not None:
        model.close()
    materials = None
    if not args.no_materials:
        print("bench: per-particle materials (reference mode)", file=sys.stderr, flush=True)
        materials = materials_leg(ctx)
    cpu = None
    if not args.no_cpu_baseline and args.steps:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu_steps = min(args.steps, args.cpu_steps)
        cpu = cpu_baseline(sc, start_state, args.warmup, cpu_steps)
    line = {
        "metric": "particle-steps/s at <=1% density error; PCISPH iters/step, plain vs learned",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / max(1, args.steps), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "particles": sc.n,
                   "grid": list(sc.params.grid.dims), "tolerance": sc.params.tolerance,
                   "min_iterations": sc.params.min_iterations,
                   "max_iterations": sc.params.max_iterations,
                   "retry_levels": sc.params.retry_levels, "warm_start": "none",
                   "parallelism": "single-gpu"},
    }
    line.update(tolerance_report(sc, args.steps, stats))
    line.update({
        "roofline": roofline,
        "cpu_baseline": cpu,
        "full_run": full,
        "learned": learned,
        "batched": batched,
        "dem": dem,
        "materials": materials,
        "state_files": extras,
        "kernels": {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"]}
                    for k, v in kernels.items()},
    })
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

"""Benchmark of the PCISPH timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

A step is one PCI_SPHSystem::update of the whole scene (grid build, neighbour
search, non-pressure forces, the tolerance-driven corrector, integration,
box collision, velocity interpolation) with all state resident in HBM.  The
default workload is the 262,144-particle PCISPH dam break (64^3 block,
s = 5 mm, 1.28 x 0.64 x 0.64 m box, 256 x 128 x 128 cells, 1 % tolerance,
3..50 iterations).  W untimed steps, then K steps timed between two stream
synchronisations; per-kernel durations come from HIP events recorded on the
library's stream over the same timed steps.

Prints one JSON line with the metric, the roofline of the dominant kernel
and the CPU baseline (the oracle on this host, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per particle per launch of each kernel:
# FP32 SoA streams, each array read and written once, neighbour gathers
# assumed cache resident ("Algorithmic bytes").
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
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP events")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
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
    sys_.set_timing(not args.no_timing)
    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    # The timed region runs in two halves; between them (untimed) the state is
    # copied out as the starting point of the CPU baseline sample.
    half = args.steps // 2
    parts = []
    start_state = None
    for k, nsteps in enumerate((half, args.steps - half)):
        ctx.sync()
        t0 = time.perf_counter()
        if nsteps:
            sys_.update(nsteps)
        ctx.sync()
        parts.append((time.perf_counter() - t0, sys_.stats()))
        if k == 0 and not args.no_cpu_baseline:
            start_state = sys_.state()
    elapsed = parts[0][0] + parts[1][0]
    s0, s1 = parts[0][1], parts[1][1]
    stats = {"total_iterations": s0["total_iterations"] + s1["total_iterations"],
             "worst_error": max(s0["worst_error"], s1["worst_error"]),
             "flags": s0["flags"] | s1["flags"]}
    timing = sys_.timing()
    sys_.set_timing(False)

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
                        "avg_launch_us": round(kernels[dom]["avg_us"], 3)}
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sc, start_state, args.warmup + half, args.cpu_steps)
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
        "kernels": {k: {"avg_us": round(v["avg_us"], 3), "launches": v["launches"]}
                    for k, v in kernels.items()},
    }
    print(json.dumps(line), flush=True)
    sys_.close()
    ctx.close()


if __name__ == "__main__":
    main()

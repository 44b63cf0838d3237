"""Learned PCISPH warm start: a per-particle MLP that predicts the converged
pressure from features the PCISPH passes already compute.

Pipeline (all on the GPU through the C-ABI):
  1. plain PCISPH runs record (features, converged pressure) of every
     `stride`-th particle per step into a device Dataset (psb_system_set_recorder);
  2. `train()` standardises the features, then runs Adam on minibatches with
     the hand-written BF16-MFMA forward/backward kernels (psb_mlp_train_step);
  3. a system with warm_start = PSB_WARM_MLP evaluates the network after the
     first density pass and the unchanged corrector enforces the tolerance.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import WARM_MLP, WARM_NONE, WARM_PREVIOUS, check, lib
from .device import Context

NUM_FEATURES = 8
FEATURE_NAMES = ("density_error_pct", "div_v_dt", "neighbours_over_32", "p_prev_scaled",
                 "speed_dt_over_h", "grad_p_x", "grad_p_y", "grad_p_z")


def num_params() -> int:
    return int(lib().psb_mlp_num_params())


class PressureMLP:
    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().psb_mlp_create(ctx.handle, int(seed), C.byref(h)))
        self.handle = h
        self.mean = np.zeros(NUM_FEATURES, np.float32)
        self.std = np.ones(NUM_FEATURES, np.float32)
        self.pressure_scale = 1.0
        ctx.adopt(self)

    def params(self) -> np.ndarray:
        w = np.empty(num_params(), np.float32)
        check(lib().psb_mlp_get_params(self.handle, w.ctypes.data, w.size))
        return w

    def set_params(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        check(lib().psb_mlp_set_params(self.handle, w.ctypes.data, w.size))

    def set_normalization(self, mean, std, pressure_scale: float) -> None:
        self.mean = np.ascontiguousarray(mean, np.float32)
        self.std = np.ascontiguousarray(std, np.float32)
        self.pressure_scale = float(pressure_scale)
        check(lib().psb_mlp_set_normalization(self.handle, self.mean.ctypes.data,
                                              self.std.ctypes.data, self.pressure_scale))

    def forward(self, feat_dev_ptr, n: int, out_dev_ptr) -> None:
        check(lib().psb_mlp_forward(self.handle, feat_dev_ptr, n, out_dev_ptr))

    def grad(self) -> np.ndarray:
        g = np.empty(num_params(), np.float32)
        check(lib().psb_mlp_get_grad(self.handle, g.ctypes.data, g.size))
        return g

    def train_step(self, feat_ptr, target_ptr, num_samples: int, idx_ptr, batch: int,
                   lr: float = 1e-3, apply_adam: bool = True) -> float:
        loss = C.c_float()
        check(lib().psb_mlp_train_step(self.handle, feat_ptr, target_ptr, int(num_samples),
                                       idx_ptr, int(batch), float(lr), 1 if apply_adam else 0,
                                       C.byref(loss)))
        return float(loss.value)

    def zero_grad(self) -> None:
        check(lib().psb_mlp_zero_grad(self.handle))

    def save(self, path) -> None:
        """Weights (FP32 master copy), feature normalisation and pressure scale."""
        np.savez(path, w=self.params(), mean=self.mean, std=self.std,
                 pressure_scale=np.float64(self.pressure_scale))

    @classmethod
    def load(cls, ctx: Context, path) -> "PressureMLP":
        z = np.load(path, allow_pickle=False)
        m = cls(ctx)
        m.set_params(z["w"])
        m.set_normalization(z["mean"], z["std"], float(z["pressure_scale"]))
        return m

    def close(self) -> None:
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().psb_mlp_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dataset:
    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().psb_dataset_create(ctx.handle, int(capacity), C.byref(h)))
        self.handle = h
        ctx.adopt(self)

    def info(self):
        f, t = C.c_void_p(), C.c_void_p()
        cnt, cap = C.c_uint64(), C.c_uint64()
        check(lib().psb_dataset_info(self.handle, C.byref(f), C.byref(t), C.byref(cnt), C.byref(cap)))
        return f.value, t.value, cnt.value, cap.value

    @property
    def count(self) -> int:
        return self.info()[2]

    def reset(self) -> None:
        check(lib().psb_dataset_reset(self.handle))

    def download(self):
        """(features [8][count], targets [count]) as numpy arrays."""
        f, t, cnt, cap = self.info()
        feat = np.empty((NUM_FEATURES, cap), np.float32)
        check(lib().psb_copy_array_from_device(self.ctx.handle, feat.ctypes.data, f, 0, feat.nbytes))
        tgt = np.empty(cap, np.float32)
        check(lib().psb_copy_array_from_device(self.ctx.handle, tgt.ctypes.data, t, 0, tgt.nbytes))
        return feat[:, :cnt].copy(), tgt[:cnt].copy()

    def close(self) -> None:
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().psb_dataset_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def record(ctx: Context, scenes_list, steps: int, dataset: Dataset, stride: int = 16,
           warmup=0):
    """Run plain PCISPH on each scene and record (features, converged
    pressure) of every `stride`-th particle per step, after `warmup` plain
    steps that are not recorded (one count for every scene, or one per
    scene).  Returns iterations/step."""
    from .system import FluidSystem
    scenes_list = list(scenes_list)
    warm = [int(warmup)] * len(scenes_list) if np.isscalar(warmup) else [int(w) for w in warmup]
    if len(warm) != len(scenes_list):
        raise ValueError(f"{len(warm)} warm-up counts for {len(scenes_list)} scenes")
    iters = []
    for sc, w in zip(scenes_list, warm):
        p = sc.params.copy()
        p.warm_start = WARM_NONE
        fs = FluidSystem(ctx, p, sc.n)
        fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
        if w:
            fs.update(w)
        check(lib().psb_system_set_recorder(fs.handle, dataset.handle, int(stride)))
        fs.update(steps)
        iters.append(fs.stats()["total_iterations"] / steps)
        check(lib().psb_system_set_recorder(fs.handle, None, 0))
        fs.close()
    return iters


def train(ctx: Context, dataset: Dataset, pressure_scale: float, epochs: int = 3,
          batch: int = 8192, lr: float = 1e-3, seed: int = 1, model: PressureMLP | None = None,
          log=None):
    """Adam on minibatches of a recorded dataset.  Returns (model, losses)."""
    fptr, tptr, count, cap = dataset.info()
    if count == 0:
        raise ValueError("empty dataset")
    feat, _ = dataset.download()
    mean = feat.mean(axis=1).astype(np.float32)
    std = feat.std(axis=1).astype(np.float32)
    std[std < 1e-6] = 1.0
    model = model or PressureMLP(ctx, seed=seed)
    model.set_normalization(mean, std, pressure_scale)
    rng = np.random.default_rng(seed)
    losses = []
    idx_dev = ctx.array(4 * batch)
    # every sample once per epoch: full minibatches, then the remainder as a
    # smaller last one (a dataset smaller than one batch still trains)
    nb = -(-count // batch)
    for ep in range(epochs):
        perm = rng.permutation(count).astype(np.uint32)
        ep_loss = 0.0
        for b in range(nb):
            part = perm[b * batch:(b + 1) * batch]
            idx_dev.upload(part)
            ep_loss += model.train_step(fptr, tptr, cap, idx_dev.ptr, len(part), lr)
        losses.append(ep_loss / nb)
        if log:
            log(f"epoch {ep}: mse {losses[-1]:.5f}")
    idx_dev.free()
    return model, losses


WARM_NAMES = {WARM_NONE: "plain", WARM_PREVIOUS: "previous_pressure", WARM_MLP: "learned"}


def run_scene(ctx: Context, sc, steps: int, warm_start: int, model: PressureMLP | None = None):
    """Advance one scene `steps` steps with the given warm start; per-step
    corrector iterations and final density errors, and the wall time."""
    import time
    from .system import FluidSystem
    p = sc.params.copy()
    p.warm_start = warm_start
    fs = FluidSystem(ctx, p, sc.n)
    try:
        if warm_start == WARM_MLP:
            if model is None:
                raise ValueError("the learned warm start needs a trained model")
            fs.set_mlp(model)
        fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
        its = np.zeros(steps, np.int32)
        err = np.zeros(steps, np.float32)
        capped = 0
        t0 = time.perf_counter()
        for k in range(steps):
            fs.update(1)
            st = fs.stats()
            its[k] = st["iterations"]
            err[k] = st["max_error"]
            capped += bool(st["iterations"] >= p.max_iterations and st["max_error"] > p.tolerance)
        wall = time.perf_counter() - t0
        return {"iterations": its, "max_error": err, "capped": capped, "seconds": wall,
                "flags": fs.stats()["flags"]}
    finally:
        fs.close()


def evaluate(ctx: Context, scenes_list, steps: int, model: PressureMLP | None,
             modes=(WARM_NONE, WARM_PREVIOUS, WARM_MLP)) -> dict:
    """Mean corrector iterations per step of each warm start over held-out
    scenes, with the share of steps each one converged within tolerance."""
    out = {}
    for mode in modes:
        if mode == WARM_MLP and model is None:
            continue
        runs = [run_scene(ctx, sc, steps, mode, model) for sc in scenes_list]
        its = np.concatenate([r["iterations"] for r in runs])
        err = np.concatenate([r["max_error"] for r in runs])
        tol = scenes_list[0].params.tolerance
        out[WARM_NAMES[mode]] = {
            "iters_per_step": float(its.mean()),
            "iters_p99": float(np.percentile(its, 99)),
            "steps_within_tolerance": float((err <= tol).mean()),
            "capped_steps": int(sum(r["capped"] for r in runs)),
            "max_error": float(err.max()),
        }
    if "plain" in out:
        base = out["plain"]["iters_per_step"]
        for k, v in out.items():
            v["iteration_reduction"] = 1.0 - v["iters_per_step"] / base
    return out


def window_starts(count: int, steps: int, horizon: int | None) -> list[int]:
    """First recorded step of each training scene: every scene from step 0
    without a horizon; with one, `count` windows of `steps` steps spread
    evenly over [0, horizon), the first at 0 and the last ending at the
    horizon (16 windows of 200 over 2000 steps start every 120 steps)."""
    if not horizon or horizon <= steps or count < 2:
        return [0] * count
    span = horizon - steps
    return [(i * span) // (count - 1) for i in range(count)]


def train_and_evaluate(ctx: Context, make_scene, train_ids, eval_ids, steps: int = 200,
                       stride: int = 16, epochs: int = 4, batch: int = 8192, lr: float = 1e-3,
                       seed: int = 1, log=None, eval_steps: int | None = None,
                       horizon: int | None = None):
    """Record plain-PCISPH converged pressures on the training scenes (`steps`
    steps each, every `stride`-th particle; with `horizon`, the scenes'
    windows spread over the first `horizon` steps, `window_starts`), train the
    network on the GPU, then compare warm starts on the held-out scenes over
    their first `eval_steps` steps (default: `steps`).  Returns (model, summary)."""
    eval_steps = steps if eval_steps is None else eval_steps
    train_scenes = [make_scene(i) for i in train_ids]
    starts = window_starts(len(train_scenes), steps, horizon)
    n_per = sum((sc.n + stride - 1) // stride for sc in train_scenes) * steps
    ds = Dataset(ctx, n_per)
    plain_iters = record(ctx, train_scenes, steps, ds, stride=stride, warmup=starts)
    from .system import FluidSystem
    probe = FluidSystem(ctx, train_scenes[0].params, train_scenes[0].n)
    pscale = probe.pressure_scale
    probe.close()
    model, losses = train(ctx, ds, pscale, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    samples = ds.count
    ds.close()
    summary = {"train_scenes": list(train_ids), "eval_scenes": list(eval_ids), "steps": steps,
               "window_starts": starts, "eval_steps": eval_steps, "stride": stride,
               "samples": int(samples), "epochs": epochs, "lr": lr, "batch": batch,
               "train_plain_iters_per_step": float(np.mean(plain_iters)),
               "mse_per_epoch": [float(x) for x in losses]}
    summary["eval"] = evaluate(ctx, [make_scene(i) for i in eval_ids], eval_steps, model)
    return model, summary


def main(argv=None) -> int:
    import argparse
    import json
    from . import scenes
    ap = argparse.ArgumentParser(description="train and evaluate the learned PCISPH warm start")
    ap.add_argument("--train-seeds", type=int, default=16)
    ap.add_argument("--eval-seeds", type=int, default=4)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--n-side", type=int, default=64)
    ap.add_argument("--stride", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--horizon", type=int, default=0,
                    help="spread the training windows over this many steps (0: all from step 0)")
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None, help="write the trained model (.npz)")
    ap.add_argument("--json", default=None, help="write the summary (.json)")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = Context(a.device)

    def make_scene(i):
        return scenes.pcisph_dam_break(n_side=a.n_side, scene_id=i)

    model, summary = train_and_evaluate(
        ctx, make_scene, range(a.train_seeds), range(a.train_seeds, a.train_seeds + a.eval_seeds),
        steps=a.steps, stride=a.stride, epochs=a.epochs, batch=a.batch, lr=a.lr,
        log=lambda m: print(m, flush=True), horizon=a.horizon or None)
    if a.out:
        model.save(a.out)
    text = json.dumps(summary, indent=1)
    print(text)
    if a.json:
        with open(a.json, "w") as f:
            f.write(text)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

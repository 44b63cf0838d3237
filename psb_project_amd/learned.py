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

from ._lib import WARM_MLP, WARM_NONE, check, lib
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
           warmup: int = 0):
    """Run plain PCISPH on each scene and record (features, converged
    pressure) of every `stride`-th particle per step.  Returns iterations/step."""
    from .system import FluidSystem
    iters = []
    for sc in scenes_list:
        p = sc.params.copy()
        p.warm_start = WARM_NONE
        fs = FluidSystem(ctx, p, sc.n)
        fs.set_particles(sc.pos4, sc.vel4, sc.vh4)
        if warmup:
            fs.update(warmup)
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
    nb = count // batch
    for ep in range(epochs):
        perm = rng.permutation(count).astype(np.uint32)
        ep_loss = 0.0
        for b in range(nb):
            idx_dev.upload(perm[b * batch:(b + 1) * batch])
            ep_loss += model.train_step(fptr, tptr, cap, idx_dev.ptr, batch, lr)
        losses.append(ep_loss / max(1, nb))
        if log:
            log(f"epoch {ep}: mse {losses[-1]:.5f}")
    idx_dev.free()
    return model, losses

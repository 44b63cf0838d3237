"""Batched scenes: many independent PCISPH scenes advanced by one launch per
phase (psb_batch_* in include/psb_sph.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BATCH_PHASES, SceneDesc, StepStats, check, lib
from .device import Context


class BatchSystem:
    def __init__(self, ctx: Context, scenes):
        """`scenes`: Scene objects sharing one material; each contributes its
        particle count, collider box and grid (params.box_*, params.grid)."""
        self.ctx = ctx
        self.scenes = list(scenes)
        S = len(self.scenes)
        descs = (SceneDesc * S)()
        for d, sc in zip(descs, self.scenes):
            d.count = sc.n
            for a in range(3):
                d.box_min[a] = sc.params.box_min[a]
                d.box_max[a] = sc.params.box_max[a]
            d.grid = sc.params.grid
        self.material = self.scenes[0].params.copy()
        h = C.c_void_p()
        check(lib().psb_batch_create(ctx.handle, C.byref(self.material), S, descs, C.byref(h)))
        self.handle = h
        ctx.adopt(self)

    @property
    def num_scenes(self) -> int:
        return len(self.scenes)

    @property
    def n(self) -> int:
        return sum(sc.n for sc in self.scenes)

    def set_particles(self, i: int, pos4, vel4=None, vh4=None) -> None:
        p = np.ascontiguousarray(pos4, np.float32)
        v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        w = None if vh4 is None else np.ascontiguousarray(vh4, np.float32)
        check(lib().psb_batch_set_particles(self.handle, i, p.ctypes.data,
                                            None if v is None else v.ctypes.data,
                                            None if w is None else w.ctypes.data))

    def set_all(self) -> None:
        for i, sc in enumerate(self.scenes):
            self.set_particles(i, sc.pos4, sc.vel4, sc.vh4)

    def update(self, steps: int = 1) -> None:
        check(lib().psb_batch_update(self.handle, int(steps)))

    def stats(self, i: int) -> dict:
        s = StepStats()
        check(lib().psb_batch_get_stats(self.handle, i, C.byref(s)))
        return {k: getattr(s, k) for k, _ in StepStats._fields_}

    def state(self, i: int) -> dict:
        n = self.scenes[i].n
        o = {"pos": np.empty((n, 4), np.float32), "vel": np.empty((n, 4), np.float32),
             "vh": np.empty((n, 4), np.float32), "density": np.empty(n, np.float32),
             "pressure": np.empty(n, np.float32), "ids": np.empty(n, np.uint32)}
        check(lib().psb_batch_get_particles(self.handle, i, o["pos"].ctypes.data,
                                            o["vel"].ctypes.data, o["vh"].ctypes.data,
                                            o["density"].ctypes.data, o["pressure"].ctypes.data,
                                            o["ids"].ctypes.data))
        return o

    def set_mlp(self, model) -> None:
        check(lib().psb_batch_set_mlp(self.handle, model.handle if model is not None else None))

    def autotune(self, steps: int = 2) -> dict:
        out = (C.c_int32 * len(BATCH_PHASES))()
        check(lib().psb_batch_autotune(self.handle, int(steps), out))
        return dict(zip(BATCH_PHASES, [int(x) for x in out]))

    def time_corrector(self, reps: int = 20) -> dict:
        """Microseconds per launch of the corrector passes over every scene
        (iteration 0 replayed back to back on the current state)."""
        d, f = C.c_double(), C.c_double()
        check(lib().psb_batch_time_corrector(self.handle, int(reps), C.byref(d), C.byref(f)))
        return {"pci_density": d.value, "pci_force": f.value}

    def set_block_sizes(self, sizes: dict) -> None:
        """Launch shapes per phase, as autotune returns them (negative: two
        threads per particle in pci_density / pci_force)."""
        arr = (C.c_int32 * len(BATCH_PHASES))(*[int(sizes[p]) for p in BATCH_PHASES])
        check(lib().psb_batch_set_block_sizes(self.handle, arr))

    def set_graphs(self, on: bool) -> None:
        check(lib().psb_batch_set_graphs(self.handle, 1 if on else 0))

    def set_latency(self, on, every: int = 1) -> None:
        """Record per-scene step latencies on every `every`-th step."""
        check(lib().psb_batch_set_latency(self.handle, int(every) if on else 0))

    def latency(self) -> dict:
        p50, p99, mx = C.c_double(), C.c_double(), C.c_double()
        cnt = C.c_uint64()
        check(lib().psb_batch_get_latency(self.handle, C.byref(p50), C.byref(p99), C.byref(mx),
                                          C.byref(cnt)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "max_ms": mx.value,
                "samples": cnt.value}

    def close(self) -> None:
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().psb_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

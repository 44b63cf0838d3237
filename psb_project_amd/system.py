"""Python host side of the fused system step (psb_system_* C-ABI).

`FluidSystem` owns one psb_system: it uploads a scene, advances it with
psb_system_update (SPHSystem::update / PCI_SPHSystem::update of the
reference) and reads the state back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Params, StepStats, SystemArrays, check, lib
from .device import Context, DeviceArray


def sphere_particle_count(params, radius: float) -> int:
    """Boundary particles per rigid sphere of this radius (lattice points inside)."""
    from ._lib import SphereParams
    sp = SphereParams(radius, 1.0, 0.0, 0.0)
    return int(lib().psb_sphere_particle_count(C.byref(params), C.byref(sp)))


class FluidSystem:
    def __init__(self, ctx: Context, params: Params, capacity: int):
        self.ctx = ctx
        self.params = params.copy()
        self.capacity = int(capacity)
        h = C.c_void_p()
        check(lib().psb_system_create(ctx.handle, C.byref(self.params), self.capacity, C.byref(h)))
        self.handle = h
        ctx.adopt(self)

    # -- state -------------------------------------------------------------
    def set_particles(self, pos4: np.ndarray, vel4: np.ndarray | None = None,
                      vh4: np.ndarray | None = None) -> None:
        """Upload (N,4) float32 positions / velocities / half-step velocities."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        n = pos4.shape[0]
        bufs = [self.ctx.to_device(pos4)]
        bufs.append(self.ctx.to_device(np.ascontiguousarray(vel4, np.float32)) if vel4 is not None else None)
        bufs.append(self.ctx.to_device(np.ascontiguousarray(vh4, np.float32)) if vh4 is not None else None)
        check(lib().psb_system_set_particles_device(
            self.handle, n, bufs[0].ptr, bufs[1].ptr if bufs[1] else None,
            bufs[2].ptr if bufs[2] else None))
        for b in bufs:
            if b is not None:
                b.free()

    def set_particles_f64(self, new_pos, old_pos=None, new_vel=None, old_vel=None,
                          vel_inter_av=None, vel_inter_ap=None) -> None:
        """Upload host FP64 xyz triples (the reference's m_h* layout)."""
        arrs = [new_pos, old_pos, new_vel, old_vel, vel_inter_av, vel_inter_ap]
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1, 3) for a in arrs]
        n = arrs[0].shape[0]
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        check(lib().psb_system_set_particles(self.handle, n, ptr(arrs[1]), ptr(arrs[0]),
                                             ptr(arrs[3]), ptr(arrs[2]), ptr(arrs[4]), ptr(arrs[5])))

    def set_materials(self, materials, index) -> None:
        """Per-particle materials (reference mode): particle i of the last
        set_particles takes materials[index[i]]; an empty list returns to
        the system's one material."""
        mats = (Params * max(len(materials), 1))(*[m.copy() for m in materials])
        idx = np.ascontiguousarray(index, np.uint8)
        check(lib().psb_system_set_materials(self.handle, len(materials), C.addressof(mats),
                                             idx.ctypes.data if idx.size else None, idx.size))

    @property
    def num_materials(self) -> int:
        return int(lib().psb_system_num_materials(self.handle))

    @property
    def n(self) -> int:
        return int(lib().psb_system_count(self.handle))

    def update(self, steps: int = 1) -> None:
        check(lib().psb_system_update(self.handle, int(steps)))

    def stats(self) -> dict:
        s = StepStats()
        check(lib().psb_system_get_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in StepStats._fields_}

    def state(self) -> dict:
        """Sorted-order FP32 state of the last step plus particle ids."""
        n = self.n
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        vh = np.empty((n, 4), np.float32)
        rho = np.empty(n, np.float32)
        prs = np.empty(n, np.float32)
        ids = np.empty(n, np.uint32)
        check(lib().psb_system_get_particles_f32(self.handle, pos.ctypes.data, vel.ctypes.data,
                                                 vh.ctypes.data, rho.ctypes.data, prs.ctypes.data,
                                                 ids.ctypes.data))
        return {"pos": pos, "vel": vel, "vh": vh, "density": rho, "pressure": prs, "ids": ids}

    def state_f64(self) -> dict:
        """Reference-layout FP64 arrays in the caller's original particle order."""
        n = self.n
        out = {k: np.empty((n, 3), np.float64) for k in
               ("old_pos", "new_pos", "old_vel", "new_vel", "vel_inter_av", "vel_inter_ap")}
        out["density"] = np.empty(n, np.float64)
        out["pressure"] = np.empty(n, np.float64)
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        check(lib().psb_system_get_particles(self.handle, P(out["old_pos"]), P(out["new_pos"]),
                                             P(out["old_vel"]), P(out["new_vel"]),
                                             P(out["vel_inter_av"]), P(out["vel_inter_ap"]),
                                             P(out["density"]), P(out["pressure"])))
        return out

    def arrays(self) -> SystemArrays:
        a = SystemArrays()
        check(lib().psb_system_get_arrays(self.handle, C.byref(a)))
        return a

    def download(self, ptr: int, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        check(lib().psb_copy_array_from_device(self.ctx.handle, out.ctypes.data, ptr, 0, out.nbytes))
        return out

    # -- learned warm start -------------------------------------------------
    @property
    def pressure_scale(self) -> float:
        return float(lib().psb_system_pressure_scale(self.handle))

    def set_spheres(self, centers, velocities=None, radius: float = 0.015,
                    density: float = 500.0, spring: float = 1e4, damping: float = 0.0) -> None:
        """Rigid DEM spheres; call before set_particles (their boundary
        particles follow the fluid's)."""
        from ._lib import SphereParams
        c = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        v = None if velocities is None else np.ascontiguousarray(velocities, np.float32).reshape(-1, 3)
        sp = SphereParams(radius, density, spring, damping)
        check(lib().psb_system_set_spheres(self.handle, c.shape[0], c.ctypes.data,
                                           None if v is None else v.ctypes.data, C.byref(sp)))

    def add_collider(self, collider) -> None:
        """A _lib.Collider (sphere or plane) applied after the box each step."""
        check(lib().psb_system_add_collider(self.handle, C.addressof(collider)))

    def clear_colliders(self) -> None:
        check(lib().psb_system_clear_colliders(self.handle))

    @property
    def num_spheres(self) -> int:
        return int(lib().psb_system_num_spheres(self.handle))

    def spheres(self) -> dict:
        m = self.num_spheres
        o = {k: np.zeros((m, 3), np.float32) for k in ("centers", "velocities", "fluid_force")}
        if m:
            check(lib().psb_system_get_spheres(self.handle, o["centers"].ctypes.data,
                                               o["velocities"].ctypes.data,
                                               o["fluid_force"].ctypes.data))
        return o

    def set_mlp(self, model) -> None:
        check(lib().psb_system_set_mlp(self.handle, model.handle if model is not None else None))

    def set_recorder(self, dataset, stride: int = 16) -> None:
        check(lib().psb_system_set_recorder(self.handle, dataset.handle if dataset else None,
                                            int(stride)))

    def features(self) -> np.ndarray:
        """Warm-start features of the last step, (8, n) in sorted order."""
        ptr = C.c_void_p()
        check(lib().psb_system_get_features(self.handle, C.byref(ptr)))
        if not ptr.value:
            raise RuntimeError("no features computed (enable the MLP warm start or a recorder)")
        return self.download(ptr.value, np.float32, (8, self.n))

    def set_state(self, pos4, vel4=None, vh4=None, pressure=None, ids=None) -> None:
        """Particles in a given storage order with their ids and pressure."""
        p = np.ascontiguousarray(pos4, np.float32)
        n = p.shape[0]
        ids = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, np.uint32)
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32)
                for a in (vel4, vh4, pressure)]
        check(lib().psb_system_set_state(self.handle, n, p.ctypes.data,
                                         *[None if a is None else a.ctypes.data for a in arrs],
                                         ids.ctypes.data))

    def save_checkpoint(self, path) -> None:
        check(lib().psb_system_save_checkpoint(self.handle, str(path).encode()))

    def load_checkpoint(self, path) -> None:
        check(lib().psb_system_load_checkpoint(self.handle, str(path).encode()))

    @staticmethod
    def checkpoint_info(path) -> dict:
        prm = Params()
        n, nf, ns = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().psb_checkpoint_info(str(path).encode(), C.byref(prm), C.byref(n), C.byref(nf),
                                        C.byref(ns)))
        return {"params": prm, "count": n.value, "fluid_count": nf.value, "num_spheres": ns.value}

    def set_halo(self, on: bool) -> None:
        """Stage each tile's neighbour rows in LDS in the corrector passes
        (default off; single lane with the compact list)."""
        check(lib().psb_system_set_halo(self.handle, 1 if on else 0))

    def set_compact_list(self, on: bool) -> None:
        """16-bit neighbour offsets in the corrector passes (default on)."""
        check(lib().psb_system_set_compact_list(self.handle, 1 if on else 0))

    def set_corrector_lanes(self, lanes: int) -> None:
        """Threads per particle in the corrector passes: 1, 2, or 0 (default:
        2 up to 1024 particles per compute unit, else 1)."""
        check(lib().psb_system_set_corrector_lanes(self.handle, int(lanes)))

    # -- timing ------------------------------------------------------------
    def set_timing(self, on, every: int = 1) -> None:
        """Per-kernel HIP events on every `every`-th step (off when `on` is false)."""
        check(lib().psb_system_set_timing(self.handle, int(every) if on else 0))

    def time_corrector(self, reps: int = 50) -> dict:
        """Per-launch durations (us) of the density and force passes,
        replayed `reps` times back to back on the current state."""
        d, f = C.c_double(), C.c_double()
        check(lib().psb_system_time_corrector(self.handle, int(reps), C.byref(d), C.byref(f)))
        return {"pci_density": d.value, "pci_force": f.value}

    def reset_timing(self) -> None:
        check(lib().psb_system_reset_timing(self.handle))

    def timing(self) -> dict:
        cap = 32
        names = (C.c_char_p * cap)()
        tot = (C.c_double * cap)()
        cnt = (C.c_int64 * cap)()
        k = lib().psb_system_get_timing(self.handle, cap, names, tot, cnt)
        return {names[i].decode(): (tot[i], cnt[i]) for i in range(k)}

    def close(self) -> None:
        # the context destroys its systems before it is finalised
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().psb_system_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

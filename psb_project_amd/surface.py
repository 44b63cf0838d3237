"""Iso-surface extraction (psb_surface_* C-ABI): SurfaceSPH::extract of the
reference (SurfaceSPH.cpp:29-190) on the GPU -- the particle field on a node
grid, marching cubes, an indexed mesh with vertex normals."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SurfaceParams, check, lib
from .device import Context


def default_params(**overrides) -> SurfaceParams:
    p = SurfaceParams()
    check(lib().psb_surface_default_params(C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def case_table():
    """(triangle counts [256], edge indices [256, 16], -1 padded)."""
    nt = np.zeros(256, np.uint8)
    tri = np.zeros((256, 16), np.int8)
    check(lib().psb_surface_case_table(nt.ctypes.data, tri.ctypes.data))
    return nt, tri


class Surface:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().psb_surface_create(ctx.handle, C.byref(h)))
        self.handle = h
        ctx.adopt(self)

    def run(self, system, params: SurfaceParams | None = None) -> tuple[int, int]:
        """Extract on the device only; (vertex count, triangle count)."""
        params = params or default_params()
        nv, nt = C.c_uint32(), C.c_uint32()
        check(lib().psb_surface_extract(self.handle, system.handle, C.byref(params), C.byref(nv),
                                        C.byref(nt)))
        return nv.value, nt.value

    def extract(self, system, params: SurfaceParams | None = None) -> dict:
        """Surface of a FluidSystem's fluid particles (current state)."""
        return self._fetch(*self.run(system, params))

    def extract_points(self, pos4, support_radius: float, params: SurfaceParams | None = None,
                       ids=None, id_limit: int = 0xFFFFFFFF) -> dict:
        """Surface of host float4 positions (uploaded for the call)."""
        params = params or default_params()
        x = np.ascontiguousarray(pos4, np.float32)
        dx = self.ctx.to_device(x)
        di = None if ids is None else self.ctx.to_device(np.ascontiguousarray(ids, np.uint32))
        nv, nt = C.c_uint32(), C.c_uint32()
        check(lib().psb_surface_extract_points(self.handle, dx.ptr, None if di is None else di.ptr,
                                               x.shape[0], id_limit, support_radius,
                                               C.byref(params), C.byref(nv), C.byref(nt)))
        return self._fetch(nv.value, nt.value)

    def field(self) -> dict:
        """Node field of the last extraction (x fastest) and its grid."""
        dims = np.zeros(3, np.uint32)
        org = np.zeros(3, np.float32)
        cell = C.c_float()
        check(lib().psb_surface_get_field(self.handle, None, dims.ctypes.data, org.ctypes.data,
                                          C.byref(cell)))
        field = np.zeros(int(np.prod(dims.astype(np.int64))), np.float32)
        check(lib().psb_surface_get_field(self.handle, field.ctypes.data, None, None, None))
        return {"field": field, "dims": dims, "origin": org, "cell": cell.value}

    def _fetch(self, nv: int, nt: int) -> dict:
        """Mesh of the last extraction; nv, nt must be the counts it returned
        (the copies fill arrays of exactly that size)."""
        out = {"vertices": np.zeros((nv, 3), np.float32), "normals": np.zeros((nv, 3), np.float32),
               "triangles": np.zeros((nt, 3), np.uint32)}
        check(lib().psb_surface_get_mesh(self.handle, out["vertices"].ctypes.data,
                                         out["normals"].ctypes.data, out["triangles"].ctypes.data))
        out.update(self.field())
        return out

    def close(self) -> None:
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().psb_surface_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

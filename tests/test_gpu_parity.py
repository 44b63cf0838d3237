"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Integer work (hash, sorted order, cell ranges, neighbour lists) must be
bit-exact.  The FP32 physics follows the same operation order as the oracle,
so state is compared bit for bit as well; the stated tolerance of the
workload (1e-5 relative per step, 1e-3 after 1000 steps) is checked on top
where a bitwise miss would still be acceptable.
"""
import ctypes as C

import numpy as np
import pytest

from psb_project_amd import scenes
from psb_project_amd._lib import (FLAG_NEIGHBOR_OVERFLOW, FLAG_NOT_CONVERGED, FLAG_OUT_OF_GRID,
                                  MODE_CONFIG, MODE_REFERENCE, Neighbors, PsbError, check, lib)

pytestmark = pytest.mark.gpu


def _grid(ctx, gd, cap):
    g = C.c_void_p()
    check(lib().psb_create_grid_desc(ctx.handle, C.byref(gd), cap, C.byref(g)))
    return g


def _grid_arrays(ctx, g, n, ncells):
    h, p, s = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nc = C.c_uint64()
    check(lib().psb_grid_get_arrays(g, C.byref(h), C.byref(p), C.byref(s), C.byref(nc)))
    assert nc.value == ncells
    out = []
    for ptr, cnt in ((h, n), (p, n), (s, ncells + 1)):
        a = np.empty(cnt, np.uint32)
        check(lib().psb_copy_array_from_device(ctx.handle, a.ctypes.data, ptr, 0, a.nbytes))
        out.append(a)
    return out


def gpu_grid_and_neighbors(ctx, x4, gd, radius, cap):
    n = x4.shape[0]
    ncells = gd.dims[0] * gd.dims[1] * gd.dims[2]
    g = _grid(ctx, gd, n)
    try:
        dpos = ctx.to_device(x4)
        check(lib().psb_store_particles(g, dpos.ptr, n))
        hash_, perm, start = _grid_arrays(ctx, g, n, ncells)
        dsorted = ctx.array(16 * n)
        check(lib().psb_grid_reorder(g, dpos.ptr, dsorted.ptr, n, 4))
        cnt_d = ctx.array(4 * n)
        lst_d = ctx.array(4 * n * cap)
        nb = Neighbors(cnt_d.addr, lst_d.addr, cap, n)
        flags = C.c_uint32()
        check(lib().psb_search_neighbors(g, dsorted.ptr, n, radius, nb, C.byref(flags)))
        cnt = cnt_d.download(np.uint32, n)
        lst = lst_d.download(np.uint32, (cap, n))
        xs = dsorted.download(np.float32, (n, 4))
        return dict(hash=hash_, perm=perm, start=start, count=cnt, list=lst, xs=xs,
                    flags=flags.value, nb=nb, bufs=(dpos, dsorted, cnt_d, lst_d))
    finally:
        lib().psb_finalize_grid(g)


SCENES = {
    "wcsph_32k": lambda: scenes.wcsph_dam_break(),
    "pcisph_262k": lambda: scenes.pcisph_dam_break(),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_grid_build_bitwise(ctx, oracle, name):
    sc = SCENES[name]()
    rng = np.random.default_rng(11)
    x = rng.permutation(sc.pos4)       # arbitrary order exercises the stable rank
    gd = sc.params.grid
    r = gpu_grid_and_neighbors(ctx, x, gd, sc.params.support_radius, sc.params.max_neighbors)
    h, perm, start, flags = oracle.grid_build(x, gd)
    assert np.array_equal(r["hash"], h)
    assert np.array_equal(r["perm"], perm)
    assert np.array_equal(r["start"], start)
    assert np.array_equal(r["xs"], x[perm])


def test_create_grid_follows_create_grid(ctx, oracle):
    """psb_create_grid takes createGrid's arguments (centre, extents, float
    cell size; UniformGrid.cu:18-31): Min = centre - extent/2 kept as float,
    nbCells = ceil(extent / (double)sizeCell).  0.005f is just below 0.005,
    so 0.64 m gives 129 cells, not the decimal 128.  The grid it makes then
    sorts particles exactly as the oracle does with that descriptor."""
    from psb_project_amd._lib import GridDesc
    cx, cy, cz, L, W, D, cell = 0.1, 0.3, -0.2, 0.64, 0.32, 0.16, np.float32(0.005)
    g = C.c_void_p()
    check(lib().psb_create_grid(ctx.handle, cx, cy, cz, L, W, D, float(cell), 8192, C.byref(g)))
    try:
        gd = GridDesc()
        check(lib().psb_grid_get_desc(g, C.byref(gd)))
        want_dims = [int(np.ceil(e / np.float64(cell))) for e in (L, W, D)]
        assert list(gd.dims) == want_dims == [129, 65, 33]
        want_min = np.float32([cx - L / 2, cy - W / 2, cz - D / 2])
        assert np.array_equal(np.float32(list(gd.min)), want_min)
        assert np.float32(gd.cell_size) == cell
        rng = np.random.default_rng(12)
        x = np.zeros((8192, 4), np.float32)
        x[:, :3] = want_min + rng.uniform(0, 1, (8192, 3)) * np.float32([L, W, D])
        dpos = ctx.to_device(x)
        check(lib().psb_store_particles(g, dpos.ptr, x.shape[0]))
        hash_, perm, start = _grid_arrays(ctx, g, x.shape[0], int(np.prod(want_dims)))
        h, p, s, _ = oracle.grid_build(x, gd)
        assert np.array_equal(hash_, h)
        assert np.array_equal(perm, p)
        assert np.array_equal(start, s)
        with pytest.raises(PsbError):
            check(lib().psb_create_grid(ctx.handle, cx, cy, cz, L, W, D, 0.0, 16, C.byref(C.c_void_p())))
    finally:
        lib().psb_finalize_grid(g)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_neighbors_bitwise(ctx, oracle, name):
    sc = SCENES[name]()
    gd = sc.params.grid
    cap = sc.params.max_neighbors
    r = gpu_grid_and_neighbors(ctx, sc.pos4, gd, sc.params.support_radius, cap)
    cnt, lst, fl = oracle.neighbors(r["xs"], gd, r["start"], sc.params.support_radius, cap)
    assert r["flags"] == fl == 0
    assert np.array_equal(r["count"], cnt)
    mask = np.arange(cap)[:, None] < cnt[None, :]
    assert np.array_equal(r["list"][mask], lst[mask])


def test_neighbor_overflow_flag(ctx, oracle):
    sc = scenes.wcsph_dam_break(n_side=12)
    r = gpu_grid_and_neighbors(ctx, sc.pos4, sc.params.grid, sc.params.support_radius, 8)
    assert r["flags"] & FLAG_NEIGHBOR_OVERFLOW
    assert r["count"].max() == 8


@pytest.mark.parametrize("mode", [MODE_REFERENCE, MODE_CONFIG])
def test_evaluate_densities_forces_bitwise(ctx, oracle, mode):
    sc = scenes.wcsph_dam_break(n_side=20, mode=mode)
    p = sc.params
    n = sc.n
    r = gpu_grid_and_neighbors(ctx, sc.pos4, p.grid, p.support_radius, p.max_neighbors)
    rng = np.random.default_rng(5)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(scale=0.2, size=(n, 3))
    dvel = ctx.to_device(vel)
    outs = {k: ctx.array(16 * n) for k in ("normals", "fP", "fV", "fS", "F")}
    rho = ctx.array(4 * n)
    prs = ctx.array(4 * n)
    dsorted = r["bufs"][1]
    check(lib().psb_evaluate_densities_forces(
        ctx.handle, C.byref(p), r["nb"], dsorted.ptr, dvel.ptr, n, rho.ptr, prs.ptr,
        outs["normals"].ptr, outs["fP"].ptr, outs["fV"].ptr, outs["fS"].ptr, outs["F"].ptr))
    ref = oracle.evaluate_densities_forces(p, r["count"], r["list"].ravel(), r["xs"], vel)
    assert np.array_equal(rho.download(np.float32, n), ref["density"])
    assert np.array_equal(prs.download(np.float32, n), ref["pressure"])
    for k in outs:
        got = outs[k].download(np.float32, (n, 4))
        assert np.array_equal(got[:, :3], ref[k][:, :3]), k


def test_pointwise_kernels_bitwise(ctx, oracle):
    n = 5000
    rng = np.random.default_rng(9)
    f32 = np.float32
    pos = np.zeros((n, 4), f32); pos[:, :3] = rng.uniform(-0.1, 0.8, (n, 3))
    vel = np.zeros((n, 4), f32); vel[:, :3] = rng.normal(size=(n, 3))
    vh = np.zeros((n, 4), f32); vh[:, :3] = rng.normal(size=(n, 3))
    F = np.zeros((n, 4), f32); F[:, :3] = rng.normal(scale=1e3, size=(n, 3))
    rho = rng.uniform(900, 1100, n).astype(f32)
    dt = f32(1e-3)
    # integrateSPHSystem
    d = {k: ctx.to_device(v) for k, v in dict(vav=np.zeros_like(vh), vap=vh, xo=np.zeros_like(pos),
                                              xn=pos, F=F, rho=rho).items()}
    check(lib().psb_integrate_sph(ctx.handle, d["vav"].ptr, d["vap"].ptr, d["xo"].ptr, d["xn"].ptr,
                                  d["F"].ptr, d["rho"].ptr, float(dt), n))
    inv = f32(1) / rho
    vap = vh[:, :3] + (F[:, :3] * inv[:, None]) * dt
    xn = pos[:, :3] + vap * dt
    assert np.array_equal(d["vap"].download(f32, (n, 4))[:, :3], vap)
    assert np.array_equal(d["vav"].download(f32, (n, 4))[:, :3], vh[:, :3])
    assert np.array_equal(d["xn"].download(f32, (n, 4))[:, :3], xn)
    assert np.array_equal(d["xo"].download(f32, (n, 4))[:, :3], pos[:, :3])
    # interpolateSPHVelocities
    dvo = ctx.to_device(np.zeros_like(vel)); dvn = ctx.to_device(vel)
    check(lib().psb_interpolate_velocities(ctx.handle, d["vav"].ptr, d["vap"].ptr, dvo.ptr, dvn.ptr, n))
    assert np.array_equal(dvn.download(f32, (n, 4))[:, :3], (vh[:, :3] + vap) * f32(0.5))
    assert np.array_equal(dvo.download(f32, (n, 4))[:, :3], vel[:, :3])
    # evaluate_ForceExt_Constante with density as weight
    dF = ctx.to_device(F)
    direc = (C.c_float * 3)(0.0, -1.0, 0.0)
    check(lib().psb_force_constante(ctx.handle, n, dF.ptr, d["rho"].ptr, direc, 9.81))
    expF = F.copy(); expF[:, 1] = F[:, 1] + f32(-9.81) * rho
    expF[:, 0] = F[:, 0] + f32(0.0) * rho; expF[:, 2] = F[:, 2] + f32(0.0) * rho
    assert np.array_equal(dF.download(f32, (n, 4))[:, :3], expF[:, :3])
    # collisionSystem_Box_CUDA
    sc = scenes.wcsph_dam_break(n_side=4)
    p = sc.params
    for a in range(3):
        p.box_min[a], p.box_max[a] = 0.0, 0.64
    for mode, elast in ((MODE_REFERENCE, 0.0), (MODE_CONFIG, 0.5)):
        p.mode, p.elast = mode, 0.5
        dx, dv = ctx.to_device(pos), ctx.to_device(vel)
        bmin = (C.c_float * 3)(*p.box_min); bmax = (C.c_float * 3)(*p.box_max)
        check(lib().psb_collision_box(ctx.handle, dx.ptr, None, dv.ptr, p.particle_radius, float(dt),
                                      n, bmin, bmax, 1, elast))
        xo, vo = oracle.collision_box(p, pos, vel)
        assert np.array_equal(dx.download(f32, (n, 4))[:, :3], xo[:, :3])
        assert np.array_equal(dv.download(f32, (n, 4))[:, :3], vo[:, :3])


# ------------------------------------------------------------------ systems
def run_pair(ctx, oracle, sc, steps, chunks=1, lanes=0, compact=True):
    from psb_project_amd.system import FluidSystem
    g = FluidSystem(ctx, sc.params, sc.n)
    g.set_corrector_lanes(lanes)
    g.set_compact_list(compact)
    o = oracle.OracleSystem(sc.params, sc.n)
    g.set_particles(sc.pos4, sc.vel4, sc.vh4)
    o.set_particles(sc.pos4, sc.vel4, sc.vh4)
    gi, oi = [], []
    for _ in range(chunks):
        g.update(steps)
        o.update(steps)
        gi.append(g.stats()); oi.append(o.stats())
    return g, o, gi, oi


def assert_state_equal(gs, os_, keys=("pos", "vel", "vh", "density", "pressure", "ids")):
    for k in keys:
        a, b = gs[k], os_[k]
        if a.ndim == 2:
            a, b = a[:, :3], b[:, :3]
        assert np.array_equal(a, b), f"{k}: max |diff| {np.abs(a.astype(np.float64) - b).max()}"


@pytest.mark.parametrize("mode", [MODE_REFERENCE, MODE_CONFIG])
@pytest.mark.parametrize("steps", [1, 10])
def test_wcsph_system_bitwise(ctx, oracle, mode, steps):
    sc = scenes.wcsph_dam_break(n_side=16, mode=mode)
    g, o, gi, oi = run_pair(ctx, oracle, sc, steps)
    assert_state_equal(g.state(), o.state())
    assert gi[0]["flags"] == oi[0]["flags"]
    assert gi[0]["max_neighbors_seen"] == oi[0]["max_neighbors_seen"] > 0


@pytest.mark.parametrize("lanes,compact", [(1, True), (2, True), (1, False)])
@pytest.mark.parametrize("mode", [MODE_REFERENCE, MODE_CONFIG])
def test_pcisph_system_bitwise(ctx, oracle, mode, lanes, compact):
    """One and two threads per particle in the corrector passes; 16-bit
    and 32-bit neighbour lists."""
    sc = scenes.pcisph_dam_break(n_side=24, mode=mode)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 1, chunks=6, lanes=lanes, compact=compact)
    for a, b in zip(gi, oi):
        assert a["iterations"] == b["iterations"]
        assert a["max_error"] == np.float32(b["max_error"])
        assert a["flags"] == b["flags"]
    # the step's largest neighbour count (spread slots, folded by the final force pass)
    assert gi[0]["max_neighbors_seen"] == oi[0]["max_neighbors_seen"] > 0
    assert_state_equal(g.state(), o.state())


def test_pcisph_clamp_walls_bitwise(ctx, oracle):
    sc = scenes.pcisph_dam_break(n_side=24, lattice_walls=False)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 1, chunks=6)
    for a, b in zip(gi, oi):
        assert a["iterations"] == b["iterations"]
        assert a["max_error"] == np.float32(b["max_error"])
    assert_state_equal(g.state(), o.state())


@pytest.mark.parametrize("lanes", [1, 2])
def test_pcisph_lattice_walls_through_impact_bitwise(ctx, oracle, lanes):
    """Coarse dam break (16^3, s = 0.02) run past the impact on the far wall
    (~step 420): particles off the lattice, predicted positions beyond the
    walls, corner and edge wall sums; bitwise against the oracle."""
    sc = scenes.pcisph_dam_break(n_side=16, s=0.02, lattice_walls=True)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 100, chunks=6, lanes=lanes)
    for a, b in zip(gi, oi):
        assert a["iterations"] == b["iterations"]
        assert a["total_iterations"] == b["total_iterations"]
        assert a["max_error"] == np.float32(b["max_error"])
    assert gi[-1]["flags"] == oi[-1]["flags"]
    assert_state_equal(g.state(), o.state())


@pytest.mark.parametrize("cap", [37, 29])
@pytest.mark.parametrize("lanes", [1, 2])
def test_pcisph_list_capacity_off_group(ctx, oracle, cap, lanes):
    """A list capacity that is not a multiple of the compact list's 8-entry
    groups (37), and one below the lattice's 33 neighbours (29: lists
    truncated in order, overflow flagged); one and two lanes; bitwise."""
    sc = scenes.pcisph_dam_break(n_side=16, s=0.02, lattice_walls=True)
    sc.params.max_neighbors = cap
    g, o, gi, oi = run_pair(ctx, oracle, sc, 20, chunks=3, lanes=lanes)
    for a, b in zip(gi, oi):
        assert a["iterations"] == b["iterations"]
        assert a["flags"] == b["flags"]
    assert bool(gi[0]["flags"] & FLAG_NEIGHBOR_OVERFLOW) == (cap == 29)
    assert_state_equal(g.state(), o.state())


@pytest.mark.parametrize("scene", ["lattice_walls_16", "random_cloud"])
def test_halo_staging_keeps_results(ctx, oracle, scene):
    """Neighbour rows staged in LDS against the global gathers and
    the oracle: a coarse dam break through its wall impact (tiles that fit),
    and a random cloud whose tiles span too many rows (global fallback)."""
    from psb_project_amd.system import FluidSystem
    if scene == "lattice_walls_16":
        sc = scenes.pcisph_dam_break(n_side=16, s=0.02, lattice_walls=True)
        steps, chunks = 60, 5
    else:
        sc = scenes.pcisph_dam_break(n_side=4)
        rng = np.random.default_rng(12)
        pos = np.zeros((12000, 4), np.float32)
        pos[:, :3] = rng.uniform([0.02, 0.02, 0.02], [0.5, 0.3, 0.3], (12000, 3))
        sc = scenes.Scene("cloud", sc.params, pos)
        steps, chunks = 4, 2
    runs = []
    for halo in (True, False):
        g = FluidSystem(ctx, sc.params, sc.n)
        g.set_corrector_lanes(1)   # staging runs with one lane per particle
        g.set_halo(halo)
        g.set_particles(sc.pos4)
        runs.append(g)
    o = oracle.OracleSystem(sc.params, sc.n)
    o.set_particles(sc.pos4)
    for _ in range(chunks):
        for g in runs:
            g.update(steps)
        o.update(steps)
        assert runs[0].stats()["total_iterations"] == o.stats()["total_iterations"]
    assert_state_equal(runs[0].state(), runs[1].state())
    assert_state_equal(runs[0].state(), o.state())
    for g in runs:
        g.close()


@pytest.mark.parametrize("lanes", [1, 2])
def test_compact_list_falls_back_on_wide_offsets(ctx, oracle, lanes):
    """A thin 180 x 100 x 4 slab on 5 mm cells: neighbours one or two cell
    layers apart in z sit more than 32767 places away in the sorted order,
    so the 16-bit list cannot hold them.  The corrector passes read the
    32-bit list in the step that overflows and the grouped 32-bit list from
    the next step on; the run equals the one without the compact list (the
    plain 32-bit list throughout) and the oracle, bit for bit."""
    from psb_project_amd.system import FluidSystem
    sc = scenes.pcisph_dam_break(n_side=4)
    pos = scenes.lattice((180, 100, 4), 0.005, seed=11)
    runs = []
    for compact in (True, False):
        g = FluidSystem(ctx, sc.params, pos.shape[0])
        g.set_compact_list(compact)
        g.set_corrector_lanes(lanes)
        g.set_particles(pos)
        g.update(2)
        g.update(4)
        runs.append(g)
    o = oracle.OracleSystem(sc.params, pos.shape[0])
    o.set_particles(pos)
    o.update(2)
    o.update(4)                   # statistics cover the last update call
    assert_state_equal(runs[0].state(), o.state())
    assert runs[0].stats()["total_iterations"] == o.stats()["total_iterations"]
    o.close()
    nb = runs[0].arrays().neighbors
    n = runs[0].n
    cnt = runs[0].download(nb.count, np.uint32, n)
    lst = runs[0].download(nb.list, np.uint32, (nb.capacity, n)).astype(np.int64)
    k = np.arange(nb.capacity)[:, None]
    off = np.where(k < cnt[None, :], lst - np.arange(n)[None, :], 0)
    assert np.abs(off).max() > 32767
    assert_state_equal(runs[0].state(), runs[1].state())
    assert runs[0].stats()["total_iterations"] == runs[1].stats()["total_iterations"]
    for g in runs:
        g.close()


def test_pcisph_warm_start_previous_bitwise(ctx, oracle):
    from psb_project_amd._lib import WARM_PREVIOUS
    sc = scenes.pcisph_dam_break(n_side=20, warm_start=WARM_PREVIOUS)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 1, chunks=5)
    for a, b in zip(gi, oi):
        assert a["iterations"] == b["iterations"]
    assert_state_equal(g.state(), o.state())


def test_pcisph_zero_tolerance_hits_cap(ctx, oracle):
    sc = scenes.pcisph_dam_break(n_side=12, tolerance=0.0, max_it=9, retry_levels=0)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 2)
    assert gi[0]["iterations"] == 9 and gi[0]["total_iterations"] == 18
    assert gi[0]["flags"] & FLAG_NOT_CONVERGED
    assert_state_equal(g.state(), o.state())


def test_wcsph_config1_full_100_steps(ctx, oracle):
    sc = scenes.wcsph_dam_break()            # 32,768 particles
    g, o, gi, oi = run_pair(ctx, oracle, sc, 100)
    assert_state_equal(g.state(), o.state())


def test_pcisph_config2_full_3_steps(ctx, oracle):
    sc = scenes.pcisph_dam_break()           # 262,144 particles
    g, o, gi, oi = run_pair(ctx, oracle, sc, 1, chunks=3)
    for a, b in zip(gi, oi):
        assert a["iterations"] == b["iterations"] and a["max_error"] == np.float32(b["max_error"])
    assert_state_equal(g.state(), o.state())


def test_wcsph_1000_steps_within_tolerance(ctx, oracle):
    """1000 steps of the 32^3 WCSPH dam break: positions, velocities and
    densities within 1e-3 relative of the FP32 oracle (bitwise expected)."""
    sc = scenes.wcsph_dam_break()
    g, o, gi, oi = run_pair(ctx, oracle, sc, 1000)
    gs, os_ = g.state(), o.state()
    for k in ("pos", "vel", "density"):
        a = gs[k][:, :3] if gs[k].ndim == 2 else gs[k]
        b = os_[k][:, :3] if os_[k].ndim == 2 else os_[k]
        scale = np.abs(b).max()
        assert np.abs(a.astype(np.float64) - b).max() <= 1e-3 * scale, k
    assert_state_equal(gs, os_)


def test_reference_layout_roundtrip(ctx, oracle):
    """psb_system_get_particles returns FP64 triples in the caller's order."""
    sc = scenes.pcisph_dam_break(n_side=10, mode=MODE_REFERENCE)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 3)
    f = g.state_f64()
    os_ = o.state()
    order = np.argsort(os_["ids"])
    assert np.array_equal(f["new_pos"], os_["pos"][order, :3].astype(np.float64))
    assert np.array_equal(f["old_pos"], os_["old_pos"][order, :3].astype(np.float64))
    assert np.array_equal(f["vel_inter_ap"], os_["vh"][order, :3].astype(np.float64))
    assert np.array_equal(f["vel_inter_av"], os_["vav"][order, :3].astype(np.float64))
    assert np.array_equal(f["new_vel"], os_["vel"][order, :3].astype(np.float64))
    assert np.array_equal(f["old_vel"], os_["old_vel"][order, :3].astype(np.float64))


def test_pci_sph_update_wrapper_chain(ctx, oracle):
    """pci_SPH_update + collide + interpolateVelocities on caller arrays
    (PciSphSystem.cpp:187-197) equals one fused system step."""
    for mode in (MODE_REFERENCE, MODE_CONFIG):
        sc = scenes.pcisph_dam_break(n_side=16, mode=mode)
        p = sc.params
        n = sc.n
        r = gpu_grid_and_neighbors(ctx, sc.pos4, p.grid, p.support_radius, p.max_neighbors)
        xs = r["xs"]
        z = np.zeros((n, 4), np.float32)
        d = {k: ctx.to_device(v) for k, v in dict(old=z, pos=xs, vel=z, vav=z, vap=z).items()}
        rho, prs = ctx.array(4 * n), ctx.array(4 * n)
        prs.zero()
        its, err = C.c_int32(), C.c_float()
        check(lib().psb_pci_sph_update(ctx.handle, C.byref(p), r["nb"], d["old"].ptr, d["pos"].ptr,
                                       d["vel"].ptr, d["vav"].ptr, d["vap"].ptr, rho.ptr, prs.ptr,
                                       None, None, None, None, None, n, C.byref(its), C.byref(err)))
        bmin = (C.c_float * 3)(*p.box_min); bmax = (C.c_float * 3)(*p.box_max)
        e = 0.0 if mode == MODE_REFERENCE else p.elast
        check(lib().psb_collision_box(ctx.handle, d["pos"].ptr, None, d["vap"].ptr,
                                      p.particle_radius, p.dt, n, bmin, bmax, 1, e))
        vold = ctx.array(16 * n)
        check(lib().psb_interpolate_velocities(ctx.handle, d["vav"].ptr, d["vap"].ptr, vold.ptr,
                                               d["vel"].ptr, n))
        o = oracle.OracleSystem(p, n)
        o.set_particles(sc.pos4)
        o.update(1)
        os_ = o.state()
        assert its.value == o.stats()["iterations"]
        assert np.array_equal(d["pos"].download(np.float32, (n, 4))[:, :3], os_["pos"][:, :3])
        assert np.array_equal(d["vap"].download(np.float32, (n, 4))[:, :3], os_["vh"][:, :3])
        assert np.array_equal(d["vel"].download(np.float32, (n, 4))[:, :3], os_["vel"][:, :3])
        assert np.array_equal(rho.download(np.float32, n), os_["density"])
        assert np.array_equal(prs.download(np.float32, n), os_["pressure"])


# ------------------------------------------------------------------ edges
def test_empty_and_single_particle(ctx, oracle):
    from psb_project_amd.system import FluidSystem
    sc = scenes.wcsph_dam_break(n_side=4)
    g = FluidSystem(ctx, sc.params, 16)
    g.set_particles(np.zeros((0, 4), np.float32))
    g.update(3)
    assert g.n == 0
    one = sc.pos4[:1].copy()
    g.set_particles(one)
    o = oracle.OracleSystem(sc.params, 16)
    o.set_particles(one)
    g.update(5)
    o.update(5)
    assert_state_equal(g.state(), o.state())


def test_out_of_grid_and_capacity_errors(ctx, oracle):
    from psb_project_amd.system import FluidSystem
    sc = scenes.wcsph_dam_break(n_side=8)
    g = FluidSystem(ctx, sc.params, sc.n)
    x = sc.pos4.copy()
    x[3, 1] = 5.0          # far outside the grid
    g.set_particles(x)
    g.update(1)
    assert g.stats()["flags"] & FLAG_OUT_OF_GRID
    with pytest.raises(PsbError) as e:
        g.set_particles(np.zeros((sc.n + 1, 4), np.float32))
    assert e.value.status == 3
    with pytest.raises(PsbError):
        g.update(-1)


@pytest.mark.parametrize("offset,nbytes", [(0, 16), (0, 1 << 20), (16, 4096 + 7), (3, 1000),
                                           (0, (64 << 20) + 12)])
def test_copy_device_to_device(ctx, offset, nbytes):
    """copyArrayDeviceToDevice: 16-byte streaming kernel plus unaligned tail."""
    rng = np.random.default_rng(nbytes)
    total = offset + nbytes
    src = rng.integers(0, 256, total, dtype=np.uint8)
    ds = ctx.to_device(src)
    dd = ctx.to_device(np.zeros(total, np.uint8))
    check(lib().psb_copy_array_device_to_device(ctx.handle, dd.ptr, ds.ptr, offset, nbytes))
    got = dd.download(np.uint8, total)
    assert np.array_equal(got[offset:], src[offset:])
    assert not got[:offset].any()


@pytest.mark.parametrize("mode", [MODE_REFERENCE, MODE_CONFIG])
def test_interaction_system_bitwise(ctx, oracle, mode):
    """interactionSystem: DEM contact forces over grid neighbour lists,
    accumulated onto existing forces, bitwise against the oracle."""
    sc = scenes.wcsph_dam_break(n_side=16)
    p = sc.params
    n = sc.n
    r = gpu_grid_and_neighbors(ctx, sc.pos4, p.grid, p.support_radius, p.max_neighbors)
    rng = np.random.default_rng(12)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(scale=0.5, size=(n, 3))
    f0 = np.zeros((n, 4), np.float32)
    f0[:, :3] = rng.normal(size=(n, 3))
    rad = rng.uniform(0.003, 0.008, n).astype(np.float32)
    k = rng.uniform(50, 150, n).astype(np.float32)
    c = rng.uniform(0.1, 1.0, n).astype(np.float32)
    d = {name: ctx.to_device(a) for name, a in dict(v=vel, f=f0, r=rad, k=k, c=c).items()}
    dsorted = r["bufs"][1]
    check(lib().psb_interaction_system(ctx.handle, mode, dsorted.ptr, d["v"].ptr, d["f"].ptr,
                                       d["r"].ptr, d["k"].ptr, d["c"].ptr, None, None, r["nb"], n))
    got = d["f"].download(np.float32, (n, 4))[:, :3]
    want = oracle.interaction_system(mode, r["xs"], vel, f0, rad, k, c, r["count"],
                                     r["list"])[:, :3]
    assert np.array_equal(got, want)
    assert np.abs(got - f0[:, :3]).max() > 0


def test_dem_spheres_bitwise(ctx, oracle):
    """Double dam break with rigid DEM spheres (boundary-particle coupling,
    lattice walls), bitwise against the oracle: fluid state, iterations and
    sphere centres / velocities / fluid forces."""
    from psb_project_amd.system import FluidSystem
    sc = scenes.double_dam_break_dem(block=(12, 12, 8), box=(0.24, 0.12, 0.12), num_spheres=6)
    m = sc.meta
    cap = sc.n + 93 * 6
    g = FluidSystem(ctx, sc.params, cap)
    o = oracle.OracleSystem(sc.params, cap)
    for s_ in (g, o):
        s_.set_spheres(m["sphere_centers"], None, m["sphere_radius"], m["sphere_density"],
                       m["spring"], m["damping"])
        s_.set_particles(sc.pos4)
    assert g.n == o.n == sc.n + 93 * 6
    for _ in range(4):
        g.update(10)
        o.update(10)
        gs, os_ = g.stats(), o.stats()
        assert gs["total_iterations"] == os_["total_iterations"]
        assert gs["max_error"] == np.float32(os_["max_error"])
    assert_state_equal(g.state(), o.state())
    a, b = g.spheres(), o.spheres()
    for k in ("centers", "velocities", "fluid_force"):
        assert np.array_equal(a[k], b[k]), k
    assert np.abs(a["velocities"]).max() > 0


RETRY_KEYS = ("iterations", "total_iterations", "flags", "retried_steps", "substeps")


@pytest.mark.parametrize("variant", ["clamp", "lattice_walls", "warm_previous", "two_lanes",
                                     "wide_list"])
def test_pcisph_retry_substeps_bitwise(ctx, oracle, variant):
    """A tight tolerance (0.4 %) with a 3-iteration cap makes most steps fail
    at dt; each is redone from its start state as two steps of dt/2, and
    those as two of dt/4 where they fail again (params.retry_levels = 2).
    State, iteration counts, errors, flags and the retry counts equal the
    oracle's bit for bit."""
    from psb_project_amd._lib import WARM_PREVIOUS
    kw = dict(n_side=16, max_it=3, tolerance=0.004, retry_levels=2)
    if variant == "lattice_walls":
        kw["lattice_walls"] = True
    if variant == "warm_previous":
        kw["warm_start"] = WARM_PREVIOUS
    sc = scenes.pcisph_dam_break(**kw)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 20, chunks=4, lanes=2 if variant == "two_lanes" else 0,
                            compact=variant != "wide_list")
    for a, b in zip(gi, oi):
        for k in RETRY_KEYS:
            assert a[k] == b[k], (k, a[k], b[k])
        assert a["max_error"] == np.float32(b["max_error"])
        assert a["worst_error"] == np.float32(b["worst_error"])
    retried = sum(a["retried_steps"] for a in gi)
    sub = sum(a["substeps"] for a in gi)
    assert retried > 0 and sub >= 2 * retried
    if variant == "clamp":
        assert sub > 2 * retried                     # some halves redone at dt/4
    assert_state_equal(g.state(), o.state())


def test_pcisph_retry_with_spheres_bitwise(ctx, oracle):
    """Retries in the DEM scene: the failed attempt's sphere update is never
    run, each substep moves the spheres by its own dt."""
    from psb_project_amd.system import FluidSystem
    sc = scenes.double_dam_break_dem(block=(12, 12, 8), box=(0.24, 0.12, 0.12), num_spheres=6,
                                     max_it=3, tolerance=0.002, dt=5e-4)
    sc.params.retry_levels = 2
    m = sc.meta
    cap = sc.n + 93 * 6
    g = FluidSystem(ctx, sc.params, cap)
    o = oracle.OracleSystem(sc.params, cap)
    for s_ in (g, o):
        s_.set_spheres(m["sphere_centers"], None, m["sphere_radius"], m["sphere_density"],
                       m["spring"], m["damping"])
        s_.set_particles(sc.pos4)
    retried = 0
    for _ in range(3):
        g.update(10)
        o.update(10)
        gs, os_ = g.stats(), o.stats()
        for k in RETRY_KEYS:
            assert gs[k] == os_[k], (k, gs[k], os_[k])
        retried += gs["retried_steps"]
    assert retried > 0
    assert_state_equal(g.state(), o.state())
    a, b = g.spheres(), o.spheres()
    for k in ("centers", "velocities", "fluid_force"):
        assert np.array_equal(a[k], b[k]), k


def test_pcisph_retry_off_flags_failures(ctx, oracle):
    """retry_levels = 0: the same scene keeps its failed steps, flagged."""
    sc = scenes.pcisph_dam_break(n_side=16, max_it=3, tolerance=0.004, retry_levels=0)
    g, o, gi, oi = run_pair(ctx, oracle, sc, 20)
    assert gi[0]["flags"] & FLAG_NOT_CONVERGED and gi[0]["retried_steps"] == 0
    for k in RETRY_KEYS:
        assert gi[0][k] == oi[0][k]
    assert_state_equal(g.state(), o.state())

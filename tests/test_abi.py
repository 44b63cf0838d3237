"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, and its parameter handling matches the reference defaults and the
closed forms of the kernel constants (no GPU needed, no compute calls)."""
import ctypes as C
import math
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    text = (ROOT / "include" / "psb_sph.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(psb_[a-z0-9_]+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 40
    for must in ("psb_store_particles", "psb_search_neighbors", "psb_evaluate_densities_forces",
                 "psb_pci_sph_update", "psb_collision_box", "psb_system_update"):
        assert must in names


def test_library_exports_every_declared_symbol(psb):
    L = psb.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_every_exported_symbol_is_declared(psb):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "psb_project_amd" / "libpsb_sph.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = sorted({l.split()[-1] for l in out.splitlines() if " T psb_" in l})
    assert set(exported) <= set(declared_functions())


def test_default_params_are_reference_defaults(psb):
    from psb_project_amd import scenes
    p = scenes.default_params()
    # SphParticle.cpp:10-18, SphSystem.cpp:9,88, PciSphSystem.cu:30
    assert p.mode == psb.MODE_REFERENCE and p.solver == psb.SOLVER_WCSPH
    assert p.support_radius == pytest.approx(0.0457)
    assert p.rest_density == pytest.approx(998.29)
    assert p.gas_stiffness == 3.0 and p.viscosity == 3.5
    assert p.surface_tension == pytest.approx(0.0728) and p.threshold == pytest.approx(7.065)
    assert p.dt == pytest.approx(0.01) and p.max_neighbors == 200
    assert p.min_iterations == 3 and p.max_iterations == 3
    # createGrid(0,0,0,1,1,1,h): Min = -0.5, ceil(1/h) cells
    assert list(p.grid.min) == [-0.5, -0.5, -0.5]
    assert list(p.grid.dims) == [math.ceil(1 / 0.0457)] * 3


def test_derived_constants_closed_form(psb):
    from psb_project_amd import scenes
    sc = scenes.pcisph_dam_break(n_side=4)
    p = sc.params
    d = scenes.derive(p)
    h = float(np.float32(p.support_radius))
    assert d.c_poly6 == np.float32(315 / (64 * math.pi * h ** 9))
    assert d.c_spiky == np.float32(-45 / (math.pi * h ** 6))
    assert d.c_visc == np.float32(45 / (math.pi * h ** 6))
    assert d.c_surf_lap == np.float32(-945 / (32 * math.pi * h ** 9))
    m, dt, rho0 = float(np.float32(p.mass)), float(np.float32(p.dt)), float(np.float32(p.rest_density))
    assert d.beta == np.float32(2 * (m * dt / rho0) ** 2)
    # template delta on the cubic lattice, h = 2s: |sum grad W|^2 = 0 and
    # sum |grad W|^2 = (45/(pi h^6))^2 s^4 (6 + 12 (2-sqrt2)^4 + 8 (2-sqrt3)^4)
    s = float(np.float32(p.lattice_spacing))
    w2 = (45 / (math.pi * h ** 6)) ** 2 * s ** 4 * (6 + 12 * (2 - math.sqrt(2)) ** 4 + 8 * (2 - math.sqrt(3)) ** 4)
    assert d.delta == pytest.approx(1.0 / (2 * (m * dt / rho0) ** 2 * w2), rel=1e-6)
    assert d.stencil == 2  # h = 2 cells


def test_oracle_and_library_constants_agree(psb, oracle):
    from psb_project_amd import scenes
    for sc in (scenes.wcsph_dam_break(n_side=4), scenes.pcisph_dam_break(n_side=4)):
        d = scenes.derive(sc.params)
        o = oracle.derive(sc.params)
        lib_vals = [d.h2, d.c_poly6, d.c_spiky, d.c_visc, d.c_surf_lap, d.c_surf_grad, d.mc_poly6,
                    d.beta, d.delta, d.tait_b, d.stencil]
        assert np.array_equal(np.float32(lib_vals), np.float32(o[:11]))


def test_invalid_params_rejected(psb):
    from psb_project_amd import scenes
    from psb_project_amd._lib import Derived
    p = scenes.default_params()
    p.mass = -1.0
    d = Derived()
    st = psb.lib().psb_derive(C.byref(p), C.byref(d))
    assert st == 1
    assert b"mass" in psb.lib().psb_last_error()
    p = scenes.default_params()
    p.max_iterations = 1000
    assert psb.lib().psb_derive(C.byref(p), C.byref(d)) == 1


def test_compute_grid_size(psb):
    nb, nt = C.c_uint32(), C.c_uint32()
    assert psb.lib().psb_compute_grid_size(1000, 256, C.byref(nb), C.byref(nt)) == 0
    assert (nb.value, nt.value) == (4, 256)
    assert psb.lib().psb_compute_grid_size(100, 256, C.byref(nb), C.byref(nt)) == 0
    assert (nb.value, nt.value) == (1, 100)


def test_init_without_device_reports_error(psb):
    from psb_project_amd.device import device_count
    if device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    st = psb.lib().psb_init(0, C.byref(h))
    assert st != 0 and not h.value

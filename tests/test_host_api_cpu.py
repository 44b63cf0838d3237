"""The C++ host API's Fluid_Material loader and saver (psb_host.hpp), run
through the example program (no GPU is touched by --print-material)."""
import json
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "psb_project_amd" / "host" / "example_dam_break"
GOLD = ROOT / "tests" / "golden" / "materials"

KEYS = {"visualizationRadius": "particle_radius", "particle_mass": "particle_mass",
        "restDensity": "rest_density", "viscosity": "viscosity",
        "surfaceTension": "surface_tension", "threshold": "threshold",
        "gasStiffness": "gas_stiffness", "kernelParticles": "kernel_particles",
        "supportRadius": "support_radius"}


@pytest.fixture(scope="module")
def exe():
    if not EXE.exists():
        from psb_project_amd.build import build_host_example
        build_host_example()
    return EXE


def material(exe, path):
    r = subprocess.run([str(exe), "--material", str(path), "--print-material"],
                       capture_output=True, text=True, check=True)
    return json.loads(r.stdout)


@pytest.mark.parametrize("name", ["Water.xml", "Mucus.xml", "Smoke.xml"])
def test_fluid_material_files_load(exe, name):
    text = (GOLD / name).read_text()
    attrs = dict(re.findall(r'(\w+)="([^"]*)"', text))
    got = material(exe, GOLD / name)
    for xml_key, key in KEYS.items():
        if xml_key in attrs:
            assert got[key] == pytest.approx(float(attrs[xml_key]), rel=1e-7), xml_key


def test_fluid_material_save_load_roundtrip(exe, tmp_path):
    out = tmp_path / "copy.xml"
    subprocess.run([str(exe), "--material", str(GOLD / "Mucus.xml"), "--save-material", str(out)],
                   check=True)
    assert material(exe, out) == material(exe, GOLD / "Mucus.xml")


def test_missing_material_file_fails(exe, tmp_path):
    r = subprocess.run([str(exe), "--material", str(tmp_path / "none.xml"), "--print-material"],
                       capture_output=True, text=True)
    assert r.returncode != 0


PARTS = ROOT / "tests" / "golden" / "particles"


def lines(exe, *args):
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, check=True)
    return [json.loads(x) for x in r.stdout.splitlines()]


def f32(x):
    """FP32 value(s) of a number or list (the program prints 9 digits)."""
    import numpy as np
    if isinstance(x, list):
        return [float(np.float32(v)) for v in x]
    return float(np.float32(x))


def as32(d):
    return {k: (v if isinstance(v, int) else f32(v)) for k, v in d.items()}


def test_particle_xml_reader(exe):
    """ParticleLoader_XML::load on a file in the reference's layout: file
    order, decimal notations (%f, %g, exponents), comments, per-particle
    material attributes."""
    out = lines(exe, "--print-xml", str(PARTS / "sph_system_3.xml"))
    assert out[0] == {"count": 3, "dt": 0.005}
    p0, p1, p2 = [as32(p) for p in out[1:]]
    assert [p["index"] for p in (p0, p1, p2)] == [0, 1, 2]
    assert p0["pos"] == [f32(0.1), f32(0.2), f32(0.3)]
    assert p0["vel"] == [1.5, -2.25, 0.0]
    assert p0["vav"] == [f32(1e-5), 0.0, -0.03125]
    assert p0["vap"] == [0.5, 0.25, 0.125]
    assert p0["mass"] == f32(0.02) and p0["support_radius"] == f32(0.0457)
    assert p0["rest_density"] == f32(998.289978) and p0["viscosity"] == 3.5
    assert p1["pos"][0] == f32(0.1457)
    assert p2["pos"] == [-1.25, 2.5, 3.0] and p2["mass"] == f32(0.03)
    assert p2["rest_density"] == 1000.0 and p2["viscosity"] == f32(0.01)


def test_particle_txt_reader(exe):
    out = lines(exe, "--print-txt", str(PARTS / "sph_system_2.txt"))
    assert out[0] == {"count": 2}
    assert as32(out[1]) == {"pos": [f32(0.1), f32(0.2), f32(0.3)], "mass": f32(0.02),
                            "density": f32(998.29)}
    assert as32(out[2]) == {"pos": [-1.5, f32(2.5e-3), 3.0], "mass": f32(0.02),
                            "density": 1000.0}


def test_particle_files_reject_other_roots(exe, tmp_path):
    bad = tmp_path / "simple.xml"
    bad.write_text('<?xml version="1.0" ?>\n<Particles name="Simple System" dt="0.01" count="0">'
                   "\n</Particles>\n")
    r = subprocess.run([str(exe), "--print-xml", str(bad)], capture_output=True, text=True)
    assert r.returncode == 1
    t = tmp_path / "x.txt"
    t.write_text("3\nparticle -p 0 0 0 -v 0 0 0 -m 1\n")
    r = subprocess.run([str(exe), "--print-txt", str(t)], capture_output=True, text=True)
    assert r.returncode == 1


def test_checkpoint_info_rejects_other_files(tmp_path):
    """psb_checkpoint_info reads the header only (no GPU): a foreign or
    truncated file is an I/O error."""
    from psb_project_amd._lib import PsbError
    from psb_project_amd.system import FluidSystem
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"PSBCKPT0" + bytes(200))
    with pytest.raises(PsbError, match="not a checkpoint"):
        FluidSystem.checkpoint_info(bad)
    short = tmp_path / "short.ckpt"
    short.write_bytes(b"PSBCKPT1")
    with pytest.raises(PsbError):
        FluidSystem.checkpoint_info(short)
    with pytest.raises(PsbError, match="cannot read"):
        FluidSystem.checkpoint_info(tmp_path / "missing.ckpt")


def test_mixed_materials_are_grouped(exe):
    """SPHSystem::init groups the particles by material (the reference packs
    one material per particle): two different files give two materials, the
    same file twice gives one."""
    r = subprocess.run([str(exe), "--check-materials", str(GOLD / "Water.xml"),
                        str(GOLD / "Mucus.xml")], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("materials 2 "), r.stdout
    r = subprocess.run([str(exe), "--check-materials", str(GOLD / "Water.xml"),
                        str(GOLD / "Water.xml")], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("materials 1 "), r.stdout

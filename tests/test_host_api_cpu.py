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

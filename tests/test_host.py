"""Host-side checks (no GPU): readers, units, builders, and that the C-ABI
libraries load and export every function their headers declare."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from moldy_amd import sysspec, system, units

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+[\s\*]+(\w+)\s*\(", txt, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("moldy_amd.h", "libmoldy_amd.so"),
                                        ("moldy_force.h", "libmoldyforce.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "moldy_amd", lib)
    if not os.path.exists(path):
        pytest.skip(f"{lib} not built")
    L = C.CDLL(path, mode=C.RTLD_GLOBAL)
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_declared_names():
    assert {"mgpu_create", "mgpu_force_calc", "mgpu_ewald", "mgpu_eval_forces",
            "mgpu_pairlist", "mgpu_cells"} <= set(_declared("moldy_amd.h"))
    assert {"force_calc", "ewald", "poteval"} <= set(_declared("moldy_force.h"))


def test_units_match_moldy_program_units():
    # 1 kJ/mol in default input units (time unit 0.1 ps) is 100 program units
    assert abs(units.unit_scale((1, 2, -2, 0), units.DEFAULT_INPUT_UNITS,
                                units.PROGRAM_UNITS) - 100.0) < 1e-9
    assert abs(units.CONV_E - 0.01) < 1e-9
    # one electron charge is 372.7 program units (the clay example's unit)
    assert abs(units.charge_to_program(1.0) - 372.74) < 0.01
    assert abs(units.density_to_program(1.0) - 0.60221367) < 1e-8


def test_sysspec_reader_tips2():
    txt = """# test
Water 64
1 0 0 0 16 0 O
2 0.7569503 0 -0.5858822 1 0.535 H
2 -0.7569503 0 -0.5858822
3 0 0 -0.15 0 -1.07 M
end
lennard-jones
1 1 0.51799 3.2407
end
"""
    sp = sysspec.parse_sysspec(txt)
    assert sp["ptype"] == 0 and len(sp["species"]) == 1
    assert sp["species"][0]["ids"] == [1, 2, 2, 3]
    assert sp["sites"][2]["charge"] == 0.535 and sp["sites"][3]["mass"] == 0.0
    ctl = sysspec.parse_control("density=1\ncutoff=6.25\nalpha=0.45\nend\n")
    s = sysspec.build_system(ctl, sp)
    assert s.nsites == 256 and s.max_id == 4
    assert abs(s.potpar[1, 1, 0] - 51.799) < 1e-9     # kJ/mol -> program units
    assert s.species[0].rdof == 3


def test_builders_shapes_and_neutrality():
    s = system.salt_water(nwater=200, nmg=4, ncl=8, seed=1)
    assert s.nsites == 812 and s.nmols == 212 and s.nmols_r == 200
    assert abs(s.site_charges().sum()) < 1e-9
    assert np.all(np.abs(s.c_of_m) <= 0.5)
    assert np.allclose(np.linalg.norm(s.quat, axis=1), 1.0, atol=1e-12)
    w = system.water_box(4096)
    assert abs(w.h[0, 0] - 49.719) < 1e-3
    big = system.salt_water()
    assert big.nsites == 132872 and abs(big.h[0, 0] - 102.44) < 0.01


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr

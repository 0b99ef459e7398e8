"""CPU checks of the product boundary (no GPU compute).

* libthp.so loads and exports every entry point declared in include/thp.h;
* the seeded system generator (host code of libthp) produces the water
  recipe: geometry, charges, moments, polarizabilities, exclusions;
* the oracle reproduces the committed golden fixtures;
* the Fortran host binding compiles and links against libthp.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import thp
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "thp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(thp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    lib = thp.load_lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", thp.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (thp_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_version_string():
    assert b"gfx950" in thp.load_lib().thp_version()


def test_pick_grid_follows_reference_table():
    lib = thp.load_lib()
    assert lib.thp_pick_grid(62.23) == 64
    assert lib.thp_pick_grid(18.6) == 20
    assert lib.thp_pick_grid(100.0) == 100
    assert lib.thp_pick_grid(220.0) == 240


def test_water_recipe_geometry_and_moments():
    s = thp.water_1944()
    assert s.n == 1944 and len(s.scaled_ik) == 1944
    o, h1, h2 = (np.stack([s.x[k::3], s.y[k::3], s.z[k::3]], 1) for k in range(3))
    b1, b2 = np.linalg.norm(h1 - o, axis=1), np.linalg.norm(h2 - o, axis=1)
    assert np.allclose(b1, 0.9572, atol=1e-12) and np.allclose(b2, 0.9572, atol=1e-12)
    ang = np.degrees(np.arccos(np.einsum("ij,ij->i", h1 - o, h2 - o) / (b1 * b2)))
    assert np.allclose(ang, 104.52, atol=1e-9)
    assert np.allclose(s.rpole[0::3, 0], -0.51966) and np.allclose(s.rpole[1::3, 0], 0.25983)
    assert abs(s.rpole[:, 0].sum()) < 1e-12
    # dipoles: 0.14 e.A on O along the bisector, 0.02 e.A on H along O->H
    bis = (h1 + h2) / 2 - o
    bis /= np.linalg.norm(bis, axis=1)[:, None]
    assert np.allclose(s.rpole[0::3, 1:4], 0.14 * bis, atol=1e-12)
    assert np.allclose(s.rpole[1::3, 1:4], 0.02 * (h1 - o) / 0.9572, atol=1e-12)
    q = s.rpole[:, 4:13].reshape(-1, 3, 3)
    assert np.allclose(q, np.transpose(q, (0, 2, 1))) and np.allclose(np.trace(q, axis1=1, axis2=2), 0)
    assert np.allclose(s.polarity[0::3], 0.837) and np.allclose(s.polarity[1::3], 0.496)
    assert np.allclose(s.thole, 0.39)
    # intramolecular pairs: d = p = 0, u = 1
    assert np.all(s.scaled_ik[:, 0] // 3 == s.scaled_ik[:, 1] // 3)
    assert np.allclose(s.scaled_dpu, [0.0, 0.0, 1.0])


def test_solute_system_is_neutral_and_inside_blob():
    s = thp.make_system(3000, 40.0, n_solute=1000, blob_radius=11.0, seed=3)
    sol = slice(3 * ((3000 - 1000) // 3), 3000)
    assert abs(s.rpole[sol, 0].sum()) < 1e-10
    r = np.sqrt((s.x[sol] - 20) ** 2 + (s.y[sol] - 20) ** 2 + (s.z[sol] - 20) ** 2)
    assert r.max() < 11.0 + 0.5
    assert np.all((s.polarity[sol] >= 1.0) & (s.polarity[sol] < 1.75))


def test_generator_is_deterministic():
    a, b = thp.water_1944(seed=7), thp.water_1944(seed=7)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.rpole, b.rpole)
    c = thp.water_1944(seed=8)
    assert not np.array_equal(a.x, c.x)


def test_oracle_reproduces_golden_water81():
    g = np.load(os.path.join(GOLD, "water81.npz"))
    s = thp.tiny_water(27, 20.3, 24)
    assert np.array_equal(s.x, g["x"]) and np.array_equal(s.rpole, g["rpole"])
    orc = O.Oracle(s)
    assert np.array_equal(orc.nblist_pairs(), g["pairs"])
    np.testing.assert_allclose(orc.efld0_direct(), g["ef_direct"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(orc.efld0_recip(), g["cphi"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(orc.tmatxb_pme(g["mu_probe"]), g["tmatxb"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(orc.tmatxbrecip(g["mu_probe"]), g["tmatxbrecip"], rtol=0, atol=1e-13)
    res = orc.induce(poleps=1e-8)
    assert res["niter"] == int(g["niter"])
    assert abs(res["ep"] - float(g["ep"])) < 1e-12 * abs(float(g["ep"]))


def test_oracle_reproduces_golden_water1944_summary():
    g = np.load(os.path.join(GOLD, "water1944_summary.npz"))
    s = thp.water_1944()
    assert np.allclose([s.x.sum(), s.y.sum(), s.z.sum()], g["xyz_sum"], rtol=1e-14)
    res = O.Oracle(s).induce(poleps=1e-8)
    assert res["niter"] == int(g["niter"])
    assert abs(res["ep"] - float(g["ep"])) < 1e-11 * abs(float(g["ep"]))


@pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"),
                    reason="amdflang not installed")
def test_fortran_host_builds_and_links(tmp_path):
    fdir = os.path.join(ROOT, "tinker-hp_amd", "fortran")
    exe = os.path.join(fdir, "polar_driver")
    assert os.path.exists(exe), "Fortran host not built (__graft_entry__.build())"
    out = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
    for name in ("thp_induce", "thp_tmatxb", "thp_efld0_direct", "thp_nblist"):
        assert name in out

"""Tinker input side (no GPU): .xyz/.key/.prm reading, kmpole's frame
assignment and unit conversion, polarization groups and pair scales, and
ewaldcof, on an AMOEBA-03 water box and on the dhfr2 example."""
import math
import os

import numpy as np
import pytest

import thp
import thp.tinker as T
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def write_water_box(tmp_path, s, grid=24):
    """the 27-water box as a Tinker .xyz (types 1 = O, 2 = H) and .key"""
    lines = [f"{s.n}  water box"]
    for m in range(s.n // 3):
        o, h1, h2 = 3 * m + 1, 3 * m + 2, 3 * m + 3
        for idx, name, typ, bonds in ((o, "O", 1, (h1, h2)), (h1, "H", 2, (o,)), (h2, "H", 2, (o,))):
            i = idx - 1
            lines.append(f"{idx:6d}  {name:2s} {s.x[i]:14.8f} {s.y[i]:14.8f} {s.z[i]:14.8f} {typ:5d} "
                         + " ".join(f"{b:5d}" for b in bonds))
    xyz = tmp_path / "water.xyz"
    xyz.write_text("\n".join(lines) + "\n")
    key = tmp_path / "water.key"
    key.write_text(f"parameters {os.path.join(GOLD, 'water03.prm')}\na-axis {s.box[0]}\n"
                   f"ewald\newald-cutoff {s.cut}\newald-alpha {s.aewald}\npme-grid {grid} {grid} {grid}\n")
    return str(xyz), str(key)


def test_water_frames_groups_and_scales(tmp_path):
    base = thp.tiny_water(27, 20.3, 24)
    s = T.load(*write_water_box(tmp_path, base))
    fr = s.frames
    n = s.n
    o, h1, h2 = np.arange(0, n, 3), np.arange(1, n, 3), np.arange(2, n, 3)
    # oxygen: bisector of its two hydrogens; hydrogen: z on the oxygen, x on the other hydrogen
    assert np.all(fr["axistype"][o] == thp.THP_AXIS_BISECTOR)
    assert np.array_equal(fr["zaxis"][o], h1) and np.array_equal(fr["xaxis"][o], h2)
    assert np.all(fr["axistype"][h1] == thp.THP_AXIS_Z_THEN_X) and np.all(fr["axistype"][h2] == thp.THP_AXIS_Z_THEN_X)
    assert np.array_equal(fr["zaxis"][h1], o) and np.array_equal(fr["xaxis"][h1], h2)
    assert np.array_equal(fr["xaxis"][h2], h1) and np.all(fr["yaxis"] == -1)
    # kmpole's conversion: dipoles times bohr, quadrupoles times bohr^2 / 3
    assert abs(fr["pole"][0, 3] - 0.14279 * T.BOHR) < 1e-15
    assert abs(fr["pole"][0, 4] - 0.37928 * T.BOHR ** 2 / 3) < 1e-15
    assert np.allclose(fr["pole"][o, 0], -0.51966) and np.allclose(fr["pole"][h1, 0], 0.25983)
    assert np.allclose(s.polarity[o], 0.837) and np.allclose(s.polarity[h1], 0.496) and np.allclose(s.thole, 0.39)
    # one polarization group per molecule; its three pairs have d = p = 0, u = 1
    assert np.array_equal(s.extra["groups"], np.repeat(np.arange(n // 3), 3))
    assert len(s.scaled_ik) == n
    assert np.all(s.scaled_ik[:, 0] // 3 == s.scaled_ik[:, 1] // 3)
    assert np.all(s.scaled_dpu == [0.0, 0.0, 1.0])
    # in the global frame the oxygen dipole lies along the H-O-H bisector
    rp = O.rotpole(s.x, s.y, s.z, fr["axistype"], fr["zaxis"], fr["xaxis"], fr["yaxis"], fr["pole"])
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    b = (xyz[h1] - xyz[o]) / np.linalg.norm(xyz[h1] - xyz[o], axis=1)[:, None] + \
        (xyz[h2] - xyz[o]) / np.linalg.norm(xyz[h2] - xyz[o], axis=1)[:, None]
    b /= np.linalg.norm(b, axis=1)[:, None]
    np.testing.assert_allclose(rp[o, 1:4], 0.14279 * T.BOHR * b, atol=1e-14)


def test_ewaldcof_meets_its_tolerance():
    for cut in (5.0, 7.0, 9.0):
        a = T.ewaldcof(cut)
        assert abs(math.erfc(a * cut) / cut - 1e-8) < 1e-16
    assert abs(T.ewaldcof(7.0) - 0.5446) < 1e-4


def test_dhfr2_every_atom_gets_parameters():
    s = T.load(os.path.join(GOLD, "dhfr2.xyz.gz"), os.path.join(GOLD, "dhfr2.key"))
    fr = s.frames
    assert s.n == 23558 and tuple(s.nfft) == (64, 64, 64) and abs(s.box[0] - 62.23) < 1e-12
    # the reference's run of this system logs its PME set-up as
    # "Polarization 0.5446 64 64 64 5" (example/dhfr.log:50)
    assert round(s.aewald, 4) == 0.5446 and s.order == 5
    # a whole system: integer net charge, every atom polarizable, frames for all
    assert abs(fr["pole"][:, 0].sum() - round(fr["pole"][:, 0].sum())) < 1e-9
    assert np.all(s.polarity > 0) and np.all(fr["axistype"] > 0)
    # 7,023 waters, each a group of three with a bisector oxygen
    grp = s.extra["groups"]
    sizes = np.bincount(grp)
    assert (sizes == 3).sum() >= 7023 and np.sum(fr["axistype"] == thp.THP_AXIS_BISECTOR) >= 7023
    # chiral sites exist (protein) and every frame atom is bonded within two bonds
    chiral = (fr["axistype"] == thp.THP_AXIS_Z_THEN_X) & (fr["yaxis"] >= 0)
    assert chiral.sum() > 50
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    for k in ("zaxis", "xaxis"):
        has = fr[k] >= 0
        d = np.linalg.norm(xyz[fr[k][has]] - xyz[has], axis=1)
        assert d.max() < 3.0

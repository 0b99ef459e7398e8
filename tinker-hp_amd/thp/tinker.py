"""Tinker input files for the polarization path.

Reads what the induced-dipole solve needs from a Tinker run's inputs and
returns a `thp.System` whose multipoles are given in their local frames
(the device rotates them each step):

- coordinates, atom types and connectivity from `.xyz` (readxyz.f);
- keywords from `.key` (getkey.f): the parameter file, box, Ewald cutoff,
  coefficient and grid, list buffer;
- from `.prm` (readprm.f): `multipole` (kmpole.f:170-255), `polarize`
  (kpolar.f) and the polarization scale keywords (initprm.f:400-474);
- the frame assignment of kmpole (kmpole.f:260-440) and its conversion of
  dipoles and quadrupoles to Angstroms (kmpole.f:529-539);
- the polarization groups of polargrp (kpolar.f:598-...) and the pair
  scales that efld0_direct and tmatxb_pme apply (newinduce_pme.f:1035-1137):
  p from the atom connectivity 1-2 .. 1-5 (the intra-group value inside a
  group), d and u from the group connectivity 1-1 .. 1-4;
- the Ewald coefficient of ewaldcof (kewald.f:575-615) when none is given.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

BOHR = 0.52917721092  # MOD_units.f:77
AXIS = {"None": 0, "Z-Only": 1, "Z-then-X": 2, "Bisector": 3, "Z-Bisect": 4, "3-Fold": 5}

# polarization scale keywords and their defaults (initprm.f:415-470)
SCALE_DEFAULTS = {
    "polar-12-scale": 0.0, "polar-13-scale": 0.0, "polar-14-scale": 1.0, "polar-15-scale": 1.0,
    "polar-12-intra": 0.0, "polar-13-intra": 0.0, "polar-14-intra": 0.5, "polar-15-intra": 1.0,
    "direct-11-scale": 0.0, "direct-12-scale": 1.0, "direct-13-scale": 1.0, "direct-14-scale": 1.0,
    "mutual-11-scale": 1.0, "mutual-12-scale": 1.0, "mutual-13-scale": 1.0, "mutual-14-scale": 1.0,
}


def _open(path):
    if path.endswith(".gz"):
        import gzip
        return gzip.open(path, "rt")
    return open(path)


@dataclass
class XyzFile:
    title: str
    names: list
    xyz: np.ndarray          # (n, 3)
    types: np.ndarray        # (n,) int
    bonds: list              # per atom, 0-based bonded atoms in file order
    box: tuple | None = None


def read_xyz(path: str) -> XyzFile:
    """Tinker .xyz: a header line (atom count, title), an optional periodic
    box line, then per atom: index, name, x, y, z, type, bonded atoms."""
    with _open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    head = lines[0].split()
    n = int(head[0])
    title = " ".join(head[1:])
    body = lines[1:]
    box = None
    first = body[0].split()
    try:
        int(first[0])
    except ValueError:
        vals = [float(v) for v in first[:6]]
        box = tuple(vals)
        body = body[1:]
    names, xyz, types, bonds = [], np.zeros((n, 3)), np.zeros(n, dtype=np.int64), []
    for i in range(n):
        t = body[i].split()
        if int(t[0]) != i + 1:
            raise ValueError(f"{path}: atom {i + 1} is numbered {t[0]}")
        names.append(t[1])
        xyz[i] = [float(t[2]), float(t[3]), float(t[4])]
        types[i] = int(t[5])
        bonds.append([int(v) - 1 for v in t[6:]])
    return XyzFile(title, names, xyz, types, bonds, box)


def read_key(path: str) -> dict:
    """keyword (lower case) -> list of value strings; the last occurrence
    wins, as in the reference's sequential keyword processing."""
    keys = {}
    with _open(path) as f:
        for ln in f:
            ln = ln.split("#")[0].strip()
            if not ln:
                continue
            t = ln.split()
            keys[t[0].lower()] = t[1:]
    return keys


@dataclass
class Params:
    multipoles: list = field(default_factory=list)   # (k, kz, kx, ky, axis type, pole[13]) in file order
    polarize: dict = field(default_factory=dict)     # type -> (alpha, thole, partner types)
    scales: dict = field(default_factory=lambda: dict(SCALE_DEFAULTS))


def read_prm(path: str, params: Params | None = None) -> Params:
    """The parameter records of the path.  A multipole record is one line
    (type, up to three frame types, charge) and four continuation lines
    (dipole; Qxx; Qyx Qyy; Qzx Qzy Qzz); the frame-type signs give the axis
    type as in kmpole.f:210-215."""
    p = params or Params()
    with _open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        t = lines[i].split("#")[0].split()
        if not t:
            i += 1
            continue
        kw = t[0].lower()
        if kw == "multipole":
            ints = [int(v) for v in t[1:-1]]
            q = float(t[-1])
            k, kz, kx, ky = (ints + [0, 0, 0, 0])[:4]
            axt = "Z-then-X"
            if kz == 0:
                axt = "None"
            if kz != 0 and kx == 0:
                axt = "Z-Only"
            if kz < 0 or kx < 0:
                axt = "Bisector"
            if kx < 0 and ky < 0:
                axt = "Z-Bisect"
            if max(kz, kx, ky) < 0:
                axt = "3-Fold"
            d = [float(v) for v in lines[i + 1].split()[:3]]
            qxx = float(lines[i + 2].split()[0])
            qyx, qyy = (float(v) for v in lines[i + 3].split()[:2])
            qzx, qzy, qzz = (float(v) for v in lines[i + 4].split()[:3])
            pole = [q, d[0], d[1], d[2], qxx, qyx, qzx, qyx, qyy, qzy, qzx, qzy, qzz]
            p.multipoles.append((k, abs(kz), abs(kx), abs(ky), axt, pole))
            i += 5
            continue
        if kw == "polarize":
            k = int(t[1])
            p.polarize[k] = (float(t[2]), float(t[3]), [int(v) for v in t[4:]])
        elif kw in SCALE_DEFAULTS:
            p.scales[kw] = float(t[1])
        i += 1
    return p


def connectivity(bonds: list):
    """1-2, 1-3, 1-4 and 1-5 lists as attach.f builds them: each list holds
    the atoms one bond further out that are not in a closer list, sorted."""
    n = len(bonds)
    n12 = [sorted(set(b)) for b in bonds]
    lists = [n12]
    for _ in range(3):
        prev = lists[-1]
        nxt = []
        for i in range(n):
            near = {i}
            for lst in lists:
                near.update(lst[i])
            s = set()
            for j in prev[i]:
                s.update(n12[j])
            nxt.append(sorted(s - near))
        lists.append(nxt)
    return lists  # [n12, n13, n14, n15]


def kmpole(types, n12, n13, params: Params):
    """Frame and multipoles of every atom (kmpole.f:260-440): the first
    parameter, in file order, whose frame types match bonded neighbours
    (z and x bonded; then z bonded and x at 1-3 through z; then z only;
    then no frame).  Dipoles and quadrupoles are converted to Angstroms,
    quadrupoles divided by 3 (kmpole.f:529-539)."""
    n = len(types)
    axt = np.zeros(n, dtype=np.int32)
    fz, fx, fy = (np.full(n, -1, dtype=np.int32) for _ in range(3))
    pole = np.zeros((n, 13))
    found = np.zeros(n, dtype=bool)
    mps = params.multipoles

    def take(i, mp, z=-1, x=-1, y=-1):
        axt[i], fz[i], fx[i], fy[i] = AXIS[mp[4]], z, x, y
        pole[i] = mp[5]
        found[i] = True

    for i in range(n):
        it = types[i]
        cands = [mp for mp in mps if mp[0] == it]
        # z and x (and y) among the 1-2 atoms
        for mp in cands:
            if found[i]:
                break
            _, zt, xt, yt, _, _ = mp
            for ji in n12[i]:
                if found[i] or types[ji] != zt:
                    continue
                for ki in n12[i]:
                    if found[i] or types[ki] != xt or ki == ji:
                        continue
                    if yt == 0:
                        take(i, mp, ji, ki)
                        break
                    for li in n12[i]:
                        if types[li] == yt and li != ji and li != ki:
                            take(i, mp, ji, ki, li)
                            break
        # z among the 1-2 atoms, x (and y) among the 1-3 atoms bonded to z
        for mp in cands:
            if found[i]:
                break
            _, zt, xt, yt, _, _ = mp
            for ji in n12[i]:
                if found[i] or types[ji] != zt:
                    continue
                for ki in n13[i]:
                    if found[i] or types[ki] != xt or ji not in n12[ki]:
                        continue
                    if yt == 0:
                        take(i, mp, ji, ki)
                        break
                    for li in n13[i]:
                        if types[li] == yt and li != ki and ji in n12[li]:
                            take(i, mp, ji, ki, li)
                            break
        # z only
        for mp in cands:
            if found[i]:
                break
            _, zt, xt, _, _, _ = mp
            for ji in n12[i]:
                if types[ji] == zt and xt == 0:
                    take(i, mp, ji)
                    break
        # no frame
        for mp in cands:
            if found[i]:
                break
            if mp[1] == 0:
                take(i, mp)
    pole[:, 1:4] *= BOHR
    pole[:, 4:13] *= BOHR ** 2 / 3.0
    return axt, fz, fx, fy, pole, found


def polargrp(types, n12, polarize: dict):
    """Polarization group of every atom (kpolar.f:598-...): an atom and a
    bonded atom whose type is among its polarize partners share a group;
    groups are the connected sets of such bonds."""
    n = len(types)
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(n):
        partners = polarize.get(int(types[i]), (0, 0, []))[2]
        for j in n12[i]:
            if int(types[j]) in partners:
                ra, rb = find(i), find(j)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    roots = {}
    grp = np.array([roots.setdefault(find(i), len(roots)) for i in range(n)], dtype=np.int64)
    return grp


def pair_scales(n12, n13, n14, n15, grp, scales: dict):
    """Scaled pairs (i < k) with their (d, p, u) scales, the pairs where any
    differs from 1 (newinduce_pme.f:1035-1137 and tmatxb_pme): p by the atom
    separation (the intra-group value when both are in one group), d and u
    by the separation of their groups (same group, bonded groups, ...)."""
    sc = scales
    n = len(grp)
    ng = int(grp.max()) + 1 if n else 0
    gbond = [set() for _ in range(ng)]
    for i in range(n):
        for j in n12[i]:
            if grp[i] != grp[j]:
                gbond[grp[i]].add(int(grp[j]))
    pval = [(sc["polar-12-scale"], sc["polar-12-intra"]), (sc["polar-13-scale"], sc["polar-13-intra"]),
            (sc["polar-14-scale"], sc["polar-14-intra"]), (sc["polar-15-scale"], sc["polar-15-intra"])]
    dval = [sc["direct-11-scale"], sc["direct-12-scale"], sc["direct-13-scale"], sc["direct-14-scale"]]
    uval = [sc["mutual-11-scale"], sc["mutual-12-scale"], sc["mutual-13-scale"], sc["mutual-14-scale"]]
    members = [[] for _ in range(ng)]
    for i in range(n):
        members[grp[i]].append(i)
    # group separation up to 3 (the reference's ip11 .. ip14)
    gdist = []
    for g in range(ng):
        dist = {g: 0}
        frontier = [g]
        for d in range(1, 4):
            nxt = []
            for h in frontier:
                for h2 in gbond[h]:
                    if h2 not in dist:
                        dist[h2] = d
                        nxt.append(h2)
            frontier = nxt
        gdist.append(dist)
    out = {}
    for i in range(n):
        for level, lst in enumerate((n12, n13, n14, n15)):
            for k in lst[i]:
                if k > i:
                    same = grp[i] == grp[k]
                    out.setdefault(k_pair(i, k), [1.0, 1.0, 1.0])[1] = pval[level][1 if same else 0]
        for g, d in gdist[grp[i]].items():
            for k in members[g]:
                if k > i:
                    e = out.setdefault(k_pair(i, k), [1.0, 1.0, 1.0])
                    e[0] = dval[d]
                    e[2] = uval[d]
    pairs = [(ik, v) for ik, v in sorted(out.items()) if v != [1.0, 1.0, 1.0]]
    ik = np.array([p[0] for p in pairs], dtype=np.int32).reshape(-1, 2)
    dpu = np.array([p[1] for p in pairs], dtype=np.float64).reshape(-1, 3)
    return ik, dpu


def k_pair(i, k):
    return (i, k) if i < k else (k, i)


def ewaldcof(cutoff: float, eps: float = 1.0e-8) -> float:
    """Ewald coefficient with erfc(a cutoff) / cutoff = eps (kewald.f:575-615)."""
    ratio, x, i = eps + 1.0, 0.5, 0
    while ratio >= eps:
        i += 1
        x *= 2.0
        ratio = math.erfc(x * cutoff) / cutoff
    lo, hi = 0.0, x
    for _ in range(i + 60):
        x = 0.5 * (lo + hi)
        if math.erfc(x * cutoff) / cutoff >= eps:
            lo = x
        else:
            hi = x
    return x


def load(xyz_path: str, key_path: str | None = None, prm_path: str | None = None, name: str = ""):
    """A thp.System with local frames from Tinker inputs.  The key file
    defaults to the .xyz name with .key; the parameter file to the key's
    `parameters` keyword (relative to the key file, '.prm' appended)."""
    import thp
    xf = read_xyz(xyz_path)
    if key_path is None:
        base = xyz_path[:-3] if xyz_path.endswith(".gz") else xyz_path
        key_path = os.path.splitext(base)[0] + ".key"
    keys = read_key(key_path)
    if prm_path is None:
        pf = keys["parameters"][0]
        if not os.path.isabs(pf):
            pf = os.path.join(os.path.dirname(os.path.abspath(key_path)), pf)
        prm_path = pf if os.path.exists(pf) else pf + ".prm"
    params = read_prm(prm_path)
    n = len(xf.types)
    n12, n13, n14, n15 = connectivity(xf.bonds)
    axt, fz, fx, fy, pole, found = kmpole(xf.types, n12, n13, params)
    if not found.all():
        raise ValueError(f"no multipole parameters for atoms {np.nonzero(~found)[0][:10] + 1} ...")
    pol = np.zeros(n)
    thole = np.zeros(n)
    for i, t in enumerate(xf.types):
        if int(t) not in params.polarize:
            raise ValueError(f"no polarize parameters for atom {i + 1} (type {t})")
        pol[i], thole[i], _ = params.polarize[int(t)]
    grp = polargrp(xf.types, n12, params.polarize)
    ik, dpu = pair_scales(n12, n13, n14, n15, grp, params.scales)
    if "a-axis" in keys:
        a = float(keys["a-axis"][0])
        box = np.array([a, float(keys.get("b-axis", [a])[0]), float(keys.get("c-axis", [a])[0])])
    elif xf.box is not None:
        box = np.array(xf.box[:3])
    else:
        raise ValueError("no periodic box: PME needs a-axis (or a box line in the .xyz)")
    cut = float(keys.get("ewald-cutoff", ["7.0"])[0])
    aewald = float(keys["ewald-alpha"][0]) if "ewald-alpha" in keys else ewaldcof(cut)
    buffer = float(keys.get("list-buffer", ["2.0"])[0])  # lbuffer (cutoffs.f:80,191)
    if "pme-grid" in keys:
        g = [int(v) for v in keys["pme-grid"]]
        nfft = tuple((g + [g[0], g[0]])[:3]) if len(g) < 3 else tuple(g[:3])
    else:
        lib = thp.load_lib()
        nfft = tuple(int(lib.thp_pick_grid(float(b))) for b in box)
    s = thp.System(n=n, box=box, x=xf.xyz[:, 0].copy(), y=xf.xyz[:, 1].copy(), z=xf.xyz[:, 2].copy(),
                   rpole=np.zeros((n, 13)), polarity=pol, thole=thole, scaled_ik=ik, scaled_dpu=dpu,
                   aewald=aewald, cut=cut, buffer=buffer, nfft=nfft, order=5,
                   name=name or os.path.basename(xyz_path))
    s.frames = dict(axistype=axt, zaxis=fz, xaxis=fx, yaxis=fy, pole=pole)
    s.extra["groups"] = grp
    s.extra["types"] = xf.types
    return s

"""Writes the Tinker-input fixtures from the reference tree (run in a
container that has it; the fixtures are data, committed with this script):

- water03.prm: the polarization records of params/water03.prm (atom,
  multipole, polarize and the polarization scale keywords);
- amoebabio09_dhfr2.prm: the same records of params/amoebabio09.prm for the
  atom types example/dhfr2.xyz uses;
- dhfr2.xyz.gz and dhfr2.key: example/dhfr2 (23,558 atoms: DHFR in 7,023
  waters), the key pointing at the fixture parameter file.
"""
import gzip
import os

REF = "/root/reference/v1.2"
HERE = os.path.dirname(os.path.abspath(__file__))
KEEP = ("polar-", "direct-", "mutual-")


def records(path, types=None):
    out = []
    lines = open(path).read().splitlines()
    i = 0
    while i < len(lines):
        t = lines[i].split()
        kw = t[0].lower() if t else ""
        if kw == "multipole":
            if types is None or int(t[1]) in types:
                out += lines[i:i + 5]
            i += 5
            continue
        if kw in ("atom", "polarize") and (types is None or int(t[1]) in types):
            out.append(lines[i])
        elif kw.startswith(KEEP) and (kw.endswith("-scale") or kw.endswith("-intra")):
            out.append(lines[i])
        i += 1
    return out


def main():
    with open(os.path.join(HERE, "water03.prm"), "w") as f:
        f.write("# polarization records of water03.prm (AMOEBA-03 water)\n")
        f.write("\n".join(records(os.path.join(REF, "params", "water03.prm"))) + "\n")
    xyz = open(os.path.join(REF, "example", "dhfr2.xyz")).read()
    types = {int(ln.split()[5]) for ln in xyz.splitlines()[1:] if len(ln.split()) > 5}
    with open(os.path.join(HERE, "amoebabio09_dhfr2.prm"), "w") as f:
        f.write("# polarization records of amoebabio09.prm for the atom types of dhfr2.xyz\n")
        f.write("\n".join(records(os.path.join(REF, "params", "amoebabio09.prm"), types)) + "\n")
    with gzip.open(os.path.join(HERE, "dhfr2.xyz.gz"), "wt") as f:
        f.write(xyz)
    key = open(os.path.join(REF, "example", "dhfr2.key")).read().splitlines()
    key = ["parameters    amoebabio09_dhfr2.prm" if ln.lower().startswith("parameters") else ln for ln in key]
    with open(os.path.join(HERE, "dhfr2.key"), "w") as f:
        f.write("\n".join(key) + "\n")


if __name__ == "__main__":
    main()

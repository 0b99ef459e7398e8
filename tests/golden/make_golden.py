"""Write the golden fixtures of the FP64 oracle (run in the build container).

water-81:   27 waters in a 20.3 A box, 24^3 grid; full outputs.
water-1944: 648 waters in an 18.643 A box, 24^3 grid; summary scalars.
The systems come from the seeded generator of libthp (thp.make_system), so
the fixtures also pin the generator's output.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tinker-hp_amd")]

import thp  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    rng = np.random.default_rng(1234)
    mu = rng.normal(size=(s.n, 2, 3)) * 0.1
    res = orc.induce(poleps=1e-8)
    np.savez_compressed(
        os.path.join(HERE, "water81.npz"),
        x=s.x, y=s.y, z=s.z, rpole=s.rpole, polarity=s.polarity, thole=s.thole,
        scaled_ik=s.scaled_ik, scaled_dpu=s.scaled_dpu, box=s.box,
        params=np.array([s.aewald, s.cut, s.buffer]), nfft=np.array(s.nfft),
        pairs=orc.nblist_pairs(), ef_direct=orc.efld0_direct(), cphi=orc.efld0_recip(),
        ef_total=res["ef"], mu_probe=mu, tmatxb=orc.tmatxb_pme(mu), tmatxbrecip=orc.tmatxbrecip(mu),
        mu=res["mu"], ep=res["ep"], niter=res["niter"])
    s = thp.water_1944()
    orc = O.Oracle(s)
    res = orc.induce(poleps=1e-8)
    np.savez_compressed(
        os.path.join(HERE, "water1944_summary.npz"),
        npairs=len(orc.nblist_pairs()), ep=res["ep"], niter=res["niter"],
        mu_rms=np.sqrt(np.mean(res["mu"] ** 2)), ef_norm=np.linalg.norm(res["ef"]),
        xyz_sum=np.array([s.x.sum(), s.y.sum(), s.z.sum()]), rpole_sum=s.rpole.sum(0))


if __name__ == "__main__":
    main()

"""Real AMOEBA inputs through the Tinker reader and the device's local
frames: an AMOEBA-03 water box and the dhfr2 example (23,558 atoms, DHFR in
water, amoebabio09 parameters) solved on the GPU against the oracle given
the same inputs rotated by its own rotpole."""
import os

import numpy as np
import pytest

import thp
import thp.tinker as T
from oracle import oracle as O
from test_tinker_io import GOLD, write_water_box

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


def _oracle(s):
    fr = s.frames
    rp = O.rotpole(s.x, s.y, s.z, fr["axistype"], fr["zaxis"], fr["xaxis"], fr["yaxis"], fr["pole"])
    return O.Oracle(s.with_multipoles(rp))


def test_water_box_from_tinker_files(ctx, tmp_path):
    s = T.load(*write_water_box(tmp_path, thp.tiny_water(27, 20.3, 24)))
    ref = _oracle(s).induce(poleps=1e-8)
    solver = thp.Solver(ctx, s, FP64)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    print(f"water: niter {st.niter}/{ref['niter']} ep {ep:.6f} vs {ref['ep']:.6f}", flush=True)
    assert st.status == thp.THP_OK and st.niter == ref["niter"]
    assert abs(ep - ref["ep"]) < 1e-11 * abs(ref["ep"]) and _rms_debye(mu, ref["mu"]) < 1e-9
    solver.close()


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_dhfr2(ctx, prec):
    s = T.load(os.path.join(GOLD, "dhfr2.xyz.gz"), os.path.join(GOLD, "dhfr2.key"))
    ref = _oracle(s).induce(poleps=1e-8)
    solver = thp.Solver(ctx, s, prec)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    rms = _rms_debye(mu, ref["mu"])
    print(f"dhfr2 prec={prec}: niter {st.niter}/{ref['niter']} ep {ep:.6f} (oracle {ref['ep']:.6f}) "
          f"erel {erel:.1e} rmsD {rms:.1e}", flush=True)
    assert st.status == thp.THP_OK
    if prec == FP64:
        assert st.niter == ref["niter"] and erel < 1e-11 and rms < 1e-9
    else:
        assert abs(st.niter - ref["niter"]) <= 1 and erel < 1e-6 and rms < 1e-5
    solver.close()

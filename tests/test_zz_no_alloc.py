"""thp_epolar1_noalloc, the gradient entry point that allocates nothing
(its forces and torques live in a file-scope device array), against
thp_epolar1 on the same solved dipoles: the dhfr2 example with local frames
of every axis type on a caller's stream, then a smaller framed water box
with a null stream, reusing the scratch the first call left behind.  The
two entry points share their kernels, and the gradient sums in a fixed
order, so the results agree bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import thp
from test_gpu_boundary import _Hip
from test_gpu_forces import _box, _device
from test_oracle_pins import _water_frames

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
D2H = 2


@pytest.fixture(scope="module")
def ctx():
    c = thp.Context(0)
    yield c
    c.close()


def _noalloc(hip, solver, stream):
    n = solver.sys.n
    dep_d, tq_d = hip.alloc(24 * n), hip.alloc(24 * n)
    rc = solver.lib.thp_epolar1_noalloc(solver.h, dep_d, tq_d, stream)
    assert rc == thp.THP_OK, solver.ctx.lib.thp_last_error(solver.ctx.h)
    dep, tq = np.zeros((n, 3)), np.zeros((n, 3))
    for host, dev in ((dep, dep_d), (tq, tq_d)):
        assert hip.h.hipMemcpyAsync(host.ctypes.data, dev, host.nbytes, D2H, stream) == 0
    assert hip.h.hipStreamSynchronize(stream) == 0
    return dep, tq


@pytest.mark.parametrize("prec", [thp.THP_PREC_FP64, thp.THP_PREC_MIXED])
def test_noalloc_gradient_equals_epolar1_on_two_calls(ctx, prec):
    import thp.tinker as tinker
    hip = _Hip()
    stream = C.c_void_p()
    assert hip.h.hipStreamCreate(C.byref(stream)) == 0
    try:
        # first call: dhfr2 (23,558 atoms, local frames) on a caller stream
        ds = tinker.load(os.path.join(GOLD, "dhfr2.xyz.gz"), os.path.join(GOLD, "dhfr2.key"))
        big = thp.Solver(ctx, ds, prec)
        _, _, st = big.newinduce_pme2(poleps=1e-8)
        assert st.status == 0
        print(f"noalloc: dhfr2 prec={prec}", flush=True)
        dep1, tq1 = _noalloc(hip, big, stream)
        ref1, rt1 = big.epolar1(with_torque=True)
        assert np.isfinite(dep1).all() and np.abs(ref1).max() > 1.0
        assert np.array_equal(dep1, ref1) and np.array_equal(tq1, rt1)
        big.close()
        # second call: a smaller framed water box (81 atoms), null stream
        s = _box(seed=5)
        small = _device(ctx, s, _water_frames(s), prec, 1e-10)
        print(f"noalloc: water-81 prec={prec}", flush=True)
        dep2, tq2 = _noalloc(hip, small, C.c_void_p())
        ref2, rt2 = small.epolar1(with_torque=True)
        assert np.array_equal(dep2, ref2) and np.array_equal(tq2, rt2)
        small.close()
    finally:
        hip.h.hipStreamDestroy(stream)
        hip.free_all()


def test_noalloc_refuses_what_it_cannot_do_without_allocating(ctx):
    hip = _Hip()
    try:
        d = hip.alloc(24 * 200000)
        # no solve of the current coordinates yet
        s = _box(seed=6)
        sv = thp.Solver(ctx, s, thp.THP_PREC_FP64)
        assert sv.lib.thp_epolar1_noalloc(sv.h, d, None, None) == -8  # THP_ERR_STATE
        sv.newinduce_pme2(poleps=1e-8)
        assert sv.lib.thp_epolar1_noalloc(sv.h, d, None, None) == thp.THP_OK
        sv.update_coords(s.x + 0.003, s.y, s.z)
        assert sv.lib.thp_epolar1_noalloc(sv.h, d, None, None) == -8  # THP_ERR_STATE
        sv.close()
        # more atoms than the scratch holds
        n = 3 * 43800
        assert n > 131072
        large = thp.Solver(ctx, thp.make_system(n, (n / 0.1) ** (1 / 3.0), seed=3), thp.THP_PREC_MIXED)
        assert large.lib.thp_epolar1_noalloc(large.h, d, None, None) == -1  # THP_ERR_ARG
        large.close()
    finally:
        hip.free_all()

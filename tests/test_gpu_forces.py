"""Polarization energy gradient on the device (thp_epolar1: epreal1c,
eprecip1, the self torque and torque.f, v1.2/source/epolar1.f:57-2543)
against the oracle's gradient, which central differences of the solved
energy pin (tests/test_oracle_pins.py)."""
import dataclasses

import numpy as np
import pytest

import thp
from oracle import oracle as O
from test_oracle_pins import _water_frames

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


@pytest.fixture(scope="module")
def ctx():
    c = thp.Context(0)
    yield c
    c.close()


def _box(nmol=27, box=10.0, nfft=12, seed=4):
    s = thp.tiny_water(nmol, box, nfft, seed=seed)
    # intramolecular p-scales of 0.5 (d-scales 0): u_d and u_p differ
    dpu = np.array(s.scaled_dpu, dtype=np.float64).reshape(-1, 3).copy()
    dpu[:, 1] = 0.5
    return dataclasses.replace(s, scaled_dpu=dpu)


def _device(ctx, s, fr, prec, poleps):
    if fr is not None:
        sf = s.with_multipoles(np.zeros((s.n, 13)))
        sf.frames = dict(axistype=fr[0], zaxis=fr[1], xaxis=fr[2], yaxis=fr[3], pole=fr[4])
    else:
        sf = s
    solver = thp.Solver(ctx, sf, prec)
    _, _, st = solver.newinduce_pme2(poleps=poleps)
    assert st.status == 0
    return solver


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (MIXED, 1e-5)])
def test_gradient_with_local_frames_matches_oracle(ctx, prec, tol):
    s = _box()
    fr = _water_frames(s)
    _, grad, _ = O.epolar1(s, fr, poleps=1e-10)
    solver = _device(ctx, s, fr, prec, 1e-10)
    dep = solver.epolar1()
    err = np.abs(dep - grad).max() / np.abs(grad).max()
    print(f"prec={prec} max|dep|={np.abs(grad).max():.4f} rel={err:.2e}", flush=True)
    assert err < tol
    solver.close()


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (MIXED, 1e-5)])
def test_gradient_with_every_axis_type_and_chiral_sites_matches_oracle(ctx, prec, tol):
    # Z-Only, Z-Bisect, 3-Fold and chiral Z-then-X sites besides the water
    # frames: every branch of the device's torque distribution against the
    # oracle's closed form, which the frame Jacobian pins
    from test_gpu_frames import _frames
    s = _box(seed=6)
    fd = _frames(s, seed=9)
    fr = (fd["axistype"], fd["zaxis"], fd["xaxis"], fd["yaxis"], fd["pole"])
    _, grad, _ = O.epolar1(s, fr, poleps=1e-10)
    solver = _device(ctx, s, fr, prec, 1e-10)
    dep = solver.epolar1()
    err = np.abs(dep - grad).max() / np.abs(grad).max()
    print(f"every axis type prec={prec} max|dep|={np.abs(grad).max():.4f} rel={err:.2e}", flush=True)
    assert err < tol
    solver.close()


def test_gradient_needs_a_solve_of_the_current_state(ctx):
    # the gradient belongs to the dipoles of the current coordinates and
    # multipoles: after a coordinate or multipole update or before any
    # solve it is refused
    s = _box(seed=8)
    solver = thp.Solver(ctx, s, FP64)
    with pytest.raises(thp.ThpError) as e:
        solver.epolar1()
    assert e.value.code == -8
    solver.newinduce_pme2(poleps=1e-8)
    solver.epolar1()
    solver.update_coords(s.x + 0.01, s.y, s.z)
    with pytest.raises(thp.ThpError) as e:
        solver.epolar1()
    assert e.value.code == -8
    solver.newinduce_pme2(poleps=1e-8)
    solver.epolar1()
    solver.update_multipoles(s.rpole * 1.01)
    with pytest.raises(thp.ThpError) as e:
        solver.epolar1()
    assert e.value.code == -8
    # after a truncated-CG solve it is that solve's gradient, whatever the
    # solver options say by the time it is asked for
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.newinduce_pme2(poleps=1e-8)
    g_tcg = solver.epolar1()
    solver.set_tcg(1, False, True, False, 0.5)
    solver.set_solver(thp.THP_POLALG_PCG)
    assert np.array_equal(solver.epolar1(), g_tcg)
    solver.close()


def test_forces_and_torques_with_global_multipoles_match_oracle(ctx):
    s = _box(seed=7)
    fr = _water_frames(s, seed=2)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    rp = O.rotpole(*xyz.T, *fr)
    s = s.with_multipoles(rp)
    mu = O.Oracle(s).induce(poleps=1e-10)["mu"]
    f1, t1 = O.epolar1_real(s, rp, mu[:, 0], mu[:, 1])
    f2, t2 = O.epolar1_recip_self(s, rp, mu[:, 0], mu[:, 1])
    solver = _device(ctx, s, None, FP64, 1e-10)
    dep, tq = solver.epolar1(with_torque=True)
    f, t = f1 + f2, t1 + t2
    assert np.abs(-dep - f).max() < 1e-9 * np.abs(f).max()
    assert np.abs(tq - t).max() < 1e-9 * np.abs(t).max()
    solver.close()


def test_gradient_on_the_1944_atom_box_matches_oracle(ctx):
    s = thp.water_1944()
    fr = _water_frames(s, seed=5)
    _, grad, _ = O.epolar1(s, fr, poleps=1e-10)
    solver = _device(ctx, s, fr, FP64, 1e-10)
    dep = solver.epolar1()
    err = np.abs(dep - grad).max() / np.abs(grad).max()
    print(f"water-1944 rel={err:.2e}", flush=True)
    assert err < 1e-9
    solver.close()


def test_fortran_host_gradient_matches_python_binding(ctx):
    # the Fortran host (polar_gradient over ISO_C_BINDING) and the Python
    # binding reach the same entry point on the same 1,944-atom box
    import os
    import re
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tinker-hp_amd", "fortran", "polar_driver")
    out = subprocess.run([exe, "1944", "18.643", "24", "1", "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = dict(re.findall(r"(\w+)=\s*(\S+)", out.stdout))
    solver = thp.Solver(ctx, thp.water_1944(), FP64)
    solver.newinduce_pme2(poleps=1e-8)
    dep = solver.epolar1()
    rms = float(np.sqrt(np.mean(dep ** 2)))
    assert abs(float(vals["GRADRMS"]) - rms) < 1e-12 * rms
    solver.close()


def test_dhfr2_gradient_mixed_matches_fp64(ctx):
    # the reference's dhfr2 example (23,558 atoms, every frame type of a
    # protein, chiral sites, polarization groups): too large for the
    # oracle's gradient here, so the mixed path is held to the FP64 path,
    # which the oracle pins on the smaller boxes
    import os
    import thp.tinker as tinker
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    s = tinker.load(os.path.join(gold, "dhfr2.xyz.gz"), os.path.join(gold, "dhfr2.key"))
    out = {}
    for prec in (FP64, MIXED):
        solver = thp.Solver(ctx, s, prec)
        _, ep, st = solver.newinduce_pme2(poleps=1e-8)
        assert st.status == 0
        out[prec] = solver.epolar1()
        solver.close()
    g64, g32 = out[FP64], out[MIXED]
    err = np.abs(g32 - g64).max() / np.abs(g64).max()
    rms = float(np.sqrt(np.mean((g32 - g64) ** 2)) / np.sqrt(np.mean(g64 ** 2)))
    print(f"dhfr2 gradient max|dep|={np.abs(g64).max():.4f} rel_max={err:.2e} rel_rms={rms:.2e}", flush=True)
    assert err < 1e-4 and rms < 1e-5


def _oracle_virial(s, fr, poleps=1e-10):
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    if fr is not None:
        rp = O.rotpole(*xyz.T, *fr)
        s = s.with_multipoles(rp)
    else:
        rp = np.asarray(s.rpole)
    mu = O.Oracle(s).induce(poleps=poleps)["mu"]
    D = O.epolar1_virial(s, rp, mu[:, 0], mu[:, 1], fr)
    return 0.5 * (D + D.T)


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (MIXED, 2e-5)])
@pytest.mark.parametrize("frames", [True, False])
def test_virial_matches_oracle(ctx, prec, tol, frames):
    # epolar1c's vir: real-space pairs, the reciprocal sum's box and moment
    # terms and the frame forces, against the oracle's strain derivative
    # (pinned by differences of the re-solved energy in a strained box)
    s = _box(nmol=27, box=10.0, nfft=16, seed=4)
    fr = _water_frames(s) if frames else None
    ref = _oracle_virial(s, fr)
    solver = _device(ctx, s, fr, prec, 1e-10)
    dep, vir = solver.epolar1_vir()
    dep0 = solver.epolar1()
    assert np.array_equal(dep, dep0)          # the gradient is unchanged
    err = np.abs(vir - ref).max() / np.abs(ref).max()
    print(f"virial frames={frames} prec={prec} rel={err:.2e}\n{vir}", flush=True)
    assert np.allclose(vir, vir.T, rtol=0, atol=1e-12 * np.abs(vir).max())
    assert err < tol
    solver.close()


def test_virial_on_dhfr2_mixed_against_fp64(ctx):
    # every frame type of a protein, chiral sites, polarization groups: the
    # mixed virial within 1e-5 of the FP64 one
    import os
    import thp.tinker as tinker
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ds = tinker.load(os.path.join(gold, "dhfr2.xyz.gz"), os.path.join(gold, "dhfr2.key"))
    virs = []
    for prec in (FP64, MIXED):
        sv = thp.Solver(ctx, ds, prec)
        sv.newinduce_pme2(poleps=1e-8)
        virs.append(sv.epolar1_vir()[1])
        sv.close()
    err = np.abs(virs[1] - virs[0]).max() / np.abs(virs[0]).max()
    print(f"dhfr2 virial (FP64)\n{virs[0]}\nmixed rel {err:.2e}", flush=True)
    assert np.abs(virs[0]).max() > 100.0 and err < 1e-5

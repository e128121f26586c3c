"""The direct pass (tiling variant 3, with 1, 2, 4 or 8 lanes per i-particle):
the whole force of a small simulation, contacts included, in one launch.  Checked against the oracle for the impact
and generator forms, ragged and one-material sizes, a state with receding
deep contacts, the potential energy, agreement with the three-pass forms,
and the graph-captured stepper.  Tolerances: tests/parity.py.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import f4
from parity import assert_force_parity, force_metrics

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 2, 4, 8], ids=lambda s: f"lanes{s}")
def direct(gpu, request):
    """The direct pass with 1, 2, 4 or 8 lanes (j-streams) per i-particle."""
    gpu.call("gi_set_tiling", request.param, 3, 0, 0, 0)
    yield request.param
    gpu.call("gi_set_tiling", 0, 0, 0, 0, 0)


def _force(gi, pos, vel, fc):
    n = len(pos)
    P, V, F = gi.DeviceArray(n, pos), gi.DeviceArray(n, vel), gi.DeviceArray(n, f4(n))
    gi.force_collision(P, V, F, n, fc)
    gi.sync()
    return F.download()


CASES = {
    # name: (n_target, n_impactor, seed, angle, speed, target spin)
    "impact_2048": (1536, 512, 20261202, 45.0, 1.2, 0.0),
    "ragged_1000_spin": (800, 200, 7, 30.0, 1.0, 0.5),
    "ragged_4097": (3000, 1097, 3, 60.0, 1.4, 0.3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_direct_force_parity(gpu, oracle, direct, case):
    nt, ni, seed, ang, spd, spin = CASES[case]
    pos, vel, fc, _, _ = gpu.make_impact(nt, ni, seed=seed, angle_deg=ang, speed_vesc=spd, target_spin=spin)
    print(f"case={case} n={nt + ni}", flush=True)
    F = _force(gpu, pos, vel, fc)
    assert gpu.plan_info()["kernel"] == "k_direct"
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    met = force_metrics(F, F64, cond, near, oracle.force_collision_f32(pos, vel, fc))
    print(met)
    assert_force_parity(met)
    assert np.all(F[:, 3] == 0.0)  # .w untouched (was zero)


@pytest.mark.parametrize("n,nfe", [(1, 0), (1, 1), (2, 1), (3, 3), (65, 0), (257, 100), (640, 640), (1100, 512)])
def test_direct_small_and_ragged_sizes(gpu, oracle, direct, n, nfe):
    rng = np.random.default_rng(n + 11)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = rng.uniform(-0.6, 0.6, (n, 3)) * max(1.0, n ** (1 / 3))
    vel[:, :3] = rng.normal(0, 2, (n, 3))
    fc = gpu.make_impact(10, 0)[2]
    fc.NFe = nfe
    print(f"n={n} nfe={nfe}", flush=True)
    F = _force(gpu, pos, vel, fc)
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    if n == 1:
        assert np.all(F[:, :3] == 0)
        return
    met = force_metrics(F, F64, cond, near, oracle.force_collision_f32(pos, vel, fc))
    assert_force_parity(met, check_ref=False)


def test_direct_force_during_impact(gpu, oracle, direct):
    """After 400 steps the bodies are in contact: receding branches, deep
    overlaps and Fe-Si pairs all occur."""
    pos, vel, fc, mc, _ = gpu.make_impact(1536, 512, seed=20261202, angle_deg=45.0, speed_vesc=1.2,
                                          gap_in_target_radii=0.3)
    n = len(pos)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.call("gi_impact_steps", P.ptr, V.ptr, F.ptr, n, C.byref(fc), C.byref(mc), 400, 0, None, None, None)
    gpu.sync()
    p, v = P.download(), V.download()
    gpu.force_collision(P, V, F, n, fc)
    gpu.sync()
    Fg = F.download()
    F64, cond, near = oracle.force_collision_f64(p, v, fc)
    met = force_metrics(Fg, F64, cond, near, oracle.force_collision_f32(p, v, fc))
    print(met)
    d = p[None, :, :3] - p[:, None, :3]
    r = np.sqrt((d**2).sum(-1))
    dot = (d * (v[None, :, :3] - v[:, None, :3])).sum(-1)
    assert np.sum((r < 0.98) & (dot >= 1.0)) > 0
    assert_force_parity(met)


def test_direct_force_separate_parity(gpu, oracle, direct):
    """Two bodies overlapping in space: the cross-body pairs must be skipped."""
    from test_gpu_force import _generator_bodies

    pos, vel, fs, _ = _generator_bodies(gpu, 1200, 500, 21)
    n = len(pos)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.force_separate(P, V, F, n, fs)
    gpu.sync()
    Fg = F.download()
    F64, cond, near = oracle.force_separate_f64(pos, vel, fs)
    met = force_metrics(Fg, F64, cond, near, oracle.force_separate_f32(pos, vel, fs))
    print(met)
    assert_force_parity(met)


def test_direct_agrees_with_the_three_pass_forms(gpu):
    pos, vel, fc, _, _ = gpu.make_impact(6000, 2000, seed=12, angle_deg=40.0, speed_vesc=1.1, target_spin=0.3)
    print("n=8000", flush=True)
    Fs = {}
    for name, tiling in (("direct", (1, 3, 0, 0, 0)), ("direct4", (4, 3, 0, 0, 0)),
                         ("one-sided", (1, 0, 0, 8192, 512)), ("mutual", (8, 2, 512, 0, 64))):
        gpu.call("gi_set_tiling", *tiling)
        try:
            Fs[name] = _force(gpu, pos, vel, fc)[:, :3].astype(np.float64)
        finally:
            gpu.call("gi_set_tiling", 0, 0, 0, 0, 0)
    ref = Fs["direct"]
    for name in ("direct4", "one-sided", "mutual"):
        d = np.linalg.norm(Fs[name] - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
        print(name, d.max(), np.percentile(d, 99.9))
        assert d.max() <= 5e-6 and np.percentile(d, 99.9) <= 1e-6
    # Newton's third law over the whole system
    assert np.abs(ref.sum(0)).max() <= 1e-5 * np.abs(ref).sum()


def test_direct_conserved_quantities_match_fp64_oracle(gpu, oracle, direct):
    pos, vel, fc, mc, _ = gpu.make_impact(1536, 512, seed=5, target_spin=0.5, gap_in_target_radii=0.3)
    n = len(pos)
    P, V = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel)
    out = (C.c_double * 9)()
    gpu.call("gi_conserved_quantities", P.ptr, V.ptr, n, C.byref(fc), out, None)
    g = np.array(list(out))
    o = oracle.conserved_f64(pos, vel, fc)
    print(g, o)
    assert g[0] == pytest.approx(o[0], rel=1e-12)
    assert g[1] == pytest.approx(o[1], rel=2e-6)
    # and in a state with contacts: the contact potential is taken exactly
    F = gpu.DeviceArray(n, f4(n))
    gpu.call("gi_impact_steps", P.ptr, V.ptr, F.ptr, n, C.byref(fc), C.byref(mc), 300, 0, None, None, None)
    gpu.call("gi_conserved_quantities", P.ptr, V.ptr, n, C.byref(fc), out, None)
    g = np.array(list(out))
    o = oracle.conserved_f64(P.download(), V.download(), fc)
    print(g, o)
    assert g[1] == pytest.approx(o[1], rel=2e-6)


def test_direct_graph_stepper_equals_the_manual_loop(gpu, direct):
    pos, vel, fc, mc, _ = gpu.make_impact(1536, 512, seed=9)
    n = len(pos)
    P1, V1, F1 = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    P2, V2, F2 = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    for _ in range(21):
        gpu.force_collision(P1, V1, F1, n, fc)
        gpu.move_collision(P1, V1, F1, n, mc)
    gpu.call("gi_impact_steps", P2.ptr, V2.ptr, F2.ptr, n, C.byref(fc), C.byref(mc), 5, 0, None, None, None)
    gpu.call("gi_impact_steps", P2.ptr, V2.ptr, F2.ptr, n, C.byref(fc), C.byref(mc), 16, 0, None, None, None)
    gpu.sync()
    assert np.array_equal(P1.download(), P2.download())
    assert np.array_equal(V1.download(), V2.download())

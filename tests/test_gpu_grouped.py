"""The mutual all-pairs pass with row groups (tiling variant 4): the four
wavefronts of a workgroup hold four consecutive rows, stream the same
j-blocks beyond the group and add their j-side sums in LDS before writing
them; pairs inside a group are single-wavefront row tasks.  Checked against
the oracle at both row sizes (groups cut by the iron/silicate border, groups
of fewer than four rows, ghost padding), against the plain mutual pass, for
the potential, in the generator's two-body form, under dispatch limits, and
in the batched engine (bitwise the same alone and in any mix).  Tolerances:
tests/parity.py.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import f4
from parity import assert_force_parity, force_metrics

pytestmark = pytest.mark.gpu


@pytest.fixture
def tiling(gpu):
    def set_tiling(r, variant, chunk_len=0, target_tasks=0, min_chunk=64):
        gpu.call("gi_set_tiling", r, variant, chunk_len, target_tasks, min_chunk)

    yield set_tiling
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
    "ragged_4097": (3000, 1097, 3, 60.0, 1.4, 0.3),
    "small_770": (700, 70, 5, 30.0, 1.0, 0.2),
    "rows_9000": (7000, 2000, 9, 40.0, 1.1, 0.5),
}


@pytest.mark.parametrize("R,L", [(4, 256), (8, 128), (8, 1024)])
@pytest.mark.parametrize("case", sorted(CASES))
def test_grouped_force_parity(gpu, oracle, tiling, case, R, L):
    nt, ni, seed, ang, spd, spin = CASES[case]
    pos, vel, fc, _, _ = gpu.make_impact(nt, ni, seed=seed, angle_deg=ang, speed_vesc=spd, target_spin=spin)
    print(f"case={case} R={R} L={L}", flush=True)
    tiling(R, 4, L)
    F = _force(gpu, pos, vel, fc)
    assert gpu.plan_info()["kernel"] == "k_mutual_grp"
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    met = force_metrics(F, F64, cond, near, oracle.force_collision_f32(pos, vel, fc))
    print(met)
    assert_force_parity(met)
    assert np.all(F[:, 3] == 0.0)


def test_grouped_agrees_with_the_mutual_pass(gpu, tiling):
    """Same pairs, same terms; only the float sums are cut differently."""
    pos, vel, fc, _, _ = gpu.make_impact(30000, 10000, seed=8, angle_deg=40.0, speed_vesc=1.1)
    print("n=40000", flush=True)
    tiling(8, 2, 1024)
    Fm = _force(gpu, pos, vel, fc)[:, :3].astype(np.float64)
    tiling(8, 4, 1024)
    Fg = _force(gpu, pos, vel, fc)[:, :3].astype(np.float64)
    d = np.linalg.norm(Fg - Fm, axis=1) / np.maximum(np.linalg.norm(Fm, axis=1), 1e-30)
    print(d.max(), np.percentile(d, 99.9))
    assert d.max() <= 5e-6 and np.percentile(d, 99.9) <= 1e-6
    tot = Fg.sum(0)
    assert np.abs(tot).max() <= 1e-5 * np.abs(Fg).sum()


@pytest.mark.parametrize("n,nfe", [(2, 1), (65, 0), (257, 100), (1100, 512), (2100, 1600), (5000, 3000)])
def test_grouped_small_and_ragged_sizes(gpu, oracle, tiling, n, nfe):
    """Ghost rows, one-material simulations, groups of one to three rows,
    the material border inside a group (nfe = 1,600: rows 0-3 iron, 4 mixed
    padding) and on a row boundary (nfe = 512)."""
    rng = np.random.default_rng(n + 11)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = rng.uniform(-0.6, 0.6, (n, 3)) * max(1.0, n ** (1 / 3))
    vel[:, :3] = rng.normal(0, 2, (n, 3))
    fc = gpu.make_impact(10, 0)[2]
    fc.NFe = nfe
    print(f"n={n} nfe={nfe}", flush=True)
    tiling(8, 4, 64)
    F = _force(gpu, pos, vel, fc)
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    met = force_metrics(F, F64, cond, near, oracle.force_collision_f32(pos, vel, fc))
    assert_force_parity(met, check_ref=False)


def test_grouped_force_separate_parity(gpu, oracle, tiling):
    from test_gpu_force import _generator_bodies

    pos, vel, fs, _ = _generator_bodies(gpu, 2400, 900, 22)
    n = len(pos)
    tiling(8, 4, 128)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.force_separate(P, V, F, n, fs)
    gpu.sync()
    Fg = F.download()
    F64, cond, near = oracle.force_separate_f64(pos, vel, fs)
    met = force_metrics(Fg, F64, cond, near, oracle.force_separate_f32(pos, vel, fs))
    print(met)
    assert_force_parity(met)


def test_grouped_conserved_quantities_match_fp64_oracle(gpu, oracle, tiling):
    pos, vel, fc, _, _ = gpu.make_impact(4500, 1500, seed=5, target_spin=0.5)
    n = len(pos)
    P, V = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel)
    tiling(8, 4, 128)
    out = (C.c_double * 9)()
    gpu.call("gi_conserved_quantities", P.ptr, V.ptr, n, C.byref(fc), out, None)
    g = np.array(list(out))
    o = oracle.conserved_f64(pos, vel, fc)
    print(g, o)
    assert g[0] == pytest.approx(o[0], rel=1e-12)
    assert g[1] == pytest.approx(o[1], rel=2e-6)


def test_grouped_dispatch_limit_is_bitwise_the_same(gpu, tiling):
    pos, vel, fc, _, _ = gpu.make_impact(3000, 1097, seed=3, angle_deg=60.0, speed_vesc=1.4)
    n = len(pos)
    tiling(8, 4, 128)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))

    def run():
        gpu.force_collision(P, V, F, n, fc)
        gpu.sync()
        return F.download(), gpu.conserved_quantities(P, V, n, fc), gpu.plan_info()["pair_tasks"]

    F0, e0, ntasks = run()
    assert ntasks > 7
    try:
        for lim in (1, 7, ntasks + 5):
            gpu.call("gi_set_dispatch_limit", lim)
            F1, e1, _ = run()
            assert np.array_equal(F0, F1) and np.array_equal(e0, e1), lim
    finally:
        gpu.call("gi_set_dispatch_limit", 0)


def _ensemble_of_one(gi, pos, vel, fc, mc, steps):
    n = len(pos)
    P, V = gi.DeviceArray(n, pos), gi.DeviceArray(n, vel)
    e = gi.Ensemble(P, V, np.array([n], np.int32), (gi.ForceCollisionConstants * 1)(fc),
                    (gi.MoveCollisionConstants * 1)(mc))
    e.steps(steps)
    gi.sync()
    out = P.download(), V.download()
    e.close()
    return out


def test_grouped_batched_is_bitwise_the_same_alone_and_in_the_mix(gpu):
    gpu.call("gi_set_batched_tiling", 8, 1024, 4)
    try:
        cfg = gpu.ensemble_configs(5, seed=14, n_range=(700, 9000))
        pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
        print("n =", list(n), flush=True)
        P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
        ens = gpu.Ensemble(P, V, n, fc, mc)
        ens.set_active([1, 0, 1, 1, 1])
        ens.steps(3)
        ens.set_active([1, 1, 1, 1, 1])
        ens.steps(4)
        gpu.sync()
        p_all, v_all = P.download(), V.download()
        off = ens.offsets
        c = ens.conserved()
        ens.close()
        for s, k in zip(range(len(n)), (7, 4, 7, 7, 7)):
            a, b = off[s], off[s + 1]
            p1, v1 = _ensemble_of_one(gpu, pos[a:b], vel[a:b], fc[s], mc[s], k)
            assert np.array_equal(p_all[a:b], p1) and np.array_equal(v_all[a:b], v1), s
        assert np.all(np.isfinite(c))
    finally:
        gpu.call("gi_set_batched_tiling", 0, 0, 0)

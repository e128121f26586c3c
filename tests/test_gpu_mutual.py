"""The mutual all-pairs pass (tiling variant 2): each pair evaluated once and
applied to both particles, rows of 64*R particles in a padded layout with
ghosts.  Checked at every row size against the oracle, against the one-sided
pass, in the generator's two-body form, for the potential, for ragged and
degenerate sizes (ghost padding, one material only), and in the batched
engine (bitwise the same alone and in any mix).  Tolerances: tests/parity.py.
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


@pytest.fixture
def batched_mutual(gpu):
    gpu.call("gi_set_batched_tiling", 8, 1024, 2)
    yield
    gpu.call("gi_set_batched_tiling", 0, 0, 0)


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
}


@pytest.mark.parametrize("R", [1, 2, 4, 8])
@pytest.mark.parametrize("case", sorted(CASES))
def test_mutual_force_parity(gpu, oracle, tiling, case, R):
    nt, ni, seed, ang, spd, spin = CASES[case]
    pos, vel, fc, _, _ = gpu.make_impact(nt, ni, seed=seed, angle_deg=ang, speed_vesc=spd, target_spin=spin)
    print(f"case={case} R={R}", flush=True)
    tiling(R, 2, 128)  # short chunks: several tasks and j-blocks per row
    F = _force(gpu, pos, vel, fc)
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    met = force_metrics(F, F64, cond, near, oracle.force_collision_f32(pos, vel, fc))
    print(met)
    assert_force_parity(met)
    assert np.all(F[:, 3] == 0.0)


def test_mutual_agrees_with_the_one_sided_pass(gpu, tiling):
    pos, vel, fc, _, _ = gpu.make_impact(15000, 5000, seed=8, angle_deg=40.0, speed_vesc=1.1)
    print("n=20000", flush=True)
    tiling(1, 0, 0, 8192, 512)
    Fp = _force(gpu, pos, vel, fc)[:, :3].astype(np.float64)
    for R, L in ((4, 1024), (8, 2048)):
        tiling(R, 2, L)
        Fm = _force(gpu, pos, vel, fc)[:, :3].astype(np.float64)
        d = np.linalg.norm(Fm - Fp, axis=1) / np.maximum(np.linalg.norm(Fp, axis=1), 1e-30)
        print(R, L, d.max(), np.percentile(d, 99.9))
        assert d.max() <= 5e-6 and np.percentile(d, 99.9) <= 1e-6
    # Newton's third law over the whole system
    tot = Fm.sum(0)
    assert np.abs(tot).max() <= 1e-5 * np.abs(Fm).sum()


@pytest.mark.parametrize("variant,R,L", [(2, 8, 128), (2, 4, 256), (0, 4, 512)])
def test_dispatch_limit_is_bitwise_the_same(gpu, tiling, variant, R, L):
    """gi_set_dispatch_limit cuts the task list into several dispatches; every
    task writes only its own partial sums, so force and potential energy are
    bitwise unchanged (limits: one task, a prime, more than there are)."""
    pos, vel, fc, _, _ = gpu.make_impact(3000, 1097, seed=3, angle_deg=60.0, speed_vesc=1.4)
    n = len(pos)
    print(f"variant={variant} R={R} L={L}", flush=True)
    tiling(R, variant, L)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))

    def run():
        gpu.force_collision(P, V, F, n, fc)
        gpu.sync()
        ntasks = gpu.plan_info()["pair_tasks"]
        e = gpu.conserved_quantities(P, V, n, fc)
        gpu.sync()
        return F.download(), e, ntasks

    F0, e0, ntasks = run()
    assert ntasks > 7
    try:
        for lim in (1, 7, ntasks + 5):
            gpu.call("gi_set_dispatch_limit", lim)
            F1, e1, _ = run()
            assert np.array_equal(F0, F1), lim
            assert np.array_equal(e0, e1), lim
    finally:
        gpu.call("gi_set_dispatch_limit", 0)


@pytest.mark.parametrize("n,nfe", [(1, 0), (2, 1), (3, 3), (65, 0), (257, 100), (640, 640), (1100, 512)])
def test_mutual_small_and_ragged_sizes(gpu, oracle, tiling, n, nfe):
    """Rows padded with ghosts, one-material simulations, a material border
    on a row boundary (nfe = 512 = 64*8)."""
    rng = np.random.default_rng(n + 7)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = rng.uniform(-0.6, 0.6, (n, 3)) * max(1.0, n ** (1 / 3))
    vel[:, :3] = rng.normal(0, 2, (n, 3))
    fc = gpu.make_impact(10, 0)[2]
    fc.NFe = nfe
    print(f"n={n} nfe={nfe}", flush=True)
    tiling(8, 2, 64)
    F = _force(gpu, pos, vel, fc)
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    if n == 1:
        assert np.all(F[:, :3] == 0)
        return
    met = force_metrics(F, F64, cond, near, oracle.force_collision_f32(pos, vel, fc))
    assert_force_parity(met, check_ref=False)


def test_mutual_force_separate_parity(gpu, oracle, tiling):
    """Two bodies overlapping in space: the cross-body pairs must be skipped."""
    from test_gpu_force import _generator_bodies

    pos, vel, fs, _ = _generator_bodies(gpu, 1200, 500, 21)
    n = len(pos)
    tiling(4, 2, 128)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.force_separate(P, V, F, n, fs)
    gpu.sync()
    Fg = F.download()
    F64, cond, near = oracle.force_separate_f64(pos, vel, fs)
    met = force_metrics(Fg, F64, cond, near, oracle.force_separate_f32(pos, vel, fs))
    print(met)
    assert_force_parity(met)


def test_mutual_conserved_quantities_match_fp64_oracle(gpu, oracle, tiling):
    pos, vel, fc, _, _ = gpu.make_impact(1536, 512, seed=5, target_spin=0.5)
    n = len(pos)
    P, V = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel)
    tiling(8, 2, 128)
    out = (C.c_double * 9)()
    gpu.call("gi_conserved_quantities", P.ptr, V.ptr, n, C.byref(fc), out, None)
    g = np.array(list(out))
    o = oracle.conserved_f64(pos, vel, fc)
    print(g, o)
    assert g[0] == pytest.approx(o[0], rel=1e-12)
    assert g[1] == pytest.approx(o[1], rel=2e-6)


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


def test_mutual_batched_is_bitwise_the_same_alone_and_in_the_mix(gpu, batched_mutual):
    cfg = gpu.ensemble_configs(5, seed=13, n_range=(700, 4100))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    print("n =", list(n), flush=True)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.set_active([1, 1, 0, 1, 1])
    ens.steps(3)
    ens.set_active([1, 1, 1, 1, 1])
    ens.steps(4)
    gpu.sync()
    p_all, v_all = P.download(), V.download()
    off = ens.offsets
    for s, k in zip(range(len(n)), (7, 7, 4, 7, 7)):
        a, b = off[s], off[s + 1]
        p1, v1 = _ensemble_of_one(gpu, pos[a:b], vel[a:b], fc[s], mc[s], k)
        assert np.array_equal(p_all[a:b], p1), s
        assert np.array_equal(v_all[a:b], v1), s
    ens.close()


def test_mixed_plan_too_large_simulation_takes_the_one_sided_pass(gpu, batched_mutual, tiling):
    """A simulation whose j-side slab would pass the per-simulation cap of the
    mutual pass with one row per slab group (N = 1.3 M) takes the one-sided
    pass inside the same batched launch as two small mutual-pass simulations;
    each small one stays bitwise the same as alone, and the large one agrees
    with its single-simulation path, which falls back too with that pass and
    keeps the mutual pass in its row-group form (a quarter of the slab)."""
    cfg = [dict(n_target=3000, n_impactor=1000, seed=1, angle_deg=40.0, speed_vesc=1.1, target_spin=0.2),
           dict(n_target=975000, n_impactor=325000, seed=2, angle_deg=45.0, speed_vesc=1.1, target_spin=0.0),
           dict(n_target=4000, n_impactor=1200, seed=3, angle_deg=50.0, speed_vesc=1.2, target_spin=0.4)]
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    print("n =", list(n), flush=True)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.steps(1)
    gpu.sync()
    p_all, v_all = P.download(), V.download()
    off = ens.offsets
    ens.close()
    for s in (0, 2):
        a, b = off[s], off[s + 1]
        p1, v1 = _ensemble_of_one(gpu, pos[a:b], vel[a:b], fc[s], mc[s], 1)
        assert np.array_equal(p_all[a:b], p1) and np.array_equal(v_all[a:b], v1), s
    a, b = off[1], off[2]
    nb = b - a
    kick = v_all[a:b, :3].astype(np.float64) - vel[a:b, :3]
    for variant, mutual in ((2, False), (4, True)):
        tiling(8, variant, 2048)
        Pb, Vb, Fb = gpu.DeviceArray(nb, pos[a:b]), gpu.DeviceArray(nb, vel[a:b]), gpu.DeviceArray(nb, f4(nb))
        gpu.impact_steps(Pb, Vb, Fb, nb, fc[1], mc[1], 1)
        gpu.sync()
        assert gpu.plan_info()["mutual"] == mutual, variant
        dv = v_all[a:b, :3].astype(np.float64) - Vb.download()[:, :3]
        assert np.abs(dv).max() <= 1e-5 * np.abs(kick).max(), variant
        del Pb, Vb, Fb


def test_large_simulation_mutual_pass_against_fp64(gpu, oracle):
    """N = 1,048,576 with the tuning table's tiling (the mutual pass): rows
    across every material against the FP64 oracle and the reference's float
    routine, and Newton's third law.  At this size float positions (|x| up to
    ~300) alone limit near-pair separations to ~1e-5 relative; the reference
    float routine's own error here is 6e-4 to 1e-2."""
    pos, vel, fc, _, _ = gpu.make_impact(786432, 262144, seed=6, angle_deg=45.0, speed_vesc=1.1)
    n = len(pos)
    print(f"n={n}", flush=True)
    F = _force(gpu, pos, vel, fc)
    assert gpu.plan_info()["mutual"]
    for a in (0, fc.NFe - 100, fc.NFe + 4000, n - 200):
        F64, cond, near = oracle.force_collision_f64(pos, vel, fc, i_begin=a, i_end=a + 128)
        Fr = oracle.force_collision_f32(pos, vel, fc, a, a + 128)
        sl = slice(a, a + 128)
        met = force_metrics(F[sl], F64[sl], cond[sl], near[sl], Fr[sl])
        print(a, met)
        assert_force_parity(met)
    tot = F[:, :3].astype(np.float64).sum(0)
    assert np.abs(tot).max() <= 1e-5 * np.abs(F[:, :3]).astype(np.float64).sum()


@pytest.fixture
def contact_overlap(gpu):
    def set_overlap(on):
        gpu.call("gi_set_contact_overlap", 1 if on else 0)

    yield set_overlap
    gpu.call("gi_set_contact_overlap", 1)


@pytest.mark.parametrize("R,L", [(4, 256), (8, 1024)])
def test_contact_overlap_is_bitwise_the_same(gpu, tiling, contact_overlap, R, L):
    """The contact corrections computed beside the all-pairs pass and added
    with the sums, against the same corrections added after the join by
    k_finish: one device function, one rounding sequence."""
    pos, vel, fc, mc, _ = gpu.make_impact(4000, 1200, seed=31, angle_deg=40.0, speed_vesc=1.1,
                                          gap_in_target_radii=0.2)
    n = len(pos)
    tiling(R, 2, L)
    # a state with contacts: step the bodies into each other first
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.call("gi_impact_steps", P.ptr, V.ptr, F.ptr, n, C.byref(fc), C.byref(mc), 300, 0, None, None, None)
    gpu.sync()
    p, v = P.download(), V.download()
    print(f"n={n} R={R} L={L}", flush=True)
    out = {}
    for on in (True, False):
        contact_overlap(on)
        out[on] = _force(gpu, p, v, fc)
        assert gpu.plan_info()["kernel"] == "k_mutual"
    from scipy.spatial import cKDTree

    assert len(cKDTree(p[:, :3].astype(np.float64)).query_pairs(1.0)) > n // 2  # contacts present
    assert np.array_equal(out[True], out[False])


def test_contact_overlap_batched_step_is_bitwise_the_same(gpu, batched_mutual, contact_overlap):
    """The batched step with the corrections beside the pair pass and the move
    fused into the sums, against the finishing kernel and a separate move."""
    cfg = gpu.ensemble_configs(4, seed=12, n_range=(900, 3500))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    print("n =", list(n), flush=True)
    out = {}
    for on in (True, False):
        contact_overlap(on)
        P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
        ens = gpu.Ensemble(P, V, n, fc, mc)
        ens.steps(5)
        ens.steps(20)
        gpu.sync()
        out[on] = (P.download(), V.download())
        ens.close()
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], out[False][1])

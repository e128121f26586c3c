"""Pins of the CPU oracle (oracle/gi_oracle.c) against what the physics and
the reference's source fix, independently of the oracle's own formulas.

The reference ships no tests, fixtures or golden vectors, so each check here
is a closed form, an invariant, a special case or a known-answer value a
plausible mistake would break (wrong sign, dropped mass factor, swapped
material, wrong receding branch, kick/drift order, ...).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import f4


def _consts(O, n=2048):
    u = O.units(n)
    fc, mc = O.impact_constants(u)
    return u, fc, mc


def _pair(p0, p1, v0=(0, 0, 0), v1=(0, 0, 0)):
    pos, vel = f4(2), f4(2)
    pos[0, :3], pos[1, :3] = p0, p1
    vel[0, :3], vel[1, :3] = v0, v1
    return pos, vel


def _mag_on_0(O, fc, r, nfe, dvx=0.0):
    """Force on particle 0 from particle 1 placed at +r along x (positive = attraction)."""
    pos, vel = _pair((0, 0, 0), (r, 0, 0), (0, 0, 0), (dvx, 0, 0))
    fc.NFe = nfe
    F = O.force_collision_f32(pos, vel, fc)
    assert F[0, 1] == 0 and F[0, 2] == 0
    return float(F[0, 0]), F


# --------------------------------------------------------------------------- gravity branch


@pytest.mark.parametrize("nfe,expect_mass", [(0, "1"), (1, "FeSi"), (2, "FeFe")])
def test_gravity_closed_form_and_third_law(oracle, nfe, expect_mass):
    _, fc, _ = _consts(oracle)
    mfe = 7.850e12 / 3.320e12  # MassFe = DensityFe / DensitySi
    gm = {"1": 1.0, "FeSi": mfe, "FeFe": mfe * mfe}[expect_mass]
    for r in (1.0, 2.0, 3.5, 10.0):
        f0, F = _mag_on_0(oracle, fc, r, nfe)
        assert f0 > 0, "gravity must pull particle 0 toward particle 1"
        assert f0 == pytest.approx(gm / r**2, rel=2e-6)
        assert F[1, 0] == -F[0, 0]  # Newton's third law, exactly


def test_self_pair_contributes_nothing(oracle):
    _, fc, _ = _consts(oracle)
    pos, vel = f4(1), f4(1)
    pos[0, :3] = (3.0, -2.0, 1.0)
    fc.NFe = 1
    assert np.all(oracle.force_collision_f32(pos, vel, fc)[:, :3] == 0)
    F64, cond, _ = oracle.force_collision_f64(pos, vel, fc)
    assert np.all(F64 == 0) and cond[0] == 0


# --------------------------------------------------------------------------- contact branch


def _potential(G, K, r):
    """Pair potential of the Greenspan law: -G/r outside, spring-like inside."""
    if r >= 1.0:
        return -G / r
    return G * (r - 2.0) + (K / 3.0) * (r - 1.0) ** 2 * (r + 2.0)


@pytest.mark.parametrize("nfe", [0, 1, 2])
def test_contact_force_is_derivative_of_potential(oracle, nfe):
    """Approaching pairs (dv = 0): f(r) = U'(r) for the stated potential, in every
    shell region, and f is continuous at r = 1."""
    _, fc, _ = _consts(oracle)
    G = [1.0, fc.GMassFeSi, fc.GMassFeFe][nfe]
    K = [fc.KSiSi, fc.KFeSi, fc.KFeFe][nfe]
    for r in (0.9999, 0.999, 0.9985, 0.99, 0.95, 0.9, 0.8, 0.6):
        f, _ = _mag_on_0(oracle, fc, r, nfe)
        h = 1e-6
        du = (_potential(G, K, r + h) - _potential(G, K, r - h)) / (2 * h)
        assert f == pytest.approx(du, rel=2e-5, abs=2e-4 * K * 1e-3)
    r_in = float(np.nextafter(np.float32(1.0), np.float32(0.0)))
    f_in, _ = _mag_on_0(oracle, fc, r_in, nfe)
    f_out, _ = _mag_on_0(oracle, fc, 1.0, nfe)
    assert f_out == pytest.approx(G, rel=1e-6)
    # continuous at r = 1: the jump over one float step is the slope 2K r times the step
    assert abs(f_in - f_out) <= 2.2 * K * (1.0 - r_in) + 1e-6 * G


def test_fe_si_stiffness_is_the_mean_of_the_pure_pairs(oracle):
    """KFeSi = KFe + KSi while KFeFe = 2 KFe and KSiSi = 2 KSi: the mixed overlap
    force is the mean of the pure ones (after removing gravity)."""
    _, fc, _ = _consts(oracle)
    for r in (0.999, 0.99, 0.95):
        ss, _ = _mag_on_0(oracle, fc, r, 0)
        fs, _ = _mag_on_0(oracle, fc, r, 1)
        ff, _ = _mag_on_0(oracle, fc, r, 2)
        assert (fs - fc.GMassFeSi) == pytest.approx(0.5 * ((ss - 1.0) + (ff - fc.GMassFeFe)), rel=1e-5)


@pytest.mark.parametrize(
    "nfe,r,kr_field,k_field",
    [
        (0, 0.9, "KRSiSi", "KSiSi"),  # Si-Si below ShellBreakSi
        (2, 0.95, "KRFeFe", "KFeFe"),  # Fe-Fe below ShellBreakFe
        (1, 0.99, "KRMix", "KFeSi"),  # Fe-Si between the two shell breaks
        (1, 0.95, "KRFeSi", "KFeSi"),  # Fe-Si below both
    ],
)
def test_receding_uses_reduced_stiffness_only_when_int_dot_positive(oracle, nfe, r, kr_field, k_field):
    """impactAddTime.cu:313-315: (int)(dp.dv) > 0 selects KR.  A positive dot
    product below 1 truncates to 0 and keeps the approach stiffness."""
    _, fc, _ = _consts(oracle)
    G = [1.0, fc.GMassFeSi, fc.GMassFeFe][nfe]
    approach, _ = _mag_on_0(oracle, fc, r, nfe, dvx=-1.0 / r)  # dot = -1
    slow_apart, _ = _mag_on_0(oracle, fc, r, nfe, dvx=0.5 / r)  # dot = 0.5 -> (int) 0
    apart, _ = _mag_on_0(oracle, fc, r, nfe, dvx=1.5 / r)  # dot = 1.5 -> (int) 1
    assert slow_apart == approach
    ratio = getattr(fc, kr_field) / getattr(fc, k_field)
    assert (apart - G) == pytest.approx((approach - G) * ratio, rel=1e-5)
    assert apart > approach  # less repulsion while receding


def test_shell_region_ignores_receding(oracle):
    """Between ShellBreak and 1 the approach stiffness applies both ways."""
    _, fc, _ = _consts(oracle)
    for nfe, r in ((0, 0.999), (2, 0.99), (1, 0.999)):
        a, _ = _mag_on_0(oracle, fc, r, nfe, dvx=-3.0)
        b, _ = _mag_on_0(oracle, fc, r, nfe, dvx=3.0)
        assert a == b


def test_trunc_to_int_matches_cvt_rzi(oracle):
    assert oracle.trunc_to_int(0.999) == 0
    assert oracle.trunc_to_int(1.0) == 1
    assert oracle.trunc_to_int(-0.5) == 0
    assert oracle.trunc_to_int(-1.5) == -1
    assert oracle.trunc_to_int(float("nan")) == 0
    assert oracle.trunc_to_int(1e10) == 2**31 - 1
    assert oracle.trunc_to_int(-1e10) == -(2**31)


# --------------------------------------------------------------------------- invariants


def _cluster(n, seed, spacing=0.93, vscale=0.5):
    rng = np.random.default_rng(seed)
    k = int(math.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)[:n]
    pos, vel = f4(n), f4(n)
    pos[:, :3] = g * spacing + rng.uniform(-0.05, 0.05, (n, 3))
    vel[:, :3] = rng.normal(0, vscale, (n, 3))
    return pos, vel


def test_third_law_and_zero_torque_with_contacts(oracle):
    _, fc, _ = _consts(oracle)
    pos, vel = _cluster(300, 1)
    fc.NFe = 90
    F64, cond, _ = oracle.force_collision_f64(pos, vel, fc)
    assert np.abs(F64.sum(0)).max() <= 1e-9 * cond.sum()
    torque = np.cross(pos[:, :3].astype(np.float64), F64).sum(0)
    assert np.abs(torque).max() <= 1e-9 * (cond * np.linalg.norm(pos[:, :3], axis=1)).sum()
    F32 = oracle.force_collision_f32(pos, vel, fc)[:, :3].astype(np.float64)
    assert np.abs(F32.sum(0)).max() <= 1e-5 * cond.sum()


def test_float_restatement_agrees_with_fp64_twin(oracle):
    _, fc, _ = _consts(oracle)
    pos, vel = _cluster(512, 2, vscale=3.0)
    fc.NFe = 150
    F32 = oracle.force_collision_f32(pos, vel, fc)[:, :3].astype(np.float64)
    F64, cond, near = oracle.force_collision_f64(pos, vel, fc)
    err = np.linalg.norm(F32 - F64, axis=1) / cond
    assert np.max(err[~near]) <= 1e-5


def _run(oracle, pos, vel, fc, mc, steps):
    p, v, _ = oracle.impact_steps_f32(pos, vel, fc, mc, steps)
    return p, v


def test_energy_conserved_with_kr_one_and_dissipated_with_kr_below_one(oracle):
    """With KR = K the law is conservative (the potential above); with the
    template KR = 0.01 K rebounds lose energy."""
    _, fc, mc = _consts(oracle)
    pos, vel = _cluster(27, 3, spacing=0.95, vscale=2.0)
    fc.NFe = mc.NFe = 8
    mc.Dt = 2e-5
    mc.DtOverMassFe = np.float32(2e-5 / fc.GMassFeSi)
    mc.DtOverMassSi = np.float32(2e-5)
    cons = oracle.FC()
    for k, _ in oracle.FC._fields_:
        setattr(cons, k, getattr(fc, k))
    cons.KRFeFe, cons.KRSiSi, cons.KRFeSi, cons.KRMix = cons.KFeFe, cons.KSiSi, cons.KFeSi, cons.KFeSi
    E0 = oracle.conserved_f64(pos, vel, cons)
    scale = abs(E0[0]) + abs(E0[1] - oracle.conserved_f64(pos * 50, vel, cons)[1])
    p, v = _run(oracle, pos, vel, cons, mc, 3000)
    E1 = oracle.conserved_f64(p, v, cons)
    assert abs(E1[2] - E0[2]) <= 2e-3 * scale
    # same start, reduced rebound stiffness: energy goes down
    p2, v2 = _run(oracle, pos, vel, fc, mc, 3000)
    E2 = oracle.conserved_f64(p2, v2, fc)
    assert E2[2] < E0[2] - 1e-2 * scale


def test_momentum_and_angular_momentum_conserved(oracle):
    _, fc, mc = _consts(oracle)
    pos, vel = _cluster(200, 4, spacing=0.97, vscale=1.0)
    c = pos[:, :3].mean(0)
    vel[:, 0] += -(pos[:, 1] - c[1]) * 0.5
    vel[:, 1] += (pos[:, 0] - c[0]) * 0.5
    fc.NFe = mc.NFe = 60
    E0 = oracle.conserved_f64(pos, vel, fc)
    p, v = _run(oracle, pos, vel, fc, mc, 200)
    E1 = oracle.conserved_f64(p, v, fc)
    mom_scale = np.abs(vel[:, :3]).sum() * 2.4
    assert np.abs(E1[3:6] - E0[3:6]).max() <= 1e-4 * mom_scale
    assert np.abs(E1[6:9] - E0[6:9]).max() <= 1e-3 * np.abs(E0[6:9]).max()


def test_kepler_circular_orbit(oracle):
    """Two silicate elements 10 diameters apart on a circular orbit keep their
    separation and complete one revolution in 2 pi sqrt(d^3 / (G M))."""
    _, fc, mc = _consts(oracle)
    d = 10.0
    v = math.sqrt(2.0 / d) / 2.0  # each body moves at half the relative speed
    pos, vel = _pair((-d / 2, 0, 0), (d / 2, 0, 0), (0, -v, 0), (0, v, 0))
    fc.NFe = mc.NFe = 0
    period = 2 * math.pi * math.sqrt(d**3 / 2.0)
    steps = int(round(period / mc.Dt))
    p, _ = _run(oracle, pos, vel, fc, mc, steps)
    sep = p[1, :3] - p[0, :3]
    assert np.linalg.norm(sep) == pytest.approx(d, rel=2e-3)
    ang = math.atan2(sep[1], sep[0])
    assert abs(ang) < 2e-2


# --------------------------------------------------------------------------- move kernels


def test_move_is_kick_then_drift_with_new_velocity(oracle):
    _, fc, mc = _consts(oracle)
    pos, vel, frc = f4(3), f4(3), f4(3)
    pos[:, :3] = [[1, 2, 3], [-1, 0.5, 2], [0, 0, 0]]
    vel[:, :3] = [[0.1, -0.2, 0.3], [1, 1, 1], [0, 0, 0]]
    frc[:, :3] = [[4, 5, 6], [-7, 8, -9], [1, 1, 1]]
    pos[:, 3], vel[:, 3], frc[:, 3] = 7.0, 8.0, 9.0
    mc.NFe = 1
    p, v = oracle.move_collision_f32(pos, vel, frc, mc)
    t = np.array([mc.DtOverMassFe, mc.DtOverMassSi, mc.DtOverMassSi], np.float32)[:, None]
    v_exp = (vel[:, :3] + frc[:, :3] * t).astype(np.float32)
    p_exp = (pos[:, :3] + v_exp * np.float32(mc.Dt)).astype(np.float32)
    assert np.array_equal(v[:, :3], v_exp) and np.array_equal(p[:, :3], p_exp)
    assert not np.array_equal(p[:2, :3], (pos[:2, :3] + vel[:2, :3] * np.float32(mc.Dt)))
    assert np.all(p[:, 3] == 7.0) and np.all(v[:, 3] == 8.0)


def test_damped_move_uses_body_damping_and_material_mass(oracle):
    u = oracle.units(2048)
    fs, ms = oracle.separate_constants(u)
    ms.boarder1, ms.boarder2, ms.boarder3 = 1, 2, 3  # one particle per [Fe1, Si1, Fe2, Si2]
    pos, vel, frc = f4(4), f4(4), f4(4)
    vel[:, :3] = 2.0
    frc[:, :3] = 3.0
    d1, d2 = 8.0, 0.5
    p, v = oracle.move_damped_separate_f32(pos, vel, frc, ms, d1, d2)
    t = [ms.DtOverMassFe, ms.DtOverMassSi, ms.DtOverMassFe, ms.DtOverMassSi]
    d = [d1, d1, d2, d2]
    for i in range(4):
        exp = np.float32(2.0) + (np.float32(3.0) - np.float32(d[i]) * np.float32(2.0)) * np.float32(t[i])
        assert v[i, 0] == np.float32(exp)


def test_separate_force_equals_each_body_alone(oracle):
    """Cross-body pairs are skipped (generateTargetImpactor.cu:529): the force on
    body 1 is that of body 1 alone, even with both bodies overlapping in space."""
    u = oracle.units(2048)
    fs, _ = oracle.separate_constants(u)
    pos1, vel1 = _cluster(150, 5, vscale=2.0)
    pos2, vel2 = _cluster(110, 6, vscale=2.0)
    pos2[:, :3] += 0.4  # interpenetrating
    b1, b2, b3 = 40, 150, 150 + 30
    fs.boarder1, fs.boarder2, fs.boarder3 = b1, b2, b3
    pos = np.concatenate([pos1, pos2])
    vel = np.concatenate([vel1, vel2])
    F = oracle.force_separate_f32(pos, vel, fs)
    fc = oracle.FC()
    for k, _ in oracle.FC._fields_[:-1]:
        setattr(fc, k, getattr(fs, k))
    fc.NFe = b1
    assert np.array_equal(F[:b2, :3], oracle.force_collision_f32(pos1, vel1, fc)[:, :3])
    fc.NFe = b3 - b2
    assert np.array_equal(F[b2:, :3], oracle.force_collision_f32(pos2, vel2, fc)[:, :3])
    # the FP64 twin masks the same pairs: each body's truth, conditioning
    # sums and near-threshold flags are those of the body alone
    F64, cond, near = oracle.force_separate_f64(pos, vel, fs)
    for sl, p_, v_, nfe in ((slice(0, b2), pos1, vel1, b1), (slice(b2, None), pos2, vel2, b3 - b2)):
        fc.NFe = nfe
        f1, c1, n1 = oracle.force_collision_f64(p_, v_, fc)
        assert np.allclose(F64[sl], f1, rtol=1e-13, atol=1e-13 * np.abs(f1).max())
        assert np.allclose(cond[sl], c1, rtol=1e-13) and np.array_equal(near[sl], n1)
    # and it is the float routine's truth: the masked float force within 1e-5 of the conditioning sum
    err = np.linalg.norm(F[:, :3].astype(np.float64) - F64, axis=1)
    assert np.max(err[~near] / cond[~near]) <= 1e-5


# --------------------------------------------------------------------------- units and host helpers


@pytest.mark.parametrize(
    "n,nfe,nsi,length_km,time_s,k_sim",
    [
        (2048, 264, 1784, 1161.65, 2935.86, 551.16),
        (65536, 8478, 57058, 365.96, 2935.86, 5553.37),
        (262144, 33914, 228230, 230.54, 2935.86, 13993.53),
    ],
)
def test_units_known_answers(oracle, n, nfe, nsi, length_km, time_s, k_sim):
    u = oracle.units(n)
    assert (u.nfe, u.nsi) == (nfe, nsi)
    assert u.unit_length == pytest.approx(length_km, abs=0.01)
    assert u.unit_time == pytest.approx(time_s, abs=0.01)
    assert u.k_fe_sim == pytest.approx(k_sim, abs=0.01)
    # closed forms: G = 1 in these units, element mass ratio = density ratio
    assert u.unit_time == pytest.approx(math.sqrt(6.0 / (3.141592654 * 6.67408e-20 * 3.320e12)), rel=1e-12)
    assert u.mass_fe == pytest.approx(7.850 / 3.320, rel=1e-15)
    vol = 3.141592654 / 6 * u.unit_length**3  # the reference's Pi (commonSetup.h:29)
    assert vol * 3.320e12 == pytest.approx(u.unit_mass, rel=1e-12)
    assert u.nfe1 + u.nfe2 == u.nfe and u.nsi1 + u.nsi2 == u.nsi


def test_percent_f_round_trip_of_impact_constants(oracle):
    assert oracle.load().or_round_f(2.3644578313253013) == 2.364458
    u = oracle.units(2048)
    fc, mc = oracle.impact_constants(u)
    assert fc.GMassFeSi == np.float32(2.364458)
    assert fc.GMassFeFe == np.float32(2.364458 * 2.364458)
    assert fc.ShellBreakFe == np.float32(0.98) and fc.ShellBreakSi == np.float32(0.998)
    assert fc.ShellBreakFeSi1 == np.float32(0.998) and fc.ShellBreakFeSi2 == np.float32(0.98)
    assert fc.KRMix == np.float32(round(u.k_fe_sim, 6) + round(u.k_si_sim, 6) * 0.01)
    assert mc.DtOverMassFe == np.float32(0.002 / 2.364458)


def test_recenter_restores_target_com_and_velocity(oracle):
    pos, vel = _cluster(500, 7, vscale=1.0)
    pos[:, :3] += 30.0
    mfe = 2.364458
    p, v = oracle.recenter_f32(pos, vel, 100, mfe, 1.0, (1.0, 2.0, 3.0), (0.1, 0.2, 0.3))
    cv = oracle.com_vel_collision_f32(p, v, 100, mfe, 1.0)
    assert np.allclose(cv[:3], [1, 2, 3], atol=1e-4)
    assert np.allclose(cv[3:], [0.1, 0.2, 0.3], atol=1e-5)
    m = np.where(np.arange(500) < 100, mfe, 1.0)
    com64 = (m[:, None] * p[:, :3].astype(np.float64)).sum(0) / m.sum()
    assert np.allclose(com64, [1, 2, 3], atol=1e-3)


def test_spin_body_is_rigid_rotation(oracle):
    """spinBody adds v = 2 pi w (n x r) about the body's centre of mass."""
    pos, vel = _cluster(343, 8, spacing=1.0)
    vel[:] = 0
    nfe, mfe = 100, 2.364458
    total = nfe * mfe + (343 - nfe)
    w = 0.05
    v = oracle.spin_body(pos, vel, 0, 343, nfe, mfe, 1.0, total, (0.0, 0.0, 2.0, w))
    com = oracle.body_com(pos, 0, nfe, 343, mfe, 1.0, total)[:3]
    r = pos[:, :3].astype(np.float64) - com
    expect = 2 * math.pi * w * np.stack([-r[:, 1], r[:, 0], np.zeros(343)], 1)
    assert np.allclose(v[:, :3], expect, rtol=1e-5, atol=1e-6)
    with pytest.raises(ValueError):
        oracle.spin_body(pos, vel, 0, 343, nfe, mfe, 1.0, total, (0.0, 0.0, 0.0, w))


# ---- Earth / Moon finder (impactBasicFunctions.h:109-217) and final sums ----

def _touch_components(pos, touch=1.5):
    """Connected components of the reference's float touch test, labelled by
    their smallest index (independent of the oracle's flood fill)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    p = np.asarray(pos, np.float32)[:, :3]
    dx = p[:, None, 0] - p[None, :, 0]
    dy = p[:, None, 1] - p[None, :, 1]
    dz = p[:, None, 2] - p[None, :, 2]
    mag = np.sqrt((dx * dx + dy * dy) + dz * dz)  # float32 throughout, no contraction
    adj = csr_matrix(mag < np.float32(touch))
    _, lab = connected_components(adj, directed=False)
    first = {}
    for i, l in enumerate(lab):
        first.setdefault(l, i)
    return np.array([first[l] for l in lab], np.int32)


@pytest.mark.parametrize("seed,n,box", [(1, 300, 9.0), (2, 400, 14.0), (3, 250, 5.0)])
def test_earth_moon_groups_are_the_touch_components(oracle, seed, n, box):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, box, (n, 3))
    r = oracle.find_earth_moon(pos)
    want = _touch_components(pos)
    assert np.array_equal(r["group"], want)
    sizes = np.bincount(want, minlength=n)
    assert r["groups"] == int((sizes > 0).sum())
    order = sorted(range(n), key=lambda g: (-sizes[g], g))
    assert r["earth_id"] == order[0] and r["earth_count"] == sizes[order[0]]
    if sizes[order[1]] > 0:
        assert r["moon_id"] == order[1] and r["moon_count"] == sizes[order[1]]


def _blob(rng, k, centre):
    """k points on a jittered unit lattice around centre (one touch group)."""
    side = int(np.ceil(k ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:k]
    return g + rng.uniform(-0.1, 0.1, (k, 3)) + np.asarray(centre)


def test_earth_moon_ties_go_to_the_smaller_group_number(oracle):
    rng = np.random.default_rng(5)
    a, b, c = _blob(rng, 25, (0, 0, 0)), _blob(rng, 40, (50, 0, 0)), _blob(rng, 40, (0, 50, 0))
    pos = np.zeros((105, 4), np.float32)
    pos[:, :3] = np.concatenate([a, b, c])
    r = oracle.find_earth_moon(pos)
    assert r["groups"] == 3
    assert (r["earth_id"], r["earth_count"]) == (25, 40)  # b, first of the two 40s
    assert (r["moon_id"], r["moon_count"]) == (65, 40)
    one = oracle.find_earth_moon(pos[:25])
    assert one["groups"] == 1 and one["earth_count"] == 25 and one["moon_id"] == -1 and one["moon_count"] == 0


def test_final_stats_of_a_rigid_rotation(oracle):
    """A drifting, rigidly rotating blob: centre, mean velocity and L = I w."""
    rng = np.random.default_rng(9)
    x = _blob(rng, 343, (3.0, -2.0, 5.0))
    n, nfe = len(x), 100
    m = np.where(np.arange(n) < nfe, 2.364458, 1.0)
    c = (m[:, None] * x).sum(0) / m.sum()
    w = np.array([0.02, -0.03, 0.05])
    u = np.array([0.3, 0.1, -0.2])
    v = np.cross(w, x - c) + u
    pos, vel = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
    pos[:, :3], vel[:, :3] = x, v
    group = np.zeros(n, np.int32)
    r = x.astype(np.float32).astype(np.float64) - c
    I = (m[:, None, None] * ((r * r).sum(1)[:, None, None] * np.eye(3) - r[:, :, None] * r[:, None, :])).sum(0)
    for fp64, tol in ((True, 1e-6), (False, 2e-3)):
        s = oracle.final_stats(pos, vel, nfe, group, 0, -1, 2.364458, fp64=fp64)
        assert s["mass"][0] == pytest.approx(m.sum(), rel=1e-6)
        assert np.allclose(s["com"][0], c, rtol=0, atol=1e-4)
        assert np.allclose(s["vel"][0], u, rtol=0, atol=tol * 0.1)
        assert np.allclose(s["angmom"][0], I @ w, rtol=tol, atol=tol * np.abs(I @ w).max())
        assert np.allclose(s["angmom"][2], s["angmom"][0], rtol=1e-6)  # the Earth is everything
        assert s["mass"][3] == 0


def test_final_stats_scope_one_is_the_union(oracle):
    """Masses and momenta add; L of the union = L(Earth) + L(Moon) + the
    orbital term sum M_s (c_s - c) x (u_s - u) (parallel-axis identity)."""
    rng = np.random.default_rng(11)
    a, b = _blob(rng, 125, (0, 0, 0)), _blob(rng, 64, (30, 4, -2))
    x = np.concatenate([a, b, rng.uniform(60, 70, (5, 3))])
    n = len(x)
    pos, vel = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
    pos[:, :3] = x
    vel[:, :3] = rng.normal(0, 1, (n, 3))
    r = oracle.find_earth_moon(pos)
    assert r["earth_id"] == 0 and r["moon_id"] == 125
    s = oracle.final_stats(pos, vel, 40, r["group"], r["earth_id"], r["moon_id"], 2.364458, fp64=True)
    M, C, U, L = s["mass"], s["com"], s["vel"], s["angmom"]
    assert M[1] == pytest.approx(M[2] + M[3], rel=1e-14)
    assert np.allclose(M[1] * U[1], M[2] * U[2] + M[3] * U[3], rtol=1e-12, atol=1e-12)
    orb = sum(M[k] * np.cross(C[k] - C[1], U[k] - U[1]) for k in (2, 3))
    assert np.allclose(L[1], L[2] + L[3] + orb, rtol=1e-10, atol=1e-9 * np.abs(L[1]).max())


# ---- generator constants (generateTargetImpactor.cu:406-451) ------------------


def test_separate_constants_known_answers_at_2048(oracle):
    """loadKernalConstantStructures of the generator at N = 2,048 with the
    template defaults: masses from the density ratio alone (MassSi = 1,
    G = 1), stiffnesses from the K_sim known answer 551.16 (units table at
    N = 2,048), shell breaks 1 - SD, and the borders of two equal bodies
    (0.55 Earth masses each, 26% iron): 264 irons and 1,784 silicates split
    in halves, so [Fe1 132 | Si1 892 | Fe2 132 | Si2 892]."""
    u = oracle.units(2048)
    fs, ms = oracle.separate_constants(u)
    mfe = 7.850 / 3.320  # unrounded: the generator never reads the %f file
    assert fs.GMassFeFe == np.float32(mfe * mfe)
    assert fs.GMassFeSi == np.float32(mfe)
    k = 551.16
    for name, want in (("KFeFe", 2 * k), ("KSiSi", 2 * k), ("KFeSi", 2 * k), ("KRFeFe", 2 * k * 0.01),
                       ("KRSiSi", 2 * k * 0.01), ("KRFeSi", 2 * k * 0.01), ("KRMix", k + k * 0.01)):
        assert getattr(fs, name) == pytest.approx(want, rel=2e-5), name
    assert fs.ShellBreakFe == np.float32(0.98) and fs.ShellBreakSi == np.float32(0.998)
    assert fs.ShellBreakFeSi1 == np.float32(0.998) and fs.ShellBreakFeSi2 == np.float32(0.98)
    assert (fs.boarder1, fs.boarder2, fs.boarder3) == (132, 1024, 1156)
    assert (ms.boarder1, ms.boarder2, ms.boarder3) == (132, 1024, 1156)
    assert ms.Dt == np.float32(0.002)
    assert ms.DtOverMassFe == np.float32(0.002 / mfe)
    assert ms.DtOverMassSi == np.float32(0.002)


def test_separate_constants_unrounded_against_rounded_impact_constants(oracle):
    """The generator's structures use the unrounded values, the impact's the
    %f-read ones: the same constants to within that rounding (6 decimals),
    and for MassFe the two differ exactly as 7.85/3.32 and 2.364458 do."""
    u = oracle.units(65536)
    fs, _ = oracle.separate_constants(u)
    fc, _ = oracle.impact_constants(u)
    for name in ("KFeFe", "KSiSi", "KFeSi", "KRFeFe", "KRSiSi", "KRFeSi", "KRMix", "GMassFeFe"):
        assert getattr(fs, name) == pytest.approx(getattr(fc, name), abs=5e-6 * 2 * u.k_fe_sim), name
    assert fs.GMassFeSi == np.float32(7.850 / 3.320) and fc.GMassFeSi == np.float32(2.364458)
    for name in ("ShellBreakFe", "ShellBreakSi", "ShellBreakFeSi1", "ShellBreakFeSi2"):
        assert getattr(fs, name) == getattr(fc, name), name


@pytest.mark.parametrize("sd", [(0.02, 0.002), (0.002, 0.02)])
def test_separate_constants_follow_the_broken_shell(oracle, sd):
    """Unequal iron and silicate stiffness (K_Si = 3 K_Fe, KR_Fe = 0.1,
    KR_Si = 0.3) expose a swapped material: the pure pairs are 2K and 2K KR
    of their own material, the mixed pair the sum of one of each, and in the
    band between the two shell breaks only the material whose shell has
    broken (the larger SD) takes its reduced stiffness."""
    u = oracle.units(2048, k=(1.0e14, 3.0e14), kr=(0.1, 0.3), sd=sd)
    fs, _ = oracle.separate_constants(u)
    kfe, ksi = u.k_fe_sim, u.k_si_sim
    assert ksi == pytest.approx(3 * kfe, rel=1e-12)
    f32 = np.float32
    assert fs.KFeFe == f32(2 * kfe) and fs.KSiSi == f32(2 * ksi) and fs.KFeSi == f32(kfe + ksi)
    assert fs.KSiSi == pytest.approx(3 * fs.KFeFe, rel=1e-6)
    assert fs.KRFeFe == pytest.approx(2 * kfe * 0.1, rel=1e-6)
    assert fs.KRSiSi == pytest.approx(2 * ksi * 0.3, rel=1e-6)
    assert fs.KRFeSi == pytest.approx(kfe * 0.1 + ksi * 0.3, rel=1e-6)
    sd_fe, sd_si = sd
    if sd_fe >= sd_si:  # iron's shell breaks first (outer band r < 1 - SD_Si only for silicate)
        assert fs.KRMix == pytest.approx(kfe + ksi * 0.3, rel=1e-6)
    else:
        assert fs.KRMix == pytest.approx(kfe * 0.1 + ksi, rel=1e-6)
    assert fs.ShellBreakFe == f32(1 - sd_fe) and fs.ShellBreakSi == f32(1 - sd_si)
    assert fs.ShellBreakFeSi1 == f32(1 - min(sd)) and fs.ShellBreakFeSi2 == f32(1 - max(sd))


def test_separate_borders_reproduce_the_mass_split(oracle):
    """Unequal bodies (0.9 and 0.1 Earth masses, 30% and 20% iron): the
    counts reproduce each material's mass share to within one element, each
    body's iron and silicate share to within one element, and the borders
    are the cumulative counts in the order [Fe1, Si1, Fe2, Si2]."""
    n = 20000
    u = oracle.units(n, fraction_earth_mass=(0.9, 0.1), fe=(0.3, 0.2))
    fs, ms = oracle.separate_constants(u)
    mfe_tot, msi_tot = 0.9 * 0.3 + 0.1 * 0.2, 0.9 * 0.7 + 0.1 * 0.8
    ratio = 7.850 / 3.320  # element mass ratio
    nfe, nsi = u.nfe, u.nsi
    assert nfe + nsi == n
    assert abs(nfe * ratio / (nfe * ratio + nsi) - mfe_tot / (mfe_tot + msi_tot)) < 2.0 / n
    assert abs(u.nfe1 - nfe * (0.9 * 0.3) / mfe_tot) < 1.0
    assert abs(u.nsi1 - nsi * (0.9 * 0.7) / msi_tot) < 1.0
    assert (fs.boarder1, fs.boarder2, fs.boarder3) == (u.nfe1, u.nfe1 + u.nsi1, u.nfe1 + u.nsi1 + u.nfe2)
    assert n - fs.boarder3 == u.nsi2
    assert (ms.boarder1, ms.boarder2, ms.boarder3) == (fs.boarder1, fs.boarder2, fs.boarder3)


# ---- body centre of mass / mean velocity (generateTargetImpactorBasicFunctions.h:29-80,119-171)


def test_body_com_and_velocity_closed_form(oracle):
    """Body 2 of a layout [Fe1 Fe1 Si1 | Fe2 Fe2 Fe2 Si2 Si2]: iron mass 2.5,
    silicate 1.0, iron of body 2 at indices 3..5 and silicate at 6..7.  The
    sums are divided by the NOMINAL body mass passed in (10.0, not the
    elements' 9.5), as the reference divides by MassOfBody2."""
    x = np.array([
        [9, 9, 9], [9, 9, 9], [9, 9, 9],          # body 1: must not contribute
        [1, 0, 0], [0, 2, 0], [0, 0, 4],          # Fe2
        [2, 2, 2], [-4, 0, 1],                    # Si2
    ], np.float32)
    v = x[:, [2, 0, 1]] * np.float32(0.5)
    pos, vel = f4(8), f4(8)
    pos[:, :3], vel[:, :3] = x, v
    # sum m x = 2.5 * (1, 2, 4) + (2, 2, 2) + (-4, 0, 1) = (0.5, 7, 13)
    c = oracle.body_com(pos, 3, 6, 8, 2.5, 1.0, 10.0)
    assert np.array_equal(c[:3], [0.05, 0.7, 1.3])
    assert c[3] == pytest.approx(math.sqrt(0.05**2 + 0.7**2 + 1.3**2), rel=1e-15)
    # velocities are (z, x, y) / 2: sum m v = (13, 0.5, 7) / 2
    w = oracle.body_vel(vel, 3, 6, 8, 2.5, 1.0, 10.0)
    assert np.allclose(w[:3], [0.65, 0.025, 0.35], rtol=1e-15, atol=0)
    # an empty (massless) body: all zero, no division
    assert np.array_equal(oracle.body_com(pos, 3, 6, 8, 2.5, 1.0, 0.0), [0, 0, 0, 0])
    # body 1 alone (iron 0..1, silicate 2): (2.5 * 18 + 9) / 6 = 9 per axis
    assert np.array_equal(oracle.body_com(pos, 0, 2, 3, 2.5, 1.0, 6.0)[:3], [9, 9, 9])


def test_body_angular_momentum_and_radius_closed_forms(oracle):
    """getAngularMomentum of a rigidly rotating, drifting body is I w about
    its centre of mass (independent of the drift), and getRadiusOfBody is the
    largest distance from that centre."""
    rng = np.random.default_rng(21)
    x = _blob(rng, 216, (4.0, -1.0, 2.0))
    n, nfe, mfe = len(x), 60, 2.364458
    m = np.where(np.arange(n) < nfe, mfe, 1.0)
    c = (m[:, None] * x).sum(0) / m.sum()
    w = np.array([0.01, 0.04, -0.02])
    v = np.cross(w, x - c) + np.array([1.0, -2.0, 0.5])
    pos, vel = f4(n), f4(n)
    pos[:, :3], vel[:, :3] = x, v
    r = pos[:, :3].astype(np.float64) - c
    I = (m[:, None, None] * ((r * r).sum(1)[:, None, None] * np.eye(3) - r[:, :, None] * r[:, None, :])).sum(0)
    L = oracle.body_angmom(pos, vel, 0, nfe, n, mfe, 1.0, m.sum())
    assert np.allclose(L[:3], I @ w, rtol=2e-5, atol=2e-5 * np.abs(I @ w).max())
    assert L[3] == pytest.approx(np.linalg.norm(L[:3]), rel=1e-14)
    d = np.sqrt(((pos[:, :3].astype(np.float64) - oracle.body_com(pos, 0, nfe, n, mfe, 1.0, m.sum())[:3]) ** 2).sum(1))
    assert oracle.body_radius(pos, 0, nfe, n, mfe, 1.0, m.sum()) == pytest.approx(d.max(), rel=1e-15)
    # a lone far element sets the radius: the centre of three unit masses at
    # (+-1, 0, 0) and (0, 0, 30) is (0, 0, 10), the far one 20 from it
    p3 = f4(3)
    p3[:, :3] = [[-1, 0, 0], [1, 0, 0], [0, 0, 30]]
    assert oracle.body_radius(p3, 0, 0, 3, 2.0, 1.0, 3.0) == 20.0


def _generator_state(rng, counts):
    n = sum(counts)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = rng.uniform(-5, 5, (n, 3))
    vel[:, :3] = rng.normal(0, 0.1, (n, 3))
    tags = np.concatenate([np.full(c, float(t), np.float32) for t, c in enumerate(counts)])
    pos[:, 3] = tags
    vel[:, 3] = np.where((tags == 0) | (tags == 2), 2.364458, 1.0)
    return pos, vel


def test_impact_initialize_reorders_offsets_and_centres(oracle):
    """impactInitialization's set-up: with zero offsets the frame is the
    generator's state permuted [Fe1, Si1, Fe2, Si2] -> [Fe1, Fe2, Si1, Si2]
    (tags 0, 2, 1, 3 in that order); offsets move each body by exactly the
    float of offset / UnitLength (velocities by offset * UnitTime /
    UnitLength); the stats are the mass-weighted centre, mean velocity and
    angular momentum about them (checked with an independent cross-product
    form)."""
    rng = np.random.default_rng(31)
    counts = (7, 20, 5, 13)  # Fe1 Si1 Fe2 Si2
    g_pos, g_vel = _generator_state(rng, counts)
    z = np.zeros(3, np.float32)
    pos, vel, st = oracle.impact_initialize(g_pos, g_vel, *counts, z, z, z, z, 100.0, 3000.0, 2.364458)
    order = np.concatenate([np.arange(0, 7), np.arange(27, 32), np.arange(7, 27), np.arange(32, 45)])
    assert np.array_equal(pos, g_pos[order]) and np.array_equal(vel, g_vel[order])
    assert list(pos[:, 3]) == [0.0] * 7 + [2.0] * 5 + [1.0] * 20 + [3.0] * 13
    ip1 = np.array([-250.0, 10.0, 3.5], np.float32)
    ip2 = np.array([400.0, -20.0, 0.0], np.float32)
    iv1 = np.array([1.5, 0.0, -0.25], np.float32)
    iv2 = np.array([-3.0, 0.5, 0.0], np.float32)
    ul, ut = 1161.65, 2935.86
    pos, vel, st = oracle.impact_initialize(g_pos, g_vel, *counts, ip1, ip2, iv1, iv2, ul, ut, 2.364458)
    body1 = np.isin(pos[:, 3], [0.0, 1.0])
    d1 = (ip1.astype(np.float64) / ul).astype(np.float32)
    d2 = (ip2.astype(np.float64) / ul).astype(np.float32)
    u1 = (iv1.astype(np.float64) * (ut / ul)).astype(np.float32)
    u2 = (iv2.astype(np.float64) * (ut / ul)).astype(np.float32)
    base_p, base_v = g_pos[order][:, :3], g_vel[order][:, :3]
    assert np.array_equal(pos[body1, :3], base_p[body1] + d1) and np.array_equal(pos[~body1, :3], base_p[~body1] + d2)
    assert np.array_equal(vel[body1, :3], base_v[body1] + u1) and np.array_equal(vel[~body1, :3], base_v[~body1] + u2)
    nfe = 12
    m = np.where(np.arange(45) < nfe, 2.364458, 1.0)
    X, V = pos[:, :3].astype(np.float64), vel[:, :3].astype(np.float64)
    c = (m[:, None] * X).sum(0) / m.sum()
    u = (m[:, None] * V).sum(0) / m.sum()
    L = (m[:, None] * np.cross(X, V)).sum(0) - m.sum() * np.cross(c, u)
    assert st[0] == pytest.approx(m.sum(), rel=1e-15)
    assert np.allclose(st[1:4], c, rtol=1e-13, atol=1e-13)
    assert np.allclose(st[4:7], u, rtol=1e-13, atol=1e-13)
    assert np.allclose(st[7:10], L, rtol=1e-10, atol=1e-10 * np.abs(L).max())

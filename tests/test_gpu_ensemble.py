"""The batched ensemble: many impacts of uneven N stepped in one launch per
kernel, captured in a HIP graph.  Each simulation's arithmetic depends only on
itself (fixed chunk length of the batched engine), so its results are bitwise
the same alone (an ensemble of one) and in any mix."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import f4

pytestmark = pytest.mark.gpu


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


def _plain(gi, pos, vel, fc, mc, steps):
    n = len(pos)
    P, V, F = gi.DeviceArray(n, pos), gi.DeviceArray(n, vel), gi.DeviceArray(n, f4(n))
    gi.impact_steps(P, V, F, n, fc, mc, steps)
    gi.sync()
    return P.download(), V.download()


def test_ensemble_result_is_bitwise_the_same_alone_and_in_the_mix(gpu):
    cfg = gpu.ensemble_configs(5, seed=3, n_range=(700, 4100))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    print("n =", list(n), flush=True)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.steps(3)
    ens.steps(4)
    gpu.sync()
    p_all, v_all = P.download(), V.download()
    off = ens.offsets
    for s in range(len(n)):
        a, b = off[s], off[s + 1]
        p1, v1 = _ensemble_of_one(gpu, pos[a:b], vel[a:b], fc[s], mc[s], 7)
        assert np.array_equal(p_all[a:b], p1), s
        assert np.array_equal(v_all[a:b], v1), s
        # the plain single-simulation path agrees to float rounding
        p2, v2 = _plain(gpu, pos[a:b], vel[a:b], fc[s], mc[s], 7)
        assert np.abs(v2[:, :3] - v1[:, :3]).max() <= 1e-4 * np.abs(v1[:, :3]).max()
        assert np.abs(p2[:, :3] - p1[:, :3]).max() <= 1e-5 * np.abs(p1[:, :3]).max()
    ens.close()


def test_inactive_simulations_are_frozen(gpu):
    cfg = gpu.ensemble_configs(4, seed=4, n_range=(600, 2000))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.set_active([1, 0, 1, 0])
    ens.steps(5)
    ens.set_active([1, 1, 1, 1])
    ens.steps(2)
    gpu.sync()
    p_all, v_all = P.download(), V.download()
    off = ens.offsets
    for s, k in zip(range(4), (7, 2, 7, 2)):
        a, b = off[s], off[s + 1]
        p1, v1 = _ensemble_of_one(gpu, pos[a:b], vel[a:b], fc[s], mc[s], k)
        assert np.array_equal(p_all[a:b], p1) and np.array_equal(v_all[a:b], v1), s
    work = ens.work()
    assert work[0] == int(n.sum()) and work[1] == float((n.astype(np.float64) ** 2).sum())
    ens.close()


def test_ensemble_force_parity_against_the_oracle(gpu, oracle):
    """One batched step from a known state: the velocity change of every
    simulation is dt/m times its force, which must match the FP64 oracle."""
    from parity import assert_force_parity, force_metrics

    cfg = gpu.ensemble_configs(3, seed=8, n_range=(800, 2500))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.steps(1)
    gpu.sync()
    v1 = V.download()
    off = ens.offsets
    for s in range(3):
        a, b = off[s], off[s + 1]
        t = np.where(np.arange(b - a) < fc[s].NFe, np.float32(mc[s].DtOverMassFe), np.float32(mc[s].DtOverMassSi))
        F_est = (v1[a:b, :3].astype(np.float64) - vel[a:b, :3]) / t[:, None]
        F64, cond, near = oracle.force_collision_f64(pos[a:b], vel[a:b], fc[s])
        met = force_metrics(F_est, F64, cond, near)
        # v + f*t rounds once in float: allow that rounding on top of the force tolerance
        slack = np.abs(vel[a:b, :3]).max() * 2 ** -23 / t.min()
        err = np.linalg.norm(F_est - F64, axis=1)
        assert np.all(err[~near] <= 1e-5 * cond[~near] + slack), (s, met)
    ens.close()


def test_ensemble_conserved_quantities_match_single(gpu):
    cfg = gpu.ensemble_configs(3, seed=5, n_range=(900, 2500))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    c = ens.conserved()
    off = ens.offsets
    for s in range(3):
        a, b = off[s], off[s + 1]
        Ps, Vs = gpu.DeviceArray(b - a, pos[a:b]), gpu.DeviceArray(b - a, vel[a:b])
        one = gpu.conserved_quantities(Ps, Vs, b - a, fc[s])
        # kinetic energy, momentum, angular momentum: FP64 sums of the same terms
        idx = [0, 3, 4, 5, 6, 7, 8]
        assert np.allclose(c[s][idx], one[idx], rtol=1e-12, atol=1e-12 * np.abs(one).max()), s
        # the potential's float chunk sums are cut differently by the two engines
        assert np.allclose(c[s][1:3], one[1:3], rtol=2e-6), s
    ens.close()


def _run_in_calls(gi, pos, vel, n, fc, mc, calls):
    P, V = gi.DeviceArray(len(pos), pos), gi.DeviceArray(len(pos), vel)
    ens = gi.Ensemble(P, V, n, fc, mc)
    for k in calls:
        ens.steps(k)
    gi.sync()
    out = P.download(), V.download()
    ens.close()
    return out


def test_step_calls_of_any_length_give_the_same_state(gpu):
    """Within a call, steps after the first carry the padded positions from one
    step to the next; the state must not depend on how the steps are cut into
    calls."""
    cfg = gpu.ensemble_configs(3, seed=8, n_range=(900, 3000))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    print("n =", list(n), flush=True)
    ref = _run_in_calls(gpu, pos, vel, n, fc, mc, [10])
    for calls in ([1] * 10, [3, 7], [2, 1, 5, 2]):
        got = _run_in_calls(gpu, pos, vel, n, fc, mc, calls)
        assert np.array_equal(got[0], ref[0]), calls
        assert np.array_equal(got[1], ref[1]), calls


def test_positions_changed_between_calls_are_used(gpu):
    cfg = gpu.ensemble_configs(3, seed=9, n_range=(900, 3000))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.steps(4)
    gpu.sync()
    p, v = P.download(), V.download()
    p[:, 0] += np.float32(0.25)  # shift every body; the caller owns the arrays
    v[:, 1] *= np.float32(0.5)
    P.upload(p)
    V.upload(v)
    ens.steps(5)
    gpu.sync()
    got = P.download(), V.download()
    ens.close()
    ref = _run_in_calls(gpu, p, v, n, fc, mc, [5])
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])


def test_contact_walk_forms_are_bitwise_the_same(gpu):
    """Plans of up to 49,152 particles walk each particle's contact cells on
    nine lanes, larger ones on one lane; both add the same nine cell-row sums
    in the same order.  Here each simulation alone takes the nine-lane walk
    and the mix of all five (about 75,000 particles) the one-lane walk."""
    cfg = gpu.ensemble_configs(5, seed=21, n_range=(14000, 16000))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    assert n.sum() > 49152 and n.max() <= 49152
    print("n =", list(n), flush=True)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.steps(4)
    gpu.sync()
    p_all, v_all = P.download(), V.download()
    off = ens.offsets
    ens.close()
    for s in range(len(n)):
        a, b = off[s], off[s + 1]
        p1, v1 = _ensemble_of_one(gpu, pos[a:b], vel[a:b], fc[s], mc[s], 4)
        assert np.array_equal(p_all[a:b], p1), s
        assert np.array_equal(v_all[a:b], v1), s


def test_many_step_calls_without_waiting_equal_one_call(gpu):
    """Twenty gi_ensemble_steps(1) calls queued without gi_ensemble_wait (each
    records an event; finished ones are retired as calls arrive) leave the
    same state, bitwise, as one gi_ensemble_steps(20)."""
    cfg = gpu.ensemble_configs(6, seed=44, n_range=(1500, 5000))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    out = []
    for split in (True, False):
        P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
        ens = gpu.Ensemble(P, V, n, fc, mc)
        if split:
            for _ in range(20):
                ens.steps(1)
        else:
            ens.steps(20)
        gpu.sync()
        out.append((P.download(), V.download()))
        ens.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_forced_contact_walk_forms_are_bitwise_the_same(gpu):
    """gi_set_contact_walk forces either walk on any plan: a 60,000-particle
    impact (one lane by its size) and a 12-simulation batched step give
    bitwise the same results with nine lanes, one lane and the default."""
    pos, vel, fc, mc, _ = gpu.make_impact(45000, 15000, seed=41, angle_deg=40.0, speed_vesc=1.1)
    n = len(pos)
    print(f"n={n}", flush=True)
    cfg = gpu.ensemble_configs(12, seed=42, n_range=(3000, 9000))
    epos, evel, en, efc, emc, _ = gpu.make_ensemble(cfg)
    out = {}
    try:
        for lanes in (0, 1, 9):
            assert gpu.call("gi_set_contact_walk", lanes) == 0
            P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
            gpu.force_collision(P, V, F, n, fc)
            EP, EV = gpu.DeviceArray(len(epos), epos), gpu.DeviceArray(len(epos), evel)
            ens = gpu.Ensemble(EP, EV, en, efc, emc)
            ens.steps(3)
            gpu.sync()
            out[lanes] = (F.download(), EP.download(), EV.download())
            ens.close()
    finally:
        gpu.call("gi_set_contact_walk", 0)
    for lanes in (1, 9):
        for a, b in zip(out[0], out[lanes]):
            assert np.array_equal(a, b), lanes


def test_resident_pool_table_changes_are_bitwise_packed_ensembles(gpu):
    """An ensemble over a pool of records whose simulation table changes
    (gi_ensemble_set_sims: gaps between simulations, one leaving, one
    arriving into the freed range) steps each simulation bitwise as an
    ensemble of one does, and never writes the records between them."""
    cfg = gpu.ensemble_configs(4, seed=21, n_range=(900, 3500))
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    src = np.concatenate([[0], np.cumsum(n)])
    off = np.array([100, 100 + n[0] + 777, 100 + n[0] + 777 + n[1], 100 + n[0] + 777 + n[1] + n[2] + 50])
    cap = int(off[-1] + n[3] + 300)
    rng = np.random.default_rng(3)
    pool_p = f4(cap)
    pool_p[:, :3] = rng.uniform(-1e4, 1e4, (cap, 3))
    pool_v = f4(cap)
    for s in range(4):
        pool_p[off[s]:off[s] + n[s]] = pos[src[s]:src[s + 1]]
        pool_v[off[s]:off[s] + n[s]] = vel[src[s]:src[s + 1]]
    P, V = gpu.DeviceArray(cap, pool_p), gpu.DeviceArray(cap, pool_v)
    ens = gpu.Ensemble.with_capacity(P, V, cap)
    FC, MC = gpu.ForceCollisionConstants, gpu.MoveCollisionConstants
    ens.set_sims(off, n, (FC * 4)(*fc), (MC * 4)(*mc))
    ens.steps(5)
    ens.wait(0)
    after5 = P.download(), V.download()
    # simulation 1 leaves; 3 moves on alone with 0 and 2 for 7 more steps
    keep = [0, 2, 3]
    ens.set_sims(off[keep], n[keep], (FC * 3)(*[fc[k] for k in keep]), (MC * 3)(*[mc[k] for k in keep]))
    ens.steps(3)
    ens.steps(4)
    ens.wait(1)
    ens.wait(0)
    p12, v12 = P.download(), V.download()
    # a new simulation (a copy of 1's initial state) enters the freed range
    ens.set_sims(off, n, (FC * 4)(*fc), (MC * 4)(*mc))
    P.upload(np.concatenate([p12[:off[1]], pos[src[1]:src[2]], p12[off[1] + n[1]:]]))
    V.upload(np.concatenate([v12[:off[1]], vel[src[1]:src[2]], v12[off[1] + n[1]:]]))
    ens.steps(2)
    gpu.sync()
    p14, v14 = P.download(), V.download()
    ens.close()
    for s, k, (pp, vv) in ((0, 12, (p12, v12)), (2, 12, (p12, v12)), (3, 12, (p12, v12)), (1, 5, after5),
                            (0, 14, (p14, v14)), (1, 2, (p14, v14))):
        p1, v1 = _ensemble_of_one(gpu, pos[src[s]:src[s + 1]], vel[src[s]:src[s + 1]], fc[s], mc[s], k)
        sl = slice(off[s], off[s] + n[s])
        assert np.array_equal(pp[sl], p1) and np.array_equal(vv[sl], v1), (s, k)
    holes = np.ones(cap, bool)
    for s in range(4):
        holes[off[s]:off[s] + n[s]] = False
    assert np.array_equal(p14[holes], pool_p[holes]) and np.array_equal(v14[holes], pool_v[holes])

"""The benchmark workloads at their real sizes, against the oracle:

  single impact, N = 262,144 (196,608 + 65,536, 45 degrees, 1.1 v_esc) with
      the tuning table's plan (the mutual pass in row groups, R = 8, 1,536 j
      per task), and a contact-dominated cloud of the same size;
  the batched ensemble of 64 impacts with N ~ U[4,096, 32,768];
  the request loop over a stream with N log-uniform in [2,048, 65,536].

The reference's force is getForcesCollisionSingleGPU
(Code/ImpactCode/impactAddTime.cu:245-394); truth is the oracle's FP64 twin."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import f4
from parity import assert_force_parity, force_metrics

pytestmark = pytest.mark.gpu


def test_roofline_impact_n262144_against_the_oracle(gpu, oracle):
    pos, vel, fc, _, _ = gpu.make_impact(196608, 65536, seed=20261205, angle_deg=45.0, speed_vesc=1.1)
    n = len(pos)
    print(f"n={n}", flush=True)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.force_collision(P, V, F, n, fc)
    gpu.sync()
    plan = gpu.plan_info()
    assert (plan["kernel"], plan["r"], plan["chunk_len"]) == ("k_mutual_grp", 8, 1536), plan
    Fg = F.download()
    nfe = fc.NFe
    # 5 x 64 rows: target iron, impactor iron (the last irons), first and
    # middle silicate, the impactor's outermost silicate
    for a in (0, nfe - 64, nfe, nfe + (n - nfe) // 2, n - 64):
        F64, cond, near = oracle.force_collision_f64(pos, vel, fc, i_begin=a, i_end=a + 64)
        Fr = oracle.force_collision_f32(pos, vel, fc, a, a + 64)
        sl = slice(a, a + 64)
        met = force_metrics(Fg[sl], F64[sl], cond[sl], near[sl], Fr[sl])
        print(a, met, flush=True)
        assert_force_parity(met)
    tot = Fg[:, :3].astype(np.float64).sum(0)
    assert np.abs(tot).max() <= 1e-5 * np.abs(Fg[:, :3]).astype(np.float64).sum()
    assert np.array_equal(Fg[:, 3], np.zeros(n, np.float32))  # .w never written (the buffer's zeros)


def test_contact_dominated_cloud_n262144_against_the_oracle(gpu, oracle):
    """N = 262,144 through the same plan, in a contact-dominated state: a
    random cloud of density ~0.95 per unit volume (about four neighbours
    within r < 1 per particle) with random velocities, so every row's force
    is mostly contact terms and each pair in contact goes through the
    all-pairs placeholder and its exact correction.  Rows of both materials
    against the FP64 truth and the reference's float routine; third law."""
    n, nfe = 262144, 78643
    rng = np.random.default_rng(20261207)
    side = (n / 0.95) ** (1.0 / 3.0)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = rng.uniform(0.0, side, (n, 3))
    vel[:, :3] = rng.normal(0.0, 2.0, (n, 3))
    fc = gpu.make_impact(10, 0)[2]
    fc.NFe = nfe
    print(f"n={n} side={side:.1f}", flush=True)
    P, V, F = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel), gpu.DeviceArray(n, f4(n))
    gpu.force_collision(P, V, F, n, fc)
    gpu.sync()
    assert gpu.plan_info()["kernel"] == "k_mutual_grp"
    Fg = F.download()
    contact_rows = 0
    for a in (0, nfe - 64, nfe, nfe + (n - nfe) // 2, n - 64):
        F64, cond, near = oracle.force_collision_f64(pos, vel, fc, i_begin=a, i_end=a + 64)
        Fr = oracle.force_collision_f32(pos, vel, fc, a, a + 64)
        sl = slice(a, a + 64)
        met = force_metrics(Fg[sl], F64[sl], cond[sl], near[sl], Fr[sl])
        print(a, met, flush=True)
        assert_force_parity(met)
        contact_rows += int(np.sum(cond[sl] > 10.0 * np.linalg.norm(F64[sl], axis=1)))
    assert contact_rows > 0  # rows whose net force is a small difference of large terms
    tot = Fg[:, :3].astype(np.float64).sum(0)
    assert np.abs(tot).max() <= 1e-5 * np.abs(Fg[:, :3]).astype(np.float64).sum()


def _alone(gi, pos, vel, fc, mc, steps):
    n = len(pos)
    P, V = gi.DeviceArray(n, pos), gi.DeviceArray(n, vel)
    ens = gi.Ensemble(P, V, np.array([n], np.int32), (gi.ForceCollisionConstants * 1)(fc),
                      (gi.MoveCollisionConstants * 1)(mc))
    ens.steps(steps)
    gi.sync()
    out = P.download(), V.download()
    ens.close()
    return out


def test_ensemble_64_uneven_impacts_against_the_oracle(gpu, oracle):
    """The 64-simulation ensemble (N ~ U[4,096, 32,768], impactor fraction
    [0.1, 0.3], angle [30, 60] degrees, speed [1.0, 1.4] v_esc, spin [0, 0.6]
    breakup): one batched step's velocity change is dt/m times the force,
    checked on rows of four simulations against the FP64 oracle; then three
    simulations run alone for the same steps end bitwise where they end in
    the mix."""
    cfg = gpu.ensemble_configs(64, seed=20261205)
    pos, vel, n, fc, mc, _ = gpu.make_ensemble(cfg)
    assert len(n) == 64 and n.min() >= 4096 and n.max() <= 32768
    print(f"64 simulations, sum N = {int(n.sum())}", flush=True)
    P, V = gpu.DeviceArray(len(pos), pos), gpu.DeviceArray(len(pos), vel)
    ens = gpu.Ensemble(P, V, n, fc, mc)
    ens.steps(1)
    gpu.sync()
    v1 = V.download()
    off = ens.offsets
    order = np.argsort(n)
    picks = [int(order[0]), int(order[31]), int(order[-1]), 17]
    for s in picks:
        a, b = off[s], off[s + 1]
        t = np.where(np.arange(b - a) < fc[s].NFe, np.float32(mc[s].DtOverMassFe), np.float32(mc[s].DtOverMassSi))
        F_est = (v1[a:b, :3].astype(np.float64) - vel[a:b, :3]) / t[:, None]
        for r0 in (0, fc[s].NFe - 32, (b - a) - 64):
            F64, cond, near = oracle.force_collision_f64(pos[a:b], vel[a:b], fc[s], i_begin=r0, i_end=r0 + 64)
            sl = slice(r0, r0 + 64)
            # v + f*t rounds once in float: allow that rounding on top of the force tolerance
            slack = np.abs(vel[a:b, :3]).max() * 2 ** -23 / t.min()
            err = np.linalg.norm(F_est[sl] - F64[sl], axis=1)
            assert int(near[sl].sum()) <= 1, (s, r0)
            assert np.all(err[~near[sl]] <= 1e-5 * cond[sl][~near[sl]] + slack), (s, r0, float(err.max()))
    ens.steps(9)
    gpu.sync()
    p_all, v_all = P.download(), V.download()
    ens.close()
    for s in picks[:3]:
        a, b = off[s], off[s + 1]
        p1, v1s = _alone(gpu, pos[a:b], vel[a:b], fc[s], mc[s], 10)
        assert np.array_equal(p_all[a:b], p1) and np.array_equal(v_all[a:b], v1s), s


def _stream(k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(2048), np.log(65536))))
        frac = float(rng.uniform(0.1, 0.3))
        ni = max(1, int(n * frac))
        out.append(dict(n_target=n - ni, n_impactor=ni, seed=7000 + i, angle_deg=float(rng.uniform(30, 60)),
                        speed_vesc=float(rng.uniform(1.0, 1.4)), target_spin=float(rng.uniform(0, 0.6)),
                        steps=int(rng.integers(10, 60)), arrival_s=0.0 if i < 3 else 0.02 * i))
    return out


def test_serving_stream_alone_and_in_the_mix_and_against_the_oracle(gpu, oracle):
    """A request stream with N log-uniform in [2,048, 65,536]: every result is
    bitwise the same served alone; the smallest request's final state follows
    the oracle's float loop (force, move) for its steps to the force
    tolerance."""
    reqs = _stream(6, seed=20261207)
    sizes = [q["n_target"] + q["n_impactor"] for q in reqs]
    assert min(sizes) >= 2048 - 1 and max(sizes) <= 65536
    print("sizes", sizes, flush=True)
    res, summ, mixed = gpu.serve(reqs, quantum_steps=16, keep_final=True)
    assert summ["completed"] == len(reqs) and summ["peak_resident_sims"] >= 2
    for i, q in enumerate(reqs):
        _, s1, fin = gpu.serve([dict(q, arrival_s=0.0)], quantum_steps=23, keep_final=True)
        assert s1["completed"] == 1
        assert np.array_equal(mixed[i][0], fin[0][0]) and np.array_equal(mixed[i][1], fin[0][1]), i
    k = int(np.argmin(sizes))
    q = reqs[k]
    pos, vel, fc, mc, _ = gpu.make_impact(q["n_target"], q["n_impactor"], seed=q["seed"], angle_deg=q["angle_deg"],
                                          speed_vesc=q["speed_vesc"], target_spin=q["target_spin"])
    po, vo, _ = oracle.impact_steps_f32(pos, vel, fc, mc, q["steps"])
    p, v = mixed[k]
    assert np.array_equal(p[:, 3], pos[:, 3]) and np.array_equal(v[:, 3], vel[:, 3])
    assert np.abs(v[:, :3] - vo[:, :3]).max() <= 1e-4 * np.abs(vo[:, :3]).max()
    assert np.abs(p[:, :3] - po[:, :3]).max() <= 1e-5 * np.abs(po[:, :3]).max()

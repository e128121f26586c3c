"""The request loop (gi_serve_run): continuous batching of impact jobs with
admission by capacity, device-resident state and per-request FP64
diagnostics.  Each result is bitwise the same alone and in the mix."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(k=6, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(rng.integers(600, 3000))
        ni = max(1, n // 4)
        out.append(dict(n_target=n - ni, n_impactor=ni, seed=1000 + i, angle_deg=float(rng.uniform(30, 60)),
                        speed_vesc=float(rng.uniform(1.0, 1.4)), target_spin=float(rng.uniform(0, 0.6)),
                        steps=int(rng.integers(20, 150)),
                        # three arrive together (they must share the device), the rest later
                        arrival_s=0.0 if i < 3 else 0.05 * (i - 2)))
    return out


def test_serve_completes_every_request_in_order(gpu):
    reqs = _requests()
    res, summ = gpu.serve(reqs, quantum_steps=25)
    print(summ, flush=True)
    assert summ["completed"] == len(reqs)
    for r, q in zip(res, reqs):
        assert r["steps"] == q["steps"] and r["n"] == q["n_target"] + q["n_impactor"]
        assert r["admit_s"] >= r["arrival_s"] - 1e-9
        assert r["finish_s"] >= r["admit_s"]
        assert r["time_to_result_s"] == pytest.approx(r["finish_s"] - r["arrival_s"])
        assert np.isfinite(r["energy0"]) and np.isfinite(r["energy1"])
        assert np.isfinite(r["angmom_rel_change"])
    assert summ["p50_time_to_result_s"] <= summ["p99_time_to_result_s"] <= summ["max_time_to_result_s"] + 1e-12
    assert summ["peak_resident_sims"] >= 2  # requests really overlapped on the device
    assert summ["sim_steps_per_s"] > 0


def test_each_result_is_bitwise_the_same_alone_and_in_the_mix(gpu):
    reqs = _requests(5, seed=12)
    _, _, mixed = gpu.serve(reqs, quantum_steps=16, keep_final=True)
    for i, q in enumerate(reqs):
        alone = dict(q, arrival_s=0.0)
        _, summ, fin = gpu.serve([alone], quantum_steps=37, keep_final=True)
        assert summ["completed"] == 1
        assert np.array_equal(mixed[i][0], fin[0][0]), i
        assert np.array_equal(mixed[i][1], fin[0][1]), i


def test_admission_waits_for_capacity(gpu):
    reqs = [dict(n_target=2000, n_impactor=500, seed=7 + i, steps=30, arrival_s=0.0) for i in range(3)]
    res, summ = gpu.serve(reqs, max_particles=4000, quantum_steps=10)
    assert summ["completed"] == 3
    assert summ["peak_resident_particles"] <= 4000
    assert summ["peak_resident_sims"] == 1
    admits = sorted(r["admit_s"] for r in res)
    assert admits[1] > admits[0] and admits[2] > admits[1]


def test_request_larger_than_capacity_is_rejected(gpu):
    with pytest.raises(gpu.GIError) as e:
        gpu.serve([dict(n_target=5000, n_impactor=1000, steps=1)], max_particles=4000)
    assert e.value.code == -4

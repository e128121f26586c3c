"""Host side of libgiant_impact.so on the CPU: units, constants, parameter
files, synthetic bodies, argument validation and the exported ABI."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO, f4

GOLDEN = REPO / "tests" / "golden"


def _bytes(s) -> bytes:
    return bytes(memoryview(s))


@pytest.mark.parametrize("n", [512, 2048, 8192, 65536, 262144, 524288])
def test_units_and_constants_match_the_oracle(gi, oracle, n):
    u = gi.units(gi.generator_setup_defaults(n))
    o = oracle.units(n)
    assert (u.nfe, u.nsi, u.nfe1, u.nsi1, u.nfe2, u.nsi2) == (o.nfe, o.nsi, o.nfe1, o.nsi1, o.nfe2, o.nsi2)
    for a, b in [("unit_length", "unit_length"), ("unit_mass", "unit_mass"), ("unit_time", "unit_time"),
                 ("mass_fe", "mass_fe"), ("k_fe", "k_fe_sim"), ("k_si", "k_si_sim"),
                 ("mass_of_body1", "mass_of_body1")]:
        assert getattr(u, a) == getattr(o, b), a
    fc, mc = gi.impact_constants(u)
    ofc, omc = oracle.impact_constants(o)
    assert _bytes(fc) == _bytes(ofc) and _bytes(mc) == _bytes(omc)
    fs, ms = gi.separate_constants(u)
    ofs, oms = oracle.separate_constants(o)
    assert _bytes(fs) == _bytes(ofs) and _bytes(ms) == _bytes(oms)


def test_struct_sizes_match_the_reference_layout(gi):
    assert C.sizeof(gi.ForceCollisionConstants) == 56
    assert C.sizeof(gi.MoveCollisionConstants) == 16
    assert C.sizeof(gi.ForceSeparateConstants) == 64
    assert C.sizeof(gi.MoveSeparateConstants) == 24


def test_fractions_must_add_to_one(gi):
    s = gi.generator_setup_defaults(1024)
    s.fraction_si_body1 = 0.7
    u = gi.Units()
    assert gi.load().gi_units_compute(C.byref(s), C.byref(u)) == -1


def test_setup_file_is_read_positionally(gi):
    """The template text of setupGeneratingTargetImpactor, key names ignored."""
    s = gi.GeneratorSetup()
    assert gi.load().gi_generator_setup_read(str(GOLDEN / "setupGeneratingTargetImpactor").encode(), C.byref(s)) == 0
    d = gi.generator_setup_defaults(1024)
    assert _bytes(s) == _bytes(d)
    assert gi.load().gi_generator_setup_read(str(GOLDEN / "missing_file").encode(), C.byref(s)) == -5
    trunc = GOLDEN / "setupGeneratingTargetImpactor.truncated"
    assert gi.load().gi_generator_setup_read(str(trunc).encode(), C.byref(s)) == -6


def test_parameter_file_round_trip(gi, tmp_path):
    u = gi.units(gi.generator_setup_defaults(2048))
    path = str(tmp_path / "targetImpactorParameters").encode()
    assert gi.load().gi_parameters_write(path, C.byref(u), 2048) == 0
    r, n = gi.Units(), C.c_int()
    assert gi.load().gi_parameters_read(path, C.byref(r), C.byref(n)) == 0
    assert n.value == 2048
    assert r.mass_fe == 2.364458 and r.mass_si == 1.0 and r.diameter == 1.0
    assert r.k_fe == round(u.k_fe, 6) and r.sd_si == 0.002
    assert (r.nfe1, r.nsi1, r.nfe2, r.nsi2) == (u.nfe1, u.nsi1, u.nfe2, u.nsi2)
    assert r.nfe == u.nfe and r.nsi == u.nsi
    # the impact constants from the file equal those computed directly
    fc1, _ = gi.impact_constants(u)
    r.mass_fe, r.mass_si = r.mass_fe, r.mass_si
    fc2, _ = gi.impact_constants(r)
    assert _bytes(fc1) == _bytes(fc2)


@pytest.mark.parametrize("nt,ni,spin", [(1536, 512, 0.0), (800, 200, 0.5), (4000, 0, 0.3)])
def test_synthetic_impact_layout_and_kinematics(gi, nt, ni, spin):
    pos, vel, fc, mc, info = gi.make_impact(nt, ni, seed=11, angle_deg=45, speed_vesc=1.2, target_spin=spin)
    n = nt + ni
    assert info.n == n and fc.NFe == info.nfe == info.nfe1 + info.nfe2 and mc.NFe == fc.NFe
    assert info.nfe1 == int(np.floor(0.3 * nt + 0.5))
    # layout [Fe1, Fe2, Si1, Si2] with the generator's tags in pos.w and masses in vel.w
    tags = np.concatenate([np.full(info.nfe1, 0), np.full(info.nfe2, 2), np.full(info.nsi1, 1), np.full(info.nsi2, 3)])
    assert np.array_equal(pos[:, 3], tags.astype(np.float32))
    assert np.all(vel[: fc.NFe, 3] == np.float32(fc.GMassFeSi)) and np.all(vel[fc.NFe :, 3] == 1.0)
    m = np.where(np.arange(n) < fc.NFe, float(fc.GMassFeSi), 1.0)
    com = (m[:, None] * pos[:, :3]).sum(0) / m.sum()
    mom = (m[:, None] * vel[:, :3]).sum(0)
    assert np.abs(com).max() < 1e-4
    assert np.abs(mom).max() < 1e-3 * m.sum()
    # jittered lattice: no pair closer than 1 - 2*sqrt(3)*jitter
    from scipy.spatial import cKDTree

    t = cKDTree(pos[:, :3])
    assert len(t.query_pairs(1 - 2 * np.sqrt(3) * 0.1 - 1e-6)) == 0
    # iron is the core of the target: innermost by radius from the body centre
    body = np.concatenate([np.arange(info.nfe1), np.arange(fc.NFe, fc.NFe + info.nsi1)])
    c = pos[body, :3].mean(0)
    r = np.linalg.norm(pos[body, :3] - c, axis=1)
    assert r[: info.nfe1].max() < np.percentile(r[info.nfe1 :], 10) + 1.0
    if spin > 0:
        mb = m[body, None]
        cm = (mb * pos[body, :3]).sum(0) / mb.sum()
        vm = (mb * vel[body, :3]).sum(0) / mb.sum()
        L = (mb * np.cross(pos[body, :3] - cm, vel[body, :3] - vm)).sum(0)
        # rigid spin about +z: L = I omega, off-axis parts only from the small
        # products of inertia of the jittered lattice
        assert L[2] > 0 and abs(L[0]) < 1e-2 * L[2] and abs(L[1]) < 1e-2 * L[2]
        r_perp2 = ((pos[body, :2] - cm[:2]) ** 2).sum(1)
        assert L[2] == pytest.approx((m[body] * r_perp2).sum() * info.target_omega, rel=1e-3)


def test_synthetic_impact_is_seeded(gi):
    a = gi.make_impact(300, 100, seed=5)[0]
    b = gi.make_impact(300, 100, seed=5)[0]
    c = gi.make_impact(300, 100, seed=6)[0]
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_invalid_arguments_are_rejected_before_touching_a_device(gi):
    lib = gi.load()
    fc, mc = gi.impact_constants(gi.units(gi.generator_setup_defaults(1024)))
    vp = C.c_void_p
    assert lib.gi_force_collision(vp(1), vp(1), vp(1), -1, C.byref(fc), None) == -1
    fc.NFe = 99
    assert lib.gi_force_collision(vp(1), vp(1), vp(1), 50, C.byref(fc), None) == -1
    assert b"NFe" in lib.gi_last_error_string()
    assert lib.gi_force_collision(None, None, None, 10, None, None) == -1
    assert lib.gi_force_collision(None, None, None, 0, C.byref(fc), None) in (0, -1)
    fs = gi.ForceSeparateConstants()
    fs.boarder1, fs.boarder2, fs.boarder3 = 5, 3, 8
    assert lib.gi_force_separate(vp(1), vp(1), vp(1), 10, C.byref(fs), None) == -1
    assert lib.gi_move_collision(vp(1), vp(1), vp(1), -3, C.byref(mc), None) == -1
    assert lib.gi_check_block_multiple(1000) == -2 and lib.gi_check_block_multiple(1024) == 0
    assert lib.gi_set_tiling(3, 0, 0, 0, 0) == -1 and lib.gi_set_tiling(4, 5, 0, 0, 0) == -1
    assert lib.gi_set_tiling(2, 4, 0, 0, 0) == -1  # the row-group form needs rows of 256 or 512
    assert lib.gi_set_tiling(0, 0, 0, 0, 0) == 0
    assert lib.gi_set_dispatch_limit(-1) == -1 and lib.gi_set_dispatch_limit(0) == 0
    assert lib.gi_set_contact_walk(4) == -1 and lib.gi_set_contact_walk(-1) == -1
    assert lib.gi_set_contact_walk(9) == 0 and lib.gi_set_contact_walk(0) == 0
    assert lib.gi_impact_steps(vp(1), vp(1), vp(1), 10, C.byref(fc), C.byref(mc), 5, 1, None, None, None) == -1
    assert lib.gi_last_error() == -1
    lib.gi_clear_error()
    assert lib.gi_last_error() == 0 and lib.gi_last_error_string() == b""
    assert lib.gi_set_tiling(3, 0, 0, 0, 0) == -1 and lib.gi_last_error() == -1
    lib.gi_clear_error()


def _declared_functions() -> list[str]:
    names = []
    for h in sorted((REPO / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(gi_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(gi):
    declared = _declared_functions()
    assert len(declared) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(gi.LIB_PATH)], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (gi_[a-z0-9_]+)", out.stdout))
    missing = [d for d in declared if d not in exported]
    assert not missing, f"declared but not exported: {missing}"
    lib = gi.load()
    for d in declared:
        assert getattr(lib, d) is not None


def test_spin_body_is_bitwise_the_oracle(gi, oracle):
    u = gi.units(gi.generator_setup_defaults(2048))
    n = u.nfe1 + u.nsi1 + u.nfe2 + u.nsi2
    pos, vel = f4(n), f4(n)
    assert gi.load().gi_make_raw_bodies(C.byref(u), 3, 0.1, pos.ctypes.data, vel.ctypes.data) == 0
    rng = np.random.default_rng(1)
    vel[:, :3] = rng.normal(0, 0.1, (n, 3))
    n1 = u.nfe1 + u.nsi1
    spin = np.array([0.2, 1.0, -0.3, 0.043], np.float32)
    v_gi = vel.copy()
    assert gi.load().gi_spin_body(pos.ctypes.data, v_gi.ctypes.data, 0, n1, u.nfe1, u.mass_fe, u.mass_si,
                                  u.mass_of_body1, spin.ctypes.data) == 0
    v_or = oracle.spin_body(pos, vel, 0, n1, u.nfe1, u.mass_fe, u.mass_si, u.mass_of_body1, spin)
    assert np.array_equal(v_gi, v_or)
    assert np.array_equal(v_gi[n1:], vel[n1:])  # body 2 untouched
    zero = np.zeros(4, np.float32)
    assert gi.load().gi_spin_body(pos.ctypes.data, v_gi.ctypes.data, 0, n1, u.nfe1, u.mass_fe, u.mass_si,
                                  u.mass_of_body1, zero.ctypes.data) == -1


def test_zero_out_drift_uses_the_nominal_body_masses(gi, oracle):
    u = gi.units(gi.generator_setup_defaults(1024))
    n = u.nfe1 + u.nsi1 + u.nfe2 + u.nsi2
    pos, vel = f4(n), f4(n)
    gi.load().gi_make_raw_bodies(C.byref(u), 4, 0.1, pos.ctypes.data, vel.ctypes.data)
    rng = np.random.default_rng(2)
    pos[:, :3] += rng.normal(0, 3.0, (1, 3)).astype(np.float32)
    vel[:, :3] = rng.normal(0.5, 0.2, (n, 3))
    p, v = pos.copy(), vel.copy()
    assert gi.load().gi_zero_out_drift(p.ctypes.data, v.ctypes.data, C.byref(u)) == 0
    b1, b2, b3 = u.nfe1, u.nfe1 + u.nsi1, u.nfe1 + u.nsi1 + u.nfe2
    for (a, f, e, M) in ((0, b1, b2, u.mass_of_body1), (b2, b3, n, u.mass_of_body2)):
        com = oracle.body_com(pos, a, f, e, u.mass_fe, u.mass_si, M)[:3]
        vm = oracle.body_com(vel, a, f, e, u.mass_fe, u.mass_si, M)[:3]
        assert np.array_equal(p[a:e, :3], (pos[a:e, :3].astype(np.float64) - com).astype(np.float32))
        assert np.array_equal(v[a:e, :3], (vel[a:e, :3].astype(np.float64) - vm).astype(np.float32))


def test_every_tuning_row_is_accepted(gi):
    """Each row of tuning/gfx950.txt reaches the planner: the tiling chosen
    for a size at a row's n_max is that row's (a rejected row would fall
    through to the next one)."""
    rows = []
    for line in (REPO / "giant-impact_amd" / "tuning" / "gfx950.txt").read_text().splitlines():
        parts = line.split()
        if not parts or parts[0].startswith("#") or parts[0] == "B":
            continue
        rows.append([int(x) for x in parts])
    assert rows
    for n_max, r, variant, chunk_len, target_tasks, min_chunk in rows:
        t = gi.tiling_for(n_max)
        assert (t["r"], t["variant"], t["chunk_len"], t["target_tasks"], t["min_chunk"]) == (
            r, variant, chunk_len, target_tasks, min_chunk), (n_max, t)
    gi.call("gi_set_tiling", 8, 2, 512, 0, 64)
    try:
        assert gi.tiling_for(1000)["variant"] == 2
        gi.call("gi_set_tiling", 1, 3, 0, 0, 0)
        assert gi.tiling_for(1000)["variant"] == 3
    finally:
        gi.call("gi_set_tiling", 0, 0, 0, 0, 0)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with a message before it
    touches a device (the contract's multi-GPU launch gets a clear answer)."""
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_mutual_plans_stay_in_bounds(gi):
    """Every index the mutual pass's kernels derive from a plan (k_mutual or
    the row-group kernel, then the per-particle sums) lies inside the buffers
    the plan sizes, and every (row, 64-block) pair holding a real particle is
    taken by exactly one task: ragged, one-material and multi-simulation
    cases at every row size and chunk length, the single-simulation plans'
    half-length tail pieces and the batched engine's per-simulation pieces
    included (host only, gi_plan_check)."""
    lib = gi.load()
    rng = np.random.default_rng(7)
    cases = [([1], [0]), ([2], [1]), ([65], [0]), ([257], [100]), ([1100], [512]), ([2100], [1600]),
             ([5000], [3000]), ([4097], [1229]), ([262144], [78643])]
    for _ in range(40):
        k = int(rng.integers(1, 6))
        ns = [int(rng.integers(1, 30000)) for _ in range(k)]
        cases.append((ns, [int(rng.integers(0, x + 1)) for x in ns]))
    for ns, nfes in cases:
        a, b = (C.c_int * len(ns))(*ns), (C.c_int * len(ns))(*nfes)
        for r, v in ((8, 4), (4, 4), (8, 2), (4, 2), (1, 2)):
            for L, batched in ((64, 0), (256, 0), (1536, 0), (1024, 1), (192, 1)):
                rc = lib.gi_plan_check(len(ns), a, b, r, v, C.c_longlong(L), batched)
                assert rc == 0, (ns, nfes, r, v, L, lib.gi_last_error_string())
    assert lib.gi_plan_check(1, (C.c_int * 1)(10), (C.c_int * 1)(3), 2, 4, C.c_longlong(64), 0) == -1  # R >= 4

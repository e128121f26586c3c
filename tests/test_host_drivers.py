"""The host side of the impact drivers, checked on the CPU against the oracle:
impactInitialization's set-up (offsets, reorder, centring record and the
initialStatsFile) and the parsing of setupImpactAddTime."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest


def _write_bodies(gi, d, u, n, pos, vel):
    d.mkdir(parents=True, exist_ok=True)
    gi.call("gi_parameters_write", str(d / "targetImpactorParameters").encode(), C.byref(u), n)
    with open(d / "targetImpactorInitialPosVel", "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.ascontiguousarray(pos, np.float32).tobytes())
        f.write(np.ascontiguousarray(vel, np.float32).tobytes())


def _values(path):
    """The number(s) after '=' on each line that has one, by label."""
    out = {}
    for line in path.read_text().splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            v = v.strip().strip("()")
            out[" ".join(k.split())] = [float(x) for x in v.split(",")]
    return out


@pytest.mark.parametrize("n,seed", [(1024, 3), (3000, 4)])
def test_impact_initialize_matches_the_oracle(gi, oracle, tmp_path, n, seed):
    u = gi.units(gi.generator_setup_defaults(n))
    counts = (u.nfe1, u.nsi1, u.nfe2, u.nsi2)
    rng = np.random.default_rng(seed)
    g_pos = np.zeros((n, 4), np.float32)
    g_vel = np.zeros((n, 4), np.float32)
    g_pos[:, :3] = rng.uniform(-6, 6, (n, 3))
    g_vel[:, :3] = rng.normal(0, 0.05, (n, 3))
    g_pos[:, 3] = np.repeat(np.arange(4, dtype=np.float32), counts)
    g_vel[:, 3] = rng.uniform(1, 3, n)
    bodies = tmp_path / "bodies"
    _write_bodies(gi, bodies, u, n, g_pos, g_vel)
    ip1, ip2 = [-9000.0, -150.5, 12.25], [9000.0, 150.5, 0.0]
    iv1, iv2 = [1.75, 0.0, -0.125], [-5.5, 0.25, 0.0]
    setup = tmp_path / "setupImpactInitialization"
    setup.write_text("".join(f"{a} {c} = {v};\n" for a, vals in (("P1", ip1), ("P2", ip2), ("V1", iv1), ("V2", iv2))
                             for c, v in zip("xyz", vals)))
    run = tmp_path / "run"
    run.mkdir()
    gi.impact_initialize(bodies, setup, run)
    # the oracle reads the same %f-rounded parameters the set-up reads
    rf = oracle.load().or_round_f
    ul, ut, um, mfe, msi = (rf(x) for x in (u.unit_length, u.unit_time, u.unit_mass, u.mass_fe, u.mass_si))
    pos, vel, st = oracle.impact_initialize(g_pos, g_vel, *counts, *(np.array(a, np.float32) for a in (ip1, ip2, iv1, iv2)),
                                            ul, ut, mfe, msi)
    frames = gi.read_frames(run / "impactPosVel", n)
    assert len(frames) == 1 and frames[0][0] == 0.0
    assert np.array_equal(frames[0][1], pos) and np.array_equal(frames[0][2], vel)
    # viewCenteringInfo: the float centre and mean velocity, printed %f
    text = (run / "viewCenteringInfo").read_text()
    want = "".join(f"\n {k} = {float(np.float32(v)):f}\n" for k, v in zip(
        ["InitialCenterOfMassX", "InitialCenterOfMassY", "InitialCenterOfMassZ",
         "InitialLinearVelocityX", "InitialLinearVelocityY", "InitialLinearVelocityZ"], st[1:7]))
    assert text == want
    # initialStatsFile: the same stats in km, kg, km/s
    vals = _values(run / "initialStatsFile")
    vc, lc = ul / ut, um * ul * ul / ut
    assert np.allclose(vals["Center of mass of the entire system"], st[1:4] * ul, rtol=0, atol=1e-6)
    assert vals["Total mass of the entire system"][0] == pytest.approx(st[0] * um, rel=1e-6)
    assert np.allclose(vals["Linear velocity of the entire system system"], st[4:7] * vc, rtol=0, atol=1e-6)
    L = st[7:10] * lc
    assert np.allclose(vals["Angular momentum of the entire system system"], L, rtol=1e-6, atol=0)
    assert vals["Magnitude of the angular momentum of the entire system"][0] == pytest.approx(np.linalg.norm(L), rel=1e-6)


def _add_time_files(gi, tmp_path, rates):
    n = 256
    u = gi.units(gi.generator_setup_defaults(n))
    d = tmp_path / "p"
    d.mkdir(parents=True)
    gi.call("gi_parameters_write", str(d / "targetImpactorParameters").encode(), C.byref(u), n)
    run = tmp_path / "r"
    run.mkdir()
    (run / "viewCenteringInfo").write_text("".join(f"\n X{k} = 0.000000\n" for k in range(6)))
    (run / "impactPosVel").write_bytes(np.zeros(1 + 8 * n, np.float32).tobytes())
    dt, record, recenter = rates
    setup = tmp_path / "setupImpactAddTime"
    setup.write_text(f"AdditionalTime = 0.001;\nDt = {dt};\nDrawRate = 10;\nRecordRate = {record};\n"
                     f"RecenterRate = {recenter};\n")
    return d, setup, run


def test_add_time_rates_of_zero_or_less_mean_never(gi, tmp_path):
    """setupImpactAddTime with RecordRate = 0 and RecenterRate = -1 is a valid
    run that never records or recentres (the reference's counters never equal
    such a rate); only Dt <= 0 is rejected, before any device work."""
    lib = gi.load()
    info = gi.ImpactRunInfo()
    d, setup, run = _add_time_files(gi, tmp_path, (0.0, 10, 10))
    assert lib.gi_impact_run(str(d).encode(), str(setup).encode(), str(run).encode(), C.byref(info)) == -1
    d, setup, run = _add_time_files(gi, tmp_path / "b", (0.002, 0, -1))
    rc = lib.gi_impact_run(str(d).encode(), str(setup).encode(), str(run).encode(), C.byref(info))
    if gi.device_count() == 0:
        assert rc not in (0, -1, -5, -6)  # parsed and accepted; stopped by the missing device
    else:
        assert rc == 0 and info.frames_written == 0

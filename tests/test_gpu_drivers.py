"""The headless drivers: generation phases (damp, rest, spin, rest, drift
removal), impact set-up and the impact run with the reference's file formats
(frames, resume, recentring)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import f4

pytestmark = pytest.mark.gpu


def _short_setup(gi, n=1024, hours=(0.1, 0.1, 0.1)):
    s = gi.generator_setup_defaults(n)
    s.damp_time, s.damp_rest_time, s.spin_rest_time = hours
    return s


def _phase_counts(u, dt):
    t, nd, nr, ns = 0.0, 0, 0, 0
    end = u.damp_time + u.damp_rest_time + u.spin_rest_time
    while True:
        t += dt
        if t < u.damp_time:
            nd += 1
        elif t < u.damp_time + u.damp_rest_time:
            nr += 1
        else:
            ns += 1
        if end < t:
            return nd, nr, ns, t


def test_generation_phases_match_the_oracle(gpu, oracle):
    s = _short_setup(gpu)
    pos, vel, u, info = gpu.generate_bodies(s, seed=5)
    nd, nr, ns, t = _phase_counts(u, s.dt)
    assert (info.damp_steps, info.rest_steps, info.spin_rest_steps) == (nd, nr, ns)
    assert info.run_time == t
    # the same phases on the CPU oracle from the same raw bodies
    n = info.n
    p0, v0 = f4(n), f4(n)
    gpu.load().gi_make_raw_bodies(C.byref(u), 5, 0.1, p0.ctypes.data, v0.ctypes.data)
    fs, ms = gpu.separate_constants(u, s.dt)
    p, v = p0, v0
    n1 = u.nfe1 + u.nsi1
    for k in range(nd + nr + ns):
        F = oracle.force_separate_f32(p, v, fs)
        if k == nd + nr:
            v = oracle.spin_body(p, v, 0, n1, u.nfe1, u.mass_fe, u.mass_si, u.mass_of_body1,
                                 [s.initial_spin1[0], s.initial_spin1[1], s.initial_spin1[2], u.spin1_w])
            v = oracle.spin_body(p, v, n1, n, n1 + u.nfe2, u.mass_fe, u.mass_si, u.mass_of_body2,
                                 [s.initial_spin2[0], s.initial_spin2[1], s.initial_spin2[2], u.spin2_w])
        d = (s.damp_rate_body1, s.damp_rate_body2) if k < nd else (0.0, 0.0)
        p, v = oracle.move_damped_separate_f32(p, v, F, ms, *d)
    gpu.load().gi_zero_out_drift(p.ctypes.data, v.ctypes.data, C.byref(u))
    scale = np.abs(p[:, :3]).max()
    print("max |dx|", np.abs(pos[:, :3] - p[:, :3]).max(), "scale", scale)
    assert np.abs(pos[:, :3] - p[:, :3]).max() <= 1e-4 * scale
    assert np.abs(vel[:, :3] - v[:, :3]).max() <= 1e-3 * np.abs(v[:, :3]).max()
    # drift removed, tags and masses passed through
    assert np.array_equal(pos[:, 3], p0[:, 3]) and np.array_equal(vel[:, 3], v0[:, 3])


def _stats_sections(text):
    """Body sections of targetImpactorStats: {body: {field: [numbers]}}."""
    import re

    out, cur = {}, None
    for line in text.splitlines():
        m = re.search(r"Stats for Body(\d)", line)
        if m:
            cur = out.setdefault(int(m.group(1)), {})
            continue
        if cur is None:
            continue
        nums = [float(x) for x in re.findall(r"[-+]?\d+\.\d+(?:e[-+]\d+)?", line)]
        for key in ("Mass =", "Radius =", "center of mass", "The average linear velocity =",
                    "The angular momentum =", "magitude of the angular"):
            if key in line:
                cur[key] = nums
    return out


def test_target_impactor_stats_file_against_the_oracle(gpu, oracle, tmp_path):
    """targetImpactorStats (generateTargetImpactorBasicFunctions.h:590-686): each
    body's mass, radius, centre of mass, mean velocity and angular momentum of
    the written final state, in the file's units, against the oracle's
    restatements of the reference's helpers."""
    s = _short_setup(gpu, n=1024)
    pos, vel, u, _ = gpu.generate_bodies(s, seed=8, out_dir=tmp_path)
    sec = _stats_sections((tmp_path / "targetImpactorStats").read_text())
    assert set(sec) == {1, 2}
    b1, b2 = u.nfe1, u.nfe1 + u.nsi1
    b3, n = b2 + u.nfe2, b2 + u.nfe2 + u.nsi2
    ranges = {1: (0, b1, b2, u.nfe1, u.nsi1, u.mass_of_body1), 2: (b2, b3, n, u.nfe2, u.nsi2, u.mass_of_body2)}
    L_c, V_c = u.unit_length, u.unit_length / u.unit_time
    M_c = u.unit_mass
    J_c = u.unit_mass * u.unit_length * u.unit_length / u.unit_time
    for b, (a0, a1, a2, nfe, nsi, nominal) in ranges.items():
        com = oracle.body_com(pos, a0, a1, a2, u.mass_fe, u.mass_si, nominal)[:3]
        v = oracle.body_vel(vel, a0, a1, a2, u.mass_fe, u.mass_si, nominal)[:3]
        L = oracle.body_angmom(pos, vel, a0, a1, a2, u.mass_fe, u.mass_si, nominal)[:3]
        r = oracle.body_radius(pos, a0, a1, a2, u.mass_fe, u.mass_si, nominal)
        got = sec[b]
        assert got["Mass ="][0] == pytest.approx((u.mass_fe * nfe + u.mass_si * nsi) * M_c, rel=1e-5)
        assert got["Radius ="][0] == pytest.approx(r * L_c, rel=1e-5, abs=1e-5)
        assert np.allclose(got["center of mass"][:3], com * L_c, rtol=1e-5, atol=2e-6 * r * L_c)
        vs = np.abs(v).max() * V_c + 1e-9
        assert np.allclose(got["The average linear velocity ="][:3], v * V_c, rtol=1e-5, atol=1e-5 * vs + 2e-6)
        Ls = np.linalg.norm(L) * J_c
        assert np.allclose(got["The angular momentum ="][:3], L * J_c, rtol=1e-5, atol=1e-5 * Ls)
        assert got["magitude of the angular"][0] == pytest.approx(Ls, rel=1e-5)


def test_impact_pipeline_files_resume_and_recenter(gpu, tmp_path):
    s = _short_setup(gpu, n=1024)
    bodies = tmp_path / "bodies"
    bodies.mkdir()
    _, _, u, _ = gpu.generate_bodies(s, seed=6, out_dir=bodies)
    n = 1024
    assert (bodies / "targetImpactorInitialPosVel").stat().st_size == 4 + 32 * n
    # place the bodies 2 body-radii apart, moving toward each other (km, km/s)
    r_km = 8.0 * u.unit_length
    setup = tmp_path / "setupImpactInitialization"
    setup.write_text(
        f"Body 1 position\nx = {-r_km};\ny = -20.0;\nz = 0.0;\n"
        f"Body 2 position\nx = {r_km};\ny = 20.0;\nz = 0.0;\n"
        "Body 1 velocity\nx = 2.0;\ny = 0.0;\nz = 0.0;\n"
        "Body 2 velocity\nx = -2.0;\ny = 0.0;\nz = 0.0;\n"
    )
    run = tmp_path / "run"
    run.mkdir()
    gpu.impact_initialize(bodies, setup, run)
    frames = gpu.read_frames(run / "impactPosVel", n)
    assert len(frames) == 1 and frames[0][0] == 0.0
    fc, _ = gpu.impact_constants(u)
    # frame 0 is in the impact layout [Fe1, Fe2, Si1, Si2]
    tags = frames[0][1][:, 3]
    nfe = u.nfe1 + u.nfe2
    assert np.all(np.isin(tags[:nfe], [0.0, 2.0])) and np.all(np.isin(tags[nfe:], [1.0, 3.0]))
    days = 600 * 0.002 * u.unit_time / 86400.0  # about 600 steps
    add = tmp_path / "setupImpactAddTime"
    add.write_text(f"AdditionTime = {days:.12f};\nDt = 0.002;\nDrawRate = 100;\nRecordRate = 100;\nRecenterRate = 1000000;\n")
    info1 = gpu.impact_run(bodies, add, run)
    assert info1.frames_written == info1.steps // 100
    f1 = gpu.read_frames(run / "impactPosVel", n)
    assert len(f1) == 1 + info1.frames_written
    # runStatsFile of the finished run: the finder's counts on the final state
    stats = (run / "runStatsFile").read_text()
    vals = {}
    for line in stats.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            vals[" ".join(k.split())] = v.strip()
    counts = [int(vals[f"{w} {m} from body {b}"]) for w in ("Earth", "Moon") for b in (1, 2)
              for m in ("iron", "silicate")]
    assert sum(counts) <= n and counts[0] + counts[1] > 0
    assert float(vals["The mass of Earth"]) > 0.0
    assert "final stats of the run when time" in stats
    times = [f[0] for f in f1]
    assert all(b > a for a, b in zip(times, times[1:]))
    # resume: the second run starts from the last frame
    info2 = gpu.impact_run(bodies, add, run)
    assert info2.start_time == pytest.approx(times[-1])
    f2 = gpu.read_frames(run / "impactPosVel", n)
    assert len(f2) == len(f1) + info2.frames_written
    # the same total in one run from frame 0 reproduces both runs' frames bitwise
    run_b = tmp_path / "run_b"
    run_b.mkdir()
    gpu.impact_initialize(bodies, setup, run_b)
    steps_total = (len(f2) - 1) * 100
    add_b = tmp_path / "setupImpactAddTime_b"
    add_b.write_text(f"AdditionTime = {(steps_total - 0.5) * 0.002 * u.unit_time / 86400.0:.12f};\nDt = 0.002;\n"
                     "DrawRate = 100;\nRecordRate = 100;\nRecenterRate = 1000000;\n")
    gpu.impact_run(bodies, add_b, run_b)
    fb = gpu.read_frames(run_b / "impactPosVel", n)
    assert len(fb) >= len(f2)
    for k in range(1, len(f2)):
        assert np.array_equal(fb[k][1], f2[k][1]) and np.array_equal(fb[k][2], f2[k][2]), k
    # recentring every 100 steps keeps the centre of mass at the recorded target
    run_c = tmp_path / "run_c"
    run_c.mkdir()
    gpu.impact_initialize(bodies, setup, run_c)
    target = [float(x.split("=")[1]) for x in (run_c / "viewCenteringInfo").read_text().split("\n") if "=" in x]
    add_c = tmp_path / "setupImpactAddTime_c"
    add_c.write_text(f"AdditionTime = {days:.12f};\nDt = 0.002;\nDrawRate = 100;\nRecordRate = 50;\nRecenterRate = 100;\n")
    gpu.impact_run(bodies, add_c, run_c)
    fcn = gpu.read_frames(run_c / "impactPosVel", n)
    m = np.where(np.arange(n) < nfe, u.mass_fe, 1.0)
    t_last, p_last, _ = fcn[-1]
    com = (m[:, None] * p_last[:, :3]).sum(0) / m.sum()
    assert np.allclose(com, target[:3], atol=0.05)

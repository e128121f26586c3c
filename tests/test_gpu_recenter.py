"""The reference's float recenter on the device (impactAddTime.cu:433-455) and
the impact run's bookkeeping of steps, frames and recentring
(impactAddTime.cu:396-481), against the oracle."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import f4

pytestmark = pytest.mark.gpu

MFE = 2.364458  # MassFe as the impact reads it back (%f)


def _upload(gi, pos, vel):
    n = len(pos)
    return gi.DeviceArray(n, pos), gi.DeviceArray(n, vel), gi.DeviceArray(n, f4(n))


@pytest.mark.parametrize("nt,ni,seed", [(1536, 512, 1), (3700, 1301, 2), (49152, 16384, 3)])
def test_float_recenter_is_bitwise_the_reference(gpu, oracle, nt, ni, seed):
    """Sums in float particle by particle, in the reference's order, with its
    double element masses: the recentred state is bitwise the oracle's (one
    tile, several tiles with a ragged last one, and 64 tiles)."""
    pos, vel, fc, _, _ = gpu.make_impact(nt, ni, seed=seed, target_spin=0.3)
    pos[:, :3] += np.float32([3.25, -1.5, 0.75])
    vel[:, :3] += np.float32([0.01, 0.02, -0.03])
    n, nfe = len(pos), fc.NFe
    com3, vel3 = (0.125, -0.25, 2.0), (0.0, 0.001, -0.002)
    P, V, _ = _upload(gpu, pos, vel)
    print(f"float recenter n={n}", flush=True)
    gpu.recenter_float(P, V, n, nfe, MFE, 1.0, com3, vel3)
    po, vo = oracle.recenter_f32(pos, vel, nfe, MFE, 1.0, com3, vel3)
    assert np.array_equal(P.download(), po) and np.array_equal(V.download(), vo)


def test_recentred_steps_track_the_reference(gpu, oracle):
    """12 steps with the reference's recenter every 4 toward a target 0.5 away
    from the current centre: the GPU run follows the oracle's loop (force,
    move, recenter) to the force tolerance, and both end with their centre
    on the target."""
    pos, vel, fc, mc, _ = gpu.make_impact(1536, 512, seed=20261202, angle_deg=45, speed_vesc=1.2)
    n = len(pos)
    com3, vel3 = (0.5, -0.25, 0.125), (0.001, 0.0, -0.002)
    P, V, F = _upload(gpu, pos, vel)
    print(f"float-recentred steps n={n}", flush=True)
    gpu.impact_steps_float_recenter(P, V, F, n, fc, mc, 12, 4, com3, vel3, MFE, 1.0)
    gpu.sync()
    po, vo, _ = oracle.impact_steps_f32(pos, vel, fc, mc, 12, 4, com3, vel3, mass_fe=MFE)
    p, v = P.download(), V.download()
    assert np.abs(v[:, :3] - vo[:, :3]).max() <= 1e-4 * np.abs(vo[:, :3]).max()
    assert np.abs(p[:, :3] - po[:, :3]).max() <= 1e-5 * np.abs(po[:, :3]).max()
    for st in ((p, v), (po, vo)):
        cv = oracle.com_vel_collision_f32(st[0], st[1], fc.NFe, MFE)
        assert np.allclose(cv[:3], com3, atol=1e-5) and np.allclose(cv[3:], vel3, atol=1e-6)
    # without the recenter the state stays ~0.5 away: the check above sees it
    P2, V2, F2 = _upload(gpu, pos, vel)
    gpu.impact_steps(P2, V2, F2, n, fc, mc, 12)
    gpu.sync()
    assert np.abs(P2.download()[:, :3] - po[:, :3]).max() > 0.4


def _bodies(gi, tmp_path, n, seed):
    u = gi.units(gi.generator_setup_defaults(n))
    pos, vel = f4(n), f4(n)
    gi.call("gi_make_raw_bodies", C.byref(u), seed, 0.1, pos.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p))
    d = tmp_path / "bodies"
    d.mkdir()
    gi.call("gi_parameters_write", str(d / "targetImpactorParameters").encode(), C.byref(u), n)
    with open(d / "targetImpactorInitialPosVel", "wb") as f:
        f.write(np.int32(n).tobytes() + pos.tobytes() + vel.tobytes())
    return u, d


def test_impact_run_frames_follow_the_reference_loop(gpu, oracle, tmp_path):
    """impactAddTime's loop for 24 steps with RecordRate 12 and RecenterRate 8
    from an initialised impact whose viewCenteringInfo target sits 0.5 away
    in x: frames at steps 12 and 24 (each recorded before that step's
    recentring), compared with the oracle's force / move / float-recenter
    sequence in the same order."""
    n = 2048
    u, bodies = _bodies(gpu, tmp_path, n, seed=11)
    r_km = 9.0 * u.unit_length
    setup = tmp_path / "setupImpactInitialization"
    setup.write_text(f"a = {-r_km};\nb = -30.0;\nc = 0.0;\nd = {r_km};\ne = 30.0;\nf = 0.0;\n"
                     "g = 1.5;\nh = 0.0;\ni = 0.0;\nj = -1.5;\nk = 0.0;\nl = 0.0;\n")
    run = tmp_path / "run"
    run.mkdir()
    gpu.impact_initialize(bodies, setup, run)
    cpath = run / "viewCenteringInfo"
    target = [float(x.split("=")[1]) for x in cpath.read_text().split("\n") if "=" in x]
    target[0] += 0.5
    names = ["InitialCenterOfMassX", "InitialCenterOfMassY", "InitialCenterOfMassZ",
             "InitialLinearVelocityX", "InitialLinearVelocityY", "InitialLinearVelocityZ"]
    cpath.write_text("".join(f"\n {k} = {v:f}\n" for k, v in zip(names, target)))
    days = 23.5 * 0.002 * u.unit_time / 86400.0  # the run ends after the step past it: 24 steps
    add = tmp_path / "setupImpactAddTime"
    add.write_text(f"AdditionalTime = {days:.12f};\nDt = 0.002;\nDrawRate = 5;\nRecordRate = 12;\nRecenterRate = 8;\n")
    print("impact run n=2048, 24 steps", flush=True)
    info = gpu.impact_run(bodies, add, run)
    assert info.steps == 24 and info.frames_written == 2
    frames = gpu.read_frames(run / "impactPosVel", n)
    assert len(frames) == 3
    # the oracle: the same constants from the same rounded parameters
    ou = oracle.units(n)
    fc, mc = oracle.impact_constants(ou)
    c3 = np.array(target[:3], np.float32)
    v3 = np.array(target[3:], np.float32)
    p, v = frames[0][1], frames[0][2]
    seq = [("steps", 8), ("recenter", 0), ("steps", 4), ("frame", 1), ("steps", 4), ("recenter", 0),
           ("steps", 8), ("frame", 2)]
    for what, k in seq:
        if what == "steps":
            p, v, _ = oracle.impact_steps_f32(p, v, fc, mc, k)
        elif what == "recenter":
            p, v = oracle.recenter_f32(p, v, fc.NFe, MFE, 1.0, c3, v3)
        else:
            t, fp, fv = frames[k]
            assert t == pytest.approx(np.float32(12 * k * 0.002), rel=1e-6)
            assert np.array_equal(fp[:, 3], p[:, 3]) and np.array_equal(fv[:, 3], v[:, 3])
            assert np.abs(fv[:, :3] - v[:, :3]).max() <= 1e-4 * np.abs(v[:, :3]).max()
            assert np.abs(fp[:, :3] - p[:, :3]).max() <= 1e-5 * np.abs(p[:, :3]).max()
    # the recentring moved the system by the 0.5 offset: the frame-2 centre is on target
    cv = oracle.com_vel_collision_f32(frames[2][1], frames[2][2], fc.NFe, MFE)
    assert abs(cv[0] - target[0]) < 0.01

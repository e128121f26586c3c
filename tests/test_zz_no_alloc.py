"""gi_force_collision_noalloc: the allocation-free form of gi_force_collision
for callers that cannot allocate (scratch in a static device array).  Its
force is bitwise gi_force_collision's wherever the tuning table gives that
call the direct pass; two calls in a row, the second smaller, share the
array."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import f4

pytestmark = pytest.mark.gpu


def _force(gi, name, pos, vel, fc):
    n = len(pos)
    P, V, F = gi.DeviceArray(n, pos), gi.DeviceArray(n, vel), gi.DeviceArray(n, f4(n))
    gi.call(name, P.ptr, V.ptr, F.ptr, n, C.byref(fc), None)
    gi.sync()
    return F.download()


def test_noalloc_force_is_bitwise_the_drop_in_force(gpu):
    gpu.call("gi_set_tiling", 0, 0, 0, 0, 0)
    cases = [(9000, 3000, 41), (1536, 512, 42), (700, 233, 43)]  # lanes 2, then 4, then 4; each smaller
    for nt, ni, seed in cases:
        pos, vel, fc, _, _ = gpu.make_impact(nt, ni, seed=seed, angle_deg=40.0, speed_vesc=1.1, gap_in_target_radii=0.3)
        print(f"noalloc n={nt + ni}", flush=True)
        a = _force(gpu, "gi_force_collision_noalloc", pos, vel, fc)
        b = _force(gpu, "gi_force_collision", pos, vel, fc)
        assert gpu.plan_info()["kernel"] == "k_direct"
        assert np.array_equal(a, b), (nt + ni, float(np.abs(a - b).max()))
        assert np.abs(a[:, :3]).max() > 0


def test_noalloc_rejects_sizes_beyond_its_maximum(gpu):
    lib = gpu.load()
    fc, _ = gpu.impact_constants(gpu.units(gpu.generator_setup_defaults(1024)))
    fc.NFe = 100
    vp = C.c_void_p
    assert lib.gi_force_collision_noalloc(vp(1), vp(1), vp(1), 12289, C.byref(fc), None) == -4
    assert b"12288" in lib.gi_last_error_string()
    lib.gi_clear_error()

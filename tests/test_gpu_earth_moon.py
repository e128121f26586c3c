"""The Earth / Moon finder on the device (gi_find_earth_moon) against the
oracle's restatement of findEarthAndMoon and of the final-statistics sums
(Code/ImpactCode/impactBasicFunctions.h:109-217,219-652): group numbers and
element counts exactly, FP64 sums against the oracle's FP64 twin (1e-9) and
against the reference's float sums (1e-4, the float restatement's own
rounding)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import f4

pytestmark = pytest.mark.gpu


def _kinds(n, nfe1, nfe2, nsi1):
    i = np.arange(n)
    return np.select([i < nfe1, i < nfe1 + nfe2, i < nfe1 + nfe2 + nsi1], [0, 1, 2], 3)


def _check(gpu, oracle, pos, vel, nfe1, nfe2, nsi1, mass_fe):
    n = len(pos)
    P, V = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel)
    g = gpu.find_earth_moon(P, V, n, nfe1, nfe2, nsi1, mass_fe, with_groups=True)
    o = oracle.find_earth_moon(pos)
    print({k: g[k] for k in ("n_groups", "earth_id", "moon_id", "earth_count", "moon_count")}, flush=True)
    assert np.array_equal(g["group"], o["group"])
    assert g["n_groups"] == o["groups"]
    assert (g["earth_id"], g["moon_id"]) == (o["earth_id"], o["moon_id"])
    assert (g["earth_count"], g["moon_count"]) == (o["earth_count"], o["moon_count"])
    kinds = _kinds(n, nfe1, nfe2, nsi1)
    for name, gid in (("earth", o["earth_id"]), ("moon", o["moon_id"])):
        want = np.bincount(kinds[o["group"] == gid], minlength=4) if gid >= 0 else np.zeros(4, int)
        got = [g[f"{name}_{k}"] for k in ("fe1", "fe2", "si1", "si2")]
        assert list(want) == got, name
    nfe = nfe1 + nfe2
    s64 = oracle.final_stats(pos, vel, nfe, o["group"], o["earth_id"], o["moon_id"], mass_fe, fp64=True)
    s32 = oracle.final_stats(pos, vel, nfe, o["group"], o["earth_id"], o["moon_id"], mass_fe, fp64=False)
    scopes = [0, 1, 2] + ([3] if o["moon_id"] >= 0 else [])
    for s in scopes:
        assert g["mass"][s] == pytest.approx(s64["mass"][s], rel=1e-12)
        L = np.abs(s64["angmom"][s]).max()
        cs = np.abs(pos[:, :3]).max()
        vs = np.abs(vel[:, :3]).max()
        assert np.allclose(g["com"][s], s64["com"][s], rtol=0, atol=1e-9 * cs)
        assert np.allclose(g["vel"][s], s64["vel"][s], rtol=0, atol=1e-9 * vs)
        # one pass: L = sum m x cross v - (sum m x) cross (sum m v) / M, whose
        # cancellation leaves ~1e-16 of the cancelled terms (M |x| |v|)
        cancel = s64["mass"][s] * cs * vs
        assert np.allclose(g["angmom"][s], s64["angmom"][s], rtol=0, atol=1e-9 * L + 1e-13 * cancel)
        # the reference's float sums agree to their own rounding
        assert np.allclose(g["com"][s], s32["com"][s], rtol=0, atol=1e-4 * cs)
        assert np.allclose(g["angmom"][s], s32["angmom"][s], rtol=0, atol=1e-3 * L + 1e-6 * cancel)
    if o["moon_id"] < 0:
        assert g["moon_count"] == 0 and g["mass"][3] == 0.0
    return g


def test_initial_impact_state_is_target_and_impactor(gpu, oracle):
    pos, vel, fc, mc, info = gpu.make_impact(1536, 512, seed=20261202, angle_deg=45.0, speed_vesc=1.2,
                                             target_spin=0.3)
    g = _check(gpu, oracle, pos, vel, info.nfe1, info.nfe2, info.nsi1, info.mass_fe)
    assert g["n_groups"] == 2 and g["earth_id"] == 0
    assert (g["earth_count"], g["moon_count"]) == (1536, 512)
    assert g["earth_fe2"] == g["earth_si2"] == 0 and g["moon_fe1"] == g["moon_si1"] == 0


@pytest.mark.parametrize("seed,n,box", [(1, 3000, 18.0), (2, 2500, 24.0), (3, 1200, 9.0)])
def test_random_clouds_many_groups(gpu, oracle, seed, n, box):
    rng = np.random.default_rng(seed)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = rng.uniform(0.0, box, (n, 3))
    vel[:, :3] = rng.normal(0.0, 1.0, (n, 3))
    print(f"seed={seed} n={n} box={box}", flush=True)
    nfe1, nfe2, nsi1 = n // 5, n // 10, n // 2
    _check(gpu, oracle, pos, vel, nfe1, nfe2, nsi1, 2.364458)


def test_touch_threshold_is_the_reference_float_test(gpu, oracle):
    """Pairs placed right at 1.5: the float sqrt test decides, bit for bit."""
    base = np.float32(1.5)
    seps = [np.nextafter(base, np.float32(0)), base, np.nextafter(base, np.float32(3))]
    pts = []
    x = 0.0
    for s in seps * 4:
        pts += [(x, 0.0, 0.0), (x + float(s), 0.0, 0.0)]
        x += 10.0
    n = len(pts)
    pos, vel = f4(n), f4(n)
    pos[:, :3] = np.array(pts, np.float32)
    _check(gpu, oracle, pos, vel, 0, 0, n, 2.364458)


def test_far_flung_elements_land_in_the_right_cells(gpu, oracle):
    """Groups near +-1e5 and 3e6: cell coordinates are formed in FP64."""
    rng = np.random.default_rng(4)
    pos, vel = f4(400), f4(400)
    for k, c in enumerate((1e5, -1e5, 3e6, -2.5e6)):
        pos[100 * k:100 * (k + 1), :3] = np.float32(c) + rng.uniform(0, 4.0, (100, 3)).astype(np.float32)
    _check(gpu, oracle, pos, vel, 50, 50, 150, 2.364458)


def test_large_impact_composition(gpu):
    """N = 65,536: the two bodies are the Earth and the Moon, with every
    element of each body and material counted."""
    pos, vel, fc, mc, info = gpu.make_impact(52429, 13107, seed=2, angle_deg=45.0, speed_vesc=1.1, target_spin=0.5)
    n = len(pos)
    P, V = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel)
    g = gpu.find_earth_moon(P, V, n, info.nfe1, info.nfe2, info.nsi1, info.mass_fe)
    assert g["n_groups"] == 2 and (g["earth_count"], g["moon_count"]) == (52429, 13107)
    assert (g["earth_fe1"], g["earth_si1"], g["moon_fe2"], g["moon_si2"]) == (info.nfe1, info.nsi1, info.nfe2,
                                                                              info.nsi2)
    # the system's momentum is zero in the centre-of-mass frame
    assert np.abs(g["vel"][0]).max() < 1e-6


def test_invalid_arguments(gpu):
    pos, vel = f4(8), f4(8)
    P, V = gpu.DeviceArray(8, pos), gpu.DeviceArray(8, vel)
    em = gpu.EarthMoon()
    lib = gpu.load()
    assert lib.gi_find_earth_moon(P.ptr, V.ptr, 8, 5, 5, 0, 2.0, 1.0, 1.5, None, C.byref(em), None) == -1
    assert lib.gi_find_earth_moon(P.ptr, V.ptr, 8, 1, 1, 1, 2.0, 1.0, 0.0, None, C.byref(em), None) == -1
    assert lib.gi_find_earth_moon(P.ptr, V.ptr, 8, 1, 1, 1, 2.0, 1.0, 1.5, None, None, None) == -1


def test_long_chains_with_shuffled_indices_are_one_group_each(gpu):
    """Deep union-find trees: 8 chains of 4,000 elements 1.4 apart (touch 1.5),
    3 apart across chains, indices shuffled.  Every element's group is the
    smallest index of its chain -- the final root pointers are never left on
    an inner node of a tree."""
    rng = np.random.default_rng(11)
    chains, per = 8, 4000
    n = chains * per
    pts = np.zeros((n, 3), np.float32)
    chain_of = np.repeat(np.arange(chains), per)
    pts[:, 0] = np.tile(np.arange(per, dtype=np.float32) * np.float32(1.4), chains)
    pts[:, 1] = chain_of.astype(np.float32) * np.float32(3.0)
    perm = rng.permutation(n)
    pos, vel = f4(n), f4(n)
    pos[perm, :3] = pts
    chain_of = chain_of[np.argsort(perm)]  # chain of each (shuffled) index
    P, V = gpu.DeviceArray(n, pos), gpu.DeviceArray(n, vel)
    g = gpu.find_earth_moon(P, V, n, 0, 0, n, 2.364458, with_groups=True)
    want = np.array([np.flatnonzero(chain_of == c).min() for c in range(chains)])[chain_of]
    assert g["n_groups"] == chains
    assert np.array_equal(g["group"], want)
    assert g["earth_count"] == per and g["moon_count"] == per

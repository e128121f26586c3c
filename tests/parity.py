"""Force-parity metrics shared by the GPU tests.

Truth is the FP64 twin of the oracle (F64); "the reference" is the oracle's
float restatement of the reference kernel (Fref).  With cond_i = sum_j |f_ij|:

  conditioned error   max_i |F_i - F64_i| / cond_i                 <= 1e-5
  relative error      per particle, |F_i - F64_i| <= |Fref_i - F64_i| + 1e-5 |F64_i|
                          (without Fref: max_i |F_i - F64_i| / |F64_i| <= 1e-5)
  against Fref        per particle, |F_i - Fref_i| / |F64_i|
                          <= 1e-5 + 2 (Fref_i's own relative error)
                      (the triangle inequality with the bound above), and
                      the 99.9th percentile <= 1e-5 + Fref's own 99.9th.

The relative figures are bounded by the reference's own error because in
contact-dominated states a particle's net force is a small difference of
large contact forces: there the reference's float routine itself differs
from FP64 truth by more than 1e-5 (measured, e.g. 1.7e-5 mid-impact at
N = 2,048), and no float evaluation can do better than its inputs' rounding
(the contact magnitude G - K (1 - r^2) is formed from the float r^2 in both).
The comparison is per particle: two float evaluations with the same rounding
floor exceed each other's worst case by a few percent at random (a maximum
against a maximum once failed at 2.12e-5 against 2.05e-5 while the
conditioned error was 1.26e-6 against the reference's 1.43e-6).  The
conditioned error has no such escape.

Particles with a pair whose branch decision lies within 1e-6 of a threshold
(r = 1, a shell break, dp.dv = 1) are excluded: there a last-bit rounding
difference legitimately selects another branch of the reference's law.  At
most 0.1% of the rows may be excluded that way, and never fewer than 8 rows
(four pairs, each excusing both its particles) are allowed: in a contact
state of 2,048 particles one or two pairs within the margin are a matter of
chance (the receding test's margin scales with |dp| |dv|).
"""
from __future__ import annotations

import numpy as np

COND_TOL = 1e-5
REL_TOL = 1e-5


def force_metrics(F, F64, cond, near, Fref=None) -> dict:
    F = np.asarray(F, dtype=np.float64)[:, :3]
    ok = ~np.asarray(near, dtype=bool)
    e = np.linalg.norm(F - F64, axis=1)
    m = np.maximum(np.linalg.norm(F64, axis=1), 1e-30)
    c = np.maximum(cond, 1e-30)
    out = {
        "n": int(len(F)),
        "near_threshold": int((~ok).sum()),
        "cond_max": float(np.max(e[ok] / c[ok])) if ok.any() else 0.0,
        "rel_max": float(np.max(e[ok] / m[ok])) if ok.any() else 0.0,
    }
    if Fref is not None:
        Fr = np.asarray(Fref, dtype=np.float64)[:, :3]
        d = np.linalg.norm(F - Fr, axis=1) / m
        er = np.linalg.norm(Fr - F64, axis=1) / m
        out["vs_ref_max"] = float(d[ok].max()) if ok.any() else 0.0
        out["vs_ref_p999"] = float(np.percentile(d[ok], 99.9)) if ok.any() else 0.0
        out["ref_rel_max"] = float(er[ok].max()) if ok.any() else 0.0
        out["ref_rel_p999"] = float(np.percentile(er[ok], 99.9)) if ok.any() else 0.0
        out["ref_cond_max"] = float(np.max(np.linalg.norm(Fr - F64, axis=1)[ok] / c[ok])) if ok.any() else 0.0
        # per particle: how far this evaluation's relative error exceeds Fref's
        out["rel_excess_max"] = float(np.max(e[ok] / m[ok] - er[ok])) if ok.any() else 0.0
        # per particle: distance from Fref beyond 1e-5 + twice Fref's own
        # error on that particle (the triangle inequality with the bound
        # above), so no particle borrows another's reference error
        out["vs_ref_excess_max"] = float(np.max(d[ok] - 2.0 * er[ok])) if ok.any() else 0.0
    return out


NEAR_MAX_FRACTION = 1e-3  # rows a last-bit branch difference may excuse
NEAR_MIN_ROWS = 8


def assert_force_parity(met: dict, check_ref: bool = True) -> None:
    assert met["near_threshold"] <= max(NEAR_MAX_FRACTION * met["n"], NEAR_MIN_ROWS), met
    assert met["cond_max"] <= COND_TOL, met
    if "rel_excess_max" in met:
        assert met["rel_excess_max"] <= REL_TOL, met
    else:
        assert met["rel_max"] <= REL_TOL, met
    if check_ref and "vs_ref_max" in met:
        assert met["vs_ref_excess_max"] <= REL_TOL, met
        assert met["vs_ref_p999"] <= REL_TOL + met["ref_rel_p999"], met

"""ctypes wrapper of the CPU oracle (oracle/gi_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.

Each function restates a reference routine; gi_oracle.c cites file:line for
each one.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "lib" / "libgi_oracle.so"

_LAW = [
    "GMassFeFe", "GMassFeSi", "KFeFe", "KSiSi", "KFeSi", "KRFeFe", "KRSiSi", "KRFeSi", "KRMix",
    "ShellBreakFe", "ShellBreakSi", "ShellBreakFeSi1", "ShellBreakFeSi2",
]


class FC(C.Structure):
    _fields_ = [(k, C.c_float) for k in _LAW] + [("NFe", C.c_int)]


class MC(C.Structure):
    _fields_ = [("Dt", C.c_float), ("DtOverMassFe", C.c_float), ("DtOverMassSi", C.c_float), ("NFe", C.c_int)]


class FS(C.Structure):
    _fields_ = [(k, C.c_float) for k in _LAW] + [("boarder1", C.c_int), ("boarder2", C.c_int), ("boarder3", C.c_int)]


class MS(C.Structure):
    _fields_ = [
        ("Dt", C.c_float), ("DtOverMassFe", C.c_float), ("DtOverMassSi", C.c_float),
        ("boarder1", C.c_int), ("boarder2", C.c_int), ("boarder3", C.c_int),
    ]


class Units(C.Structure):
    _fields_ = [
        ("fraction_earth_mass_body1", C.c_double), ("fraction_earth_mass_body2", C.c_double),
        ("fraction_fe_body1", C.c_double), ("fraction_si_body1", C.c_double),
        ("fraction_fe_body2", C.c_double), ("fraction_si_body2", C.c_double),
        ("total_number_of_elements", C.c_int),
        ("density_fe", C.c_double), ("density_si", C.c_double),
        ("k_fe", C.c_double), ("k_si", C.c_double),
        ("kr_fe", C.c_double), ("kr_si", C.c_double), ("sd_fe", C.c_double), ("sd_si", C.c_double),
        ("dt", C.c_double),
        ("nfe", C.c_int), ("nsi", C.c_int), ("nfe1", C.c_int), ("nsi1", C.c_int), ("nfe2", C.c_int), ("nsi2", C.c_int),
        ("unit_length", C.c_double), ("unit_mass", C.c_double), ("unit_time", C.c_double),
        ("mass_fe", C.c_double), ("mass_si", C.c_double), ("diameter", C.c_double),
        ("k_fe_sim", C.c_double), ("k_si_sim", C.c_double),
        ("mass_of_body1", C.c_double), ("mass_of_body2", C.c_double),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-C", str(ORACLE_DIR), "-B"], check=True, stdout=subprocess.DEVNULL)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    P = C.POINTER
    sig = {
        "or_units_compute": (i, [P(Units)]),
        "or_impact_constants": (None, [P(Units), P(FC), P(MC)]),
        "or_separate_constants": (None, [P(Units), P(FS), P(MS)]),
        "or_round_f": (d, [d]),
        "or_force_collision_f32": (None, [vp, vp, vp, i, P(FC), i, i, i]),
        "or_move_collision_f32": (None, [vp, vp, vp, i, P(MC)]),
        "or_force_separate_f32": (None, [vp, vp, vp, i, P(FS), i]),
        "or_move_damped_separate_f32": (None, [vp, vp, vp, i, P(MS), C.c_float, C.c_float]),
        "or_force_collision_f64": (None, [vp, vp, vp, vp, vp, d, i, P(FC), i, i, i]),
        "or_force_separate_f64": (None, [vp, vp, vp, vp, vp, d, i, P(FS), i]),
        "or_conserved_f64": (None, [vp, vp, i, P(FC), vp, i]),
        "or_find_earth_moon": (i, [vp, i, C.c_float, vp, P(i), P(i), P(i), P(i)]),
        "or_final_stats_f32": (None, [vp, vp, i, i, vp, i, i, C.c_double, C.c_double, vp, vp, vp, vp]),
        "or_final_stats_f64": (None, [vp, vp, i, i, vp, i, i, C.c_double, C.c_double, vp, vp, vp, vp]),
        "or_com_vel_collision_f32": (None, [vp, vp, i, i, d, d, vp]),
        "or_recenter_f32": (None, [vp, vp, i, i, d, d, vp, vp]),
        "or_body_com": (None, [vp, i, i, i, d, d, d, vp]),
        "or_body_vel": (None, [vp, i, i, i, d, d, d, vp]),
        "or_spin_body": (i, [vp, vp, i, i, i, d, d, d, vp]),
        "or_body_angmom": (None, [vp, vp, i, i, i, d, d, d, vp]),
        "or_body_radius": (d, [vp, i, i, i, d, d, d]),
        "or_impact_initialize": (None, [vp, vp, i, i, i, i, vp, vp, vp, vp, d, d, d, d, vp, vp, vp]),
        "or_impact_steps_f32": (None, [vp, vp, vp, i, P(FC), P(MC), i, i, vp, vp, d, d, i]),
        "or_trunc_to_int": (i, [C.c_float]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def threads() -> int:
    return int(os.environ.get("GI_ORACLE_THREADS", str(os.cpu_count() or 1)))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def _f4(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def as_fc(c) -> FC:
    out = FC()
    for k in _LAW + ["NFe"]:
        setattr(out, k, getattr(c, k))
    return out


def as_mc(c) -> MC:
    out = MC()
    for k in ["Dt", "DtOverMassFe", "DtOverMassSi", "NFe"]:
        setattr(out, k, getattr(c, k))
    return out


def as_fs(c) -> FS:
    out = FS()
    for k in _LAW + ["boarder1", "boarder2", "boarder3"]:
        setattr(out, k, getattr(c, k))
    return out


def as_ms(c) -> MS:
    out = MS()
    for k in ["Dt", "DtOverMassFe", "DtOverMassSi", "boarder1", "boarder2", "boarder3"]:
        setattr(out, k, getattr(c, k))
    return out


def force_collision_f32(pos, vel, fc, i_begin=0, i_end=None, nthreads=None, force=None) -> np.ndarray:
    pos, vel = _f4(pos), _f4(vel)
    n = pos.shape[0]
    out = np.zeros((n, 4), dtype=np.float32) if force is None else force
    load().or_force_collision_f32(
        _ptr(pos), _ptr(vel), _ptr(out), n, C.byref(as_fc(fc)), int(i_begin), int(n if i_end is None else i_end),
        int(nthreads or threads()),
    )
    return out


def move_collision_f32(pos, vel, force, mc):
    pos, vel, force = _f4(pos).copy(), _f4(vel).copy(), _f4(force)
    load().or_move_collision_f32(_ptr(pos), _ptr(vel), _ptr(force), pos.shape[0], C.byref(as_mc(mc)))
    return pos, vel


def force_separate_f32(pos, vel, fs, nthreads=None) -> np.ndarray:
    pos, vel = _f4(pos), _f4(vel)
    out = np.zeros_like(pos)
    load().or_force_separate_f32(_ptr(pos), _ptr(vel), _ptr(out), pos.shape[0], C.byref(as_fs(fs)), int(nthreads or threads()))
    return out


def move_damped_separate_f32(pos, vel, force, ms, d1, d2):
    pos, vel, force = _f4(pos).copy(), _f4(vel).copy(), _f4(force)
    load().or_move_damped_separate_f32(_ptr(pos), _ptr(vel), _ptr(force), pos.shape[0], C.byref(as_ms(ms)), d1, d2)
    return pos, vel


def force_collision_f64(pos, vel, fc, margin=1e-6, nthreads=None, i_begin=0, i_end=None):
    """FP64 truth: (force (n,3), cond (n,) = sum_j |f_ij|, near (n,) bool).

    Only rows [i_begin, i_end) are computed; the other rows are zero."""
    pos, vel = _f4(pos), _f4(vel)
    n = pos.shape[0]
    f = np.zeros((n, 3), dtype=np.float64)
    cond = np.zeros(n, dtype=np.float64)
    near = np.zeros(n, dtype=np.int32)
    load().or_force_collision_f64(
        _ptr(pos), _ptr(vel), _ptr(f), _ptr(cond), _ptr(near), margin, n, C.byref(as_fc(fc)),
        int(i_begin), int(n if i_end is None else i_end), int(nthreads or threads())
    )
    return f, cond, near.astype(bool)


def force_separate_f64(pos, vel, fs, margin=1e-6, nthreads=None):
    pos, vel = _f4(pos), _f4(vel)
    n = pos.shape[0]
    f = np.zeros((n, 3), dtype=np.float64)
    cond = np.zeros(n, dtype=np.float64)
    near = np.zeros(n, dtype=np.int32)
    load().or_force_separate_f64(
        _ptr(pos), _ptr(vel), _ptr(f), _ptr(cond), _ptr(near), margin, n, C.byref(as_fs(fs)), int(nthreads or threads())
    )
    return f, cond, near.astype(bool)


def find_earth_moon(pos, touch: float = 1.5) -> dict:
    """findEarthAndMoon restated: group numbers (index of each group's first
    element), the Earth and Moon group numbers and sizes."""
    pos = _f4(pos)
    n = pos.shape[0]
    group = np.zeros(n, dtype=np.int32)
    e, m, ne, nm = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    groups = load().or_find_earth_moon(_ptr(pos), n, float(touch), _ptr(group), C.byref(e), C.byref(m), C.byref(ne),
                                       C.byref(nm))
    return {"group": group, "groups": groups, "earth_id": e.value, "moon_id": m.value, "earth_count": ne.value,
            "moon_count": nm.value}


def final_stats(pos, vel, nfe, group, earth, moon, mass_fe, mass_si=1.0, fp64=False) -> dict:
    """get*Collision(scope 0..3) restated (float, or the FP64 twin)."""
    pos, vel = _f4(pos), _f4(vel)
    group = np.ascontiguousarray(group, dtype=np.int32)
    dt = np.float64 if fp64 else np.float32
    mass, com, v, L = np.zeros(4, dt), np.zeros((4, 3), dt), np.zeros((4, 3), dt), np.zeros((4, 3), dt)
    f = load().or_final_stats_f64 if fp64 else load().or_final_stats_f32
    f(_ptr(pos), _ptr(vel), pos.shape[0], int(nfe), _ptr(group), int(earth), int(moon), float(mass_fe), float(mass_si),
      _ptr(mass), _ptr(com), _ptr(v), _ptr(L))
    return {"mass": mass, "com": com, "vel": v, "angmom": L}


def conserved_f64(pos, vel, fc, nthreads=None) -> np.ndarray:
    """[kinetic, potential, total, Px, Py, Pz, Lx, Ly, Lz] in FP64."""
    pos, vel = _f4(pos), _f4(vel)
    out = np.zeros(9, dtype=np.float64)
    load().or_conserved_f64(_ptr(pos), _ptr(vel), pos.shape[0], C.byref(as_fc(fc)), _ptr(out), int(nthreads or threads()))
    return out


def com_vel_collision_f32(pos, vel, nfe, mass_fe, mass_si=1.0) -> np.ndarray:
    pos, vel = _f4(pos), _f4(vel)
    out = np.zeros(6, dtype=np.float32)
    load().or_com_vel_collision_f32(_ptr(pos), _ptr(vel), pos.shape[0], int(nfe), mass_fe, mass_si, _ptr(out))
    return out


def recenter_f32(pos, vel, nfe, mass_fe, mass_si, com3, vel3):
    pos, vel = _f4(pos).copy(), _f4(vel).copy()
    c = np.asarray(com3, dtype=np.float32)
    v = np.asarray(vel3, dtype=np.float32)
    load().or_recenter_f32(_ptr(pos), _ptr(vel), pos.shape[0], int(nfe), mass_fe, mass_si, _ptr(c), _ptr(v))
    return pos, vel


def spin_body(pos, vel, start, stop, stop_fe, mass_fe, mass_si, total_mass, spin4):
    pos, vel = _f4(pos), _f4(vel).copy()
    s = np.asarray(spin4, dtype=np.float32)
    rc = load().or_spin_body(_ptr(pos), _ptr(vel), start, stop, stop_fe, mass_fe, mass_si, total_mass, _ptr(s))
    if rc != 0:
        raise ValueError("spin direction vector is zero")
    return vel


def body_com(pos, start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass) -> np.ndarray:
    pos = _f4(pos)
    out = np.zeros(4, dtype=np.float64)
    load().or_body_com(_ptr(pos), start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass, _ptr(out))
    return out


def body_vel(vel, start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass) -> np.ndarray:
    vel = _f4(vel)
    out = np.zeros(4, dtype=np.float64)
    load().or_body_vel(_ptr(vel), start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass, _ptr(out))
    return out


def body_angmom(pos, vel, start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass) -> np.ndarray:
    pos, vel = _f4(pos), _f4(vel)
    out = np.zeros(4, dtype=np.float64)
    load().or_body_angmom(_ptr(pos), _ptr(vel), start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass, _ptr(out))
    return out


def body_radius(pos, start, stop_fe, stop, mass_fe, mass_si, total_mass) -> float:
    return float(load().or_body_radius(_ptr(_f4(pos)), start, stop_fe, stop, mass_fe, mass_si, total_mass))


def impact_initialize(gpos, gvel, nfe1, nsi1, nfe2, nsi2, ip1, ip2, iv1, iv2, unit_length, unit_time,
                      mass_fe, mass_si=1.0):
    """impactInitialization's set-up restated: (pos, vel) in the impact
    layout and stats10 = [mass, com xyz, vel xyz, L xyz] (FP64, sim units)."""
    gpos, gvel = _f4(gpos), _f4(gvel)
    pos, vel = np.zeros_like(gpos), np.zeros_like(gvel)
    st = np.zeros(10, dtype=np.float64)
    f3 = [np.ascontiguousarray(a, dtype=np.float32) for a in (ip1, ip2, iv1, iv2)]
    load().or_impact_initialize(_ptr(gpos), _ptr(gvel), int(nfe1), int(nsi1), int(nfe2), int(nsi2),
                                *[_ptr(a) for a in f3], float(unit_length), float(unit_time), float(mass_fe),
                                float(mass_si), _ptr(pos), _ptr(vel), _ptr(st))
    return pos, vel, st


def impact_steps_f32(pos, vel, fc, mc, n_steps, recenter_rate=0, com3=(0, 0, 0), vel3=(0, 0, 0),
                     mass_fe=None, mass_si=1.0, nthreads=None):
    pos, vel = _f4(pos).copy(), _f4(vel).copy()
    force = np.zeros_like(pos)
    c = np.asarray(com3, dtype=np.float32)
    v = np.asarray(vel3, dtype=np.float32)
    mfe = float(fc.GMassFeSi) if mass_fe is None else mass_fe
    load().or_impact_steps_f32(
        _ptr(pos), _ptr(vel), _ptr(force), pos.shape[0], C.byref(as_fc(fc)), C.byref(as_mc(mc)), int(n_steps),
        int(recenter_rate), _ptr(c), _ptr(v), mfe, mass_si, int(nthreads or threads()),
    )
    return pos, vel, force


def units(n, *, fraction_earth_mass=(0.55, 0.55), fe=(0.26, 0.26), density=(7.850e12, 3.320e12),
          k=(1.5e14, 1.5e14), kr=(0.01, 0.01), sd=(0.02, 0.002), dt=0.002) -> Units:
    """Units of setTargetImpactorParameters with the template defaults."""
    u = Units()
    u.fraction_earth_mass_body1, u.fraction_earth_mass_body2 = fraction_earth_mass
    u.fraction_fe_body1, u.fraction_fe_body2 = fe
    u.fraction_si_body1, u.fraction_si_body2 = 1.0 - fe[0], 1.0 - fe[1]
    u.total_number_of_elements = n
    u.density_fe, u.density_si = density
    u.k_fe, u.k_si = k
    u.kr_fe, u.kr_si = kr
    u.sd_fe, u.sd_si = sd
    u.dt = dt
    rc = load().or_units_compute(C.byref(u))
    if rc != 0:
        raise ValueError(f"or_units_compute: {rc}")
    return u


def impact_constants(u: Units):
    fc, mc = FC(), MC()
    load().or_impact_constants(C.byref(u), C.byref(fc), C.byref(mc))
    return fc, mc


def separate_constants(u: Units):
    fs, ms = FS(), MS()
    load().or_separate_constants(C.byref(u), C.byref(fs), C.byref(ms))
    return fs, ms


def trunc_to_int(v: float) -> int:
    return load().or_trunc_to_int(v)

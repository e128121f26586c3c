"""ctypes bindings of libgiant_impact.so (the C ABI in include/giant_impact.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  This
module only marshals arguments; there is no Python or CPU fallback.  Loading
fails loudly when the shared library has not been built.

Device memory is handled through the library's own gi_malloc / gi_copy_*
(``DeviceArray``), so no other GPU framework is needed.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "lib" / "libgiant_impact.so"
HEADER_PATH = PKG_DIR.parent / "include" / "giant_impact.h"


class GIError(RuntimeError):
    def __init__(self, func: str, code: int, message: str):
        super().__init__(f"{func} failed with code {code}: {message}")
        self.code = code


# ---------------------------------------------------------------------------
# ABI structures (layout-identical to the reference's kernel structs)
# ---------------------------------------------------------------------------

_LAW_FIELDS = [
    ("GMassFeFe", C.c_float),
    ("GMassFeSi", C.c_float),
    ("KFeFe", C.c_float),
    ("KSiSi", C.c_float),
    ("KFeSi", C.c_float),
    ("KRFeFe", C.c_float),
    ("KRSiSi", C.c_float),
    ("KRFeSi", C.c_float),
    ("KRMix", C.c_float),
    ("ShellBreakFe", C.c_float),
    ("ShellBreakSi", C.c_float),
    ("ShellBreakFeSi1", C.c_float),
    ("ShellBreakFeSi2", C.c_float),
]


class ForceCollisionConstants(C.Structure):
    """forceCollisionKernalConstantsStruct (Code/ImpactCode/impactAddTime.cu:4-20)."""

    _fields_ = _LAW_FIELDS + [("NFe", C.c_int)]


class MoveCollisionConstants(C.Structure):
    """moveCollisionKernalConstantsStruct (Code/ImpactCode/impactAddTime.cu:22-28)."""

    _fields_ = [("Dt", C.c_float), ("DtOverMassFe", C.c_float), ("DtOverMassSi", C.c_float), ("NFe", C.c_int)]


class ForceSeparateConstants(C.Structure):
    """forceSeperateKernalConstantsStruct (generateTargetImpactor.cu:3-21)."""

    _fields_ = _LAW_FIELDS + [("boarder1", C.c_int), ("boarder2", C.c_int), ("boarder3", C.c_int)]


class MoveSeparateConstants(C.Structure):
    """moveSeperateKernalConstantsStruct (generateTargetImpactor.cu:23-31)."""

    _fields_ = [
        ("Dt", C.c_float),
        ("DtOverMassFe", C.c_float),
        ("DtOverMassSi", C.c_float),
        ("boarder1", C.c_int),
        ("boarder2", C.c_int),
        ("boarder3", C.c_int),
    ]


class GeneratorSetup(C.Structure):
    _fields_ = [
        ("initial_spin1", C.c_float * 4),
        ("initial_spin2", C.c_float * 4),
        ("fraction_earth_mass_body1", C.c_double),
        ("fraction_earth_mass_body2", C.c_double),
        ("fraction_fe_body1", C.c_double),
        ("fraction_si_body1", C.c_double),
        ("fraction_fe_body2", C.c_double),
        ("fraction_si_body2", C.c_double),
        ("damp_rate_body1", C.c_float),
        ("damp_rate_body2", C.c_float),
        ("total_number_of_elements", C.c_int),
        ("damp_time", C.c_double),
        ("damp_rest_time", C.c_double),
        ("spin_rest_time", C.c_double),
        ("dt", C.c_double),
        ("density_fe", C.c_double),
        ("density_si", C.c_double),
        ("k_fe", C.c_double),
        ("k_si", C.c_double),
        ("kr_fe", C.c_double),
        ("kr_si", C.c_double),
        ("sd_fe", C.c_double),
        ("sd_si", C.c_double),
        ("draw_rate", C.c_int),
    ]


class Units(C.Structure):
    _fields_ = [
        ("nfe", C.c_int),
        ("nsi", C.c_int),
        ("nfe1", C.c_int),
        ("nsi1", C.c_int),
        ("nfe2", C.c_int),
        ("nsi2", C.c_int),
        ("unit_length", C.c_double),
        ("unit_mass", C.c_double),
        ("unit_time", C.c_double),
        ("mass_fe", C.c_double),
        ("mass_si", C.c_double),
        ("diameter", C.c_double),
        ("gravity", C.c_double),
        ("k_fe", C.c_double),
        ("k_si", C.c_double),
        ("kr_fe", C.c_double),
        ("kr_si", C.c_double),
        ("sd_fe", C.c_double),
        ("sd_si", C.c_double),
        ("mass_of_body1", C.c_double),
        ("mass_of_body2", C.c_double),
        ("damp_time", C.c_double),
        ("damp_rest_time", C.c_double),
        ("spin_rest_time", C.c_double),
        ("spin1_w", C.c_float),
        ("spin2_w", C.c_float),
    ]


class ImpactConfig(C.Structure):
    _fields_ = [
        ("n_target", C.c_int),
        ("n_impactor", C.c_int),
        ("fe_fraction", C.c_double),
        ("jitter", C.c_double),
        ("seed", C.c_ulonglong),
        ("angle_deg", C.c_double),
        ("speed_vesc", C.c_double),
        ("target_spin", C.c_double),
        ("impactor_spin", C.c_double),
        ("gap_in_target_radii", C.c_double),
    ]


class ImpactInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nfe", C.c_int),
        ("nfe1", C.c_int),
        ("nfe2", C.c_int),
        ("nsi1", C.c_int),
        ("nsi2", C.c_int),
        ("mass_fe", C.c_double),
        ("mass_si", C.c_double),
        ("radius_target", C.c_double),
        ("radius_impactor", C.c_double),
        ("mass_target", C.c_double),
        ("mass_impactor", C.c_double),
        ("v_escape", C.c_double),
        ("v_relative", C.c_double),
        ("impact_parameter", C.c_double),
        ("target_omega", C.c_double),
        ("impactor_omega", C.c_double),
        ("com", C.c_double * 3),
        ("vel", C.c_double * 3),
    ]


# ---------------------------------------------------------------------------
# Library loading
# ---------------------------------------------------------------------------

_lib = None


class EarthMoon(C.Structure):  # gi_earth_moon
    _fields_ = [(k, C.c_int) for k in ("n_groups", "earth_id", "moon_id", "earth_count", "moon_count",
                                          "earth_fe1", "earth_fe2", "earth_si1", "earth_si2",
                                          "moon_fe1", "moon_fe2", "moon_si1", "moon_si2")] + [
        ("mass", C.c_double * 4), ("com", (C.c_double * 3) * 4), ("vel", (C.c_double * 3) * 4),
        ("angmom", (C.c_double * 3) * 4)]


def lib_path() -> Path:
    return LIB_PATH


def load() -> C.CDLL:
    """Load libgiant_impact.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
            f"or make -C {PKG_DIR})"
        )
    lib = C.CDLL(str(LIB_PATH))
    vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
    P = C.POINTER
    sig = {
        "gi_abi_version": (i, []),
        "gi_last_error_string": (C.c_char_p, []),
        "gi_last_error": (i, []),
        "gi_clear_error": (None, []),
        "gi_sync": (i, []),
        "gi_device_count": (i, []),
        "gi_set_device": (i, [i]),
        "gi_check_block_multiple": (i, [i]),
        "gi_malloc": (i, [P(vp), sz]),
        "gi_free": (i, [vp]),
        "gi_copy_h2d": (i, [vp, vp, sz]),
        "gi_copy_d2h": (i, [vp, vp, sz]),
        "gi_copy_d2d": (i, [vp, vp, sz, vp]),
        "gi_stream_create": (i, [P(vp)]),
        "gi_stream_destroy": (i, [vp]),
        "gi_stream_sync": (i, [vp]),
        "gi_device_sync": (i, []),
        "gi_mem_info": (i, [P(sz), P(sz)]),
        "gi_force_collision": (i, [vp, vp, vp, i, P(ForceCollisionConstants), vp]),
        "gi_move_collision": (i, [vp, vp, vp, i, P(MoveCollisionConstants), vp]),
        "gi_force_separate": (i, [vp, vp, vp, i, P(ForceSeparateConstants), vp]),
        "gi_move_damped_separate": (
            i,
            [vp, vp, vp, i, P(MoveSeparateConstants), C.c_float, C.c_float, vp],
        ),
        "gi_set_kernel_timing": (i, [i]),
        "gi_kernel_timing": (i, [C.c_char_p, P(C.c_longlong), P(C.c_double)]),
        "gi_reset_kernel_timing": (i, []),
        "gi_generator_setup_defaults": (None, [P(GeneratorSetup), i]),
        "gi_generator_setup_read": (i, [C.c_char_p, P(GeneratorSetup)]),
        "gi_units_compute": (i, [P(GeneratorSetup), P(Units)]),
        "gi_separate_constants": (None, [P(Units), C.c_double, P(ForceSeparateConstants), P(MoveSeparateConstants)]),
        "gi_impact_constants": (None, [P(Units), C.c_double, P(ForceCollisionConstants), P(MoveCollisionConstants)]),
        "gi_parameters_write": (i, [C.c_char_p, P(Units), i]),
        "gi_parameters_read": (i, [C.c_char_p, P(Units), P(i)]),
        "gi_make_impact": (
            i,
            [P(ImpactConfig), vp, vp, P(ForceCollisionConstants), P(MoveCollisionConstants), P(ImpactInfo)],
        ),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _optional = {
        "gi_center_of_mass_velocity": (i, [vp, vp, i, i, C.c_double, C.c_double, P(C.c_double), vp]),
        "gi_recenter": (i, [vp, vp, i, i, C.c_double, C.c_double, P(C.c_double), P(C.c_double), vp]),
        "gi_recenter_float": (i, [vp, vp, i, i, C.c_double, C.c_double, P(C.c_float), P(C.c_float), vp]),
        "gi_impact_steps_float_recenter": (
            i,
            [vp, vp, vp, i, P(ForceCollisionConstants), P(MoveCollisionConstants), i, i, P(C.c_float), P(C.c_float),
             C.c_double, C.c_double, vp],
        ),
        "gi_conserved_quantities": (i, [vp, vp, i, P(ForceCollisionConstants), P(C.c_double), vp]),
        "gi_impact_steps": (
            i,
            [vp, vp, vp, i, P(ForceCollisionConstants), P(MoveCollisionConstants), i, i, P(C.c_double), P(C.c_double), vp],
        ),
        "gi_ensemble_create": (i, [P(vp), i, P(i), vp, vp, vp, vp, vp]),
        "gi_ensemble_create_capacity": (i, [P(vp), i, vp, vp, vp]),
        "gi_ensemble_set_sims": (i, [vp, i, P(i), P(i), vp, vp]),
        "gi_ensemble_wait": (i, [vp, i]),
        "gi_force_collision_noalloc": (i, [vp, vp, vp, i, P(ForceCollisionConstants), vp]),
        "gi_ensemble_destroy": (i, [vp]),
        "gi_ensemble_steps": (i, [vp, i]),
        "gi_ensemble_set_active": (i, [vp, P(i)]),
        "gi_ensemble_conserved": (i, [vp, P(C.c_double)]),
        "gi_ensemble_work": (i, [vp, P(C.c_longlong), P(C.c_double)]),
        "gi_set_tiling": (i, [i, i, C.c_longlong, C.c_longlong, C.c_longlong]),
        "gi_plan_info": (i, [P(i), P(i), P(C.c_longlong), P(C.c_longlong)]),
        "gi_tiling_for": (i, [C.c_longlong, P(i), P(i), P(C.c_longlong), P(C.c_longlong), P(C.c_longlong)]),
        "gi_find_earth_moon": (i, [vp, vp, i, i, i, i, C.c_double, C.c_double, C.c_float, vp, P(EarthMoon), vp]),
        "gi_write_run_stats": (i, [C.c_char_p, P(EarthMoon), P(Units), C.c_double]),
        "gi_set_batched_tiling": (i, [i, C.c_longlong, i]),
        "gi_set_dispatch_limit": (i, [C.c_longlong]),
        "gi_set_contact_overlap": (i, [C.c_int]),
        "gi_set_contact_walk": (i, [C.c_int]),
        "gi_serve_run": (i, [vp, i, vp, vp, vp]),
        "gi_spin_body": (i, [vp, vp, i, i, i, C.c_double, C.c_double, C.c_double, vp]),
        "gi_zero_out_drift": (i, [vp, vp, vp]),
        "gi_make_raw_bodies": (i, [vp, C.c_ulonglong, C.c_double, vp, vp]),
        "gi_generate_bodies": (i, [vp, C.c_ulonglong, C.c_double, vp, vp, vp, vp, C.c_char_p]),
        "gi_impact_initialize": (i, [C.c_char_p, C.c_char_p, C.c_char_p]),
        "gi_impact_run": (i, [C.c_char_p, C.c_char_p, C.c_char_p, vp]),
        "gi_separate_steps": (i, [vp, vp, vp, i, P(ForceSeparateConstants), P(MoveSeparateConstants), C.c_float,
                                  C.c_float, i, vp]),
    }
    for name, (res, args) in _optional.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    _lib = lib
    return lib


def check(func: str, rc: int) -> None:
    if rc != 0:
        msg = load().gi_last_error_string()
        raise GIError(func, rc, msg.decode() if msg else "")


def call(name: str, *args):
    rc = getattr(load(), name)(*args)
    check(name, rc)
    return rc


def device_count() -> int:
    return load().gi_device_count()


# ---------------------------------------------------------------------------
# Device arrays of float4 records
# ---------------------------------------------------------------------------


class DeviceArray:
    """n float4 records in device memory (gi_malloc), freed on close()/GC."""

    def __init__(self, n: int, init: np.ndarray | None = None):
        self.n = int(n)
        self.nbytes = max(16 * self.n, 16)
        self.ptr = C.c_void_p()
        call("gi_malloc", C.byref(self.ptr), self.nbytes)
        if init is not None:
            self.upload(init)

    @property
    def value(self) -> int:
        return self.ptr.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 4)
        assert a.shape[0] == self.n
        if self.n:
            call("gi_copy_h2d", self.ptr, a.ctypes.data_as(C.c_void_p), 16 * self.n)

    def download(self) -> np.ndarray:
        out = np.empty((self.n, 4), dtype=np.float32)
        if self.n:
            call("gi_copy_d2h", out.ctypes.data_as(C.c_void_p), self.ptr, 16 * self.n)
        return out

    def close(self) -> None:
        if self.ptr is not None and self.ptr.value:
            load().gi_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Host-side helpers
# ---------------------------------------------------------------------------


def generator_setup_defaults(n: int) -> GeneratorSetup:
    s = GeneratorSetup()
    load().gi_generator_setup_defaults(C.byref(s), n)
    return s


def units(setup: GeneratorSetup) -> Units:
    u = Units()
    call("gi_units_compute", C.byref(setup), C.byref(u))
    return u


def impact_constants(u: Units, dt: float = 0.002):
    fc, mc = ForceCollisionConstants(), MoveCollisionConstants()
    load().gi_impact_constants(C.byref(u), dt, C.byref(fc), C.byref(mc))
    return fc, mc


def separate_constants(u: Units, dt: float = 0.002):
    fc, mc = ForceSeparateConstants(), MoveSeparateConstants()
    load().gi_separate_constants(C.byref(u), dt, C.byref(fc), C.byref(mc))
    return fc, mc


def make_impact(
    n_target: int,
    n_impactor: int,
    *,
    fe_fraction: float = 0.3,
    jitter: float = 0.1,
    seed: int = 20261202,
    angle_deg: float = 45.0,
    speed_vesc: float = 1.2,
    target_spin: float = 0.0,
    impactor_spin: float = 0.0,
    gap_in_target_radii: float = 1.0,
):
    """Seeded synthetic impact in the reference impact layout [Fe1, Fe2, Si1, Si2].

    Returns (pos, vel, fc, mc, info) with pos/vel float32 arrays of shape (n, 4).
    """
    cfg = ImpactConfig(
        n_target, n_impactor, fe_fraction, jitter, seed, angle_deg, speed_vesc, target_spin, impactor_spin,
        gap_in_target_radii,
    )
    n = n_target + n_impactor
    pos = np.zeros((n, 4), dtype=np.float32)
    vel = np.zeros((n, 4), dtype=np.float32)
    fc, mc, info = ForceCollisionConstants(), MoveCollisionConstants(), ImpactInfo()
    call(
        "gi_make_impact",
        C.byref(cfg),
        pos.ctypes.data_as(C.c_void_p),
        vel.ctypes.data_as(C.c_void_p),
        C.byref(fc),
        C.byref(mc),
        C.byref(info),
    )
    return pos, vel, fc, mc, info


# ---------------------------------------------------------------------------
# The four step-loop entry points (device pointers)
# ---------------------------------------------------------------------------


def _p(x) -> C.c_void_p:
    if isinstance(x, DeviceArray):
        return x.ptr
    return C.c_void_p(int(x)) if x is not None else C.c_void_p()


def force_collision(pos, vel, force, n: int, fc: ForceCollisionConstants, stream=None) -> None:
    call("gi_force_collision", _p(pos), _p(vel), _p(force), int(n), C.byref(fc), _p(stream))


def move_collision(pos, vel, force, n: int, mc: MoveCollisionConstants, stream=None) -> None:
    call("gi_move_collision", _p(pos), _p(vel), _p(force), int(n), C.byref(mc), _p(stream))


def force_separate(pos, vel, force, n: int, fc: ForceSeparateConstants, stream=None) -> None:
    call("gi_force_separate", _p(pos), _p(vel), _p(force), int(n), C.byref(fc), _p(stream))


def move_damped_separate(pos, vel, force, n: int, mc: MoveSeparateConstants, d1: float, d2: float, stream=None):
    call("gi_move_damped_separate", _p(pos), _p(vel), _p(force), int(n), C.byref(mc), float(d1), float(d2), _p(stream))


def find_earth_moon(P, V, n: int, nfe1: int, nfe2: int, nsi1: int, mass_fe: float, mass_si: float = 1.0,
                    touch: float = 1.5, with_groups: bool = False, stream=None) -> dict:
    """The Earth / Moon finder on device arrays; scopes 0 system, 1 Earth-Moon,
    2 Earth, 3 Moon."""
    em = EarthMoon()
    G = DeviceArray(max(n, 1) // 4 + 1) if with_groups else None  # n ints = n/4 float4 records
    call("gi_find_earth_moon", _p(P), _p(V), int(n), int(nfe1), int(nfe2), int(nsi1), float(mass_fe), float(mass_si),
         float(touch), _p(G), C.byref(em), _p(stream))
    out = {k: getattr(em, k) for k, _ in EarthMoon._fields_[:13]}
    out["mass"] = np.array(em.mass[:])
    for k in ("com", "vel", "angmom"):
        out[k] = np.array([list(r) for r in getattr(em, k)])
    if with_groups:
        out["group"] = G.download().view(np.int32).reshape(-1)[:n].copy()
    return out


def tiling_for(n: int) -> dict:
    """The tiling a single simulation of n particles gets (host only)."""
    r, v = C.c_int(), C.c_int()
    L, t, m = C.c_longlong(), C.c_longlong(), C.c_longlong()
    call("gi_tiling_for", int(n), C.byref(r), C.byref(v), C.byref(L), C.byref(t), C.byref(m))
    return {"r": r.value, "variant": v.value, "chunk_len": L.value, "target_tasks": t.value, "min_chunk": m.value}


def plan_info() -> dict:
    """The plan of the last drop-in force call on the current device."""
    m, r, t, L = C.c_int(), C.c_int(), C.c_longlong(), C.c_longlong()
    call("gi_plan_info", C.byref(m), C.byref(r), C.byref(t), C.byref(L))
    return {"mutual": m.value in (1, 3), "direct": m.value == 2, "r": r.value, "pair_tasks": t.value,
            "chunk_len": L.value, "kernel": ("k_gravity", "k_mutual", "k_direct", "k_mutual_grp")[m.value]}


def sync() -> None:
    call("gi_device_sync")


# ---------------------------------------------------------------------------
# Batched ensemble
# ---------------------------------------------------------------------------


def ensemble_configs(n_sims: int, seed: int = 20261205, n_range=(4096, 32768), impactor_frac=(0.1, 0.3),
                     angle=(30.0, 60.0), speed=(1.0, 1.4), spin=(0.0, 0.6), log_uniform_n: bool = False) -> list:
    """Seeded impact parameters of an uneven ensemble (one dict per simulation)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_sims):
        if log_uniform_n:
            n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        else:
            n = int(rng.integers(n_range[0], n_range[1] + 1))
        f = float(rng.uniform(*impactor_frac))
        ni = max(1, int(round(f * n)))
        out.append(
            dict(
                n_target=n - ni,
                n_impactor=ni,
                seed=int(seed * 1000 + s),
                angle_deg=float(rng.uniform(*angle)),
                speed_vesc=float(rng.uniform(*speed)),
                target_spin=float(rng.uniform(*spin)),
            )
        )
    return out


def make_ensemble(configs: list):
    """Pack the impacts of `configs` back to back.

    Returns (pos, vel, n (int32 array), fc array, mc array, infos)."""
    parts = [make_impact(**c) for c in configs]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    n = np.array([len(p[0]) for p in parts], dtype=np.int32)
    fc = (ForceCollisionConstants * len(parts))(*[p[2] for p in parts])
    mc = (MoveCollisionConstants * len(parts))(*[p[3] for p in parts])
    return pos, vel, n, fc, mc, [p[4] for p in parts]


class Ensemble:
    """gi_ensemble_*: many impacts stepped together, one launch per kernel per step."""

    def __init__(self, P: DeviceArray, V: DeviceArray, n, fc, mc, stream=None):
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.n_sims = len(self.n)
        self.offsets = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
        self.handle = C.c_void_p()
        call(
            "gi_ensemble_create",
            C.byref(self.handle),
            self.n_sims,
            self.n.ctypes.data_as(C.POINTER(C.c_int)),
            C.cast(fc, C.c_void_p),
            C.cast(mc, C.c_void_p),
            P.ptr,
            V.ptr,
            _p(stream),
        )

    @classmethod
    def with_capacity(cls, P: DeviceArray, V: DeviceArray, capacity: int, stream=None) -> "Ensemble":
        """gi_ensemble_create_capacity: a pool of `capacity` records with no
        simulations yet (place them with set_sims)."""
        self = cls.__new__(cls)
        self.n = np.zeros(0, np.int32)
        self.n_sims = 0
        self.offsets = np.zeros(1, np.int64)
        self.handle = C.c_void_p()
        call("gi_ensemble_create_capacity", C.byref(self.handle), int(capacity), P.ptr, V.ptr, _p(stream))
        return self

    def set_sims(self, offsets, n, fc, mc) -> None:
        """gi_ensemble_set_sims: simulation s at records [offsets[s], offsets[s] + n[s])."""
        o = np.ascontiguousarray(offsets, dtype=np.int32)
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.n_sims = len(self.n)
        call("gi_ensemble_set_sims", self.handle, self.n_sims, o.ctypes.data_as(C.POINTER(C.c_int)),
             self.n.ctypes.data_as(C.POINTER(C.c_int)), C.cast(fc, C.c_void_p), C.cast(mc, C.c_void_p))

    def wait(self, max_pending: int = 0) -> None:
        call("gi_ensemble_wait", self.handle, int(max_pending))

    def steps(self, k: int) -> None:
        call("gi_ensemble_steps", self.handle, int(k))

    def set_active(self, active) -> None:
        a = np.ascontiguousarray(active, dtype=np.int32)
        call("gi_ensemble_set_active", self.handle, a.ctypes.data_as(C.POINTER(C.c_int)))

    def conserved(self) -> np.ndarray:
        out = np.zeros((self.n_sims, 9), dtype=np.float64)
        call("gi_ensemble_conserved", self.handle, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def work(self):
        p, q = C.c_longlong(), C.c_double()
        call("gi_ensemble_work", self.handle, C.byref(p), C.byref(q))
        return p.value, q.value

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            load().gi_ensemble_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def impact_steps(P, V, F, n: int, fc, mc, steps: int, recenter_rate: int = 0, com3=None, vel3=None, stream=None):
    c3 = (C.c_double * 3)(*com3) if com3 is not None else None
    v3 = (C.c_double * 3)(*vel3) if vel3 is not None else None
    call("gi_impact_steps", _p(P), _p(V), _p(F), int(n), C.byref(fc), C.byref(mc), int(steps), int(recenter_rate),
         c3, v3, _p(stream))


def recenter_float(P, V, n: int, nfe: int, mass_fe: float, mass_si: float, com3, vel3, stream=None) -> None:
    c3 = (C.c_float * 3)(*com3)
    v3 = (C.c_float * 3)(*vel3)
    call("gi_recenter_float", _p(P), _p(V), int(n), int(nfe), float(mass_fe), float(mass_si), c3, v3, _p(stream))


def impact_steps_float_recenter(P, V, F, n: int, fc, mc, steps: int, recenter_rate: int, com3, vel3,
                                mass_fe: float, mass_si: float = 1.0, stream=None) -> None:
    c3 = (C.c_float * 3)(*com3)
    v3 = (C.c_float * 3)(*vel3)
    call("gi_impact_steps_float_recenter", _p(P), _p(V), _p(F), int(n), C.byref(fc), C.byref(mc), int(steps),
         int(recenter_rate), c3, v3, float(mass_fe), float(mass_si), _p(stream))


def conserved_quantities(P, V, n: int, fc, stream=None) -> np.ndarray:
    out = (C.c_double * 9)()
    call("gi_conserved_quantities", _p(P), _p(V), int(n), C.byref(fc), out, _p(stream))
    return np.array(list(out))


# ---------------------------------------------------------------------------
# Request loop
# ---------------------------------------------------------------------------


class ServeRequest(C.Structure):
    _fields_ = [
        ("impact", ImpactConfig),
        ("steps", C.c_int),
        ("arrival_s", C.c_double),
        ("final_pos", C.c_void_p),
        ("final_vel", C.c_void_p),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("steps", C.c_int),
        ("arrival_s", C.c_double),
        ("admit_s", C.c_double),
        ("finish_s", C.c_double),
        ("time_to_result_s", C.c_double),
        ("energy0", C.c_double),
        ("energy1", C.c_double),
        ("angmom0", C.c_double * 3),
        ("angmom1", C.c_double * 3),
        ("energy_rel_change", C.c_double),
        ("angmom_rel_change", C.c_double),
    ]


class ServeOptions(C.Structure):
    _fields_ = [
        ("max_particles", C.c_longlong),
        ("max_sims", C.c_int),
        ("quantum_steps", C.c_int),
        ("hbm_fraction", C.c_double),
    ]


class ServeSummary(C.Structure):
    _fields_ = [
        ("completed", C.c_int),
        ("wall_s", C.c_double),
        ("total_sim_steps", C.c_double),
        ("sim_steps_per_s", C.c_double),
        ("pair_interactions_per_s", C.c_double),
        ("p50_time_to_result_s", C.c_double),
        ("p99_time_to_result_s", C.c_double),
        ("max_time_to_result_s", C.c_double),
        ("capacity_particles", C.c_longlong),
        ("peak_resident_particles", C.c_longlong),
        ("peak_resident_sims", C.c_int),
        ("rounds", C.c_int),
        ("max_energy_rel_change", C.c_double),
        ("max_angmom_rel_change", C.c_double),
    ]


def _struct_dict(s) -> dict:
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = list(v) if isinstance(v, C.Array) else v
    return out


def serve(requests: list, *, max_particles: int = 0, max_sims: int = 0, quantum_steps: int = 0,
          hbm_fraction: float = 0.0, keep_final: bool = False):
    """Run the request loop.  requests: dicts with the make_impact keywords plus
    'steps' and 'arrival_s'.  Returns (results, summary[, finals])."""
    k = len(requests)
    reqs = (ServeRequest * k)()
    finals = []
    for i, r in enumerate(requests):
        cfg = ImpactConfig(
            r["n_target"], r["n_impactor"], r.get("fe_fraction", 0.3), r.get("jitter", 0.1), r.get("seed", 1),
            r.get("angle_deg", 45.0), r.get("speed_vesc", 1.2), r.get("target_spin", 0.0),
            r.get("impactor_spin", 0.0), r.get("gap_in_target_radii", 1.0),
        )
        reqs[i].impact = cfg
        reqs[i].steps = int(r["steps"])
        reqs[i].arrival_s = float(r.get("arrival_s", 0.0))
        if keep_final:
            n = r["n_target"] + r["n_impactor"]
            fp, fv = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
            finals.append((fp, fv))
            reqs[i].final_pos = fp.ctypes.data
            reqs[i].final_vel = fv.ctypes.data
    res = (ServeResult * k)()
    summ = ServeSummary()
    opt = ServeOptions(max_particles, max_sims, quantum_steps, hbm_fraction)
    call("gi_serve_run", reqs, k, C.byref(opt), res, C.byref(summ))
    out = ([_struct_dict(x) for x in res], _struct_dict(summ))
    return out + (finals,) if keep_final else out


# ---------------------------------------------------------------------------
# Host drivers (generation, impact set-up, impact run)
# ---------------------------------------------------------------------------


class GenerateInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("damp_steps", C.c_longlong),
        ("rest_steps", C.c_longlong),
        ("spin_rest_steps", C.c_longlong),
        ("run_time", C.c_double),
    ]


class ImpactRunInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("steps", C.c_longlong),
        ("frames_written", C.c_int),
        ("start_time", C.c_double),
        ("end_time", C.c_double),
    ]


def generate_bodies(setup: GeneratorSetup, seed: int = 1, jitter: float = 0.1, out_dir=None):
    """Run the generation phases; returns (pos, vel, units, info) in the [Fe1, Si1, Fe2, Si2] layout."""
    n = setup.total_number_of_elements
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    u, info = Units(), GenerateInfo()
    call("gi_generate_bodies", C.byref(setup), seed, jitter, pos.ctypes.data_as(C.c_void_p),
         vel.ctypes.data_as(C.c_void_p), C.byref(u), C.byref(info),
         None if out_dir is None else str(out_dir).encode())
    return pos, vel, u, info


def impact_initialize(in_dir, setup_path, out_dir) -> None:
    call("gi_impact_initialize", str(in_dir).encode(), str(setup_path).encode(), str(out_dir).encode())


def impact_run(param_dir, setup_path, run_dir) -> ImpactRunInfo:
    info = ImpactRunInfo()
    call("gi_impact_run", str(param_dir).encode(), str(setup_path).encode(), str(run_dir).encode(), C.byref(info))
    return info


def read_frames(path, n: int):
    """Frames of an impactPosVel file: list of (time, pos (n,4), vel (n,4))."""
    raw = Path(path).read_bytes()
    fb = 4 + 32 * n
    if len(raw) % fb:
        raise ValueError(f"{path}: size {len(raw)} is not a multiple of the frame size {fb}")
    out = []
    for k in range(len(raw) // fb):
        b = raw[k * fb : (k + 1) * fb]
        t = np.frombuffer(b[:4], np.float32)[0]
        a = np.frombuffer(b[4:], np.float32).reshape(2, n, 4)
        out.append((float(t), a[0].copy(), a[1].copy()))
    return out

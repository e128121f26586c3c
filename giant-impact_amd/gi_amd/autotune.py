"""Autotuner of the all-pairs pass (run on the GPU).

For each size class it times the whole force evaluation (HIP events, via
gi_kernel_timing("force"): contact grid, all-pairs pass and contact
correction, or the direct pass alone) over a grid of tilings -- the
one-sided pass with R (i-particles per lane), the r^-3 variant and the
number of wavefront tasks, the mutual pass (variant 2) and its row-group form
(variant 4) with R (rows of 64*R) and their chunk length, and for the small
classes the direct pass (variant 3)
-- and writes the fastest to tuning/gfx950.txt, which libgiant_impact.so
reads on first use:

    n_max R variant chunk_len target_tasks min_chunk

Usage (on an MI355X):  python -m gi_amd.autotune [--quick] [--out PATH]
"""
from __future__ import annotations

import argparse
import ctypes as C
import itertools
import sys
import time
from pathlib import Path

import numpy as np

from . import DeviceArray, call, device_count, force_collision, load, make_impact, sync

TUNING_DIR = Path(__file__).resolve().parent.parent / "tuning"

# (n_max of the class, n_target, n_impactor of the representative impact)
CLASSES = [
    (3072, 1536, 512),
    (6144, 3072, 1024),
    (12288, 6144, 2048),
    (16384, 10752, 3584),
    (24576, 15360, 5120),
    (49152, 24576, 8192),
    (98304, 49152, 16384),
    (196608, 98304, 32768),
    (393216, 196608, 65536),
    (2097152, 786432, 262144),
]


# classes up to this size also try the direct pass (one launch per force;
# it wins where the three-pass form is bound by its launches)
DIRECT_MAX_CLASS = 49152


def time_gravity(n_target: int, n_impactor: int, reps: int):
    pos, vel, fc, _, _ = make_impact(n_target, n_impactor, seed=77)
    n = len(pos)
    P, V, F = DeviceArray(n, pos), DeviceArray(n, vel), DeviceArray(n, np.zeros((n, 4), np.float32))
    lib = load()

    def run() -> float:
        force_collision(P, V, F, n, fc)  # plan + warm
        sync()
        lib.gi_set_kernel_timing(1)
        lib.gi_reset_kernel_timing()
        for _ in range(reps):
            force_collision(P, V, F, n, fc)
        sync()
        lib.gi_set_kernel_timing(0)
        L, T = C.c_longlong(), C.c_double()
        call("gi_kernel_timing", b"force", C.byref(L), C.byref(T))
        return T.value / max(L.value, 1)

    return run, n


def tune_batched(quick: bool):
    """Pass, chunk length and R of the batched engine, timed on the
    64-simulation uneven ensemble (N uniform in [4096, 32768])."""
    from . import Ensemble, ensemble_configs, make_ensemble

    pos, vel, n, fc, mc, _ = make_ensemble(ensemble_configs(64))
    ntot = int(n.sum())
    pairs = float((n.astype(np.float64) ** 2).sum())
    P, V = DeviceArray(ntot, pos), DeviceArray(ntot, vel)
    best = None
    Ls = [1024, 2048, 4096, 8192, 16384] if not quick else [4096]
    cands = [(R, L, 0) for R, L in itertools.product([1, 2], Ls)]
    cands += [(R, L, v) for R, L, v in itertools.product([4, 8], [256, 384, 512, 768, 1024, 1536, 2048], [2, 4])]
    for R, L, v in cands:
        call("gi_set_batched_tiling", R, L, v)
        P.upload(pos)
        V.upload(vel)
        e = Ensemble(P, V, n, fc, mc)
        e.steps(2)
        sync()
        dt = float("inf")
        for _ in range(3):  # the best of three segments of 10 steps
            t0 = time.perf_counter()
            e.steps(10)
            sync()
            dt = min(dt, (time.perf_counter() - t0) / 10)
        e.close()
        print(f"batched R={R} v={v} L={L:6d}: {dt * 1e3:8.3f} ms/step  {pairs / dt:.3e} pair/s", flush=True)
        if best is None or dt < best[0]:
            best = (dt, L, R, v)
    call("gi_set_batched_tiling", 0, 0, 0)
    print(f"# batched: best chunk_len={best[1]} R={best[2]} variant={best[3]} ({pairs / best[0]:.3e} pair/s)",
          flush=True)
    return best[1], best[2], best[3]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="fewer candidates (smoke run)")
    ap.add_argument("--out", default=str(TUNING_DIR / "gfx950.txt"))
    ap.add_argument("--max-class", type=int, default=len(CLASSES))
    ap.add_argument("--classes", default="", help="comma-separated class indices to tune (default: all up to "
                    "--max-class); rows of the other classes are kept from the current table")
    ap.add_argument("--variants", action="store_true", help="also sweep the rsq*rcp variant")
    ap.add_argument("--batched", action="store_true", help="also tune the batched engine (64-simulation ensemble)")
    args = ap.parse_args(argv)
    if device_count() < 1:
        print("autotune: no HIP device", file=sys.stderr)
        return 1
    lib = load()
    Rs = [1, 2, 4, 8]
    # variant 1 (rsq * rcp) lost to variant 0 at every size on gfx950 (20% slower
    # at N = 262,144); --variants sweeps it again
    variants = [0, 1] if args.variants else [0]
    targets = [4096, 8192, 16384, 32768, 65536, 131072] if not args.quick else [32768]
    mutual_Ls = [192, 256, 384, 512, 768, 1024, 1536, 2048] if not args.quick else [512]
    rows = []
    only = {int(x) for x in args.classes.split(",") if x.strip()}
    kept = {}
    cur = TUNING_DIR / "gfx950.txt"
    if cur.exists():
        for line in cur.read_text().splitlines():
            parts = line.split()
            if parts and not parts[0].startswith("#") and parts[0] != "B":
                kept[int(parts[0])] = tuple(int(x) for x in parts)
    for ci, (n_max, nt, ni) in enumerate(CLASSES[: args.max_class]):
        if only and ci not in only:
            if n_max in kept:
                rows.append(kept[n_max])
            continue
        run, n = time_gravity(nt, ni, reps=3 if nt + ni > 100000 else 10)
        mc = 128 if n < 4096 else 512
        # (R, variant, chunk_len, target_tasks, min_chunk)
        cands = [(R, v, 0, tt, mc) for R, v, tt in itertools.product(Rs, variants, targets)]
        cands += [(R, v, L, 0, 64) for R, L, v in itertools.product([4, 8], mutual_Ls, [2, 4])]
        if n_max <= DIRECT_MAX_CLASS:
            cands += [(S, 3, 0, 0, 0) for S in (1, 2, 4, 8)]  # R = lanes per i-particle
        best = None
        for cand in cands:
            call("gi_set_tiling", *cand)
            ms = run()
            pairs = n * n / (ms / 1e3)
            print(f"n={n:8d} R={cand[0]} v={cand[1]} L={cand[2]:5d} tasks={cand[3]:6d}: {ms:9.4f} ms  "
                  f"{pairs:.3e} pair/s", flush=True)
            if best is None or ms < best[0]:
                best = (ms, cand)
        ms, cand = best
        print(f"# class n<={n_max}: best {cand} ({n * n / (ms / 1e3):.3e} pair/s)", flush=True)
        rows.append((n_max,) + cand)
    call("gi_set_tiling", 0, 0, 0, 0, 0)
    batched = tune_batched(args.quick) if args.batched else None
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    with open(out, "w") as f:
        f.write("# all-pairs tiling for gfx950, written by gi_amd.autotune on " + time.strftime("%Y-%m-%d") + "\n")
        f.write("# n_max R variant chunk_len target_tasks min_chunk\n")
        for r in rows:
            f.write(" ".join(str(x) for x in r) + "\n")
        if batched:
            f.write("# batched engine: B chunk_len R variant\n")
            f.write(f"B {batched[0]} {batched[1]} {batched[2]}\n")
        elif cur.exists():  # keep the current batched row
            for line in cur.read_text().splitlines():
                if line.startswith("# batched") or line.startswith("B "):
                    f.write(line + "\n")
    print(f"wrote {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

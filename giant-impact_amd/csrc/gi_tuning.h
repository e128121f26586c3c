// gi_tuning.h -- tiling of the all-pairs pass and its tuning table.
//
// R            i-particles held in registers per lane (a wavefront covers 64*R)
// variant      0 = one-sided pass, r^-3 as rsq cubed; 1 = one-sided, rsq *
//              rcp (1 multiply, 2 transcendental ops); 2 = mutual pass (each
//              pair once for both particles, rows of 64*R, rsq cubed);
//              3 = direct pass (small simulations: the whole force in one
//              launch; R = lanes per i-particle, 1, 2 or 4 (8 acts as 4),
//              64 / R i-particles per workgroup; the sizes unused);
//              4 = mutual pass with row groups (a workgroup's four
//              wavefronts hold four rows and add their j-side sums in LDS
//              before writing them: a quarter of the j-side slab; R = 4 or 8)
// chunk_len    j-particles per task (0 = derive from target_tasks)
// target_tasks wavefront tasks to aim for (several per SIMD of the 256 CUs x
//              4 SIMDs so the tail of the launch stays short)
// min_chunk    lower bound on chunk_len (each chunk costs one partial slot
//              per particle)
//
// The table is written by the autotuner (gi_amd.autotune) to
// tuning/<arch>.txt next to the library's lib/ directory, one row per size
// class: "n_max R variant chunk_len target_tasks min_chunk" (particles of
// the whole launch, rows in increasing n_max).  Without a table the defaults
// below apply.
#pragma once

#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace gi {

struct Tiling {
    int R;
    int variant;
    long long chunk_len;
    long long target_tasks;
    long long min_chunk;
};

struct TilingOverride {
    int active = 0;
    Tiling t{4, 0, 0, 8192, 512};
};

inline TilingOverride& tiling_override() {
    static TilingOverride o;
    return o;
}

struct TuningRow {
    long long n_max;
    Tiling t;
};

inline void tuning_anchor() {}

// Rows of tuning/gfx950.txt, located relative to this shared library.
inline const std::vector<TuningRow>& tuning_table() {
    static std::vector<TuningRow> rows;
    static std::once_flag once;
    std::call_once(once, [] {
        Dl_info info;
        if (!dladdr((void*)&tuning_anchor, &info) || !info.dli_fname) return;
        std::string path(info.dli_fname);
        size_t cut = path.rfind('/');
        if (cut == std::string::npos) return;
        path = path.substr(0, cut);  // .../lib
        cut = path.rfind('/');
        path = (cut == std::string::npos ? std::string(".") : path.substr(0, cut)) + "/tuning/gfx950.txt";
        FILE* f = std::fopen(path.c_str(), "r");
        if (!f) return;
        char line[256];
        while (std::fgets(line, sizeof line, f)) {
            if (line[0] == '#') continue;
            TuningRow r;
            if (std::sscanf(line, "%lld %d %d %lld %lld %lld", &r.n_max, &r.t.R, &r.t.variant, &r.t.chunk_len,
                            &r.t.target_tasks, &r.t.min_chunk) == 6 &&
                (r.t.R == 1 || r.t.R == 2 || r.t.R == 4 || r.t.R == 8) && r.t.variant >= 0 && r.t.variant <= 4 &&
                (r.t.variant != 4 || r.t.R >= 4) &&
                r.t.chunk_len >= 0 && r.t.target_tasks >= 0 &&
                (r.t.target_tasks > 0 || r.t.chunk_len > 0 || r.t.variant == 3))
                rows.push_back(r);
        }
        std::fclose(f);
    });
    return rows;
}

// Tiling of the batched engine (ensembles and the request loop): one j-chunk
// length for every simulation, whatever its size or company, so that a
// simulation's sums are the same in any batch.  Row "B chunk_len R [variant]"
// of the tuning file; gi_set_batched_tiling overrides it.
struct BatchedTiling {
    long long chunk_len;
    int R;
    int variant;
};

inline BatchedTiling& batched_override() {
    static BatchedTiling b{0, 0, 0};
    return b;
}

inline BatchedTiling batched_tiling() {
    if (batched_override().chunk_len > 0) return batched_override();
    static BatchedTiling from_file{0, 0, 0};
    static std::once_flag once;
    std::call_once(once, [] {
        Dl_info info;
        if (!dladdr((void*)&tuning_anchor, &info) || !info.dli_fname) return;
        std::string path(info.dli_fname);
        size_t cut = path.rfind('/');
        if (cut == std::string::npos) return;
        path = path.substr(0, cut);
        cut = path.rfind('/');
        path = (cut == std::string::npos ? std::string(".") : path.substr(0, cut)) + "/tuning/gfx950.txt";
        FILE* f = std::fopen(path.c_str(), "r");
        if (!f) return;
        char line[256];
        while (std::fgets(line, sizeof line, f)) {
            long long L;
            int R, V = 0;
            const int got = std::sscanf(line, "B %lld %d %d", &L, &R, &V);
            if (got >= 2 && L >= 64 && (R == 1 || R == 2 || R == 4 || R == 8) &&
                (V == 0 || V == 1 || V == 2 || (V == 4 && R >= 4)))
                from_file = BatchedTiling{L, R, V};
        }
        std::fclose(f);
    });
    if (from_file.chunk_len > 0) return from_file;
    return BatchedTiling{4096, 1, 0};
}

inline Tiling choose_tiling(long long ntot, long long pairs, int nsims) {
    (void)pairs;
    (void)nsims;
    if (tiling_override().active) return tiling_override().t;
    for (const TuningRow& r : tuning_table())
        if (ntot <= r.n_max) return r.t;
    Tiling t{4, 0, 0, 8192, 512};
    if (ntot < 4096) t.min_chunk = 128;
    return t;
}

}  // namespace gi

// gi_tuning.h -- tile sizes of the all-pairs pass.
//
// R           i-particles held in registers per lane (a wavefront covers 64*R)
// chunk_len   j-particles per task (0 = derive from target_tasks)
// target_tasks number of wavefront tasks to aim for (several per SIMD of the
//             256 CUs x 4 SIMDs so the tail of the launch stays short)
// min_chunk   lower bound on chunk_len (each chunk costs one partial slot
//             per particle)
#pragma once

namespace gi {

struct Tiling {
    int R;
    long long chunk_len;
    long long target_tasks;
    long long min_chunk;
};

// Overridable at run time by gi_set_tiling (the autotuner's entry point).
struct TilingOverride {
    int active = 0;
    Tiling t{4, 0, 8192, 512};
};

inline TilingOverride& tiling_override() {
    static TilingOverride o;
    return o;
}

inline Tiling choose_tiling(long long ntot, long long pairs, int nsims) {
    if (tiling_override().active) return tiling_override().t;
    (void)pairs;
    (void)nsims;
    Tiling t{4, 0, 8192, 512};
    if (ntot < 4096) t.min_chunk = 128;
    return t;
}

}  // namespace gi

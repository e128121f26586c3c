// gi_abi.hip -- C ABI (include/giant_impact.h) over the gfx950 kernels.
//
// Host-side responsibilities: per-device scratch (partial sums, contact grid,
// task tables), the work plan of the all-pairs pass (tiling chosen by the
// autotuned table in gi_tuning.h), kernel launches in stream order, HIP-event
// timing of the all-pairs kernel, and HIP-graph capture of the step loop.
#include "gi_kernels.hip"
#include "gi_earth_moon.hip"

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/giant_impact.h"
#include "gi_graph_cache.h"
#include "gi_tuning.h"

static_assert(sizeof(gi_float4) == 16, "float4 layout");
static_assert(sizeof(gi_force_collision_constants) == 56, "forceCollisionKernalConstantsStruct is 56 bytes");
static_assert(sizeof(gi_move_collision_constants) == 16, "moveCollisionKernalConstantsStruct is 16 bytes");
static_assert(sizeof(gi_force_separate_constants) == 64, "forceSeperateKernalConstantsStruct is 64 bytes");
static_assert(sizeof(gi_move_separate_constants) == 24, "moveSeperateKernalConstantsStruct is 24 bytes");

namespace gi {

static thread_local std::string g_last_error;
static thread_local int g_last_code = 0;

static int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    g_last_code = code;
    return code;
}

#define GI_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error((int)e_, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                            \
    } while (0)

// ---------------------------------------------------------------------------
// Device scratch
// ---------------------------------------------------------------------------

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    int ensure(size_t n) {
        if (n <= cap && p) return 0;
        release();
        size_t want = n < 64 ? 64 : n;
        hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_error((int)e, "hipMalloc(%zu bytes): %s", want * sizeof(T), hipGetErrorString(e));
        }
        cap = want;
        return 0;
    }
};

// HIP events around one timed launch; kind 0 = all-pairs gravity, 1 = move.
// kind 2 = the whole all-pairs pass (for the mutual form: padding, the
// all-pairs kernel and the per-particle sums).
constexpr int kTimedKinds = 4;
struct TimingPair {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0;
};

// The plan of one force evaluation over a set of packed simulations.
struct Plan {
    std::vector<SimDesc> caller_sims;  // the simulations as requested
    std::vector<SimDesc> sims;         // with the planner's fields filled in
    std::vector<GravTask> tasks;
    std::vector<int> block_sim, block_first, sim_first_block;
    Tiling tiling{4, 0, 0, 0, 0};
    int plain_R = 4;  // i-particles per lane of the one-sided kernel's tasks
    int ntot = 0;
    int partial_len = 0;  // float4 elements
    int buckets = 0;      // contact-grid buckets over all simulations
    // mutual all-pairs pass (tiling.variant == 2, or 4 with row groups)
    bool mutual = false;
    int group_rows = 1;                // W: rows per group (4 with row groups, else 1)
    std::vector<GroupTask> gtasks;     // row groups: one per workgroup
    std::vector<MutualTask> mtasks;
    std::vector<RowDesc> rows;
    std::vector<SumBlock> sblocks;  // k_mutual_sum's workgroups
    long long gtot = 0;       // padded particles
    long long iside_len = 0;  // floats
    long long jside_len = 0;  // floats
    long long jintra_len = 0;  // floats (row groups: pairs inside a group)
    long long chunk_len = 0;  // j per task of the mutual pass (0 if none)
    // every simulation takes the mutual pass: the contact corrections run
    // beside it (k_contact_corr) and k_mutual_sum writes the force
    bool split_finish = false;
    bool overlap_allowed = true;  // gi_set_contact_overlap when planned
    int contact_lanes = 1;        // lanes per particle of the contact walk (1 or 9)
    int contact_walk = 0;         // gi_set_contact_walk when planned
    // direct pass (tiling.variant == 3): one workgroup per 64 / direct_lanes
    // i-particles
    bool direct = false;
    int direct_lanes = 1;
    std::vector<DirectBlock> dblocks;
};

// Scratch and plan of one owner (the drop-in entry points of a device, or
// one ensemble).
struct Workspace {
    DevBuf<float4> partial;
    DevBuf<float4> gpos, gsum;     // mutual pass: padded positions, per-particle sums
    DevBuf<float> iside, jside;    // mutual pass: row slots, j-side slabs
    DevBuf<float> jintra;          // row groups: j-side entries inside a group
    DevBuf<GroupTask> gtasks;
    DevBuf<MutualTask> mtasks;
    DevBuf<RowDesc> rows;
    DevBuf<SumBlock> sblocks;
    DevBuf<DirectBlock> dblocks;
    DevBuf<int> counts, starts, sorted, cellhash, block_sums;
    DevBuf<float4> spos, svel;  // positions / velocities in bucket order
    DevBuf<float> corr;         // contact corrections: per particle (sum, compensation) of x, y, z
    DevBuf<GravTask> tasks;
    DevBuf<SimDesc> sims;
    DevBuf<int> block_sim, block_first, sim_first_block;
    DevBuf<double> moments, sums, pe;
    Plan plan;
    bool plan_valid = false;
    bool batched = false;               // plans with the batched engine's tiling
    unsigned long long generation = 0;  // bumped on every re-plan
};

// A captured step sequence (the graph cache's payload).
struct GraphPayload {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void destroy() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};
using GraphCacheHip = GraphCache<GraphPayload>;

// Scratch of the Earth / Moon finder.
struct CCWork {
    DevBuf<int> counts, starts, cellhash, sorted, block_sums, parent, size, mcounts;
    DevBuf<unsigned long long> best;
    DevBuf<double> moments;
};

struct DevCtx {
    int device = -1;
    Workspace ws;
    CCWork cc;
    DevBuf<double> rc_blocks, rc_sums;
    DevBuf<float> rc_shift;  // the reference-float recenter's six shifts
    hipStream_t capture_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // side stream of a force evaluation (default priority): the contact grid
    // is built on it while the all-pairs pass runs
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // second state of the steppers' multi-step graphs with the move fused
    // into the direct pass (steps alternate between the caller's arrays and
    // these)
    DevBuf<float4> alt_pos, alt_vel;
    GraphCacheHip graphs;
    std::vector<TimingPair> pending, pool;
    long long timed_launches[kTimedKinds] = {0, 0, 0, 0};
    double timed_ms[kTimedKinds] = {0.0, 0.0, 0.0, 0.0};
};

static std::mutex g_ctx_mutex;
static DevCtx* g_ctx[64];
static bool g_timing = false;

static int current_ctx(DevCtx** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(GI_E_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= 64) return set_error(GI_E_NO_DEVICE, "device id %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_ctx_mutex);
    if (!g_ctx[dev]) {
        g_ctx[dev] = new DevCtx();
        g_ctx[dev]->device = dev;
    }
    *out = g_ctx[dev];
    return 0;
}

// ---------------------------------------------------------------------------
// Planning
// ---------------------------------------------------------------------------

static PairLaw law_of(const gi_force_collision_constants* c) {
    PairLaw L;
    std::memcpy(&L, c, sizeof(PairLaw));  // the 13 leading floats, in struct order
    return L;
}

static PairLaw law_of(const gi_force_separate_constants* c) {
    PairLaw L;
    std::memcpy(&L, c, sizeof(PairLaw));
    return L;
}
static_assert(sizeof(PairLaw) == 13 * sizeof(float), "PairLaw mirrors the 13 leading floats");

static bool same_sims(const std::vector<SimDesc>& a, const std::vector<SimDesc>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(SimDesc)) == 0;
}

static bool same_tiling(const Tiling& a, const Tiling& b) {
    return a.R == b.R && a.variant == b.variant && a.chunk_len == b.chunk_len &&
           a.target_tasks == b.target_tasks && a.min_chunk == b.min_chunk;
}

// One-sided j-chunk length of a simulation of n particles planned alone.  It
// depends on that simulation's own size only (through its row of the tuning
// table), never on the rest of the launch: chunk boundaries fix the order of
// every float sum.  A simulation whose row asks for the mutual pass but is
// too large for it takes the one-sided pass with the default task count.
static long long sim_chunk_len(long long n) {
    Tiling t = choose_tiling(n, n * n, 1);
    if (t.variant >= 2) t = Tiling{1, 0, 0, 32768, 512};
    long long L = t.chunk_len;
    if (L <= 0) {
        const long long ib = 64LL * t.R;
        L = ((n + ib - 1) / ib) * n / (t.target_tasks > 0 ? t.target_tasks : 1);
    }
    if (L < t.min_chunk) L = t.min_chunk;
    if (L < 64) L = 64;
    return (L + 63) / 64 * 64;
}

// j-side slabs of the mutual pass grow as 3 N^2 / (128 R W) floats per
// simulation (W rows per slab group: 1, or 4 in the row-group form); a
// simulation whose slab would exceed this many floats (16 GiB: about
// N > 1.2 M at R = 8 and W = 1, 2.4 M at R = 8 and W = 4) takes the one-sided
// pass.  The choice depends on the simulation and the tiling alone, so its
// sums are the same in any batch.
constexpr long long kJsideMaxFloatsPerSim = 1LL << 32;
constexpr long long kMutualMaxChunk = 4096;

// Mutual pass: padded layout, rows of B = 64 R, one diagonal task per row and
// the row's j-range beyond it cut (at the iron/silicate border) into chunks
// of at most L in whole 64-blocks.  Simulations left with gn = 0 (inactive,
// empty or too large) are for the one-sided pass.
// The row-group form of the mutual pass (variant 4): per simulation, groups
// of W rows; each group's j-range beyond its last row is cut (at the
// iron/silicate border) into pieces of at most L as whole 64-blocks, one
// workgroup task each; every row has a diagonal task and, unless it is the
// last of its group, tasks against the rest of its group (per material
// segment, cut into pieces of at most L too: uncut, a row against up to
// 3 x 512 j was the longest task of a small simulation -- one 8,192-particle
// simulation in the batched engine stepped in 0.41 ms uncut, 0.24 cut).  The
// diagonal is cut likewise, into ranges of the row's 64-blocks of about the
// work of a piece of L.  A row's i-side slots are its group's pieces in
// order, then its diagonal pieces, then its in-group tasks: contiguous, as
// the sums expect.  In the batched engine each simulation's L is the
// engine's, shortened to N / kBatchedPieceDiv (whole 64-blocks, at least
// one) for simulations too small to fill the GPU with pieces of the full
// length: a function of the simulation alone, so its sums are the same in
// any batch.  In a single-simulation plan the last groups (the shortest
// j-ranges, whose pieces sort last) take pieces of L / 2 until they hold
// kTailPieces pieces of L -- one round of the grouped kernel's resident
// workgroups on MI355X (256 CUs x 6) -- so the launch drains through
// half-length tasks: 0.4% faster at N = 262,144 (pieces of 1,536) than every
// group at L, and also faster than cutting 768 or 3,072 pieces' worth, or
// to L / 3 or L / 4.  In a batched plan the count must depend on the
// simulation alone: its last groups holding an eighth of its pieces take
// L / 2 (one 30,000-particle simulation stepped 16% faster, eight mixed ones
// 2%, the 64-simulation ensemble the same within 0.2%; a quarter and a
// sixteenth gained less).
constexpr long long kBatchedPieceDiv = 32;
constexpr long long kTailPieces = 1536;
constexpr long long kBatchedTailDiv = 8;

static int plan_mutual_groups(const std::vector<SimDesc>& sims, const Tiling& tl, long long L0, int W, Plan& p,
                              std::vector<MutualTask>& tasks, std::vector<RowDesc>& rows, int& slot, bool batched) {
    const int B = 64 * tl.R;
    for (size_t si = 0; si < sims.size(); ++si) {
        const SimDesc& s = sims[si];
        if (s.gn == 0) continue;
        const long long L = batched ? std::min(L0, std::max(64LL, ((s.end - s.start) / kBatchedPieceDiv + 63) / 64 * 64))
                                    : L0;
        const int nrows = s.gn / B;
        const int fe_real = (s.fe_end - s.start + 63) / 64 * 64;
        const int si_real = s.gfe + (s.end - s.fe_end + 63) / 64 * 64;
        // pieces of at most len over [a0, a1), whole 64-blocks
        auto npieces = [](int a0, int a1, long long len) {
            return a1 <= a0 ? 0LL : std::min<long long>((a1 - a0) / 64, (a1 - a0 + len - 1) / len);
        };
        // pieces of L of group g
        auto group_pieces = [&](int g) {
            const int gend = std::min((g + 1) * W, nrows) * B;
            return npieces(gend, fe_real, L) + npieces(std::max(gend, s.gfe), si_real, L);
        };
        // groups [g_tail, ...) take pieces of Lt
        int g_tail = (nrows + W - 1) / W;
        const long long Lt = std::max(64LL, (L / 2 + 63) / 64 * 64);
        long long want = kTailPieces;
        if (batched) {
            long long total = 0;
            for (int g = 0; g * W < nrows; ++g) total += group_pieces(g);
            want = (total + kBatchedTailDiv - 1) / kBatchedTailDiv;
        }
        for (long long held = 0; g_tail > 0 && held < want;) held += group_pieces(--g_tail);
        for (int g = 0; g * W < nrows; ++g) {
            const int r0 = g * W, r1 = std::min(r0 + W, nrows), gend = r1 * B;
            const long long Lg = g >= g_tail ? Lt : L;
            struct Piece {
                int a, b, jfe;
            };
            std::vector<Piece> pieces;
            const int seg0[2] = {gend, std::max(gend, s.gfe)};
            const int seg1[2] = {fe_real, si_real};
            for (int sg = 0; sg < 2; ++sg) {
                const int a0 = seg0[sg], a1 = seg1[sg];
                if (a1 <= a0) continue;
                const long long blocks = (a1 - a0) / 64;
                const long long np = npieces(a0, a1, Lg);
                for (long long k = 0; k < np; ++k)
                    pieces.push_back(Piece{a0 + (int)(64 * (blocks * k / np)), a0 + (int)(64 * (blocks * (k + 1) / np)),
                                           sg == 0 ? 1 : 0});
            }
            for (int r = r0; r < r1; ++r) {
                RowDesc rd{slot, 0};
                slot += (int)pieces.size();
                const int rb = r * B;
                const bool rowfe = rb < s.gfe;
                // the row against itself: block m costs m + 1 sub-block
                // passes; cut in pieces of about the work of a piece of L
                const int re = std::min(rb + B, rowfe ? fe_real : si_real);
                const long long nb = (re - rb) / 64, total = nb * (nb + 1) / 2;
                const long long target = std::max(1LL, L / 64 * tl.R);
                const long long nd = std::min(nb, (total + target - 1) / target);
                long long m = 0, done = 0;  // blocks [0, m) taken, their work
                for (long long k = 1; k <= nd && m < nb; ++k) {
                    const long long m_begin = m, goal = total * k / nd;
                    do done += ++m;  // block m - 1, work m
                    while (m < nb && done + (m + 1) <= goal);
                    if (k == nd)
                        while (m < nb) done += ++m;
                    tasks.push_back(MutualTask{(int)si, r, rb + (int)(64 * m_begin), rb + (int)(64 * m), slot++,
                                               rowfe ? 1 : 0, 1, 0});
                }
                // the rest of the group: iron part, then silicate part
                const int a = rb + B, b = gend;
                auto push_pieces = [&](int a0, int a1, int jfe) {
                    const long long blocks = (a1 - a0) / 64;
                    long long np = (a1 - a0 + L - 1) / L;
                    if (np > blocks) np = blocks;
                    for (long long k = 0; k < np; ++k)
                        tasks.push_back(MutualTask{(int)si, r, a0 + (int)(64 * (blocks * k / np)),
                                                   a0 + (int)(64 * (blocks * (k + 1) / np)), slot++, jfe, 0, 0});
                };
                const int ia = a, ib = std::min(b, fe_real);
                if (ia < ib) push_pieces(ia, ib, 1);
                const int sa = std::max(a, s.gfe), sb = std::min(b, si_real);
                if (sa < sb) push_pieces(sa, sb, 0);
                rd.nslot = slot - rd.slot0;
                rows.push_back(rd);
            }
            for (size_t k = 0; k < pieces.size(); ++k)
                p.gtasks.push_back(GroupTask{0, (int)si, g, pieces[k].a, pieces[k].b, (int)k, pieces[k].jfe, 0});
        }
    }
    return 0;
}

// Entries of a simulation's j-side slab (per component) with groups of W
// rows: group g holds one entry per j beyond its last row.
static long long group_slab_entries(long long gn, long long nrows, int B, int W) {
    long long e = 0;
    for (long long g = 0; g * W < nrows; ++g) e += gn - std::min((g + 1) * W, nrows) * B;
    return e;
}

static int plan_mutual(std::vector<SimDesc>& sims, const Tiling& tl, bool batched, Plan& p) {
    const int B = 64 * tl.R;
    const int W = tl.variant == 4 ? kGroupRows : 1;
    long long gb = 0, jb = 0, ji = 0, off_total = 0, row = 0;
    for (auto& s : sims) {
        s.gbase = (int)gb;
        s.row0 = (int)row;
        s.jbase = jb;
        s.jibase = ji;
        s.gfe = s.gn = 0;
        if (s.end == s.start || !s.active) continue;
        const long long nfe = s.fe_end - s.start, nsi = s.end - s.fe_end;
        const long long gfe = (nfe + B - 1) / B * B;
        const long long gn = gfe + (nsi + B - 1) / B * B;
        const long long nrows = gn / B;
        const long long slab = 3 * group_slab_entries(gn, nrows, B, W);
        if (slab > kJsideMaxFloatsPerSim) continue;
        s.gfe = (int)gfe;
        s.gn = (int)gn;
        gb += gn;
        row += nrows;
        off_total += gn * (nrows - 1) / 2;
        jb += slab;
        if (W > 1) ji += 3LL * (W - 1) * B * nrows;
    }
    if (gb == 0) return 0;
    if (jb > (1LL << 35)) return set_error(GI_E_TOO_LARGE, "mutual pass: j-side slabs too large (%lld floats)", jb);
    long long L = tl.chunk_len;
    if (L <= 0 && !batched) {
        const long long tt = tl.target_tasks > 0 ? tl.target_tasks : 8192;
        L = (off_total + tt - 1) / tt;
        if (L < tl.min_chunk) L = tl.min_chunk;
    }
    // at most 4096 j per task: its float i-side sums stay short (see k_mutual)
    L = std::min(kMutualMaxChunk, std::max(64LL, (L + 63) / 64 * 64));
    std::vector<MutualTask> tasks;
    std::vector<RowDesc> rows;
    int slot = 0;
    if (W > 1) {
        int rc = plan_mutual_groups(sims, tl, L, W, p, tasks, rows, slot, batched);
        if (rc) return rc;
    }
    for (size_t si = 0; si < sims.size() && W == 1; ++si) {
        const SimDesc& s = sims[si];
        if (s.gn == 0) continue;
        const int nrows = s.gn / B;
        // End of each segment's last 64-block holding a real particle: j
        // beyond it are ghosts only, whose terms are exactly zero, so no task
        // streams them (the rows themselves keep their ghost slots).
        const int fe_real = (s.fe_end - s.start + 63) / 64 * 64;
        const int si_real = s.gfe + (s.end - s.fe_end + 63) / 64 * 64;
        for (int I = 0; I < nrows; ++I) {
            RowDesc rd{slot, 0};
            const int rb = I * B;
            const bool rowfe = rb < s.gfe;
            tasks.push_back(
                MutualTask{(int)si, I, rb, std::min(rb + B, rowfe ? fe_real : si_real), slot++, rowfe ? 1 : 0, 1, 0});
            const int seg0[2] = {rb + B, std::max(rb + B, s.gfe)};
            const int seg1[2] = {fe_real, si_real};
            for (int g = 0; g < 2; ++g) {
                const int a0 = seg0[g], a1 = seg1[g];
                if (a1 <= a0) continue;
                const long long blocks = (a1 - a0) / 64;
                long long pieces = (a1 - a0 + L - 1) / L;
                if (pieces > blocks) pieces = blocks;
                for (long long k = 0; k < pieces; ++k) {
                    const int a = a0 + (int)(64 * (blocks * k / pieces));
                    const int b = a0 + (int)(64 * (blocks * (k + 1) / pieces));
                    tasks.push_back(MutualTask{(int)si, I, a, b, slot++, g == 0 ? 1 : 0, 0, 0});
                }
            }
            rd.nslot = slot - rd.slot0;
            rows.push_back(rd);
        }
    }
    // k_mutual_sum's workgroups: the 64-blocks of every row that hold a real
    // particle, heaviest first (a particle adds its row's slots and one slab
    // entry per earlier row).
    std::vector<SumBlock> sblocks;
    std::vector<int> swork;
    for (size_t si = 0; si < sims.size(); ++si) {
        const SimDesc& s = sims[si];
        if (s.gn == 0) continue;
        const int nfe = s.fe_end - s.start, nsi = s.end - s.fe_end;
        for (int q0 = 0; q0 < s.gn; q0 += 64) {
            if (q0 < s.gfe ? q0 >= nfe : q0 - s.gfe >= nsi) continue;  // ghosts only
            const int I = q0 / B;
            sblocks.push_back(SumBlock{(int)si, q0});
            swork.push_back(rows[s.row0 + I].nslot + I / W + I % W);
        }
    }
    std::vector<int> order(sblocks.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return swork[a] > swork[b]; });
    p.sblocks.resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) p.sblocks[k] = sblocks[order[k]];
    // Longest first (VALU ops per lane and j: R*15+6; a diagonal task's
    // block m passes over m + 1 of the row's sub-blocks).
    auto cost = [&](const MutualTask& t) {
        if (!t.diag) return (long long)(t.j1 - t.j0) * (15 * tl.R + 6);
        const long long m0 = (t.j0 - (long long)t.row * B) / 64, m1 = (t.j1 - (long long)t.row * B) / 64;
        return 64 * (15 * (m1 * (m1 + 1) - m0 * (m0 + 1)) / 2 + 6 * (m1 - m0));
    };
    std::stable_sort(tasks.begin(), tasks.end(),
                     [&](const MutualTask& a, const MutualTask& b) { return cost(a) > cost(b); });
    if (W > 1) {
        // the row tasks, W to a workgroup in cost order (a filler pads the
        // last), then every workgroup longest first
        while (tasks.size() % W) tasks.push_back(MutualTask{0, 0, 0, 0, -1, 0, 0, 0});
        for (size_t k = 0; k < tasks.size(); k += W)
            p.gtasks.push_back(GroupTask{1, tasks[k].sim, 0, 0, 0, 0, 0, (int)k});
        auto gcost = [&](const GroupTask& g) {
            return g.kind == 1 ? cost(tasks[g.rtask0]) : (long long)(g.j1 - g.j0) * (15 * tl.R + 6);
        };
        std::stable_sort(p.gtasks.begin(), p.gtasks.end(),
                         [&](const GroupTask& a, const GroupTask& b) { return gcost(a) > gcost(b); });
    }
    p.mtasks = std::move(tasks);
    p.rows = std::move(rows);
    p.group_rows = W;
    p.gtot = gb;
    p.iside_len = (long long)slot * 3 * B;
    p.jside_len = jb;
    p.jintra_len = ji;
    p.chunk_len = L;
    p.mutual = W > 1 ? !p.gtasks.empty() : !p.mtasks.empty();
    return 0;
}

// The plan of one force evaluation over `sims`: the mutual pass's tasks and
// rows for the simulations that take it, then the one-sided pass's chunks
// and tasks for the rest, the contact-grid bucket ranges and the moment
// blocks.  batched: every simulation uses the fixed tiling of the batched
// engine, so its sums are the same alone and in any batch.
static std::atomic<int> g_contact_overlap{1};  // see gi_set_contact_overlap
static std::atomic<int> g_contact_walk{0};     // see gi_set_contact_walk

// The contact walk for plans of up to this many particles uses nine lanes per
// particle: their all-pairs pass is too short to hide a one-lane walk.
constexpr int kContactRowLanesMaxN = 49152;

static int build_plan(std::vector<SimDesc> sims, const Tiling& tl, bool batched, Plan* out) {
    Plan p;
    p.caller_sims = sims;
    p.tiling = tl;
    long long ntot = 0;
    // One-sided tasks use the tiling's R, except beside the mutual pass, where
    // the one-sided pass is a fallback for simulations too large for it and
    // takes R = 1 (the one-sided optimum on gfx950).
    p.plain_R = (tl.variant == 2 || tl.variant == 4) ? 1 : tl.R;
    const int ib = 64 * p.plain_R;
    for (auto& s : sims) {
        long long n = s.end - s.start;
        if (n < 0 || s.fe_end < s.start || s.fe_end > s.end || s.start != ntot)
            return set_error(GI_E_INVALID_ARG, "simulation ranges must be packed and ordered");
        ntot = s.end;
    }
    // The contact-grid scan handles up to 4096 * 1024 buckets = 2 * 2^21.
    if (ntot > (1LL << 21)) return set_error(GI_E_TOO_LARGE, "too many particles (%lld > 2^21)", ntot);
    if (tl.variant == 2 || tl.variant == 4) {
        int rc = plan_mutual(sims, tl, batched, p);
        if (rc < 0) return rc;
    }
    if (tl.variant == 3) {
        // R = lanes per i-particle (j-streams per wavefront): 1, 2, 4 or 8
        p.direct_lanes = std::min(tl.R, 8);
        const int ib = 64 / p.direct_lanes;
        for (size_t si = 0; si < sims.size(); ++si)
            if (sims[si].active)
                for (int i0 = sims[si].start; i0 < sims[si].end; i0 += ib)
                    p.dblocks.push_back(DirectBlock{(int)si, i0});
        p.direct = true;
    }
    long long pbase = 0;
    for (size_t si = 0; si < sims.size(); ++si) {
        SimDesc& s = sims[si];
        const int n = s.end - s.start;
        s.partial_base = (int)pbase;
        s.nchunk = 0;
        if (n == 0 || !s.active || s.gn > 0 || p.direct) continue;  // gn > 0: mutual pass
        const long long L = batched ? ((tl.variant == 2 || tl.variant == 4) ? 16384 : tl.chunk_len) : sim_chunk_len(n);
        // j segments: iron [start, fe_end), silicate [fe_end, end); each split
        // into near-equal chunks of at most L.
        const int seg0[2] = {s.start, s.fe_end};
        const int seg1[2] = {s.fe_end, s.end};
        int chunk = 0;
        for (int g = 0; g < 2; ++g) {
            const int len = seg1[g] - seg0[g];
            if (len <= 0) continue;
            const int pieces = (int)((len + L - 1) / L);
            for (int q = 0; q < pieces; ++q) {
                const int a = seg0[g] + (int)((long long)len * q / pieces);
                const int b = seg0[g] + (int)((long long)len * (q + 1) / pieces);
                for (int i0 = s.start; i0 < s.end; i0 += ib) {
                    GravTask t;
                    t.i0 = i0;
                    t.sim = (int)si;
                    t.j0 = a;
                    t.j1 = b;
                    t.chunk = chunk;
                    t.jfe = (g == 0);
                    t.pad0 = t.pad1 = 0;
                    p.tasks.push_back(t);
                }
                ++chunk;
            }
        }
        s.nchunk = chunk;
        pbase += (long long)chunk * n;
        if (pbase > (1LL << 31) - 1) return set_error(GI_E_TOO_LARGE, "partial buffer too large");
    }
    // Longest tasks first, so the tail of the launch is made of short ones.
    std::stable_sort(p.tasks.begin(), p.tasks.end(),
                     [](const GravTask& a, const GravTask& b) { return (a.j1 - a.j0) > (b.j1 - b.j0); });
    p.partial_len = (int)pbase;
    p.ntot = (int)ntot;
    // Contact grid: each simulation owns a power-of-two range of buckets of at
    // least its particle count (cells are visited by Morton code modulo it).
    long long hb = 0;
    for (auto& s : sims) {
        unsigned sz = 64;
        while (sz < (unsigned)(s.end - s.start)) sz <<= 1;
        s.hbase = (int)hb;
        s.hmask = sz - 1;
        hb += sz;
    }
    if ((hb + 1023) / 1024 > 4096) return set_error(GI_E_TOO_LARGE, "contact grid too large (%lld buckets)", hb);
    p.buckets = (int)hb;
    // moment blocks: 256 particles each, never straddling simulations
    p.sim_first_block.push_back(0);
    for (size_t si = 0; si < sims.size(); ++si) {
        for (int f = sims[si].start; f < sims[si].end; f += 256) {
            p.block_sim.push_back((int)si);
            p.block_first.push_back(f);
        }
        p.sim_first_block.push_back((int)p.block_sim.size());
    }
    p.overlap_allowed = g_contact_overlap.load() != 0;
    p.contact_walk = g_contact_walk.load();
    p.contact_lanes = p.contact_walk ? p.contact_walk : (p.ntot <= kContactRowLanesMaxN ? 9 : 1);
    p.split_finish = p.mutual && p.tasks.empty() && !p.direct && p.overlap_allowed;
    p.sims = std::move(sims);
    *out = std::move(p);
    return 0;
}

// Every index the mutual pass's kernels (k_mutual or k_mutual_grp, then
// k_mutual_sum) derive from a plan, checked against the buffers the plan
// sizes, on the host: a plan that would read or write out of bounds is an
// error before anything is launched (gi_plan_check runs it for the tests).
static int check_plan_indexes(const Plan& p) {
    if (!p.mutual) return 0;
    const int R = p.tiling.R, B = 64 * R, W = p.group_rows;
    const long long slots = p.iside_len / (3LL * B);
#define GI_CHECK(c, ...)                                            \
    do {                                                            \
        if (!(c)) return set_error(GI_E_STATE, "plan check: " __VA_ARGS__); \
    } while (0)
    long long jside_end = 0, jintra_end = 0;
    for (size_t si = 0; si < p.sims.size(); ++si) {
        const SimDesc& sd = p.sims[si];
        if (sd.gn == 0) continue;
        const long long nrows = sd.gn / B;
        GI_CHECK(sd.gn % B == 0 && sd.gfe % B == 0 && sd.gfe <= sd.gn, "sim %zu padded sizes", si);
        GI_CHECK(sd.jbase >= jside_end && sd.jibase >= jintra_end, "sim %zu slab regions overlap", si);
        long long ent = 0;
        for (long long g = 0; g * W < nrows; ++g) ent += sd.gn - std::min((g + 1) * W, nrows) * B;
        jside_end = sd.jbase + 3 * ent;
        jintra_end = sd.jibase + (W > 1 ? 3LL * (W - 1) * B * nrows : 0);
        GI_CHECK(jside_end <= p.jside_len && jintra_end <= std::max(p.jintra_len, 0LL), "sim %zu slabs exceed buffers", si);
        GI_CHECK(sd.gbase + sd.gn <= p.gtot && sd.row0 + nrows <= (long long)p.rows.size(), "sim %zu rows", si);
        for (long long rr = 0; rr < nrows; ++rr) {
            const RowDesc& rd = p.rows[sd.row0 + rr];
            GI_CHECK(rd.slot0 >= 0 && rd.nslot >= 1 && rd.slot0 + rd.nslot <= slots, "sim %zu row %lld slots", si, rr);
        }
    }
    // one wavefront's row task: k_mutual's J (W = 1) or the in-group slab (W > 1)
    auto check_row_task = [&](const MutualTask& t) -> int {
        GI_CHECK(t.sim >= 0 && t.sim < (int)p.sims.size(), "task sim");
        const SimDesc& sd = p.sims[t.sim];
        const long long nrows = sd.gn / B;
        GI_CHECK(t.row >= 0 && t.row < nrows, "task row %d of %lld", t.row, nrows);
        GI_CHECK(t.slot >= 0 && t.slot < slots, "task slot %d", t.slot);
        const RowDesc& rd = p.rows[sd.row0 + t.row];
        GI_CHECK(t.slot >= rd.slot0 && t.slot < rd.slot0 + rd.nslot, "task slot outside its row's");
        GI_CHECK(t.j0 % 64 == 0 && t.j1 % 64 == 0 && 0 <= t.j0 && t.j0 < t.j1 && t.j1 <= sd.gn, "task j-range");
        const long long ib = (long long)t.row * B;
        if (t.diag) {
            GI_CHECK(t.j0 >= ib && t.j0 < t.j1 && t.j1 <= ib + B, "diagonal task j-range");
            return 0;
        }
        const long long first = ib + B;
        GI_CHECK(t.j0 >= first, "row task reaches back into its row");
        GI_CHECK((t.jfe != 0) == (t.j1 <= sd.gfe) && (t.jfe != 0 || t.j0 >= sd.gfe), "task material");
        if (W == 1) {
            GI_CHECK(sd.jbase + 3 * jside_row_offset(t.row, sd.gn, B) + 3 * (sd.gn - first) <= p.jside_len, "J");
        } else {
            const long long ilen = (long long)(W - 1) * B;
            GI_CHECK(t.j1 - first <= ilen, "in-group task beyond its group");
            GI_CHECK(t.j1 <= std::min(((long long)t.row / W + 1) * W, nrows) * B, "in-group task beyond its group");
            GI_CHECK(sd.jibase + 3 * ilen * (t.row + 1) <= p.jintra_len, "in-group slab");
        }
        return 0;
    };
    int rc;
    if (W == 1) {
        for (const MutualTask& t : p.mtasks)
            if ((rc = check_row_task(t))) return rc;
    } else {
        for (const GroupTask& g : p.gtasks) {
            if (g.kind == 1) {
                GI_CHECK(g.rtask0 >= 0 && g.rtask0 % W == 0 && g.rtask0 + W <= (int)p.mtasks.size(), "row bundle");
                for (int w = 0; w < W; ++w) {
                    const MutualTask& t = p.mtasks[g.rtask0 + w];
                    if (t.j1 > t.j0 && (rc = check_row_task(t))) return rc;
                }
                continue;
            }
            GI_CHECK(g.kind == 0 && g.sim >= 0 && g.sim < (int)p.sims.size(), "group task kind/sim");
            const SimDesc& sd = p.sims[g.sim];
            const long long nrows = sd.gn / B;
            GI_CHECK(g.group >= 0 && (long long)g.group * W < nrows, "group %d of %lld rows", g.group, nrows);
            const long long gend = std::min(((long long)g.group + 1) * W, nrows) * B;
            GI_CHECK(g.j0 % 64 == 0 && g.j1 % 64 == 0 && gend <= g.j0 && g.j0 < g.j1 && g.j1 <= sd.gn,
                     "group task j-range [%d, %d) after %lld", g.j0, g.j1, gend);
            GI_CHECK((g.jfe != 0) == (g.j1 <= sd.gfe) && (g.jfe != 0 || g.j0 >= sd.gfe), "group task material");
            const long long len = sd.gn - gend;
            GI_CHECK(sd.jbase + 3 * jside_row_offset(g.group, sd.gn, W * B) + 3 * len <= p.jside_len, "group J");
            for (long long rr = (long long)g.group * W; rr < std::min(((long long)g.group + 1) * W, nrows); ++rr) {
                const RowDesc& rd = p.rows[sd.row0 + rr];
                GI_CHECK(g.piece >= 0 && g.piece < rd.nslot, "group piece %d, row %lld has %d slots", g.piece, rr,
                         rd.nslot);
            }
        }
    }
    // k_mutual_sum: every real particle's slots and slab entries
    for (const SumBlock& b : p.sblocks) {
        GI_CHECK(b.sim >= 0 && b.sim < (int)p.sims.size(), "sum block sim");
        const SimDesc& sd = p.sims[b.sim];
        GI_CHECK(b.q0 % 64 == 0 && b.q0 + 64 <= sd.gn, "sum block places");
        const int row = b.q0 / B, group = row / W;
        const long long WB = (long long)W * B, ilen = (long long)(W - 1) * B;
        for (int q = b.q0; q < b.q0 + 64; ++q) {
            for (int g = 0; g < group; ++g) {
                const long long len = sd.gn - (g + 1) * WB;
                const long long o = q - (g + 1) * WB;
                GI_CHECK(o >= 0 && o < len, "sum: group entry");
                GI_CHECK(sd.jbase + 3 * (jside_row_offset(g, sd.gn, WB) + o) + 2 < p.jside_len, "sum: J");
            }
            for (int r2 = group * W; r2 < row; ++r2) {
                const long long o = q - (long long)(r2 + 1) * B;
                GI_CHECK(o >= 0 && o < ilen && sd.jibase + 3 * (ilen * r2 + o) + 2 < p.jintra_len,
                         "sum: in-group entry");
            }
        }
    }
    // Coverage: every (row, 64-block) pair with a real particle in the block,
    // at or beyond the row's first block, is taken by exactly one task (its
    // diagonal pieces, its row or in-group tasks, its group's pieces), and no
    // task takes any other (simulations of at most 2^26 such pairs).
    for (size_t si = 0; si < p.sims.size(); ++si) {
        const SimDesc& sd = p.sims[si];
        if (sd.gn == 0) continue;
        const long long nrows = sd.gn / B, nb = sd.gn / 64;
        if (nrows * nb > (1LL << 26)) continue;
        const long long fe_real = (sd.fe_end - sd.start + 63) / 64 * 64;
        const long long si_real = sd.gfe + (sd.end - sd.fe_end + 63) / 64 * 64;
        std::vector<unsigned char> cov((size_t)(nrows * nb), 0);
        auto mark = [&](long long row, long long j0, long long j1) -> int {
            for (long long b = j0 / 64; b < j1 / 64; ++b)
                GI_CHECK(++cov[(size_t)(row * nb + b)] == 1, "sim %zu row %lld block %lld taken twice", si, row, b);
            return 0;
        };
        for (const MutualTask& t : p.mtasks)
            if (t.sim == (int)si && t.j1 > t.j0 && (rc = mark(t.row, t.j0, t.j1))) return rc;
        for (const GroupTask& g : p.gtasks)
            if (g.kind == 0 && g.sim == (int)si)
                for (long long rr = (long long)g.group * W; rr < std::min(((long long)g.group + 1) * W, nrows); ++rr)
                    if ((rc = mark(rr, g.j0, g.j1))) return rc;
        for (long long rr = 0; rr < nrows; ++rr)
            for (long long b = 0; b < nb; ++b) {
                const long long j = 64 * b;
                const bool want = (j < fe_real || (j >= sd.gfe && j < si_real)) && j >= rr * B;
                GI_CHECK(cov[(size_t)(rr * nb + b)] == (want ? 1 : 0), "sim %zu row %lld block %lld: taken %d times",
                         si, rr, b, (int)cov[(size_t)(rr * nb + b)]);
            }
    }
#undef GI_CHECK
    return 0;
}

template <typename T>
static int upload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    int rc = b.ensure(v.size());
    if (rc) return rc;
    if (!v.empty()) GI_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return 0;
}

// Make ws->plan describe `sims` (re-planning and uploading only on change).
static int ensure_plan(Workspace* ws, const std::vector<SimDesc>& sims, hipStream_t st) {
    long long ntot = sims.empty() ? 0 : sims.back().end;
    long long pairs = 0;
    for (auto& s : sims)
        if (s.active) pairs += (long long)(s.end - s.start) * (s.end - s.start);
    Tiling tl = choose_tiling(ntot, pairs, (int)sims.size());
    if (ws->batched) {
        const BatchedTiling bt = batched_tiling();
        tl = Tiling{bt.R, bt.variant, bt.chunk_len, 0, 0};
    }
    if (ws->plan_valid && same_tiling(ws->plan.tiling, tl) && same_sims(ws->plan.caller_sims, sims) &&
        ws->plan.overlap_allowed == (g_contact_overlap.load() != 0) &&
        ws->plan.contact_walk == g_contact_walk.load())
        return 0;
    Plan p;
    int rc = build_plan(sims, tl, ws->batched, &p);
    if (rc) return rc;
    if ((rc = ws->partial.ensure((size_t)p.partial_len + 1))) return rc;
    if (p.mutual) {
        if ((rc = ws->gpos.ensure((size_t)p.gtot + 1))) return rc;
        if ((rc = ws->gsum.ensure((size_t)p.ntot + 1))) return rc;
        if ((rc = ws->iside.ensure((size_t)p.iside_len + 1))) return rc;
        if ((rc = ws->jside.ensure((size_t)p.jside_len + 1))) return rc;
        if ((rc = ws->jintra.ensure((size_t)p.jintra_len + 1))) return rc;
        if ((rc = upload(ws->gtasks, p.gtasks, st))) return rc;
        if ((rc = upload(ws->mtasks, p.mtasks, st))) return rc;
        if ((rc = upload(ws->rows, p.rows, st))) return rc;
        if ((rc = upload(ws->sblocks, p.sblocks, st))) return rc;
        if (p.split_finish && (rc = ws->corr.ensure((size_t)6 * p.ntot + 1))) return rc;
    }
    if (p.direct && (rc = upload(ws->dblocks, p.dblocks, st))) return rc;
    if ((rc = ws->counts.ensure((size_t)p.buckets + 1))) return rc;
    if ((rc = ws->starts.ensure((size_t)p.buckets + 2))) return rc;
    if ((rc = ws->spos.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ws->svel.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ws->sorted.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ws->cellhash.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ws->block_sums.ensure((size_t)p.buckets / 1024 + 8))) return rc;
    if ((rc = upload(ws->tasks, p.tasks, st))) return rc;
    if ((rc = upload(ws->sims, p.sims, st))) return rc;
    if ((rc = upload(ws->block_sim, p.block_sim, st))) return rc;
    if ((rc = upload(ws->block_first, p.block_first, st))) return rc;
    if ((rc = upload(ws->sim_first_block, p.sim_first_block, st))) return rc;
    if ((rc = ws->moments.ensure(p.block_sim.size() * kMoments + 1))) return rc;
    if ((rc = ws->sums.ensure(p.sims.size() * kMoments + 1))) return rc;
    if ((rc = ws->pe.ensure((size_t)p.ntot + 1))) return rc;
    // The uploads read host vectors that live on in ws->plan.
    ws->plan = std::move(p);
    ws->plan_valid = true;
    ws->generation++;
    return 0;
}

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// Launch sequences
// ---------------------------------------------------------------------------

// Contact grid of the current plan: bucket counts, offsets (scan), particle
// ids per bucket sorted by id, and positions / velocities in that order
// (vel may be null when only positions are needed).
static int launch_grid(Workspace* ws, const float4* pos, const float4* vel, hipStream_t st) {
    const Plan& p = ws->plan;
    const int m = p.buckets;
    const int nsims = (int)p.sims.size();
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(m, 256)), dim3(256), 0, st, ws->counts.p, m, 0);
    hipLaunchKernelGGL(k_grid_count, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, pos, p.ntot, ws->sims.p,
                       nsims, ws->counts.p, ws->cellhash.p);
    const int nb = nblk(m, 1024);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, ws->counts.p, m, ws->starts.p,
                       ws->block_sums.p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTopThreads), 0, st, ws->block_sums.p, nb, ws->starts.p + m);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk(m, 256)), dim3(256), 0, st, ws->starts.p, m, ws->block_sums.p);
    hipLaunchKernelGGL(k_grid_scatter, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, p.ntot, ws->cellhash.p,
                       ws->starts.p, ws->counts.p, ws->sorted.p);
    hipLaunchKernelGGL(k_grid_sort_buckets, dim3(nblk(m, 256)), dim3(256), 0, st, m, ws->starts.p,
                       ws->sorted.p);
    hipLaunchKernelGGL(k_grid_gather, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, p.ntot,
                       (const int*)ws->sorted.p, pos, vel, ws->spos.p, vel ? ws->svel.p : nullptr);
    GI_HIP(hipGetLastError());
    return 0;
}

static TimingPair take_pair(DevCtx* ctx, int kind) {
    TimingPair t;
    if (!ctx->pool.empty()) {
        t = ctx->pool.back();
        ctx->pool.pop_back();
    } else {
        (void)hipEventCreate(&t.a);
        (void)hipEventCreate(&t.b);
    }
    t.kind = kind;
    return t;
}

// Wavefront tasks per all-pairs dispatch (0: all in one); see
// gi_set_dispatch_limit.
static std::atomic<long long> g_dispatch_limit{0};

// Tasks [t0, t0 + ntasks) of the plan's one-sided list.
template <bool POT, int V>
static void launch_gravity_kernel(int R, hipStream_t st, const float4* pos, const Workspace* ws, int t0,
                                  int ntasks) {
    const dim3 grid(nblk(ntasks, 4)), block(256);
    const GravTask* tasks = ws->tasks.p + t0;
    if (R == 8)
        hipLaunchKernelGGL((k_gravity<8, 128, POT, V>), grid, block, 0, st, pos, tasks, ntasks, ws->sims.p,
                           ws->partial.p);
    else if (R == 2)
        hipLaunchKernelGGL((k_gravity<2, 128, POT, V>), grid, block, 0, st, pos, tasks, ntasks, ws->sims.p,
                           ws->partial.p);
    else if (R == 1)
        hipLaunchKernelGGL((k_gravity<1, 128, POT, V>), grid, block, 0, st, pos, tasks, ntasks, ws->sims.p,
                           ws->partial.p);
    else
        hipLaunchKernelGGL((k_gravity<4, 128, POT, V>), grid, block, 0, st, pos, tasks, ntasks, ws->sims.p,
                           ws->partial.p);
}

// Workgroup tasks [t0, t0 + ntasks) of the plan's row-group list.
template <bool POT>
static void launch_mutual_group_kernel(int R, hipStream_t st, const Workspace* ws, int t0, int ntasks) {
    const dim3 grid(ntasks), block(64 * kMutualWaves);
    const GroupTask* tasks = ws->gtasks.p + t0;
    if (R == 8)
        hipLaunchKernelGGL((k_mutual_grp<8, POT>), grid, block, 0, st, ws->gpos.p, tasks, ntasks, ws->mtasks.p,
                           ws->sims.p, ws->rows.p, ws->iside.p, ws->jside.p, ws->jintra.p);
    else
        hipLaunchKernelGGL((k_mutual_grp<4, POT>), grid, block, 0, st, ws->gpos.p, tasks, ntasks, ws->mtasks.p,
                           ws->sims.p, ws->rows.p, ws->iside.p, ws->jside.p, ws->jintra.p);
}

// Tasks [t0, t0 + ntasks) of the plan's mutual list.
template <bool POT>
static void launch_mutual_kernel(int R, hipStream_t st, const Workspace* ws, int t0, int ntasks) {
    const dim3 grid(nblk(ntasks, kMutualWaves)), block(64 * kMutualWaves);
    const MutualTask* tasks = ws->mtasks.p + t0;
    if (R == 8)
        hipLaunchKernelGGL((k_mutual<8, POT>), grid, block, 0, st, ws->gpos.p, tasks, ntasks, ws->sims.p,
                           ws->iside.p, ws->jside.p);
    else if (R == 4)
        hipLaunchKernelGGL((k_mutual<4, POT>), grid, block, 0, st, ws->gpos.p, tasks, ntasks, ws->sims.p,
                           ws->iside.p, ws->jside.p);
    else if (R == 2)
        hipLaunchKernelGGL((k_mutual<2, POT>), grid, block, 0, st, ws->gpos.p, tasks, ntasks, ws->sims.p,
                           ws->iside.p, ws->jside.p);
    else
        hipLaunchKernelGGL((k_mutual<1, POT>), grid, block, 0, st, ws->gpos.p, tasks, ntasks, ws->sims.p,
                           ws->iside.p, ws->jside.p);
}

// The all-pairs pass of the current plan: the one-sided kernel for the
// simulations planned one-sided, and for the others the mutual pass (padded
// copy, pair kernel, per-particle sums into ws->gsum).  Timed: kind 0 spans
// the all-pairs kernel(s), kind 2 the whole pass.  With force (plans with
// split_finish), st waits for `join` (the contact corrections in ws->corr)
// and the sums go to force with the corrections added (and with mpos / mvel
// the batched move is applied).  With fork, the event is recorded on st
// right after the padded copy, and side_work -- the work that waits for it
// on another stream -- is enqueued after the pair kernels, so that a caller
// enqueueing call by call does not hold them back (work forked from there
// does not compete with the copy).
static int launch_gravity(DevCtx* ctx, Workspace* ws, const float4* pos, bool pot, hipStream_t st,
                          bool timed, float4* force = nullptr, hipEvent_t join = nullptr, hipEvent_t fork = nullptr,
                          const std::function<int()>& side_work = nullptr, float4* mpos = nullptr,
                          float4* mvel = nullptr, const SumMove& mv = SumMove{0.0f, 0.0f, 0.0f, 1}) {
    const Plan& p = ws->plan;
    const int nplain = (int)p.tasks.size();
    const bool grouped = p.group_rows > 1;
    const int nmut = grouped ? (int)p.gtasks.size() : (int)p.mtasks.size();
    int rc;
    if (nplain == 0 && nmut == 0) {
        if (fork) GI_HIP(hipEventRecord(fork, st));
        if (side_work && (rc = side_work())) return rc;
        if (join) GI_HIP(hipStreamWaitEvent(st, join, 0));
        return 0;
    }
    const int nsims = (int)p.sims.size();
    TimingPair pass, tp;
    if (timed) {
        pass = take_pair(ctx, 2);
        GI_HIP(hipEventRecord(pass.a, st));
    }
    if (nmut > 0)
        hipLaunchKernelGGL(k_mutual_pad, dim3(nblk(p.gtot, 256)), dim3(256), 0, st, pos, (int)p.gtot, ws->sims.p,
                           nsims, ws->gpos.p);
    if (fork) GI_HIP(hipEventRecord(fork, st));
    if (timed) {
        tp = take_pair(ctx, 0);
        GI_HIP(hipEventRecord(tp.a, st));
    }
    // Every task writes only its own outputs, so cutting a task list into
    // several dispatches changes nothing in the results.
    const long long lim = g_dispatch_limit.load();
    const int step_plain = lim > 0 && lim < nplain ? (int)lim : std::max(nplain, 1);
    for (int t0 = 0; t0 < nplain; t0 += step_plain) {
        const int cnt = std::min(step_plain, nplain - t0);
        if (pot)
            launch_gravity_kernel<true, 0>(p.plain_R, st, pos, ws, t0, cnt);
        else if (p.tiling.variant == 1)
            launch_gravity_kernel<false, 1>(p.plain_R, st, pos, ws, t0, cnt);
        else
            launch_gravity_kernel<false, 0>(p.plain_R, st, pos, ws, t0, cnt);
    }
    const int step_mut = lim > 0 && lim < nmut ? (int)lim : std::max(nmut, 1);
    for (int t0 = 0; t0 < nmut; t0 += step_mut) {
        const int cnt = std::min(step_mut, nmut - t0);
        if (grouped && pot)
            launch_mutual_group_kernel<true>(p.tiling.R, st, ws, t0, cnt);
        else if (grouped)
            launch_mutual_group_kernel<false>(p.tiling.R, st, ws, t0, cnt);
        else if (pot)
            launch_mutual_kernel<true>(p.tiling.R, st, ws, t0, cnt);
        else
            launch_mutual_kernel<false>(p.tiling.R, st, ws, t0, cnt);
    }
    GI_HIP(hipGetLastError());
    if (timed) {
        GI_HIP(hipEventRecord(tp.b, st));
        ctx->pending.push_back(tp);
    }
    if (side_work && (rc = side_work())) return rc;
    if (join) GI_HIP(hipStreamWaitEvent(st, join, 0));
    if (nmut > 0 && !p.sblocks.empty()) {
        const dim3 grid((unsigned)p.sblocks.size()), block(64 * kSumSplit);
        const float* corr = force ? (const float*)ws->corr.p : nullptr;
        if (pot)
            hipLaunchKernelGGL(k_mutual_sum<true>, grid, block, 0, st, ws->sblocks.p, ws->sims.p, ws->rows.p,
                               64 * p.tiling.R, p.group_rows, ws->iside.p, ws->jside.p, ws->jintra.p, ws->gsum.p,
                               corr, p.ntot, force, mpos, mvel, mv);
        else
            hipLaunchKernelGGL(k_mutual_sum<false>, grid, block, 0, st, ws->sblocks.p, ws->sims.p, ws->rows.p,
                               64 * p.tiling.R, p.group_rows, ws->iside.p, ws->jside.p, ws->jintra.p, ws->gsum.p,
                               corr, p.ntot, force, mpos, mvel, mv);
        GI_HIP(hipGetLastError());
    }
    if (timed) {
        GI_HIP(hipEventRecord(pass.b, st));
        ctx->pending.push_back(pass);
    }
    return 0;
}

template <bool POT, int MOVE>
static void direct_kernel(int lanes, dim3 grid, dim3 block, hipStream_t st, const float4* pos, const float4* vel,
                          const Workspace* ws, float4* force, double* pe, const FusedMove& mv) {
    const DirectBlock* bl = ws->dblocks.p;
    if (lanes == 8)
        hipLaunchKernelGGL((k_direct<POT, 8, MOVE>), grid, block, 0, st, pos, vel, bl, ws->sims.p, force, pe, mv);
    else if (lanes == 4)
        hipLaunchKernelGGL((k_direct<POT, 4, MOVE>), grid, block, 0, st, pos, vel, bl, ws->sims.p, force, pe, mv);
    else if (lanes == 2)
        hipLaunchKernelGGL((k_direct<POT, 2, MOVE>), grid, block, 0, st, pos, vel, bl, ws->sims.p, force, pe, mv);
    else
        hipLaunchKernelGGL((k_direct<POT, 1, MOVE>), grid, block, 0, st, pos, vel, bl, ws->sims.p, force, pe, mv);
}

// The direct pass of the current plan: the force, or with pot the
// per-particle potential into ws->pe, or with move = 1 / 2 the force and the
// collision / damped move into mv's arrays.  Timed as both kind 0 and kind 2.
static int launch_direct(DevCtx* ctx, Workspace* ws, const float4* pos, const float4* vel, float4* force, bool pot,
                         hipStream_t st, bool timed, int move = 0, const FusedMove& mv = FusedMove{}) {
    const Plan& p = ws->plan;
    const int nb = (int)p.dblocks.size();
    if (nb == 0) return 0;
    TimingPair t0, t2;
    if (timed) {
        t2 = take_pair(ctx, 2);
        t0 = take_pair(ctx, 0);
        GI_HIP(hipEventRecord(t2.a, st));
        GI_HIP(hipEventRecord(t0.a, st));
    }
    const dim3 grid(nb), block(64 * kDirectWaves);
    const int lanes = p.direct_lanes;
    if (pot)
        direct_kernel<true, 0>(lanes, grid, block, st, pos, vel, ws, nullptr, ws->pe.p, mv);
    else if (move == 1)
        direct_kernel<false, 1>(lanes, grid, block, st, pos, vel, ws, force, nullptr, mv);
    else if (move == 2)
        direct_kernel<false, 2>(lanes, grid, block, st, pos, vel, ws, force, nullptr, mv);
    else
        direct_kernel<false, 0>(lanes, grid, block, st, pos, vel, ws, force, nullptr, mv);
    GI_HIP(hipGetLastError());
    if (timed) {
        GI_HIP(hipEventRecord(t0.b, st));
        GI_HIP(hipEventRecord(t2.b, st));
        ctx->pending.push_back(t0);
        ctx->pending.push_back(t2);
    }
    return 0;
}

template <int LANES>
static void launch_contact_lanes(const Workspace* ws, float4* force, hipStream_t st) {
    const Plan& p = ws->plan;
    const dim3 grid(contact_blocks<LANES>(p.ntot)), block(64 * kContactWaves);
    if (force)
        hipLaunchKernelGGL(k_finish<LANES>, grid, block, 0, st, force, p.ntot, ws->sims.p, (int)p.sims.size(),
                           ws->partial.p, p.mutual ? ws->gsum.p : nullptr, ws->starts.p, ws->sorted.p, ws->spos.p,
                           ws->svel.p, p.tiling.variant);
    else
        hipLaunchKernelGGL(k_contact_corr<LANES>, grid, block, 0, st, p.ntot, ws->sims.p, (int)p.sims.size(),
                           ws->starts.p, ws->sorted.p, ws->spos.p, ws->svel.p, p.tiling.variant, ws->corr.p);
}

// With force: k_finish (sums + corrections into force); without: the
// corrections alone into ws->corr (k_contact_corr).
static void launch_contact(const Workspace* ws, float4* force, hipStream_t st) {
    if (ws->plan.contact_lanes == kCellRows)
        launch_contact_lanes<kCellRows>(ws, force, st);
    else
        launch_contact_lanes<1>(ws, force, st);
}

// Full force evaluation for the current plan: grid, all pairs, finish (or
// the direct pass alone).  With move_pos / move_vel (the arrays pos / vel
// point to) and a plan with split_finish, the move is fused into the last
// kernel and *moved is set; otherwise the caller moves.  The move takes each
// simulation's own constants (the batched engine), or mv's when given.
static int run_force(DevCtx* ctx, Workspace* ws, const float4* pos, const float4* vel, float4* force,
                     hipStream_t st, bool timed, float4* move_pos = nullptr, float4* move_vel = nullptr,
                     bool* moved = nullptr, const SumMove* mv = nullptr) {
    const Plan& p = ws->plan;
    if (p.ntot == 0) return 0;
    TimingPair whole;
    if (timed) {
        whole = take_pair(ctx, 3);
        GI_HIP(hipEventRecord(whole.a, st));
    }
    int rc;
    if (p.direct) {
        if ((rc = launch_direct(ctx, ws, pos, vel, force, false, st, timed))) return rc;
    } else {
        // The contact grid reads only pos / vel: it is built on the side
        // stream while the all-pairs pass runs on st (under graph capture the
        // fork and join become edges).  When every simulation takes the
        // mutual pass, the contact corrections follow on the side stream and
        // k_mutual_sum adds them; otherwise k_finish adds them after the join.
        // The fork follows the mutual pass's padded copy, which then has the
        // device to itself.
        if (!ctx->side) GI_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        if (!ctx->ev_fork) GI_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        if (!ctx->ev_join) GI_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
        const std::function<int()> side_work = [&]() -> int {
            int r;
            GI_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
            if ((r = launch_grid(ws, pos, vel, ctx->side))) return r;
            if (p.split_finish) {
                launch_contact(ws, nullptr, ctx->side);
                GI_HIP(hipGetLastError());
            }
            GI_HIP(hipEventRecord(ctx->ev_join, ctx->side));
            return 0;
        };
        if (p.split_finish) {
            const bool fuse = move_pos && move_vel && moved;
            if ((rc = launch_gravity(ctx, ws, pos, false, st, timed, force, ctx->ev_join, ctx->ev_fork, side_work,
                                     fuse ? move_pos : nullptr, fuse ? move_vel : nullptr,
                                     mv ? *mv : SumMove{0.0f, 0.0f, 0.0f, 1})))
                return rc;
            if (fuse) *moved = true;
        } else {
            if ((rc = launch_gravity(ctx, ws, pos, false, st, timed, nullptr, nullptr, ctx->ev_fork, side_work)))
                return rc;
            GI_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
            launch_contact(ws, force, st);
            GI_HIP(hipGetLastError());
        }
    }
    if (timed) {
        GI_HIP(hipEventRecord(whole.b, st));
        ctx->pending.push_back(whole);
    }
    return 0;
}

// Per-simulation FP64 moments (with the potential) into ws->sums.
static int run_conserved(DevCtx* ctx, Workspace* ws, const float4* pos, const float4* vel, hipStream_t st) {
    const Plan& p = ws->plan;
    if (p.ntot == 0) return 0;
    int rc;
    if (p.direct) {
        if ((rc = launch_direct(ctx, ws, pos, vel, nullptr, true, st, false))) return rc;
    } else {
        if ((rc = launch_grid(ws, pos, nullptr, st))) return rc;
        if ((rc = launch_gravity(ctx, ws, pos, true, st, false))) return rc;
        hipLaunchKernelGGL(k_finish_potential, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, p.ntot, ws->sims.p,
                           (int)p.sims.size(), ws->partial.p, p.mutual ? ws->gsum.p : nullptr, ws->starts.p,
                           ws->sorted.p, ws->spos.p, ws->pe.p);
    }
    const int nb = (int)p.block_sim.size();
    hipLaunchKernelGGL(k_moments, dim3(nb), dim3(256), 0, st, pos, vel, (const double*)ws->pe.p,
                       ws->block_sim.p, ws->block_first.p, ws->sims.p, 0.0, 0.0, ws->moments.p);
    const int ns = (int)p.sims.size();
    hipLaunchKernelGGL(k_moments_final, dim3(nblk(ns, 64)), dim3(64), 0, st, (const double*)ws->moments.p,
                       ws->sim_first_block.p, ns, ws->sums.p);
    GI_HIP(hipGetLastError());
    return 0;
}

// kinetic, potential, total, P(3), L about the centre of mass (3) from the 14 moments.
static void moments_to_conserved(const double* m, double* out9) {
    const double M = m[0];
    out9[0] = m[10];
    out9[1] = m[11];
    out9[2] = m[10] + m[11];
    out9[3] = m[4];
    out9[4] = m[5];
    out9[5] = m[6];
    // L_com = sum m x cross v - (sum m x) cross (sum m v) / M
    const double X[3] = {m[1], m[2], m[3]}, P[3] = {m[4], m[5], m[6]};
    out9[6] = m[7] - (X[1] * P[2] - X[2] * P[1]) / M;
    out9[7] = m[8] - (X[2] * P[0] - X[0] * P[2]) / M;
    out9[8] = m[9] - (X[0] * P[1] - X[1] * P[0]) / M;
}

static int launch_recenter(DevCtx* ctx, float4* pos, float4* vel, int n, int nfe, double mfe, double msi,
                           const double* tc, const double* tv, hipStream_t st) {
    const int nb = nblk(n, 256);
    int rc;
    if ((rc = ctx->rc_blocks.ensure((size_t)nb * kMoments1 + 1))) return rc;
    if ((rc = ctx->rc_sums.ensure(kMoments1 + 1))) return rc;
    hipLaunchKernelGGL(k_moments_one, dim3(nb), dim3(256), 0, st, pos, vel, n, nfe, mfe, msi, ctx->rc_blocks.p);
    hipLaunchKernelGGL(k_moments_one_final, dim3(1), dim3(64), 0, st, (const double*)ctx->rc_blocks.p, nb,
                       ctx->rc_sums.p);
    hipLaunchKernelGGL(k_recenter_apply, dim3(nb), dim3(256), 0, st, pos, vel, n, (const double*)ctx->rc_sums.p,
                       tc[0], tc[1], tc[2], tv[0], tv[1], tv[2]);
    GI_HIP(hipGetLastError());
    return 0;
}

// The reference's float recenter (k_recenter_ref_sum: its host sums in its
// order, one workgroup) and the shift of every particle.
static int launch_recenter_ref(DevCtx* ctx, float4* pos, float4* vel, int n, int nfe, double mfe, double msi,
                               const float* tc, const float* tv, hipStream_t st) {
    int rc;
    if ((rc = ctx->rc_shift.ensure(8))) return rc;
    hipLaunchKernelGGL(k_recenter_ref_sum, dim3(1), dim3(256), 0, st, (const float4*)pos, (const float4*)vel, n, nfe,
                       mfe, msi, tc[0], tc[1], tc[2], tv[0], tv[1], tv[2], ctx->rc_shift.p);
    hipLaunchKernelGGL(k_recenter_ref_apply, dim3(nblk(n, 256)), dim3(256), 0, st, pos, vel, n,
                       (const float*)ctx->rc_shift.p);
    GI_HIP(hipGetLastError());
    return 0;
}

static hipStream_t as_stream(gi_stream_t s) { return (hipStream_t)s; }

static SimDesc make_sim(int start, int fe_end, int end, const PairLaw& L) {
    SimDesc s;
    std::memset(&s, 0, sizeof s);
    s.start = start;
    s.fe_end = fe_end;
    s.end = end;
    s.active = 1;
    s.law = L;
    return s;
}

// ---------------------------------------------------------------------------
// HIP graphs of the step loop
// ---------------------------------------------------------------------------

template <typename T>
static void key_append(std::vector<unsigned char>& k, const T& v) {
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&v);
    k.insert(k.end(), b, b + sizeof(T));
}

// Capture `body` on a private stream into a graph cached under `key` (the
// lookup and eviction policy is GraphCache's: the kGraphCacheSize most
// recently used graphs are kept; `keep` lists graphs the caller already holds,
// which are never evicted; the device is synchronised before an evicted graph
// is destroyed, since it may still have launches in flight from an earlier
// call on any stream).
constexpr size_t kGraphCacheSize = 8;

// With update, a graph of the same key captured for an earlier plan
// generation is updated in place (hipGraphExecUpdate: the executable graph
// stays, its nodes take the new plan's parameters) when the new capture has
// the same topology; otherwise a new one is instantiated.
template <typename F>
static int get_graph(DevCtx* ctx, GraphCacheHip& cache, const std::vector<unsigned char>& key,
                     unsigned long long generation, F body, hipGraphExec_t* out,
                     std::initializer_list<hipGraphExec_t> keep = {}, bool update = false) {
    cache.capacity = kGraphCacheSize;
    auto record = [&](hipGraph_t* g) -> int {
        if (!ctx->capture_stream) GI_HIP(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
        GI_HIP(hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
        int rc = body(ctx->capture_stream);
        *g = nullptr;
        hipError_t ce = hipStreamEndCapture(ctx->capture_stream, g);
        if (rc) {
            if (*g) (void)hipGraphDestroy(*g);
            *g = nullptr;
            return rc;
        }
        if (ce != hipSuccess) return set_error((int)ce, "hipStreamEndCapture: %s", hipGetErrorString(ce));
        return 0;
    };
    auto refresh = [&](GraphPayload& p, bool* done) -> int {
        *done = false;
        if (!update) return 0;
        hipGraph_t g = nullptr;
        int rc = record(&g);
        if (rc) return rc;
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(p.exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            if (p.graph) (void)hipGraphDestroy(p.graph);
            p.graph = g;
            *done = true;
        } else {
            (void)hipGraphDestroy(g);
            (void)hipGetLastError();  // a refused update is not an error of the call
        }
        return 0;
    };
    auto capture = [&](GraphPayload* e) -> int {
        hipGraph_t g = nullptr;
        int rc = record(&g);
        if (rc) return rc;
        e->graph = g;
        hipError_t ie = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            e->graph = nullptr;
            return set_error((int)ie, "hipGraphInstantiate: %s", hipGetErrorString(ie));
        }
        return 0;
    };
    auto sync = [&]() -> int {
        GI_HIP(hipDeviceSynchronize());
        return 0;
    };
    auto destroy = [](GraphPayload& p) { p.destroy(); };
    auto held = [&](const GraphPayload& p) {
        for (hipGraphExec_t h : keep)
            if (p.exec == h) return true;
        return false;
    };
    GraphPayload got;
    int rc = cache.get(key, generation, capture, refresh, sync, destroy, held, &got);
    if (rc) return rc;
    *out = got.exec;
    return 0;
}

// Steps per graph of the multi-step form: a whole graph launch per step is
// a fixed cost that small simulations feel.
constexpr int kStepsPerGraph = 8;

// kStepsPerGraph steps of a direct-pass plan with the move fused into the
// pass: step k reads one state and writes the other (the caller's arrays P, V
// and ctx->alt_*), so after an even number of steps the state is back in P, V.
static_assert(kStepsPerGraph % 2 == 0, "fused steps must end in the caller's arrays");
static int fused_direct_steps(DevCtx* ctx, float4* P, float4* V, float4* F, FusedMove mv, int move, hipStream_t cs) {
    float4 *sp = P, *sv = V, *dp = ctx->alt_pos.p, *dv = ctx->alt_vel.p;
    for (int k = 0; k < kStepsPerGraph; ++k) {
        mv.pout = dp;
        mv.vout = dv;
        int r = launch_direct(ctx, &ctx->ws, sp, sv, F, false, cs, false, move, mv);
        if (r) return r;
        std::swap(sp, dp);
        std::swap(sv, dv);
    }
    return 0;
}

// Launch `count` steps: the kStepsPerGraph-step graph as often as it fits,
// then the one-step graph.
static int launch_step_run(hipGraphExec_t one, hipGraphExec_t many, int count, hipStream_t st) {
    if (many)
        for (; count >= kStepsPerGraph; count -= kStepsPerGraph) GI_HIP(hipGraphLaunch(many, st));
    for (; count > 0; --count) GI_HIP(hipGraphLaunch(one, st));
    return 0;
}

// Run graph launches on `st`; the legacy default stream cannot be captured
// or used with a non-blocking capture stream without explicit ordering, so
// work for st == NULL goes to the capture stream between two events.
struct StreamBridge {
    DevCtx* ctx;
    hipStream_t user, run;
    int begin(DevCtx* c, hipStream_t u) {
        ctx = c;
        user = u;
        run = u;
        if (u) return 0;
        if (!c->capture_stream) GI_HIP(hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking));
        if (!c->ev_in) GI_HIP(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
        if (!c->ev_out) GI_HIP(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
        GI_HIP(hipEventRecord(c->ev_in, nullptr));
        GI_HIP(hipStreamWaitEvent(c->capture_stream, c->ev_in, 0));
        run = c->capture_stream;
        return 0;
    }
    int end() {
        if (user) return 0;
        GI_HIP(hipEventRecord(ctx->ev_out, run));
        GI_HIP(hipStreamWaitEvent(nullptr, ctx->ev_out, 0));
        return 0;
    }
};

}  // namespace gi

// ===========================================================================
// Ensemble
// ===========================================================================

namespace gi {

// ---------------------------------------------------------------------------
// The allocation-free force entry point: its whole scratch (the simulation's
// descriptor and the direct pass's block table) lives in one file-scope
// device array, so a call allocates nothing.
// ---------------------------------------------------------------------------

constexpr int kNoAllocMaxN = 12288;                    // particles per call, at most
constexpr int kNoAllocMaxBlocks = kNoAllocMaxN / 8;  // >= 8 i-particles per workgroup
struct NoAllocScratch {
    SimDesc sim;
    DirectBlock blocks[kNoAllocMaxBlocks];
};

// Host code must never pass this name to a kernel (on the host it is a
// stand-in, not a device address): gi_force_collision_noalloc resolves each
// device's own address with hipGetSymbolAddress and passes that.
static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

static std::mutex g_noalloc_mutex;
static NoAllocScratch* g_noalloc_ptr[64];  // per device, set once

// Writes the descriptor and the block table of one simulation of n
// particles, ib i-particles per workgroup.
__global__ void k_noalloc_plan(NoAllocScratch* __restrict__ s, SimDesc sd, int nb, int ib) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) s->sim = sd;
    if (b < nb) s->blocks[b] = DirectBlock{0, b * ib};
}

}  // namespace gi

struct gi_ensemble {
    int device = -1;
    gi::Workspace ws;       // the stepping plan (active simulations)
    gi::Workspace ws_diag;  // all simulations, for the conserved quantities
    // The simulation table over [0, end of the last simulation): the caller's
    // simulations in offset order, with frozen fillers over the gaps between
    // them (gi_ensemble_set_sims), so that the table covers its particles
    // back to back.  caller[k] = the caller's index of sims[k], -1 for a filler.
    std::vector<gi::SimDesc> sims;
    std::vector<int> caller;
    int n_caller = 0;
    long long capacity = 0;  // records in pos / vel
    float4* pos = nullptr;
    float4* vel = nullptr;
    gi::DevBuf<float4> force;
    hipStream_t stream = nullptr;
    gi::GraphCacheHip graphs;
    // one event per gi_ensemble_steps call still possibly running (oldest first)
    std::vector<hipEvent_t> pending, spare;
    ~gi_ensemble() {
        // launches of its graphs may still be in flight on its device
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && device >= 0 && cur != device) (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        graphs.clear([](gi::GraphPayload& p) { p.destroy(); });
        for (hipEvent_t ev : pending) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : spare) (void)hipEventDestroy(ev);
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    }
};

using namespace gi;

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int gi_abi_version(void) { return GI_ABI_VERSION; }

const char* gi_last_error_string(void) { return g_last_error.c_str(); }

int gi_last_error(void) { return g_last_code; }

void gi_clear_error(void) {
    g_last_code = 0;
    g_last_error.clear();
}

int gi_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gi_set_device(int device) {
    GI_HIP(hipSetDevice(device));
    return 0;
}

int gi_check_block_multiple(int n) {
    if (n % 256 != 0)
        return set_error(GI_E_NOT_MULTIPLE_OF_BLOCK, "Number of Particles is not a multiple of the block size");
    return 0;
}

int gi_malloc(void** dptr, size_t bytes) {
    if (!dptr) return set_error(GI_E_INVALID_ARG, "gi_malloc: null out pointer");
    GI_HIP(hipMalloc(dptr, bytes ? bytes : 16));
    return 0;
}

int gi_free(void* dptr) {
    if (dptr) GI_HIP(hipFree(dptr));
    return 0;
}

int gi_copy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes) GI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int gi_copy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes) GI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int gi_copy_d2d(void* dst, const void* src, size_t bytes, gi_stream_t stream) {
    if (bytes) GI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return 0;
}

int gi_stream_create(gi_stream_t* stream) {
    if (!stream) return set_error(GI_E_INVALID_ARG, "gi_stream_create: null out pointer");
    hipStream_t s;
    GI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (gi_stream_t)s;
    return 0;
}

int gi_stream_destroy(gi_stream_t stream) {
    if (stream) GI_HIP(hipStreamDestroy(as_stream(stream)));
    return 0;
}

int gi_stream_sync(gi_stream_t stream) {
    GI_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int gi_device_sync(void) {
    GI_HIP(hipDeviceSynchronize());
    return 0;
}

int gi_sync(void) { return gi_device_sync(); }

int gi_tiling_for(long long n, int* r, int* variant, long long* chunk_len, long long* target_tasks,
                  long long* min_chunk) {
    if (n < 0) return set_error(GI_E_INVALID_ARG, "gi_tiling_for: n < 0");
    const Tiling t = choose_tiling(n, n * n, 1);
    if (r) *r = t.R;
    if (variant) *variant = t.variant;
    if (chunk_len) *chunk_len = t.chunk_len;
    if (target_tasks) *target_tasks = t.target_tasks;
    if (min_chunk) *min_chunk = t.min_chunk;
    return 0;
}

int gi_plan_check(int n_sims, const int* n, const int* nfe, int r, int variant, long long chunk_len, int batched) {
    if (n_sims <= 0 || !n || !nfe || (r != 1 && r != 2 && r != 4 && r != 8) || (variant != 2 && variant != 4) ||
        (variant == 4 && r < 4) || chunk_len < 0)
        return set_error(GI_E_INVALID_ARG, "gi_plan_check: invalid arguments");
    std::vector<SimDesc> sims;
    int off = 0;
    for (int s = 0; s < n_sims; ++s) {
        if (n[s] < 0 || nfe[s] < 0 || nfe[s] > n[s]) return set_error(GI_E_INVALID_ARG, "gi_plan_check: sim %d", s);
        sims.push_back(make_sim(off, off + nfe[s], off + n[s], PairLaw{}));
        off += n[s];
    }
    Plan p;
    int rc = build_plan(sims, Tiling{r, variant, chunk_len, 0, 64}, batched != 0, &p);
    if (rc) return rc;
    return check_plan_indexes(p);
}

int gi_plan_info(int* mutual, int* r, long long* pair_tasks, long long* chunk_len) {
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    const Plan& p = ctx->ws.plan;
    const bool valid = ctx->ws.plan_valid;
    if (mutual) *mutual = !valid ? 0 : (p.direct ? 2 : (p.mutual ? (p.group_rows > 1 ? 3 : 1) : 0));
    if (r) *r = valid ? p.tiling.R : 0;
    if (pair_tasks)
        *pair_tasks = !valid     ? 0
                      : p.direct ? (long long)p.dblocks.size()
                                 : (long long)(p.tasks.size() + (p.group_rows > 1 ? p.gtasks.size() : p.mtasks.size()));
    if (chunk_len) *chunk_len = valid ? p.chunk_len : 0;
    return 0;
}

int gi_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return set_error(GI_E_INVALID_ARG, "gi_mem_info: null pointer");
    GI_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return 0;
}

int gi_set_batched_tiling(int r, long long chunk_len, int variant) {
    if (r == 0) {
        batched_override() = BatchedTiling{0, 0, 0};
        return 0;
    }
    if ((r != 1 && r != 2 && r != 4 && r != 8) || chunk_len < 64 || variant < 0 || variant == 3 || variant > 4 ||
        (variant == 4 && r < 4))
        return set_error(GI_E_INVALID_ARG,
                         "gi_set_batched_tiling: r in {1,2,4,8}, chunk_len >= 64, variant in {0,1,2,4} (4: r >= 4)");
    batched_override() = BatchedTiling{(chunk_len + 63) / 64 * 64, r, variant};
    return 0;
}

int gi_set_contact_overlap(int enable) {
    g_contact_overlap.store(enable ? 1 : 0);
    return 0;
}

int gi_set_contact_walk(int lanes) {
    if (lanes != 0 && lanes != 1 && lanes != kCellRows)
        return set_error(GI_E_INVALID_ARG, "gi_set_contact_walk: lanes in {0 (by plan size), 1, 9}");
    g_contact_walk.store(lanes);
    return 0;
}

int gi_set_dispatch_limit(long long max_tasks) {
    if (max_tasks < 0) return set_error(GI_E_INVALID_ARG, "gi_set_dispatch_limit: max_tasks < 0");
    g_dispatch_limit.store(max_tasks);
    return 0;
}

int gi_set_tiling(int r, int variant, long long chunk_len, long long target_tasks, long long min_chunk) {
    if (r == 0) {
        tiling_override().active = 0;
        return 0;
    }
    if (r != 1 && r != 2 && r != 4 && r != 8)
        return set_error(GI_E_INVALID_ARG, "gi_set_tiling: r must be 1, 2, 4 or 8");
    if (variant < 0 || variant > 4 || (variant == 4 && r < 4))
        return set_error(GI_E_INVALID_ARG, "gi_set_tiling: variant must be 0, 1, 2, 3 or 4 (4 needs r >= 4)");
    if (chunk_len < 0 || target_tasks < 0 || min_chunk < 0)
        return set_error(GI_E_INVALID_ARG, "gi_set_tiling: negative size");
    tiling_override().active = 1;
    tiling_override().t = Tiling{r, variant, chunk_len, target_tasks > 0 ? target_tasks : 8192, min_chunk};
    return 0;
}

static int check_particles(const char* fn, const void* pos, const void* vel, const void* force, int n) {
    if (n < 0) return set_error(GI_E_INVALID_ARG, "%s: n < 0", fn);
    if (n > 0 && (!pos || !vel || !force)) return set_error(GI_E_INVALID_ARG, "%s: null device pointer", fn);
    return 0;
}

int gi_force_collision(const gi_float4* pos, const gi_float4* vel, gi_float4* force, int n,
                       const gi_force_collision_constants* c, gi_stream_t stream) {
    int rc = check_particles("gi_force_collision", pos, vel, force, n);
    if (rc) return rc;
    if (!c) return set_error(GI_E_INVALID_ARG, "gi_force_collision: null constants");
    if (c->NFe < 0 || c->NFe > n) return set_error(GI_E_INVALID_ARG, "gi_force_collision: NFe out of [0,n]");
    if (n == 0) return 0;
    DevCtx* ctx;
    if ((rc = current_ctx(&ctx))) return rc;
    hipStream_t st = as_stream(stream);
    std::vector<SimDesc> sims{make_sim(0, c->NFe, n, law_of(c))};
    if ((rc = ensure_plan(&ctx->ws, sims, st))) return rc;
    return run_force(ctx, &ctx->ws, (const float4*)pos, (const float4*)vel, (float4*)force, st, g_timing);
}

int gi_force_collision_noalloc(const gi_float4* pos, const gi_float4* vel, gi_float4* force, int n,
                               const gi_force_collision_constants* c, gi_stream_t stream) {
    int rc = check_particles("gi_force_collision_noalloc", pos, vel, force, n);
    if (rc) return rc;
    if (!c) return set_error(GI_E_INVALID_ARG, "gi_force_collision_noalloc: null constants");
    if (c->NFe < 0 || c->NFe > n) return set_error(GI_E_INVALID_ARG, "gi_force_collision_noalloc: NFe out of [0,n]");
    if (n > kNoAllocMaxN)
        return set_error(GI_E_TOO_LARGE, "gi_force_collision_noalloc: n = %d exceeds its maximum of %d particles", n,
                         kNoAllocMaxN);
    if (n == 0) return 0;
    int dev = 0;
    GI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_error(GI_E_NO_DEVICE, "device id %d out of range", dev);
    NoAllocScratch* sp;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mutex);
        if (!g_noalloc_ptr[dev]) {
            void* p = nullptr;
            GI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_noalloc_ptr[dev] = (NoAllocScratch*)p;
        }
        sp = g_noalloc_ptr[dev];
    }
    // The direct pass with the tuning table's lanes per particle for this n
    // (the plan gi_force_collision makes where the table picks that pass).
    const Tiling t = choose_tiling(n, (long long)n * n, 1);
    const int lanes = t.variant == 3 ? std::min(t.R, 8) : 1;
    const int ib = 64 / lanes;
    const int nb = nblk(n, ib);
    if (nb > kNoAllocMaxBlocks)
        return set_error(GI_E_TOO_LARGE, "gi_force_collision_noalloc: %d workgroups exceed its maximum of %d", nb,
                         kNoAllocMaxBlocks);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_noalloc_plan, dim3(nblk(nb, 256)), dim3(256), 0, st, sp, make_sim(0, c->NFe, n, law_of(c)),
                       nb, ib);
    const dim3 grid(nb), block(64 * kDirectWaves);
    const float4 *P = (const float4*)pos, *V = (const float4*)vel;
    float4* F = (float4*)force;
    if (lanes == 8)
        hipLaunchKernelGGL((k_direct<false, 8, 0>), grid, block, 0, st, P, V, sp->blocks, &sp->sim, F, nullptr,
                           FusedMove{});
    else if (lanes == 4)
        hipLaunchKernelGGL((k_direct<false, 4, 0>), grid, block, 0, st, P, V, sp->blocks, &sp->sim, F, nullptr,
                           FusedMove{});
    else if (lanes == 2)
        hipLaunchKernelGGL((k_direct<false, 2, 0>), grid, block, 0, st, P, V, sp->blocks, &sp->sim, F, nullptr,
                           FusedMove{});
    else
        hipLaunchKernelGGL((k_direct<false, 1, 0>), grid, block, 0, st, P, V, sp->blocks, &sp->sim, F, nullptr,
                           FusedMove{});
    GI_HIP(hipGetLastError());
    return 0;
}

int gi_force_separate(const gi_float4* pos, const gi_float4* vel, gi_float4* force, int n,
                      const gi_force_separate_constants* c, gi_stream_t stream) {
    int rc = check_particles("gi_force_separate", pos, vel, force, n);
    if (rc) return rc;
    if (!c) return set_error(GI_E_INVALID_ARG, "gi_force_separate: null constants");
    if (!(0 <= c->boarder1 && c->boarder1 <= c->boarder2 && c->boarder2 <= c->boarder3 && c->boarder3 <= n))
        return set_error(GI_E_INVALID_ARG, "gi_force_separate: borders must satisfy 0<=b1<=b2<=b3<=n");
    if (n == 0) return 0;
    DevCtx* ctx;
    if ((rc = current_ctx(&ctx))) return rc;
    hipStream_t st = as_stream(stream);
    const PairLaw L = law_of(c);
    // Body 1 = [0, b2) with iron [0, b1); body 2 = [b2, n) with iron [b2, b3):
    // two independent simulations, so cross-body pairs never meet
    // (generateTargetImpactor.cu:529-532).
    std::vector<SimDesc> sims;
    sims.push_back(make_sim(0, c->boarder1, c->boarder2, L));
    sims.push_back(make_sim(c->boarder2, c->boarder3, n, L));
    if ((rc = ensure_plan(&ctx->ws, sims, st))) return rc;
    return run_force(ctx, &ctx->ws, (const float4*)pos, (const float4*)vel, (float4*)force, st, g_timing);
}

int gi_move_collision(gi_float4* pos, gi_float4* vel, const gi_float4* force, int n,
                      const gi_move_collision_constants* c, gi_stream_t stream) {
    int rc = check_particles("gi_move_collision", pos, vel, force, n);
    if (rc) return rc;
    if (!c) return set_error(GI_E_INVALID_ARG, "gi_move_collision: null constants");
    if (n == 0) return 0;
    hipStream_t st = as_stream(stream);
    DevCtx* ctx = nullptr;
    TimingPair tp;
    if (g_timing) {
        if ((rc = current_ctx(&ctx))) return rc;
        tp = take_pair(ctx, 1);
        GI_HIP(hipEventRecord(tp.a, st));
    }
    hipLaunchKernelGGL(k_move_collision, dim3(nblk(n, 256)), dim3(256), 0, st, (float4*)pos, (float4*)vel,
                       (const float4*)force, n, c->Dt, c->DtOverMassFe, c->DtOverMassSi, c->NFe);
    GI_HIP(hipGetLastError());
    if (g_timing) {
        GI_HIP(hipEventRecord(tp.b, st));
        ctx->pending.push_back(tp);
    }
    return 0;
}

int gi_move_damped_separate(gi_float4* pos, gi_float4* vel, const gi_float4* force, int n,
                            const gi_move_separate_constants* c, float damp_rate_body1, float damp_rate_body2,
                            gi_stream_t stream) {
    int rc = check_particles("gi_move_damped_separate", pos, vel, force, n);
    if (rc) return rc;
    if (!c) return set_error(GI_E_INVALID_ARG, "gi_move_damped_separate: null constants");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_move_damped_separate, dim3(nblk(n, 256)), dim3(256), 0, as_stream(stream),
                       (float4*)pos, (float4*)vel, (const float4*)force, n, c->Dt, c->DtOverMassFe,
                       c->DtOverMassSi, c->boarder1, c->boarder2, c->boarder3, damp_rate_body1, damp_rate_body2);
    GI_HIP(hipGetLastError());
    return 0;
}

int gi_center_of_mass_velocity(const gi_float4* pos, const gi_float4* vel, int n, int nfe, double mass_fe,
                               double mass_si, double* out6, gi_stream_t stream) {
    if (n <= 0 || !pos || !vel || !out6 || nfe < 0 || nfe > n)
        return set_error(GI_E_INVALID_ARG, "gi_center_of_mass_velocity: invalid arguments");
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    const int nb = nblk(n, 256);
    if ((rc = ctx->rc_blocks.ensure((size_t)nb * kMoments1 + 1))) return rc;
    if ((rc = ctx->rc_sums.ensure(kMoments1 + 1))) return rc;
    hipLaunchKernelGGL(k_moments_one, dim3(nb), dim3(256), 0, st, (const float4*)pos, (const float4*)vel, n, nfe,
                       mass_fe, mass_si, ctx->rc_blocks.p);
    hipLaunchKernelGGL(k_moments_one_final, dim3(1), dim3(64), 0, st, (const double*)ctx->rc_blocks.p, nb,
                       ctx->rc_sums.p);
    GI_HIP(hipGetLastError());
    double s[kMoments1];
    GI_HIP(hipMemcpyAsync(s, ctx->rc_sums.p, sizeof s, hipMemcpyDeviceToHost, st));
    GI_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 6; ++k) out6[k] = s[k + 1] / s[0];
    return 0;
}

int gi_recenter(gi_float4* pos, gi_float4* vel, int n, int nfe, double mass_fe, double mass_si,
                const double* target_com3, const double* target_vel3, gi_stream_t stream) {
    if (n <= 0 || !pos || !vel || !target_com3 || !target_vel3 || nfe < 0 || nfe > n)
        return set_error(GI_E_INVALID_ARG, "gi_recenter: invalid arguments");
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    return launch_recenter(ctx, (float4*)pos, (float4*)vel, n, nfe, mass_fe, mass_si, target_com3, target_vel3,
                           as_stream(stream));
}

int gi_recenter_float(gi_float4* pos, gi_float4* vel, int n, int nfe, double mass_fe, double mass_si,
                      const float* initial_com3, const float* initial_vel3, gi_stream_t stream) {
    if (n <= 0 || !pos || !vel || !initial_com3 || !initial_vel3 || nfe < 0 || nfe > n)
        return set_error(GI_E_INVALID_ARG, "gi_recenter_float: invalid arguments");
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    return launch_recenter_ref(ctx, (float4*)pos, (float4*)vel, n, nfe, mass_fe, mass_si, initial_com3, initial_vel3,
                               as_stream(stream));
}

int gi_find_earth_moon(const gi_float4* pos, const gi_float4* vel, int n, int nfe1, int nfe2, int nsi1,
                       double mass_fe, double mass_si, float touch, int* group, gi_earth_moon* out,
                       gi_stream_t stream) {
    if (n <= 0 || !pos || !vel || !out || nfe1 < 0 || nfe2 < 0 || nsi1 < 0 || (long long)nfe1 + nfe2 + nsi1 > n ||
        !(touch > 0.0f) || !(touch < 1e30f))
        return set_error(GI_E_INVALID_ARG, "gi_find_earth_moon: invalid arguments");
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    CCWork& w = ctx->cc;
    hipStream_t st = as_stream(stream);
    unsigned m = 64;
    while (m < (unsigned)n) m <<= 1;
    if ((m + 1023) / 1024 > 4096) return set_error(GI_E_TOO_LARGE, "gi_find_earth_moon: n too large (%d)", n);
    const int nbs = nblk(m, 1024);
    if ((rc = w.counts.ensure(m + 1)) || (rc = w.starts.ensure(m + 2)) || (rc = w.cellhash.ensure(n)) ||
        (rc = w.sorted.ensure(n)) || (rc = w.block_sums.ensure(nbs + 8)) || (rc = w.parent.ensure(n)) ||
        (rc = w.size.ensure(n + 1)) || (rc = w.best.ensure(2)))
        return rc;
    // cells a sixteenth larger than touch: a touching pair is never two cells apart
    const double inv_c = 1.0 / ((double)touch * 1.0625);
    const float4* P = (const float4*)pos;
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(m, 256)), dim3(256), 0, st, w.counts.p, (int)m, 0);
    hipLaunchKernelGGL(k_cc_count, dim3(nblk(n, 256)), dim3(256), 0, st, P, n, inv_c, m - 1, w.counts.p,
                       w.cellhash.p);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nbs), dim3(256), 0, st, w.counts.p, (int)m, w.starts.p, w.block_sums.p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTopThreads), 0, st, w.block_sums.p, nbs, w.starts.p + m);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk(m, 256)), dim3(256), 0, st, w.starts.p, (int)m, w.block_sums.p);
    hipLaunchKernelGGL(k_grid_scatter, dim3(nblk(n, 256)), dim3(256), 0, st, n, w.cellhash.p, w.starts.p,
                       w.counts.p, w.sorted.p);
    hipLaunchKernelGGL(k_cc_init, dim3(nblk(n, 256)), dim3(256), 0, st, w.parent.p, w.size.p, n);
    hipLaunchKernelGGL(k_cc_union, dim3(nblk(n, 256)), dim3(256), 0, st, P, n, inv_c, m - 1, touch, w.starts.p,
                       w.sorted.p, w.parent.p);
    hipLaunchKernelGGL(k_cc_compress, dim3(nblk(n, 256)), dim3(256), 0, st, w.parent.p, w.size.p, n);
    const unsigned long long zero[2] = {0ull, 0ull};
    GI_HIP(hipMemcpyAsync(w.best.p, zero, sizeof zero, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cc_top, dim3(nblk(n, 256)), dim3(256), 0, st, w.size.p, n, -1, w.best.p);
    GI_HIP(hipGetLastError());
    unsigned long long best[2];
    GI_HIP(hipMemcpyAsync(best, w.best.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    GI_HIP(hipStreamSynchronize(st));
    const int earth = (int)~(unsigned)(best[0] & 0xffffffffull);
    hipLaunchKernelGGL(k_cc_top, dim3(nblk(n, 256)), dim3(256), 0, st, w.size.p, n, earth, w.best.p + 1);
    GI_HIP(hipGetLastError());
    GI_HIP(hipMemcpyAsync(best + 1, w.best.p + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    GI_HIP(hipStreamSynchronize(st));
    const int moon = best[1] ? (int)~(unsigned)(best[1] & 0xffffffffull) : -1;
    // moments of the system, the Earth and the Moon
    const int nb = std::min(1024, nblk(n, kCCBlock));
    if ((rc = w.moments.ensure((size_t)nb * kCCClasses * kCCMoments)) ||
        (rc = w.mcounts.ensure((size_t)nb * kCCCounts)))
        return rc;
    hipLaunchKernelGGL(k_cc_moments, dim3(nb), dim3(kCCBlock), 0, st, P, (const float4*)vel, n, w.parent.p, earth,
                       moon, nfe1, nfe2, nsi1, mass_fe, mass_si, w.moments.p, w.mcounts.p);
    GI_HIP(hipGetLastError());
    std::vector<double> mom((size_t)nb * kCCClasses * kCCMoments);
    std::vector<int> cnt((size_t)nb * kCCCounts);
    std::vector<int> sizes(n);
    GI_HIP(hipMemcpyAsync(mom.data(), w.moments.p, mom.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    GI_HIP(hipMemcpyAsync(cnt.data(), w.mcounts.p, cnt.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    GI_HIP(hipMemcpyAsync(sizes.data(), w.size.p, sizes.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    if (group) GI_HIP(hipMemcpyAsync(group, w.parent.p, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, st));
    GI_HIP(hipStreamSynchronize(st));
    // sums over blocks in block order; scope 1 (Earth-Moon) = Earth + Moon
    double M[4][kCCMoments] = {};
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < kCCClasses; ++c)
            for (int k = 0; k < kCCMoments; ++k)
                M[c == 0 ? 0 : c + 1][k] += mom[((size_t)b * kCCClasses + c) * kCCMoments + k];
    for (int k = 0; k < kCCMoments; ++k) M[1][k] = M[2][k] + M[3][k];
    std::memset(out, 0, sizeof *out);
    int c8[kCCCounts] = {};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < kCCCounts; ++k) c8[k] += cnt[(size_t)b * kCCCounts + k];
    out->n_groups = 0;
    for (int i = 0; i < n; ++i) out->n_groups += sizes[i] > 0;
    out->earth_id = earth;
    out->moon_id = moon;
    out->earth_count = sizes[earth];
    out->moon_count = moon >= 0 ? sizes[moon] : 0;
    out->earth_fe1 = c8[0];
    out->earth_fe2 = c8[1];
    out->earth_si1 = c8[2];
    out->earth_si2 = c8[3];
    out->moon_fe1 = c8[4];
    out->moon_fe2 = c8[5];
    out->moon_si1 = c8[6];
    out->moon_si2 = c8[7];
    for (int s = 0; s < 4; ++s) {
        const double* a = M[s];
        const double mass = a[0];
        out->mass[s] = mass;
        for (int d = 0; d < 3; ++d) {
            out->com[s][d] = a[1 + d] / mass;  // 0/0 = NaN for an empty scope, as the reference divides
            out->vel[s][d] = a[4 + d] / mass;
        }
        // L about the scope's centre of mass relative to its mean velocity:
        // sum m x cross v - (sum m x) cross (sum m v) / M
        const double* X = a + 1;
        const double* Pm = a + 4;
        const double cx = (X[1] * Pm[2] - X[2] * Pm[1]) / mass;
        const double cy = (X[2] * Pm[0] - X[0] * Pm[2]) / mass;
        const double cz = (X[0] * Pm[1] - X[1] * Pm[0]) / mass;
        out->angmom[s][0] = a[7] - cx;
        out->angmom[s][1] = a[8] - cy;
        out->angmom[s][2] = a[9] - cz;
    }
    return 0;
}

int gi_conserved_quantities(const gi_float4* pos, const gi_float4* vel, int n, const gi_force_collision_constants* c,
                            double* out9, gi_stream_t stream) {
    if (n <= 0 || !pos || !vel || !c || !out9 || c->NFe < 0 || c->NFe > n)
        return set_error(GI_E_INVALID_ARG, "gi_conserved_quantities: invalid arguments");
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    std::vector<SimDesc> sims{make_sim(0, c->NFe, n, law_of(c))};
    if ((rc = ensure_plan(&ctx->ws, sims, st))) return rc;
    if ((rc = run_conserved(ctx, &ctx->ws, (const float4*)pos, (const float4*)vel, st))) return rc;
    double m[kMoments];
    GI_HIP(hipMemcpyAsync(m, ctx->ws.sums.p, sizeof m, hipMemcpyDeviceToHost, st));
    GI_HIP(hipStreamSynchronize(st));
    moments_to_conserved(m, out9);
    return 0;
}

// The recenter of an impact step run: the reference's float one (its host
// sums restated on the device, with its double element masses and float
// targets), or FP64 device moments toward double targets.
struct RecenterSpec {
    bool ref_float = false;
    double mass_fe = 0.0, mass_si = 0.0;  // ref_float: the element masses
    float tcf[3] = {0, 0, 0}, tvf[3] = {0, 0, 0};
    double tcd[3] = {0, 0, 0}, tvd[3] = {0, 0, 0};
};

static int impact_steps_impl(gi_float4* pos, gi_float4* vel, gi_float4* force, int n,
                             const gi_force_collision_constants* fc, const gi_move_collision_constants* mc,
                             int n_steps, int recenter_rate, const RecenterSpec& rs, gi_stream_t stream) {
    int rc;
    if (n == 0 || n_steps == 0) return 0;
    DevCtx* ctx;
    if ((rc = current_ctx(&ctx))) return rc;
    StreamBridge br;
    if ((rc = br.begin(ctx, as_stream(stream)))) return rc;
    std::vector<SimDesc> sims{make_sim(0, fc->NFe, n, law_of(fc))};
    if ((rc = ensure_plan(&ctx->ws, sims, br.run))) return rc;
    // Make sure the plan uploads are complete before they are captured around.
    GI_HIP(hipStreamSynchronize(br.run));
    float4 *P = (float4*)pos, *V = (float4*)vel, *F = (float4*)force;
    const gi_move_collision_constants m = *mc;
    std::vector<unsigned char> key;
    key_append(key, 1);
    key_append(key, P);
    key_append(key, V);
    key_append(key, F);
    key_append(key, n);
    key_append(key, *fc);
    key_append(key, m);
    // The move is fused into the force's last kernel when the plan allows
    // (k_move_collision's arithmetic, with these constants; nfe = NFe).
    const SumMove smv{m.Dt, m.DtOverMassFe, m.DtOverMassSi, 0};
    auto body = [&](hipStream_t cs, int steps) -> int {
        for (int k = 0; k < steps; ++k) {
            bool moved = false;
            int r = run_force(ctx, &ctx->ws, P, V, F, cs, false, P, V, &moved, &smv);
            if (r) return r;
            if (!moved)
                hipLaunchKernelGGL(k_move_collision, dim3(nblk(n, 256)), dim3(256), 0, cs, P, V, (const float4*)F,
                                   n, m.Dt, m.DtOverMassFe, m.DtOverMassSi, m.NFe);
            GI_HIP(hipGetLastError());
        }
        return 0;
    };
    hipGraphExec_t step, steps = nullptr;
    rc = get_graph(ctx, ctx->graphs, key, ctx->ws.generation, [&](hipStream_t cs) { return body(cs, 1); }, &step);
    if (rc) return rc;
    if (n_steps >= kStepsPerGraph && (recenter_rate == 0 || recenter_rate >= kStepsPerGraph)) {
        key_append(key, kStepsPerGraph);
        if (ctx->ws.plan.direct) {
            // the move fused into the direct pass, alternating states
            if ((rc = ctx->alt_pos.ensure((size_t)n)) || (rc = ctx->alt_vel.ensure((size_t)n))) return rc;
            key_append(key, ctx->alt_pos.p);
            key_append(key, ctx->alt_vel.p);
            const FusedMove mv{nullptr, nullptr, m.Dt, m.DtOverMassFe, m.DtOverMassSi, m.NFe, 0, 0, 0.0f, 0.0f};
            rc = get_graph(ctx, ctx->graphs, key, ctx->ws.generation,
                           [&](hipStream_t cs) { return fused_direct_steps(ctx, P, V, F, mv, 1, cs); }, &steps,
                           {step});
        } else {
            rc = get_graph(ctx, ctx->graphs, key, ctx->ws.generation,
                           [&](hipStream_t cs) { return body(cs, kStepsPerGraph); }, &steps, {step});
        }
        if (rc) return rc;
    }
    hipGraphExec_t rec = nullptr;
    if (recenter_rate > 0) {
        std::vector<unsigned char> k2;
        key_append(k2, rs.ref_float ? 5 : 2);
        key_append(k2, P);
        key_append(k2, V);
        key_append(k2, n);
        key_append(k2, fc->NFe);
        // The recenter's scratch exists before capture, and its address is
        // part of the key: a graph holding a buffer that was since
        // reallocated is never launched again.
        if (rs.ref_float) {
            key_append(k2, rs.mass_fe);
            key_append(k2, rs.mass_si);
            for (int a = 0; a < 3; ++a) {
                key_append(k2, rs.tcf[a]);
                key_append(k2, rs.tvf[a]);
            }
            if ((rc = ctx->rc_shift.ensure(8))) return rc;
            key_append(k2, ctx->rc_shift.p);
            rc = get_graph(ctx, ctx->graphs, k2, ctx->ws.generation,
                           [&](hipStream_t cs) -> int {
                               return launch_recenter_ref(ctx, P, V, n, fc->NFe, rs.mass_fe, rs.mass_si, rs.tcf,
                                                          rs.tvf, cs);
                           },
                           &rec, {step, steps});
        } else {
            // The FP64 recenter uses the move's masses: m = Dt / (Dt/m).
            const double mfe = (double)m.Dt / (double)m.DtOverMassFe;
            const double msi = (double)m.Dt / (double)m.DtOverMassSi;
            for (int a = 0; a < 3; ++a) {
                key_append(k2, rs.tcd[a]);
                key_append(k2, rs.tvd[a]);
            }
            if ((rc = ctx->rc_blocks.ensure((size_t)nblk(n, 256) * kMoments1 + 1))) return rc;
            if ((rc = ctx->rc_sums.ensure(kMoments1 + 1))) return rc;
            key_append(k2, ctx->rc_blocks.p);
            key_append(k2, ctx->rc_sums.p);
            rc = get_graph(ctx, ctx->graphs, k2, ctx->ws.generation,
                           [&](hipStream_t cs) -> int {
                               return launch_recenter(ctx, P, V, n, fc->NFe, mfe, msi, rs.tcd, rs.tvd, cs);
                           },
                           &rec, {step, steps});
        }
        if (rc) return rc;
    }
    // as steps 1..n_steps with a recenter after every step s with
    // s % recenter_rate == 0
    for (int s = 0; s < n_steps;) {
        const int k = rec ? std::min(n_steps - s, recenter_rate - s % recenter_rate) : n_steps - s;
        if ((rc = launch_step_run(step, steps, k, br.run))) return rc;
        s += k;
        if (rec && s % recenter_rate == 0) GI_HIP(hipGraphLaunch(rec, br.run));
    }
    return br.end();
}

int gi_impact_steps(gi_float4* pos, gi_float4* vel, gi_float4* force, int n, const gi_force_collision_constants* fc,
                    const gi_move_collision_constants* mc, int n_steps, int recenter_rate, const double* target_com3,
                    const double* target_vel3, gi_stream_t stream) {
    int rc = check_particles("gi_impact_steps", pos, vel, force, n);
    if (rc) return rc;
    if (!fc || !mc || n_steps < 0 || recenter_rate < 0 || fc->NFe < 0 || fc->NFe > n || mc->NFe != fc->NFe)
        return set_error(GI_E_INVALID_ARG, "gi_impact_steps: invalid arguments");
    if (recenter_rate > 0 && (!target_com3 || !target_vel3))
        return set_error(GI_E_INVALID_ARG, "gi_impact_steps: recenter needs target_com3 and target_vel3");
    RecenterSpec rs;
    if (recenter_rate > 0)
        for (int a = 0; a < 3; ++a) {
            rs.tcd[a] = target_com3[a];
            rs.tvd[a] = target_vel3[a];
        }
    return impact_steps_impl(pos, vel, force, n, fc, mc, n_steps, recenter_rate, rs, stream);
}

int gi_impact_steps_float_recenter(gi_float4* pos, gi_float4* vel, gi_float4* force, int n,
                                   const gi_force_collision_constants* fc, const gi_move_collision_constants* mc,
                                   int n_steps, int recenter_rate, const float* initial_com3,
                                   const float* initial_vel3, double mass_fe, double mass_si, gi_stream_t stream) {
    int rc = check_particles("gi_impact_steps_float_recenter", pos, vel, force, n);
    if (rc) return rc;
    if (!fc || !mc || n_steps < 0 || recenter_rate < 0 || fc->NFe < 0 || fc->NFe > n || mc->NFe != fc->NFe)
        return set_error(GI_E_INVALID_ARG, "gi_impact_steps_float_recenter: invalid arguments");
    if (recenter_rate > 0 && (!initial_com3 || !initial_vel3))
        return set_error(GI_E_INVALID_ARG, "gi_impact_steps_float_recenter: recenter needs initial_com3 and initial_vel3");
    RecenterSpec rs;
    rs.ref_float = true;
    rs.mass_fe = mass_fe;
    rs.mass_si = mass_si;
    if (recenter_rate > 0)
        for (int a = 0; a < 3; ++a) {
            rs.tcf[a] = initial_com3[a];
            rs.tvf[a] = initial_vel3[a];
        }
    return impact_steps_impl(pos, vel, force, n, fc, mc, n_steps, recenter_rate, rs,
                             stream);
}

int gi_separate_steps(gi_float4* pos, gi_float4* vel, gi_float4* force, int n, const gi_force_separate_constants* fc,
                      const gi_move_separate_constants* mc, float damp_rate_body1, float damp_rate_body2, int n_steps,
                      gi_stream_t stream) {
    int rc = check_particles("gi_separate_steps", pos, vel, force, n);
    if (rc) return rc;
    if (!fc || !mc || n_steps < 0)
        return set_error(GI_E_INVALID_ARG, "gi_separate_steps: invalid arguments");
    if (!(0 <= fc->boarder1 && fc->boarder1 <= fc->boarder2 && fc->boarder2 <= fc->boarder3 && fc->boarder3 <= n))
        return set_error(GI_E_INVALID_ARG, "gi_separate_steps: borders must satisfy 0<=b1<=b2<=b3<=n");
    if (n == 0 || n_steps == 0) return 0;
    DevCtx* ctx;
    if ((rc = current_ctx(&ctx))) return rc;
    StreamBridge br;
    if ((rc = br.begin(ctx, as_stream(stream)))) return rc;
    const PairLaw L = law_of(fc);
    std::vector<SimDesc> sims;
    sims.push_back(make_sim(0, fc->boarder1, fc->boarder2, L));
    sims.push_back(make_sim(fc->boarder2, fc->boarder3, n, L));
    if ((rc = ensure_plan(&ctx->ws, sims, br.run))) return rc;
    GI_HIP(hipStreamSynchronize(br.run));
    float4 *P = (float4*)pos, *V = (float4*)vel, *F = (float4*)force;
    const gi_move_separate_constants m = *mc;
    std::vector<unsigned char> key;
    key_append(key, 4);
    key_append(key, P);
    key_append(key, V);
    key_append(key, F);
    key_append(key, n);
    key_append(key, *fc);
    key_append(key, m);
    key_append(key, damp_rate_body1);
    key_append(key, damp_rate_body2);
    auto body = [&](hipStream_t cs, int steps) -> int {
        for (int k = 0; k < steps; ++k) {
            int r = run_force(ctx, &ctx->ws, P, V, F, cs, false);
            if (r) return r;
            hipLaunchKernelGGL(k_move_damped_separate, dim3(nblk(n, 256)), dim3(256), 0, cs, P, V,
                               (const float4*)F, n, m.Dt, m.DtOverMassFe, m.DtOverMassSi, m.boarder1, m.boarder2,
                               m.boarder3, damp_rate_body1, damp_rate_body2);
            GI_HIP(hipGetLastError());
        }
        return 0;
    };
    hipGraphExec_t step, steps = nullptr;
    rc = get_graph(ctx, ctx->graphs, key, ctx->ws.generation, [&](hipStream_t cs) { return body(cs, 1); }, &step);
    if (rc) return rc;
    if (n_steps >= kStepsPerGraph) {
        key_append(key, kStepsPerGraph);
        if (ctx->ws.plan.direct) {
            // the damped move fused into the direct pass, alternating states
            if ((rc = ctx->alt_pos.ensure((size_t)n)) || (rc = ctx->alt_vel.ensure((size_t)n))) return rc;
            key_append(key, ctx->alt_pos.p);
            key_append(key, ctx->alt_vel.p);
            const FusedMove mv{nullptr,    nullptr,    m.Dt,       m.DtOverMassFe,  m.DtOverMassSi,
                               m.boarder1, m.boarder2, m.boarder3, damp_rate_body1, damp_rate_body2};
            rc = get_graph(ctx, ctx->graphs, key, ctx->ws.generation,
                           [&](hipStream_t cs) { return fused_direct_steps(ctx, P, V, F, mv, 2, cs); }, &steps,
                           {step});
        } else {
            rc = get_graph(ctx, ctx->graphs, key, ctx->ws.generation,
                           [&](hipStream_t cs) { return body(cs, kStepsPerGraph); }, &steps, {step});
        }
        if (rc) return rc;
    }
    if ((rc = launch_step_run(step, steps, n_steps, br.run))) return rc;
    return br.end();
}

// ---- ensemble ---------------------------------------------------------------

int gi_ensemble_create(gi_ensemble** out, int n_sims, const int* n, const gi_force_collision_constants* fc,
                       const gi_move_collision_constants* mc, gi_float4* pos, gi_float4* vel, gi_stream_t stream) {
    if (!out || n_sims <= 0 || !n || !fc || !mc || !pos || !vel)
        return set_error(GI_E_INVALID_ARG, "gi_ensemble_create: invalid arguments");
    *out = nullptr;
    int dev = 0;
    GI_HIP(hipGetDevice(&dev));
    gi_ensemble* e = new gi_ensemble();
    e->device = dev;
    e->ws.batched = true;
    e->ws_diag.batched = true;
    int off = 0;
    for (int s = 0; s < n_sims; ++s) {
        if (n[s] <= 0 || fc[s].NFe < 0 || fc[s].NFe > n[s] || mc[s].NFe != fc[s].NFe) {
            delete e;
            return set_error(GI_E_INVALID_ARG, "gi_ensemble_create: simulation %d has invalid n/NFe", s);
        }
        if ((long long)off + n[s] > (1LL << 21)) {
            delete e;
            return set_error(GI_E_TOO_LARGE, "gi_ensemble_create: more than 2^21 particles");
        }
        SimDesc d = make_sim(off, off + fc[s].NFe, off + n[s], law_of(&fc[s]));
        d.dt = mc[s].Dt;
        d.dt_over_mass_fe = mc[s].DtOverMassFe;
        d.dt_over_mass_si = mc[s].DtOverMassSi;
        e->sims.push_back(d);
        e->caller.push_back(s);
        off += n[s];
    }
    e->n_caller = n_sims;
    e->capacity = off;
    e->pos = (float4*)pos;
    e->vel = (float4*)vel;
    e->stream = as_stream(stream);
    int rc = e->force.ensure((size_t)off);
    if (!rc) rc = ensure_plan(&e->ws, e->sims, e->stream);
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return 0;
}

int gi_ensemble_create_capacity(gi_ensemble** out, int capacity, gi_float4* pos, gi_float4* vel,
                                gi_stream_t stream) {
    if (!out || capacity <= 0 || !pos || !vel)
        return set_error(GI_E_INVALID_ARG, "gi_ensemble_create_capacity: invalid arguments");
    if (capacity > (1 << 21)) return set_error(GI_E_TOO_LARGE, "gi_ensemble_create_capacity: more than 2^21 particles");
    *out = nullptr;
    int dev = 0;
    GI_HIP(hipGetDevice(&dev));
    gi_ensemble* e = new gi_ensemble();
    e->device = dev;
    e->ws.batched = true;
    e->ws_diag.batched = true;
    e->capacity = capacity;
    e->pos = (float4*)pos;
    e->vel = (float4*)vel;
    e->stream = as_stream(stream);
    int rc = e->force.ensure((size_t)capacity);
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return 0;
}

int gi_ensemble_set_sims(gi_ensemble* e, int n_sims, const int* offset, const int* n,
                         const gi_force_collision_constants* fc, const gi_move_collision_constants* mc) {
    if (!e || n_sims < 0 || (n_sims > 0 && (!offset || !n || !fc || !mc)))
        return set_error(GI_E_INVALID_ARG, "gi_ensemble_set_sims: invalid arguments");
    std::vector<SimDesc> sims;
    std::vector<int> caller;
    long long cursor = 0;
    for (int s = 0; s < n_sims; ++s) {
        if (n[s] <= 0 || offset[s] < cursor || (long long)offset[s] + n[s] > e->capacity || fc[s].NFe < 0 ||
            fc[s].NFe > n[s] || mc[s].NFe != fc[s].NFe)
            return set_error(GI_E_INVALID_ARG,
                             "gi_ensemble_set_sims: simulation %d must have n > 0, NFe in [0, n], and lie after the "
                             "previous one inside the capacity",
                             s);
        if (offset[s] > cursor) {
            // a frozen filler over the gap: never stepped, its records never written
            SimDesc f = make_sim((int)cursor, (int)cursor, offset[s], PairLaw{});
            f.active = 0;
            sims.push_back(f);
            caller.push_back(-1);
        }
        SimDesc d = make_sim(offset[s], offset[s] + fc[s].NFe, offset[s] + n[s], law_of(&fc[s]));
        d.dt = mc[s].Dt;
        d.dt_over_mass_fe = mc[s].DtOverMassFe;
        d.dt_over_mass_si = mc[s].DtOverMassSi;
        sims.push_back(d);
        caller.push_back(s);
        cursor = (long long)offset[s] + n[s];
    }
    e->sims = std::move(sims);
    e->caller = std::move(caller);
    e->n_caller = n_sims;
    return 0;
}

int gi_ensemble_destroy(gi_ensemble* e) {
    delete e;
    return 0;
}

int gi_ensemble_set_active(gi_ensemble* e, const int* active) {
    if (!e || !active) return set_error(GI_E_INVALID_ARG, "gi_ensemble_set_active: invalid arguments");
    for (size_t s = 0; s < e->sims.size(); ++s)
        if (e->caller[s] >= 0) e->sims[s].active = active[e->caller[s]] ? 1 : 0;
    return 0;
}

int gi_ensemble_wait(gi_ensemble* e, int max_pending) {
    if (!e || max_pending < 0) return set_error(GI_E_INVALID_ARG, "gi_ensemble_wait: invalid arguments");
    while ((int)e->pending.size() > max_pending) {
        hipEvent_t ev = e->pending.front();
        GI_HIP(hipEventSynchronize(ev));
        e->pending.erase(e->pending.begin());
        e->spare.push_back(ev);
    }
    return 0;
}

int gi_ensemble_steps(gi_ensemble* e, int n_steps) {
    if (!e || n_steps < 0) return set_error(GI_E_INVALID_ARG, "gi_ensemble_steps: invalid arguments");
    if (n_steps == 0) return 0;
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    if (ctx->device != e->device) return set_error(GI_E_STATE, "gi_ensemble_steps: wrong current device");
    bool any = false;
    for (const SimDesc& sd : e->sims) any = any || (sd.active && sd.end > sd.start);
    if (!any) return 0;
    StreamBridge br;
    if ((rc = br.begin(ctx, e->stream))) return rc;
    const unsigned long long gen0 = e->ws.generation;
    // A new simulation table may reallocate scratch that earlier launches on
    // this stream still use: let them finish first.
    if (!e->ws.plan_valid || !same_sims(e->ws.plan.caller_sims, e->sims)) GI_HIP(hipStreamSynchronize(br.run));
    if ((rc = ensure_plan(&e->ws, e->sims, br.run))) return rc;
    if (e->ws.generation != gen0) GI_HIP(hipStreamSynchronize(br.run));
    const int ntot = e->ws.plan.ntot;
    const int ns = (int)e->ws.plan.sims.size();
    std::vector<unsigned char> key;
    key_append(key, 3);
    hipGraphExec_t step;
    rc = get_graph(ctx, e->graphs, key, e->ws.generation,
                   [&](hipStream_t cs) -> int {
                       bool moved = false;
                       int r = run_force(ctx, &e->ws, e->pos, e->vel, e->force.p, cs, false, e->pos, e->vel,
                                         &moved);
                       if (r) return r;
                       if (!moved)
                           hipLaunchKernelGGL(k_move, dim3(nblk(ntot, 256)), dim3(256), 0, cs, e->pos, e->vel,
                                              (const float4*)e->force.p, ntot, e->ws.sims.p, ns);
                       GI_HIP(hipGetLastError());
                       return 0;
                   },
                   &step, {}, true);
    if (rc) return rc;
    for (int s = 0; s < n_steps; ++s) GI_HIP(hipGraphLaunch(step, br.run));
    // retire the calls that have finished, so a caller that never waits keeps
    // a bounded list
    while (!e->pending.empty() && hipEventQuery(e->pending.front()) == hipSuccess) {
        e->spare.push_back(e->pending.front());
        e->pending.erase(e->pending.begin());
    }
    hipEvent_t ev;
    if (!e->spare.empty()) {
        ev = e->spare.back();
        e->spare.pop_back();
    } else {
        GI_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    GI_HIP(hipEventRecord(ev, br.run));
    e->pending.push_back(ev);
    return br.end();
}

int gi_ensemble_conserved(gi_ensemble* e, double* out) {
    if (!e || !out) return set_error(GI_E_INVALID_ARG, "gi_ensemble_conserved: invalid arguments");
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    std::vector<SimDesc> all = e->sims;
    for (size_t s = 0; s < all.size(); ++s) all[s].active = e->caller[s] >= 0 ? 1 : 0;  // fillers stay frozen
    if (all.empty()) return 0;
    StreamBridge br;
    if ((rc = br.begin(ctx, e->stream))) return rc;
    if ((rc = ensure_plan(&e->ws_diag, all, br.run))) return rc;
    if ((rc = run_conserved(ctx, &e->ws_diag, e->pos, e->vel, br.run))) return rc;
    const size_t ns = all.size();
    std::vector<double> m(ns * kMoments);
    GI_HIP(hipMemcpyAsync(m.data(), e->ws_diag.sums.p, m.size() * sizeof(double), hipMemcpyDeviceToHost, br.run));
    GI_HIP(hipStreamSynchronize(br.run));
    for (size_t s = 0; s < ns; ++s)
        if (e->caller[s] >= 0) moments_to_conserved(&m[s * kMoments], out + 9 * e->caller[s]);
    return br.end();
}

int gi_ensemble_work(gi_ensemble* e, long long* particles, double* pairs) {
    if (!e) return set_error(GI_E_INVALID_ARG, "gi_ensemble_work: null ensemble");
    long long p = 0;
    double q = 0.0;
    for (auto& s : e->sims)
        if (s.active) {
            p += s.end - s.start;
            q += (double)(s.end - s.start) * (double)(s.end - s.start);
        }
    if (particles) *particles = p;
    if (pairs) *pairs = q;
    return 0;
}

// ---- timing -------------------------------------------------------------------

int gi_set_kernel_timing(int enable) {
    g_timing = enable != 0;
    return 0;
}

int gi_kernel_timing(const char* kernel, long long* launches, double* total_ms) {
    int kind = 0;
    if (kernel && std::strcmp(kernel, "move") == 0)
        kind = 1;
    else if (kernel && std::strcmp(kernel, "gravity_pass") == 0)
        kind = 2;
    else if (kernel && std::strcmp(kernel, "force") == 0)
        kind = 3;
    else if (kernel && std::strcmp(kernel, "gravity") != 0)
        return set_error(GI_E_INVALID_ARG, "unknown kernel '%s' (timed: gravity, gravity_pass, force, move)",
                         kernel);
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    for (auto& tp : ctx->pending) {
        GI_HIP(hipEventSynchronize(tp.b));
        float ms = 0.0f;
        GI_HIP(hipEventElapsedTime(&ms, tp.a, tp.b));
        ctx->timed_ms[tp.kind] += ms;
        ctx->timed_launches[tp.kind]++;
        ctx->pool.push_back(tp);
    }
    ctx->pending.clear();
    if (launches) *launches = ctx->timed_launches[kind];
    if (total_ms) *total_ms = ctx->timed_ms[kind];
    return 0;
}

int gi_reset_kernel_timing(void) {
    long long l;
    double t;
    int rc = gi_kernel_timing(nullptr, &l, &t);
    if (rc) return rc;
    DevCtx* ctx;
    if ((rc = current_ctx(&ctx))) return rc;
    for (int k = 0; k < kTimedKinds; ++k) {
        ctx->timed_launches[k] = 0;
        ctx->timed_ms[k] = 0.0;
    }
    return 0;
}

}  // extern "C"

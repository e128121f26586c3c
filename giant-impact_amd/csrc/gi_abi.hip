// gi_abi.hip -- C ABI (include/giant_impact.h) over the gfx950 kernels.
//
// Host-side responsibilities: per-device scratch (partial sums, contact grid,
// task tables), the work plan of the all-pairs pass (tiling chosen by the
// autotuned table in gi_tuning.h), kernel launches in stream order, HIP-event
// timing of the all-pairs kernel, and HIP-graph capture of the step loop.
#include "gi_kernels.hip"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/giant_impact.h"
#include "gi_tuning.h"

static_assert(sizeof(gi_float4) == 16, "float4 layout");
static_assert(sizeof(gi_force_collision_constants) == 56, "forceCollisionKernalConstantsStruct is 56 bytes");
static_assert(sizeof(gi_move_collision_constants) == 16, "moveCollisionKernalConstantsStruct is 16 bytes");
static_assert(sizeof(gi_force_separate_constants) == 64, "forceSeperateKernalConstantsStruct is 64 bytes");
static_assert(sizeof(gi_move_separate_constants) == 24, "moveSeperateKernalConstantsStruct is 24 bytes");

namespace gi {

static thread_local std::string g_last_error;

static int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
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
// Per-device scratch
// ---------------------------------------------------------------------------

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    int ensure(size_t n) {
        if (n <= cap && p) return 0;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return set_error((int)e, "hipFree: %s", hipGetErrorString(e));
        }
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

struct TimingPair {
    hipEvent_t a = nullptr, b = nullptr;
};

// The plan of one force evaluation over a set of packed simulations.
struct Plan {
    std::vector<SimDesc> sims;
    std::vector<GravTask> tasks;
    std::vector<int> block_sim, block_first, sim_first_block;
    int ntot = 0;
    int R = 4;
    int partial_len = 0;  // float4 elements
    unsigned mask = 0;    // hash table size - 1
};

struct DevCtx {
    int device = -1;
    DevBuf<float4> partial;
    DevBuf<int> counts, starts, sorted, cellhash, block_sums;
    DevBuf<GravTask> tasks;
    DevBuf<SimDesc> sims;
    DevBuf<int> block_sim, block_first, sim_first_block;
    DevBuf<double> moments, sums, pe, target;
    // cached plan (uploaded copy)
    Plan plan;
    bool plan_valid = false;
    // timing
    std::vector<TimingPair> pending;
    std::vector<TimingPair> pool;
    long long timed_launches = 0;
    double timed_ms = 0.0;
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
    L.GMassFeFe = c->GMassFeFe;
    L.GMassFeSi = c->GMassFeSi;
    L.KFeFe = c->KFeFe;
    L.KSiSi = c->KSiSi;
    L.KFeSi = c->KFeSi;
    L.KRFeFe = c->KRFeFe;
    L.KRSiSi = c->KRSiSi;
    L.KRFeSi = c->KRFeSi;
    L.KRMix = c->KRMix;
    L.ShellBreakFe = c->ShellBreakFe;
    L.ShellBreakSi = c->ShellBreakSi;
    L.ShellBreakFeSi1 = c->ShellBreakFeSi1;
    L.ShellBreakFeSi2 = c->ShellBreakFeSi2;
    return L;
}

static PairLaw law_of(const gi_force_separate_constants* c) {
    gi_force_collision_constants t;
    std::memcpy(&t, c, 13 * sizeof(float));
    t.NFe = 0;
    return law_of(&t);
}

static bool same_sims(const std::vector<SimDesc>& a, const std::vector<SimDesc>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(SimDesc)) == 0;
}

// Fill nchunk / partial_base of the simulations and build the task table.
static int build_plan(std::vector<SimDesc> sims, const Tiling& tl, Plan* out) {
    Plan p;
    p.R = tl.R;
    long long ntot = 0;
    long long work = 0;  // sum over sims of iblocks * n (pair-slots)
    const int ib = 64 * tl.R;
    for (auto& s : sims) {
        long long n = s.end - s.start;
        if (n < 0 || s.fe_end < s.start || s.fe_end > s.end || s.start != ntot)
            return set_error(GI_E_INVALID_ARG, "simulation ranges must be packed and ordered");
        ntot = s.end;
        if (s.active) work += ((n + ib - 1) / ib) * n;
    }
    // The contact-grid scan handles up to 4096 * 1024 buckets = 2 * 2^21.
    if (ntot > (1LL << 21)) return set_error(GI_E_TOO_LARGE, "too many particles (%lld > 2^21)", ntot);
    long long L = tl.chunk_len > 0 ? tl.chunk_len : work / (long long)tl.target_tasks;
    if (L < tl.min_chunk) L = tl.min_chunk;
    L = (L + 63) / 64 * 64;
    long long pbase = 0;
    for (size_t si = 0; si < sims.size(); ++si) {
        SimDesc& s = sims[si];
        const int n = s.end - s.start;
        s.partial_base = (int)pbase;
        s.nchunk = 0;
        if (n == 0 || !s.active) continue;
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
    // Interleave tasks so consecutive wavefronts share a j-chunk (L2 reuse on
    // each XCD) while different i-blocks spread over the 8 XCDs.
    p.partial_len = (int)pbase;
    p.ntot = (int)ntot;
    unsigned m = 1024;
    while (m < 2u * (unsigned)ntot) m <<= 1;
    p.mask = m - 1;
    // moment blocks: 256 particles each, never straddling simulations
    p.sim_first_block.push_back(0);
    for (size_t si = 0; si < sims.size(); ++si) {
        for (int f = sims[si].start; f < sims[si].end; f += 256) {
            p.block_sim.push_back((int)si);
            p.block_first.push_back(f);
        }
        p.sim_first_block.push_back((int)p.block_sim.size());
    }
    p.sims = std::move(sims);
    *out = std::move(p);
    return 0;
}

template <typename T>
static int upload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    int rc = b.ensure(v.size());
    if (rc) return rc;
    if (!v.empty()) GI_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return 0;
}

// Make ctx->plan describe `sims` (re-planning and uploading only on change).
static int ensure_plan(DevCtx* ctx, const std::vector<SimDesc>& sims, hipStream_t st) {
    long long ntot = sims.empty() ? 0 : sims.back().end;
    long long pairs = 0;
    for (auto& s : sims) pairs += (long long)(s.end - s.start) * (s.end - s.start);
    Tiling tl = choose_tiling(ntot, pairs, (int)sims.size());
    if (ctx->plan_valid && same_sims(ctx->plan.sims, sims) && ctx->plan.R == tl.R) return 0;
    Plan p;
    int rc = build_plan(sims, tl, &p);
    if (rc) return rc;
    if ((rc = ctx->partial.ensure((size_t)p.partial_len + 1))) return rc;
    if ((rc = ctx->counts.ensure((size_t)p.mask + 1))) return rc;
    if ((rc = ctx->starts.ensure((size_t)p.mask + 2))) return rc;
    if ((rc = ctx->sorted.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ctx->cellhash.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ctx->block_sums.ensure(((size_t)p.mask + 1) / 1024 + 8))) return rc;
    if ((rc = upload(ctx->tasks, p.tasks, st))) return rc;
    if ((rc = upload(ctx->sims, p.sims, st))) return rc;
    if ((rc = upload(ctx->block_sim, p.block_sim, st))) return rc;
    if ((rc = upload(ctx->block_first, p.block_first, st))) return rc;
    if ((rc = upload(ctx->sim_first_block, p.sim_first_block, st))) return rc;
    if ((rc = ctx->moments.ensure(p.block_sim.size() * kMoments + 1))) return rc;
    if ((rc = ctx->sums.ensure(p.sims.size() * kMoments + 1))) return rc;
    if ((rc = ctx->pe.ensure((size_t)p.ntot + 1))) return rc;
    if ((rc = ctx->target.ensure(8))) return rc;
    ctx->plan = std::move(p);
    ctx->plan_valid = true;
    return 0;
}

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// Launch sequences
// ---------------------------------------------------------------------------

static int launch_grid(DevCtx* ctx, const float4* pos, hipStream_t st) {
    const Plan& p = ctx->plan;
    const int m = (int)p.mask + 1;
    const int nsims = (int)p.sims.size();
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(m, 256)), dim3(256), 0, st, ctx->counts.p, m, 0);
    hipLaunchKernelGGL(k_grid_count, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, pos, p.ntot,
                       ctx->sims.p, nsims, p.mask, ctx->counts.p, ctx->cellhash.p);
    const int nb = nblk(m, 1024);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, ctx->counts.p, m, ctx->starts.p,
                       ctx->block_sums.p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, ctx->block_sums.p, nb, ctx->starts.p + m);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk(m, 256)), dim3(256), 0, st, ctx->starts.p, m,
                       ctx->block_sums.p);
    hipLaunchKernelGGL(k_grid_scatter, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, p.ntot,
                       ctx->cellhash.p, ctx->starts.p, ctx->counts.p, ctx->sorted.p);
    hipLaunchKernelGGL(k_grid_sort_buckets, dim3(nblk(m, 256)), dim3(256), 0, st, m, ctx->starts.p,
                       ctx->sorted.p);
    GI_HIP(hipGetLastError());
    return 0;
}

static TimingPair take_pair(DevCtx* ctx) {
    TimingPair t;
    if (!ctx->pool.empty()) {
        t = ctx->pool.back();
        ctx->pool.pop_back();
        return t;
    }
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    return t;
}

static int launch_gravity(DevCtx* ctx, const float4* pos, bool pot, hipStream_t st, bool timed) {
    const Plan& p = ctx->plan;
    const int ntasks = (int)p.tasks.size();
    if (ntasks == 0) return 0;
    const int wpb = 4;
    dim3 grid(nblk(ntasks, wpb)), block(64 * wpb);
    TimingPair tp;
    if (timed) {
        tp = take_pair(ctx);
        (void)hipEventRecord(tp.a, st);
    }
    if (pot) {
        hipLaunchKernelGGL((k_gravity<4, 128, true>), grid, block, 0, st, pos, ctx->tasks.p, ntasks,
                           ctx->sims.p, ctx->partial.p);
    } else if (p.R == 8) {
        hipLaunchKernelGGL((k_gravity<8, 128, false>), grid, block, 0, st, pos, ctx->tasks.p, ntasks,
                           ctx->sims.p, ctx->partial.p);
    } else if (p.R == 2) {
        hipLaunchKernelGGL((k_gravity<2, 128, false>), grid, block, 0, st, pos, ctx->tasks.p, ntasks,
                           ctx->sims.p, ctx->partial.p);
    } else {
        hipLaunchKernelGGL((k_gravity<4, 128, false>), grid, block, 0, st, pos, ctx->tasks.p, ntasks,
                           ctx->sims.p, ctx->partial.p);
    }
    GI_HIP(hipGetLastError());
    if (timed) {
        (void)hipEventRecord(tp.b, st);
        ctx->pending.push_back(tp);
    }
    return 0;
}

// Full force evaluation for the current plan: grid, all pairs, finish.
static int run_force(DevCtx* ctx, const float4* pos, const float4* vel, float4* force,
                     hipStream_t st, bool timed) {
    const Plan& p = ctx->plan;
    if (p.ntot == 0) return 0;
    int rc = launch_grid(ctx, pos, st);
    if (rc) return rc;
    // A plan with the R of the pot variant must match: the pot variant uses R=4 tasks.
    if ((rc = launch_gravity(ctx, pos, false, st, timed))) return rc;
    hipLaunchKernelGGL(k_finish, dim3(nblk(p.ntot, 256)), dim3(256), 0, st, pos, vel, force, p.ntot,
                       ctx->sims.p, (int)p.sims.size(), ctx->partial.p, ctx->starts.p,
                       ctx->sorted.p, p.mask);
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

}  // namespace gi

using namespace gi;

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int gi_abi_version(void) { return GI_ABI_VERSION; }

const char* gi_last_error_string(void) { return g_last_error.c_str(); }

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
        return set_error(GI_E_NOT_MULTIPLE_OF_BLOCK,
                         "Number of Particles is not a multiple of the block size");
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

int gi_mem_info(size_t* free_bytes, size_t* total_bytes) {
    GI_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return 0;
}

int gi_force_collision(const gi_float4* pos, const gi_float4* vel, gi_float4* force, int n,
                       const gi_force_collision_constants* c, gi_stream_t stream) {
    if (n < 0 || !c || (n > 0 && (!pos || !vel || !force)))
        return set_error(GI_E_INVALID_ARG, "gi_force_collision: invalid arguments");
    if (c->NFe < 0 || c->NFe > n) return set_error(GI_E_INVALID_ARG, "gi_force_collision: NFe out of [0,n]");
    if (n == 0) return 0;
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    std::vector<SimDesc> sims{make_sim(0, c->NFe, n, law_of(c))};
    if ((rc = ensure_plan(ctx, sims, st))) return rc;
    return run_force(ctx, (const float4*)pos, (const float4*)vel, (float4*)force, st, g_timing);
}

int gi_force_separate(const gi_float4* pos, const gi_float4* vel, gi_float4* force, int n,
                      const gi_force_separate_constants* c, gi_stream_t stream) {
    if (n < 0 || !c || (n > 0 && (!pos || !vel || !force)))
        return set_error(GI_E_INVALID_ARG, "gi_force_separate: invalid arguments");
    if (!(0 <= c->boarder1 && c->boarder1 <= c->boarder2 && c->boarder2 <= c->boarder3 && c->boarder3 <= n))
        return set_error(GI_E_INVALID_ARG, "gi_force_separate: borders must satisfy 0<=b1<=b2<=b3<=n");
    if (n == 0) return 0;
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    PairLaw L = law_of(c);
    // Body 1 = [0, b2) with iron [0, b1); body 2 = [b2, n) with iron [b2, b3):
    // two independent simulations, so cross-body pairs never meet
    // (generateTargetImpactor.cu:529-532).
    std::vector<SimDesc> sims;
    sims.push_back(make_sim(0, c->boarder1, c->boarder2, L));
    sims.push_back(make_sim(c->boarder2, c->boarder3, n, L));
    if ((rc = ensure_plan(ctx, sims, st))) return rc;
    return run_force(ctx, (const float4*)pos, (const float4*)vel, (float4*)force, st, g_timing);
}

int gi_move_collision(gi_float4* pos, gi_float4* vel, const gi_float4* force, int n,
                      const gi_move_collision_constants* c, gi_stream_t stream) {
    if (n < 0 || !c || (n > 0 && (!pos || !vel || !force)))
        return set_error(GI_E_INVALID_ARG, "gi_move_collision: invalid arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_move_collision, dim3(nblk(n, 256)), dim3(256), 0, as_stream(stream),
                       (float4*)pos, (float4*)vel, (const float4*)force, n, c->Dt, c->DtOverMassFe,
                       c->DtOverMassSi, c->NFe);
    GI_HIP(hipGetLastError());
    return 0;
}

int gi_move_damped_separate(gi_float4* pos, gi_float4* vel, const gi_float4* force, int n,
                            const gi_move_separate_constants* c, float damp_rate_body1,
                            float damp_rate_body2, gi_stream_t stream) {
    if (n < 0 || !c || (n > 0 && (!pos || !vel || !force)))
        return set_error(GI_E_INVALID_ARG, "gi_move_damped_separate: invalid arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_move_damped_separate, dim3(nblk(n, 256)), dim3(256), 0, as_stream(stream),
                       (float4*)pos, (float4*)vel, (const float4*)force, n, c->Dt, c->DtOverMassFe,
                       c->DtOverMassSi, c->boarder1, c->boarder2, c->boarder3, damp_rate_body1,
                       damp_rate_body2);
    GI_HIP(hipGetLastError());
    return 0;
}

int gi_set_kernel_timing(int enable) {
    g_timing = enable != 0;
    return 0;
}

int gi_kernel_timing(const char* kernel, long long* launches, double* total_ms) {
    if (kernel && std::strcmp(kernel, "gravity") != 0)
        return set_error(GI_E_INVALID_ARG, "unknown kernel '%s' (timed kernel: gravity)", kernel);
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    for (auto& tp : ctx->pending) {
        GI_HIP(hipEventSynchronize(tp.b));
        float ms = 0.0f;
        GI_HIP(hipEventElapsedTime(&ms, tp.a, tp.b));
        ctx->timed_ms += ms;
        ctx->timed_launches++;
        ctx->pool.push_back(tp);
    }
    ctx->pending.clear();
    if (launches) *launches = ctx->timed_launches;
    if (total_ms) *total_ms = ctx->timed_ms;
    return 0;
}

int gi_reset_kernel_timing(void) {
    DevCtx* ctx;
    int rc = current_ctx(&ctx);
    if (rc) return rc;
    long long l;
    double t;
    if ((rc = gi_kernel_timing(nullptr, &l, &t))) return rc;
    ctx->timed_launches = 0;
    ctx->timed_ms = 0.0;
    return 0;
}

}  // extern "C"

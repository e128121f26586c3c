// gi_serve.cpp -- the request loop: continuous batching of impact jobs on one
// GPU over one resident batched engine (gi_ensemble_*).
//
// Requests arrive over wall-clock time.  The loop owns one pool of particle
// records in HBM (pos / vel arrays of `capacity` records) and one ensemble over
// it, created once.  An arrived request is admitted as soon as a simulation
// slot is free, a contiguous range of the pool fits it (first fit) and its
// all-pairs scratch fits the HBM budget; requests are considered in arrival
// order, and one that does not fit lets later ones past it (first fit) until
// it has been passed over kMaxBypass times, after which nothing later is
// admitted before it.  Admission writes the request's initial state into its
// range and its row into the simulation table; completion removes the row.
// The arrays never move and the step graph is updated in place, so changes
// of the resident set cost a table update, not a rebuild.  Resident requests
// advance together a quantum of steps at a time; the host keeps one quantum
// queued behind the running one and synchronises only when a request
// finishes.  Because the batched engine cuts every simulation's sums the same
// way whatever its company, a request's result is bitwise the same alone and
// in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/giant_impact_host.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int idx = -1;  // request index
    int n = 0;
    int steps_left = 0;
    int offset = 0;  // first record in the pool
    gi_force_collision_constants fc;
    gi_move_collision_constants mc;
};

// Device bytes per pool record: pos and vel, the ensemble's force, and the
// engine's per-particle scratch (padded copy, sums, contact grid,
// bucket-ordered copies, corrections, potentials).
constexpr long long kBytesPerParticle = 32 + 16 + 16 + 16 + 4 * 4 + 32 + 24 + 8 + 16;

// The mutual all-pairs pass also keeps, per simulation, a j-side slab and
// i-side slots that grow as N^2: about 0.012 (N + 1024)^2 bytes each at rows
// of 512 and chunks of 1,024 (padded N).  Admission counts 0.03 (N + 1024)^2.
double quadratic_bytes(int n) {
    const double g = (double)n + 1024.0;
    return 0.03 * g * g;
}

// A request passed over this many times blocks the admission of later ones.
constexpr int kMaxBypass = 4;

// First-fit allocation of contiguous record ranges of the pool.
struct RangePool {
    std::vector<std::pair<int, int>> free;  // (start, length), by start
    explicit RangePool(int capacity) { free.push_back({0, capacity}); }
    int take(int n) {
        for (size_t k = 0; k < free.size(); ++k)
            if (free[k].second >= n) {
                const int at = free[k].first;
                free[k].first += n;
                free[k].second -= n;
                if (free[k].second == 0) free.erase(free.begin() + (long)k);
                return at;
            }
        return -1;
    }
    void give(int at, int n) {
        auto it = std::lower_bound(free.begin(), free.end(), std::make_pair(at, 0));
        it = free.insert(it, {at, n});
        size_t k = (size_t)(it - free.begin());
        if (k + 1 < free.size() && free[k].first + free[k].second == free[k + 1].first) {
            free[k].second += free[k + 1].second;
            free.erase(free.begin() + (long)k + 1);
        }
        if (k > 0 && free[k - 1].first + free[k - 1].second == free[k].first) {
            free[k - 1].second += free[k].second;
            free.erase(free.begin() + (long)k);
        }
    }
};

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

double norm3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

}  // namespace

extern "C" int gi_serve_run(const gi_serve_request *reqs, int n_requests, const gi_serve_options *opt,
                            gi_serve_result *results, gi_serve_summary *summary) {
    if (!reqs || n_requests <= 0 || !results) return GI_E_INVALID_ARG;
    gi_serve_options o{};
    if (opt) o = *opt;
    if (o.max_sims <= 0) o.max_sims = 256;
    if (o.quantum_steps <= 0) o.quantum_steps = 50;
    if (o.hbm_fraction <= 0.0 || o.hbm_fraction > 1.0) o.hbm_fraction = 0.5;

    long long biggest = 0;
    for (int k = 0; k < n_requests; ++k) {
        const gi_impact_config &c = reqs[k].impact;
        if (c.n_target <= 0 || c.n_impactor < 0 || reqs[k].steps < 0) return GI_E_INVALID_ARG;
        biggest = std::max<long long>(biggest, (long long)c.n_target + c.n_impactor);
    }
    size_t free_b = 0, total_b = 0;
    int rc = gi_mem_info(&free_b, &total_b);
    if (rc) return rc;
    long long capacity = o.max_particles;
    if (capacity <= 0) capacity = (long long)(0.5 * o.hbm_fraction * (double)free_b) / kBytesPerParticle;
    capacity = std::min<long long>(capacity, 1LL << 21);
    if (capacity < biggest) return GI_E_TOO_LARGE;
    // The N^2 scratch's budget: a share of what is free after the pool, taken
    // again at every admission (the engine's scratch only grows, so what it
    // already holds stays available to it).
    const double pool_bytes = (double)capacity * kBytesPerParticle;
    if (quadratic_bytes((int)biggest) > o.hbm_fraction * ((double)free_b - pool_bytes)) return GI_E_TOO_LARGE;

    void *P = nullptr, *V = nullptr;
    gi_stream_t st = nullptr;
    gi_ensemble *ens = nullptr;
    auto cleanup = [&](int code) {
        if (ens) gi_ensemble_destroy(ens);
        if (st) gi_stream_destroy(st);
        if (P) gi_free(P);
        if (V) gi_free(V);
        return code;
    };
    if ((rc = gi_malloc(&P, (size_t)capacity * 16)) || (rc = gi_malloc(&V, (size_t)capacity * 16))) return cleanup(rc);
    {
        // Records no request holds are never stepped, but the contact grid
        // bins every record of the table's span: keep them one per cell.
        std::vector<gi_float4> lattice((size_t)capacity);
        for (long long k = 0; k < capacity; ++k)
            lattice[(size_t)k] = gi_float4{2.0f * (float)(k % 128), 2.0f * (float)((k / 128) % 128),
                                           2.0f * (float)(k / 16384), 0.0f};
        if ((rc = gi_copy_h2d(P, lattice.data(), (size_t)capacity * 16))) return cleanup(rc);
        for (auto &r : lattice) r = gi_float4{0.0f, 0.0f, 0.0f, 0.0f};
        if ((rc = gi_copy_h2d(V, lattice.data(), (size_t)capacity * 16))) return cleanup(rc);
    }
    if ((rc = gi_stream_create(&st))) return cleanup(rc);
    if ((rc = gi_ensemble_create_capacity(&ens, (int)capacity, (gi_float4 *)P, (gi_float4 *)V, st)))
        return cleanup(rc);
    gi_float4 *pool_pos = (gi_float4 *)P, *pool_vel = (gi_float4 *)V;

    std::vector<int> order(n_requests);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::memset(results, 0, sizeof(gi_serve_result) * n_requests);

    RangePool pool((int)capacity);
    std::vector<Job> active;
    std::vector<int> waiting;  // request indices in arrival order, not yet admitted
    std::vector<int> bypassed(n_requests, 0);
    size_t next = 0;
    int completed = 0, rounds = 0, peak_sims = 0;
    long long resident = 0, peak_resident = 0;
    double resident_quad = 0.0, quad_high = 0.0;
    double total_steps = 0.0, total_pairs = 0.0;
    bool table_changed = false;
    const Clock::time_point t0 = Clock::now();

    auto push_table = [&]() -> int {
        std::sort(active.begin(), active.end(), [](const Job &a, const Job &b) { return a.offset < b.offset; });
        std::vector<int> off, ns;
        std::vector<gi_force_collision_constants> fcs;
        std::vector<gi_move_collision_constants> mcs;
        for (const Job &j : active) {
            off.push_back(j.offset);
            ns.push_back(j.n);
            fcs.push_back(j.fc);
            mcs.push_back(j.mc);
        }
        table_changed = false;
        return gi_ensemble_set_sims(ens, (int)active.size(), off.data(), ns.data(), fcs.data(), mcs.data());
    };

    while (completed < n_requests) {
        // ---- arrivals and admission (first fit, in arrival order) ----
        const double now = seconds_since(t0);
        while (next < order.size() && reqs[order[next]].arrival_s <= now) waiting.push_back(order[next++]);
        double free_now = -1.0;
        for (size_t w = 0; w < waiting.size();) {
            if ((int)active.size() >= o.max_sims) break;
            const int idx = waiting[w];
            const gi_serve_request &r = reqs[idx];
            const int n = r.impact.n_target + r.impact.n_impactor;
            bool fits = true;
            if (resident_quad + quadratic_bytes(n) > quad_high) {
                if (free_now < 0.0) {
                    size_t fb = 0, tb = 0;
                    if ((rc = gi_mem_info(&fb, &tb))) return cleanup(rc);
                    free_now = (double)fb;
                }
                fits = resident_quad + quadratic_bytes(n) <= quad_high + o.hbm_fraction * free_now;
            }
            const int at = fits ? pool.take(n) : -1;
            if (at < 0) {
                if (bypassed[idx] >= kMaxBypass) break;  // nothing later goes before it
                ++w;
                continue;
            }
            for (size_t k = 0; k < w; ++k) ++bypassed[waiting[k]];  // passed over by this admission
            Job j;
            j.idx = idx;
            j.n = n;
            j.steps_left = r.steps;
            j.offset = at;
            {
                std::vector<gi_float4> p0((size_t)n), v0((size_t)n);
                if ((rc = gi_make_impact(&r.impact, p0.data(), v0.data(), &j.fc, &j.mc, nullptr))) return cleanup(rc);
                if ((rc = gi_copy_h2d(pool_pos + at, p0.data(), (size_t)n * 16)) ||
                    (rc = gi_copy_h2d(pool_vel + at, v0.data(), (size_t)n * 16)))
                    return cleanup(rc);
            }
            double c9[9];
            if ((rc = gi_conserved_quantities(pool_pos + at, pool_vel + at, n, &j.fc, c9, nullptr))) return cleanup(rc);
            gi_serve_result &res = results[idx];
            res.n = n;
            res.arrival_s = r.arrival_s;
            res.admit_s = seconds_since(t0);
            res.energy0 = c9[2];
            std::memcpy(res.angmom0, c9 + 6, sizeof res.angmom0);
            resident += n;
            resident_quad += quadratic_bytes(n);
            quad_high = std::max(quad_high, resident_quad);
            active.push_back(j);
            waiting.erase(waiting.begin() + (long)w);
            table_changed = true;
            free_now = -1.0;
        }
        if (table_changed) {
            if ((rc = push_table())) return cleanup(rc);
            peak_resident = std::max(peak_resident, resident);
            peak_sims = std::max(peak_sims, (int)active.size());
        }
        if (active.empty()) {
            if (next < order.size()) {
                const double wait = reqs[order[next]].arrival_s - seconds_since(t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            }
            continue;
        }
        // ---- one quantum of batched steps, queued behind the running one ----
        int q = o.quantum_steps;
        for (const Job &j : active) q = std::min(q, std::max(j.steps_left, 0));
        if (q > 0) {
            if ((rc = gi_ensemble_steps(ens, q))) return cleanup(rc);
            ++rounds;
            for (Job &j : active) {
                j.steps_left -= q;
                total_steps += q;
                total_pairs += (double)q * j.n * (double)j.n;
            }
        }
        bool any_done = false;
        for (const Job &j : active) any_done |= j.steps_left <= 0;
        if (!any_done) {
            if ((rc = gi_ensemble_wait(ens, 1))) return cleanup(rc);
            continue;
        }
        // ---- completion: the finished requests' state is final once the queue drains ----
        if ((rc = gi_ensemble_wait(ens, 0))) return cleanup(rc);
        const double tdone = seconds_since(t0);
        std::vector<Job> keep;
        for (Job &j : active) {
            if (j.steps_left > 0) {
                keep.push_back(j);
                continue;
            }
            gi_serve_result &res = results[j.idx];
            const gi_serve_request &r = reqs[j.idx];
            double c9[9];
            if ((rc = gi_conserved_quantities(pool_pos + j.offset, pool_vel + j.offset, j.n, &j.fc, c9, nullptr)))
                return cleanup(rc);
            res.steps = r.steps;
            res.energy1 = c9[2];
            std::memcpy(res.angmom1, c9 + 6, sizeof res.angmom1);
            res.finish_s = tdone;
            res.time_to_result_s = tdone - res.arrival_s;
            res.energy_rel_change = std::fabs(res.energy1 - res.energy0) / std::max(std::fabs(res.energy0), 1e-300);
            const double d[3] = {res.angmom1[0] - res.angmom0[0], res.angmom1[1] - res.angmom0[1],
                                 res.angmom1[2] - res.angmom0[2]};
            res.angmom_rel_change = norm3(d) / std::max(norm3(res.angmom0), 1e-300);
            const size_t bytes = (size_t)j.n * 16;
            if (r.final_pos && (rc = gi_copy_d2h(r.final_pos, pool_pos + j.offset, bytes))) return cleanup(rc);
            if (r.final_vel && (rc = gi_copy_d2h(r.final_vel, pool_vel + j.offset, bytes))) return cleanup(rc);
            pool.give(j.offset, j.n);
            resident -= j.n;
            resident_quad -= quadratic_bytes(j.n);
            ++completed;
        }
        active = std::move(keep);
        table_changed = true;
    }
    const double wall = seconds_since(t0);
    if (summary) {
        std::memset(summary, 0, sizeof *summary);
        std::vector<double> ttr;
        for (int k = 0; k < n_requests; ++k) {
            ttr.push_back(results[k].time_to_result_s);
            summary->max_energy_rel_change = std::max(summary->max_energy_rel_change, results[k].energy_rel_change);
            summary->max_angmom_rel_change = std::max(summary->max_angmom_rel_change, results[k].angmom_rel_change);
        }
        std::sort(ttr.begin(), ttr.end());
        auto pct = [&](double p) {
            const double x = p * (ttr.size() - 1);
            const size_t a = (size_t)std::floor(x), b = std::min(a + 1, ttr.size() - 1);
            return ttr[a] + (ttr[b] - ttr[a]) * (x - a);
        };
        summary->completed = completed;
        summary->wall_s = wall;
        summary->total_sim_steps = total_steps;
        summary->sim_steps_per_s = total_steps / wall;
        summary->pair_interactions_per_s = total_pairs / wall;
        summary->p50_time_to_result_s = pct(0.50);
        summary->p99_time_to_result_s = pct(0.99);
        summary->max_time_to_result_s = ttr.back();
        summary->capacity_particles = capacity;
        summary->peak_resident_particles = peak_resident;
        summary->peak_resident_sims = peak_sims;
        summary->rounds = rounds;
    }
    return cleanup(0);
}

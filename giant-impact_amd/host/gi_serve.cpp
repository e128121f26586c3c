// gi_serve.cpp -- the request loop: continuous batching of impact jobs on one
// GPU over the batched engine (gi_ensemble_*).
//
// Requests arrive over wall-clock time.  A request is admitted (in arrival
// order) as soon as it has arrived, a simulation slot is free and its
// particles fit in the resident capacity, which is sized from the device's
// free HBM.  Resident requests live back to back in one packed pos/vel array
// in HBM and advance together, a quantum of steps per round, in the batched
// engine's launches.  A finished request leaves at the end of its round; the
// packed array is rebuilt by device-to-device copies whenever the resident
// set changes, so a request's state never leaves the device between its
// admission and its completion.  Because the batched engine cuts every
// simulation's sums the same way whatever its company, a request's result is
// bitwise the same alone and in any mix.
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
    int offset = 0;  // in the current packed layout
    bool fresh = true;
    gi_force_collision_constants fc;
    gi_move_collision_constants mc;
    std::vector<gi_float4> pos0, vel0;
};

// Device bytes per resident particle: two packed pos/vel copies (double
// buffer), the ensemble's force array, and the engine's scratch (partial sums
// of at most a few chunks, contact grid, bucket-ordered copies, potentials).
constexpr long long kBytesPerParticle = 2 * 32 + 16 + 4 * 16 + 16 + 32 + 8 + 16;

// The mutual all-pairs pass also keeps, per simulation, a j-side slab and
// i-side slots that grow as N^2: about 0.012 (N + 1024)^2 bytes each at rows
// of 512 and chunks of 512 (padded N).  Admission counts 0.03 (N + 1024)^2.
double quadratic_bytes(int n) {
    const double g = (double)n + 1024.0;
    return 0.03 * g * g;
}

struct DevArrays {
    void *pos = nullptr, *vel = nullptr;
};

void free_arrays(DevArrays &a) {
    if (a.pos) gi_free(a.pos);
    if (a.vel) gi_free(a.vel);
    a.pos = a.vel = nullptr;
}

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace

extern "C" int gi_serve_run(const gi_serve_request *reqs, int n_requests, const gi_serve_options *opt,
                            gi_serve_result *results, gi_serve_summary *summary) {
    if (!reqs || n_requests <= 0 || !results) return GI_E_INVALID_ARG;
    gi_serve_options o{};
    if (opt) o = *opt;
    if (o.max_sims <= 0) o.max_sims = 256;
    if (o.quantum_steps <= 0) o.quantum_steps = 100;
    if (o.hbm_fraction <= 0.0 || o.hbm_fraction > 1.0) o.hbm_fraction = 0.5;

    long long biggest = 0;
    for (int k = 0; k < n_requests; ++k) {
        const gi_impact_config &c = reqs[k].impact;
        if (c.n_target <= 0 || c.n_impactor < 0 || reqs[k].steps < 0) return GI_E_INVALID_ARG;
        biggest = std::max<long long>(biggest, (long long)c.n_target + c.n_impactor);
    }
    size_t free_b = 0, total_b = 0;
    {
        int rc = gi_mem_info(&free_b, &total_b);
        if (rc) return rc;
    }
    const double budget = o.hbm_fraction * (double)free_b;
    long long capacity = o.max_particles;
    if (capacity <= 0) capacity = (long long)(0.5 * budget) / kBytesPerParticle;
    capacity = std::min<long long>(capacity, 1LL << 21);
    // what is left of the budget for the per-simulation N^2 scratch
    const double quad_budget = budget - (double)capacity * kBytesPerParticle;
    if (capacity < biggest || quadratic_bytes((int)biggest) > quad_budget) return GI_E_TOO_LARGE;

    DevArrays buf[2];
    int rc = 0;
    for (auto &b : buf) {
        if ((rc = gi_malloc(&b.pos, (size_t)capacity * 16)) || (rc = gi_malloc(&b.vel, (size_t)capacity * 16))) {
            free_arrays(buf[0]);
            free_arrays(buf[1]);
            return rc;
        }
    }
    int cur = 0;
    gi_ensemble *ens = nullptr;

    std::vector<int> order(n_requests);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::memset(results, 0, sizeof(gi_serve_result) * n_requests);

    std::vector<Job> active;
    size_t next = 0;
    int completed = 0, rounds = 0, peak_sims = 0;
    long long resident = 0, peak_resident = 0;
    double resident_quad = 0.0;
    double total_steps = 0.0, total_pairs = 0.0;
    const Clock::time_point t0 = Clock::now();

    auto fail = [&](int code) {
        if (ens) gi_ensemble_destroy(ens);
        free_arrays(buf[0]);
        free_arrays(buf[1]);
        return code;
    };

    // Pack `active` into the other buffer, recreate the ensemble, and record
    // the starting conserved quantities of the freshly admitted jobs.
    auto rebuild = [&]() -> int {
        const int dst = 1 - cur;
        int off = 0;
        for (Job &j : active) {
            const size_t bytes = (size_t)j.n * 16;
            char *dp = (char *)buf[dst].pos + (size_t)off * 16;
            char *dv = (char *)buf[dst].vel + (size_t)off * 16;
            int r;
            if (j.fresh) {
                if ((r = gi_copy_h2d(dp, j.pos0.data(), bytes)) || (r = gi_copy_h2d(dv, j.vel0.data(), bytes)))
                    return r;
            } else {
                const char *sp = (const char *)buf[cur].pos + (size_t)j.offset * 16;
                const char *sv = (const char *)buf[cur].vel + (size_t)j.offset * 16;
                if ((r = gi_copy_d2d(dp, sp, bytes, nullptr)) || (r = gi_copy_d2d(dv, sv, bytes, nullptr))) return r;
            }
            j.offset = off;
            off += j.n;
        }
        if (ens) {
            gi_ensemble_destroy(ens);
            ens = nullptr;
        }
        cur = dst;
        if (active.empty()) return gi_device_sync();
        std::vector<int> ns;
        std::vector<gi_force_collision_constants> fcs;
        std::vector<gi_move_collision_constants> mcs;
        for (const Job &j : active) {
            ns.push_back(j.n);
            fcs.push_back(j.fc);
            mcs.push_back(j.mc);
        }
        int r = gi_ensemble_create(&ens, (int)active.size(), ns.data(), fcs.data(), mcs.data(),
                                   (gi_float4 *)buf[cur].pos, (gi_float4 *)buf[cur].vel, nullptr);
        if (r) return r;
        bool any_fresh = false;
        for (const Job &j : active) any_fresh |= j.fresh;
        if (any_fresh) {
            std::vector<double> c(9 * active.size());
            if ((r = gi_ensemble_conserved(ens, c.data()))) return r;
            for (size_t s = 0; s < active.size(); ++s) {
                Job &j = active[s];
                if (!j.fresh) continue;
                gi_serve_result &res = results[j.idx];
                res.energy0 = c[9 * s + 2];
                std::memcpy(res.angmom0, &c[9 * s + 6], sizeof res.angmom0);
                j.fresh = false;
                j.pos0.clear();
                j.pos0.shrink_to_fit();
                j.vel0.clear();
                j.vel0.shrink_to_fit();
            }
        }
        return gi_device_sync();
    };

    while (completed < n_requests) {
        // ---- admission ----
        bool changed = false;
        double now = seconds_since(t0);
        while (next < order.size() && reqs[order[next]].arrival_s <= now) {
            const gi_serve_request &r = reqs[order[next]];
            const int n = r.impact.n_target + r.impact.n_impactor;
            if ((int)active.size() >= o.max_sims || resident + n > capacity ||
                resident_quad + quadratic_bytes(n) > quad_budget)
                break;  // wait for room, FIFO
            Job j;
            j.idx = order[next];
            j.n = n;
            j.steps_left = r.steps;
            j.pos0.resize(n);
            j.vel0.resize(n);
            if ((rc = gi_make_impact(&r.impact, j.pos0.data(), j.vel0.data(), &j.fc, &j.mc, nullptr)))
                return fail(rc);
            results[j.idx].n = n;
            results[j.idx].arrival_s = r.arrival_s;
            results[j.idx].admit_s = now;
            resident += n;
            resident_quad += quadratic_bytes(n);
            active.push_back(std::move(j));
            ++next;
            changed = true;
        }
        if (changed) {
            if ((rc = rebuild())) return fail(rc);
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
        // ---- one round of batched steps ----
        int q = o.quantum_steps;
        for (const Job &j : active) q = std::min(q, std::max(j.steps_left, 0));
        if (q > 0) {
            if ((rc = gi_ensemble_steps(ens, q)) || (rc = gi_device_sync())) return fail(rc);
            ++rounds;
            for (Job &j : active) {
                j.steps_left -= q;
                total_steps += q;
                total_pairs += (double)q * j.n * (double)j.n;
            }
        }
        // ---- completion ----
        bool any_done = false;
        for (const Job &j : active) any_done |= j.steps_left <= 0;
        if (!any_done) continue;
        std::vector<double> c(9 * active.size());
        if ((rc = gi_ensemble_conserved(ens, c.data()))) return fail(rc);
        const double tdone = seconds_since(t0);
        std::vector<Job> keep;
        for (size_t s = 0; s < active.size(); ++s) {
            Job &j = active[s];
            if (j.steps_left > 0) {
                keep.push_back(std::move(j));
                continue;
            }
            gi_serve_result &res = results[j.idx];
            const gi_serve_request &r = reqs[j.idx];
            res.steps = r.steps;
            res.energy1 = c[9 * s + 2];
            std::memcpy(res.angmom1, &c[9 * s + 6], sizeof res.angmom1);
            res.finish_s = tdone;
            res.time_to_result_s = tdone - res.arrival_s;
            res.energy_rel_change = std::fabs(res.energy1 - res.energy0) / std::max(std::fabs(res.energy0), 1e-300);
            const double dl = std::sqrt(std::pow(res.angmom1[0] - res.angmom0[0], 2) +
                                        std::pow(res.angmom1[1] - res.angmom0[1], 2) +
                                        std::pow(res.angmom1[2] - res.angmom0[2], 2));
            const double l0 = std::sqrt(res.angmom0[0] * res.angmom0[0] + res.angmom0[1] * res.angmom0[1] +
                                        res.angmom0[2] * res.angmom0[2]);
            res.angmom_rel_change = dl / std::max(l0, 1e-300);
            const size_t bytes = (size_t)j.n * 16;
            if (r.final_pos && (rc = gi_copy_d2h(r.final_pos, (char *)buf[cur].pos + (size_t)j.offset * 16, bytes)))
                return fail(rc);
            if (r.final_vel && (rc = gi_copy_d2h(r.final_vel, (char *)buf[cur].vel + (size_t)j.offset * 16, bytes)))
                return fail(rc);
            resident -= j.n;
            resident_quad -= quadratic_bytes(j.n);
            ++completed;
        }
        active = std::move(keep);
        if ((rc = rebuild())) return fail(rc);
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
    if (ens) gi_ensemble_destroy(ens);
    free_arrays(buf[0]);
    free_arrays(buf[1]);
    return 0;
}

// gi_kernels.hip -- HIP kernels of the Giant-Impact step loop for gfx950.
//
// The reference evaluates the Greenspan Fe/Si force in one all-pairs kernel
// whose inner loop branches on material, distance and receding
// (getForcesCollisionSingleGPU, Code/ImpactCode/impactAddTime.cu:245-377).
// Here the same force is produced by three passes:
//
//   1. all pairs, branch free: sum_j (G mi mj) dp / r^3 with a hardware rsq,
//      in one of two forms chosen per simulation by the tuning table:
//      k_gravity (one-sided: i-particles in registers, R per lane,
//      j-particles through wave-uniform scalar loads, Kahan-folded float
//      tiles, one j-chunk per wavefront, partial sums to a slot buffer), or
//      the mutual pass (k_mutual_pad, k_mutual, k_mutual_sum: each pair
//      evaluated once for both particles, per-particle compensated sums).
//   2. the contact grid -- a spatial hash of cells of size 1 (the element
//      diameter) built by counting sort (k_grid_*).
//   3. k_finish   -- per particle: its all-pairs sum plus, for every pair
//      with r < 1 found through the grid, the reference's exact piecewise
//      contact force minus what pass 1 added for that pair.  When every
//      simulation takes the mutual pass, the same correction is computed by
//      k_contact_corr beside pass 1 (it needs only the grid) and added by
//      k_mutual_sum.
//
// Only pairs with r < 1 can differ from gravity (impactAddTime.cu:300,320,349),
// so the sum equals the reference's force up to float rounding, and the
// contact branch decisions are made with the reference's own arithmetic.
#include "gi_device.h"

#include <type_traits>

namespace gi {

// ---------------------------------------------------------------------------
// Pass 1: all-pairs gravity (and the potential variant for the energy check)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float v) {
    float y = v - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// R: i-particles per lane.  T: j per float tile between Kahan folds.
// POT: accumulate sum_j 1/r (self excluded) instead of sum_j dp/r^3.
// V: how r^-3 is formed (0: rsq cubed, 1: rsq times rcp).
template <int R, int T, bool POT, int V>
__global__ __launch_bounds__(256) void k_gravity(const float4* __restrict__ pos,
                                                 const GravTask* __restrict__ tasks, int ntasks,
                                                 const SimDesc* __restrict__ sims,
                                                 float4* __restrict__ partial) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (w >= ntasks) return;
    const GravTask t = tasks[w];
    const int start = sims[t.sim].start;
    const int fe_end = sims[t.sim].fe_end;
    const int end = sims[t.sim].end;
    const int pbase = sims[t.sim].partial_base;
    const float gFeFe = sims[t.sim].law.GMassFeFe;
    const float gFeSi = sims[t.sim].law.GMassFeSi;

    float xi[R], yi[R], zi[R];
    float ax[R], ay[R], az[R], cx[R], cy[R], cz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int i = t.i0 + r * 64 + lane;
        int ic = i < end ? i : end - 1;
        float4 p = pos[ic];
        xi[r] = p.x;
        yi[r] = p.y;
        zi[r] = p.z;
        ax[r] = ay[r] = az[r] = 0.0f;
        cx[r] = cy[r] = cz[r] = 0.0f;
    }

    for (int jb = t.j0; jb < t.j1; jb += T) {
        const int je = min(jb + T, t.j1);
        float tx[R], ty[R], tz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) tx[r] = ty[r] = tz[r] = 0.0f;
#pragma unroll 4
        for (int j = jb; j < je; ++j) {
            const float4 pj = pos[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float dx = pj.x - xi[r];
                float dy = pj.y - yi[r];
                float dz = pj.z - zi[r];
                if constexpr (!POT) {
                    float s = inv_r3<V>(fast_r2(dx, dy, dz));
                    tx[r] = __builtin_fmaf(s, dx, tx[r]);
                    ty[r] = __builtin_fmaf(s, dy, ty[r]);
                    tz[r] = __builtin_fmaf(s, dz, tz[r]);
                } else {
                    float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    float ri = r2 > 0.0f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
                    tx[r] += ri;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            kahan_add(ax[r], cx[r], tx[r]);
            if constexpr (!POT) {
                kahan_add(ay[r], cy[r], ty[r]);
                kahan_add(az[r], cz[r], tz[r]);
            }
        }
    }

    const int n = end - start;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int i = t.i0 + r * 64 + lane;
        if (i < end) {
            const bool ife = i < fe_end;
            const float g = t.jfe ? (ife ? gFeFe : gFeSi) : (ife ? gFeSi : 1.0f);
            float4 o;
            o.x = (ax[r] - cx[r]) * g;
            o.y = (ay[r] - cy[r]) * g;
            o.z = (az[r] - cz[r]) * g;
            o.w = 0.0f;
            partial[pbase + t.chunk * n + (i - start)] = o;
        }
    }
}

// sim_of: for each particle its simulation index (particle ranges are given
// by sims[]); computed by a binary search over the sorted starts.
__device__ __forceinline__ int find_sim(const SimDesc* __restrict__ sims, int nsims, int i) {
    int lo = 0, hi = nsims - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sims[mid].start <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Pass 1, mutual form: each pair evaluated once, applied to both particles
// ---------------------------------------------------------------------------
//
// The pair term (G mi mj) dp / r^3 is odd in dp and r^2 is even, so the term
// j receives from i is exactly the negation of the one i receives from j.
// A wavefront holds a row of 64*R particles in registers (R per lane) and
// streams 64-particle blocks j of the same simulation that lie beyond the
// row.  Each lane holds the force accumulator of one j of the block; after
// every step the accumulators rotate by one lane (DPP wave_ror), so in 64
// steps every lane has met every j of the block and each j has collected the
// terms of all 64*R row particles.  The positions stay put in wave-private
// LDS and each step reads the one its lane's accumulator now belongs to.
// Per pair: 15 VALU ops including one rsq plus 3/R lane rotations, for two
// interactions, against 12 ops per interaction one-sided.  Measured on
// gfx950: rotating the positions with DPP as well is 6% slower (the
// rotation sits on every step's dependency chain); accumulating the j-side
// terms with LDS float adds instead of rotating is about 5x slower.
//
// Simulations are copied into a padded layout (iron then silicate, each
// rounded up to whole rows with ghosts far away) so that every row and every
// 64-block is of one material and the gravity product is one factor per task.
// The row against itself ("diag" task) takes each of its pairs once too (a
// one-sided form at R = 8 made the 64-simulation batched step 1.5% slower).
// Outputs: per task an i-side slot (an (x, y, z) record per place of the
// row) and per (row, j-block) a j-side slab entry (an (x, y, z) record per
// j), so that k_mutual_sum reads 768 contiguous bytes per wavefront and
// entry; k_mutual_sum adds them in
// compensated float sums in a fixed order.  Both sides accumulate in plain
// float: a j-side entry sums
// 64*R terms, an i-side slot at most one chunk (<= 4096 j, the planner's
// cap), so the rounding error stays at a few 1e-6 of the sum of the terms'
// magnitudes; leaving out tile sums and compensation keeps the register
// count low enough for five (k_mutual) or six (k_mutual_grp) wavefronts per
// SIMD at R = 8.

__device__ __forceinline__ float rotate_lanes(float v) {
    // wave_ror:1 -- lane l reads lane l-1 (mod 64).
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x13C, 0xF, 0xF, false));
}

// Orders this wavefront's LDS accesses before and after (its lanes exchange
// data through LDS).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Offset (in entries, per plane) of row `row`'s part of a simulation's
// j-side slab: rows before it hold gn - (r+1)*B entries each.
__host__ __device__ __forceinline__ long long jside_row_offset(int row, int gn, int B) {
    return (long long)row * gn - (long long)B * row * (row + 1) / 2;
}

constexpr int kMutualWaves = 4;  // wavefronts per block of k_mutual

// One wavefront's task t of simulation sd (P: its padded positions, q: the
// wavefront's 128 float4 of LDS): a row against itself (t.diag) or against
// j beyond the row, whose j-side sums go to J (one (x, y, z) record per
// entry, entry j - (t.row + 1) * B); the row's i-side sums go to slot t.slot
// (one (x, y, z) record per place of the row).  POT: x only.
template <int R, bool POT>
__device__ __forceinline__ void mutual_wave_task(const MutualTask& t, const SimDesc& sd,
                                                 const float4* __restrict__ P, float4* __restrict__ q,
                                                 float* __restrict__ iside, float* __restrict__ J, int lane) {
    constexpr int B = 64 * R;
    const int ib = t.row * B;
    const bool ife = ib < sd.gfe;
    const float g = t.jfe ? (ife ? sd.law.GMassFeFe : sd.law.GMassFeSi) : (ife ? sd.law.GMassFeSi : 1.0f);

    float xi[R], yi[R], zi[R];
    float ax[R], ay[R], az[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float4 p = P[ib + r * 64 + lane];
        xi[r] = p.x;
        yi[r] = p.y;
        zi[r] = p.z;
        ax[r] = ay[r] = az[r] = 0.0f;
    }

    // The block's positions, twice, in wave-private LDS: q[m] holds
    // j_((127 - m) mod 64), so step k reads q[63 - lane + k] =
    // j_((lane - k) mod 64) -- the j whose accumulator lane `lane` holds
    // after k rotations -- at the immediate offset 16 k.
    if (t.diag) {
        // The row against itself, each pair once: its 64-blocks m holding a
        // real particle against the row's sub-blocks r < m (both sides, as
        // below), and against sub-block m itself over rotations 1..32 (lanes
        // 0..31 only at 32): every pair of the block exactly once, the self
        // pair never.  The j-side sums return to their lanes after 64
        // rotations and belong to sub-block m's particles.  A task takes the
        // blocks m of [j0, j1) (the planner may cut a row's diagonal in
        // pieces of about equal work, each with its own slot).
        const int m0 = (t.j0 - ib) >> 6, m1 = (t.j1 - ib) >> 6;
        for (int m = m0; m < m1; ++m) {
            const float4 own = P[ib + m * 64 + lane];
            wave_lds_sync();  // the previous block's reads are done
            q[63 - lane] = own;
            q[127 - lane] = own;
            wave_lds_sync();
            float xj = own.x, yj = own.y, zj = own.z;
            const float4* __restrict__ qb = q + (63 - lane);
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll 4
            for (int k = 0; k < 64; ++k) {
                const float4 pn = qb[k + 1 < 64 ? k + 1 : 63];  // next step's j
                const bool self_on = (k >= 1 && k < 32) || (k == 32 && lane < 32);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r > m || (r == m && !self_on)) continue;
                    const float dx = xj - xi[r];
                    const float dy = yj - yi[r];
                    const float dz = zj - zi[r];
                    if constexpr (!POT) {
                        const float s = fast_inv_r3(fast_r2(dx, dy, dz));
                        ax[r] = __builtin_fmaf(s, dx, ax[r]);
                        ay[r] = __builtin_fmaf(s, dy, ay[r]);
                        az[r] = __builtin_fmaf(s, dz, az[r]);
                        fx = __builtin_fmaf(s, dx, fx);
                        fy = __builtin_fmaf(s, dy, fy);
                        fz = __builtin_fmaf(s, dz, fz);
                    } else {
                        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        const float ri = (r2 > 0.0f && r2 < kGhostR2) ? __builtin_amdgcn_rsqf(r2) : 0.0f;
                        ax[r] += ri;
                        fx += ri;
                    }
                }
                xj = pn.x;
                yj = pn.y;
                zj = pn.z;
                fx = rotate_lanes(fx);
                if constexpr (!POT) {
                    fy = rotate_lanes(fy);
                    fz = rotate_lanes(fz);
                }
            }
            // The force on j is the negated sum; the potential is symmetric.
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r == m) {
                    if constexpr (!POT) {
                        ax[r] -= fx;
                        ay[r] -= fy;
                        az[r] -= fz;
                    } else {
                        ax[r] += fx;
                    }
                }
        }
    } else {
        const int first = (t.row + 1) * B;  // first j of the row's slab
        float4 nx = P[t.j0 + lane];
        for (int J0 = t.j0; J0 < t.j1; J0 += 64) {
            float xj = nx.x, yj = nx.y, zj = nx.z;
            wave_lds_sync();  // the previous block's reads are done
            q[63 - lane] = nx;
            q[127 - lane] = nx;
            wave_lds_sync();
            // next block's positions, in flight during this block
            nx = P[(J0 + 64 < t.j1 ? J0 + 64 : J0) + lane];
            const float4* __restrict__ qb = q + (63 - lane);
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
            // unroll 16: 1-2% faster than 8 at 14k-98k particles (measured)
#pragma unroll 16
            for (int k = 0; k < 64; ++k) {
                const float4 pn = qb[k + 1 < 64 ? k + 1 : 63];  // next step's j
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float dx = xj - xi[r];
                    const float dy = yj - yi[r];
                    const float dz = zj - zi[r];
                    if constexpr (!POT) {
                        const float s = fast_inv_r3(fast_r2(dx, dy, dz));
                        ax[r] = __builtin_fmaf(s, dx, ax[r]);
                        ay[r] = __builtin_fmaf(s, dy, ay[r]);
                        az[r] = __builtin_fmaf(s, dz, az[r]);
                        fx = __builtin_fmaf(s, dx, fx);
                        fy = __builtin_fmaf(s, dy, fy);
                        fz = __builtin_fmaf(s, dz, fz);
                    } else {
                        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        const float ri = (r2 > 0.0f && r2 < kGhostR2) ? __builtin_amdgcn_rsqf(r2) : 0.0f;
                        ax[r] += ri;
                        fx += ri;
                    }
                }
                xj = pn.x;
                yj = pn.y;
                zj = pn.z;
                fx = rotate_lanes(fx);
                if constexpr (!POT) {
                    fy = rotate_lanes(fy);
                    fz = rotate_lanes(fz);
                }
            }
            // After 64 rotations lane l holds j = J0 + l again.  The force on
            // j is the negated sum (dp points from j to the row particle);
            // the potential is symmetric.
            const long long o = (long long)(J0 - first) + lane;
            if constexpr (!POT) {
                J[3 * o] = -fx * g;
                J[3 * o + 1] = -fy * g;
                J[3 * o + 2] = -fz * g;
            } else {
                J[3 * o] = fx * g;
            }
        }
    }

    float* __restrict__ I = iside + (size_t)t.slot * (3 * B);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        I[3 * (r * 64 + lane)] = ax[r] * g;
        if constexpr (!POT) {
            I[3 * (r * 64 + lane) + 1] = ay[r] * g;
            I[3 * (r * 64 + lane) + 2] = az[r] * g;
        }
    }
}

template <int R, bool POT>
__global__ __launch_bounds__(64 * kMutualWaves) void k_mutual(const float4* __restrict__ gpos,
                                                              const MutualTask* __restrict__ tasks,
                                                              int ntasks, const SimDesc* __restrict__ sims,
                                                              float* __restrict__ iside,
                                                              float* __restrict__ jside) {
    constexpr int B = 64 * R;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * kMutualWaves + wv);
    if (w >= ntasks) return;
    const MutualTask t = tasks[w];
    const SimDesc& sd = sims[t.sim];
    __shared__ float4 lq[kMutualWaves][128];
    float* __restrict__ J = jside + sd.jbase + 3 * jside_row_offset(t.row, sd.gn, B);
    mutual_wave_task<R, POT>(t, sd, gpos + sd.gbase, lq[wv], iside, J, lane);
}

// ---------------------------------------------------------------------------
// Pass 1, mutual form with row groups: the j-side sums of kMutualWaves rows
// added in LDS before they are written
// ---------------------------------------------------------------------------
//
// In k_mutual every (row, 64-block) pair leaves a j-side slab entry, so the
// slab holds ~N^2 / (128 R) floats per component -- more than the whole pass
// writes otherwise.  Here a workgroup's W = kMutualWaves wavefronts hold W
// consecutive rows (a group) and stream the same j-blocks beyond the group:
// the block's positions sit once in LDS for all of them, and after each block
// the W j-side sums (each times its own row's gravity product) are added in a
// fixed order and written as one entry per (group, block), a W-th of the
// entries.  Pairs inside a group (a row against itself and against the later
// rows of its group) are single-wavefront row tasks as in k_mutual, packed W
// to a workgroup, whose j-side sums go to a small per-row slab.
constexpr int kGroupRows = kMutualWaves;

template <int R, bool POT>
__device__ __forceinline__ void mutual_group_task(const GroupTask& gt, const SimDesc& sd,
                                                  const RowDesc* __restrict__ rows, const float4* __restrict__ P,
                                                  float* __restrict__ iside, float* __restrict__ jside, int wv,
                                                  int lane, float4 (*__restrict__ sq)[128],
                                                  float (*__restrict__ red)[3][kGroupRows][64]) {
    constexpr int B = 64 * R;
    constexpr int W = kGroupRows;
    const int nrows = sd.gn / B;
    const int row = gt.group * W + wv;
    const bool live = row < nrows;  // wave-uniform
    const int ib = row * B;
    const bool ife = ib < sd.gfe;
    const float g = !live ? 0.0f
                          : gt.jfe ? (ife ? sd.law.GMassFeFe : sd.law.GMassFeSi) : (ife ? sd.law.GMassFeSi : 1.0f);
    float xi[R], yi[R], zi[R];
    float ax[R], ay[R], az[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float4 p = live ? P[ib + r * 64 + lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        xi[r] = p.x;
        yi[r] = p.y;
        zi[r] = p.z;
        ax[r] = ay[r] = az[r] = 0.0f;
    }
    const int gend = min((gt.group + 1) * W, nrows) * B;  // first j of the group's slab
    float* __restrict__ J = jside + sd.jbase + 3 * jside_row_offset(gt.group, sd.gn, W * B);
    // Each wavefront stages a quarter of every block: lanes 0..15 of wave wv
    // own j = 16 wv + lane, stored twice as k_mutual lays a block out.
    const int jl = wv * 16 + (lane & 15);
    const bool stager = lane < 16;
    float4 nx = stager ? P[gt.j0 + jl] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (stager) {
        sq[0][63 - jl] = nx;
        sq[0][127 - jl] = nx;
    }
    __syncthreads();
    int buf = 0;
    for (int J0 = gt.j0; J0 < gt.j1; J0 += 64) {
        const bool more = J0 + 64 < gt.j1;
        if (more && stager) nx = P[J0 + 64 + jl];  // the next block's quarter, in flight
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        if (live) {
            const float4* __restrict__ qb = sq[buf] + (63 - lane);
            const float4 c0 = qb[0];
            float xj = c0.x, yj = c0.y, zj = c0.z;
            // unroll 2: 1.7% faster than 8 at N = 262,144 and 2% on the
            // batched step; 4 and 16 in between, 1 slower (measured)
#pragma unroll 2
            for (int k = 0; k < 64; ++k) {
                const float4 pn = qb[k + 1 < 64 ? k + 1 : 63];  // next step's j
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float dx = xj - xi[r];
                    const float dy = yj - yi[r];
                    const float dz = zj - zi[r];
                    if constexpr (!POT) {
                        const float s = fast_inv_r3(fast_r2(dx, dy, dz));
                        ax[r] = __builtin_fmaf(s, dx, ax[r]);
                        ay[r] = __builtin_fmaf(s, dy, ay[r]);
                        az[r] = __builtin_fmaf(s, dz, az[r]);
                        fx = __builtin_fmaf(s, dx, fx);
                        fy = __builtin_fmaf(s, dy, fy);
                        fz = __builtin_fmaf(s, dz, fz);
                    } else {
                        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        const float ri = (r2 > 0.0f && r2 < kGhostR2) ? __builtin_amdgcn_rsqf(r2) : 0.0f;
                        ax[r] += ri;
                        fx += ri;
                    }
                }
                xj = pn.x;
                yj = pn.y;
                zj = pn.z;
                fx = rotate_lanes(fx);
                if constexpr (!POT) {
                    fy = rotate_lanes(fy);
                    fz = rotate_lanes(fz);
                }
            }
        }
        // This wavefront's j-side sums, times its own gravity product (the
        // force on j is the negated sum; the potential is symmetric).
        if constexpr (!POT) {
            red[buf][0][wv][lane] = -fx * g;
            red[buf][1][wv][lane] = -fy * g;
            red[buf][2][wv][lane] = -fz * g;
        } else {
            red[buf][0][wv][lane] = fx * g;
        }
        if (more && stager) {
            sq[buf ^ 1][63 - jl] = nx;
            sq[buf ^ 1][127 - jl] = nx;
        }
        __syncthreads();
        // wavefront c adds component c over the group's rows in row order
        if (wv < (POT ? 1 : 3)) {
            float a = red[buf][wv][0][lane];
#pragma unroll
            for (int w = 1; w < W; ++w) a += red[buf][wv][w][lane];
            J[3 * ((J0 - gend) + lane) + wv] = a;
        }
        buf ^= 1;
    }
    if (live) {
        float* __restrict__ I = iside + (size_t)(rows[sd.row0 + row].slot0 + gt.piece) * (3 * B);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            I[3 * (r * 64 + lane)] = ax[r] * g;
            if constexpr (!POT) {
                I[3 * (r * 64 + lane) + 1] = ay[r] * g;
                I[3 * (r * 64 + lane) + 2] = az[r] * g;
            }
        }
    }
}

// Six wavefronts per SIMD (at most 80 VGPRs; R = 8 needs a few more
// unconstrained, and the compiler spills two or three values outside the
// inner loop): 2.5% faster at N = 262,144 than five.
template <int R, bool POT>
__global__ __launch_bounds__(64 * kMutualWaves) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_mutual_grp(
    const float4* __restrict__ gpos, const GroupTask* __restrict__ gtasks, int ntasks,
    const MutualTask* __restrict__ rtasks, const SimDesc* __restrict__ sims, const RowDesc* __restrict__ rows,
    float* __restrict__ iside, float* __restrict__ jside, float* __restrict__ jintra) {
    constexpr int B = 64 * R;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if ((int)blockIdx.x >= ntasks) return;
    const GroupTask gt = gtasks[blockIdx.x];
    __shared__ float4 lq[kMutualWaves][128];
    __shared__ float red[2][3][kGroupRows][64];
    if (gt.kind == 1) {
        const MutualTask t = rtasks[gt.rtask0 + wv];
        if (t.j1 <= t.j0) return;  // a filler of the last workgroup
        const SimDesc& sd = sims[t.sim];
        float* __restrict__ J = jintra + sd.jibase + 3LL * (kGroupRows - 1) * B * t.row;
        mutual_wave_task<R, POT>(t, sd, gpos + sd.gbase, lq[wv], iside, J, lane);
        return;
    }
    const SimDesc& sd = sims[gt.sim];
    mutual_group_task<R, POT>(gt, sd, rows, gpos + sd.gbase, iside, jside, wv, lane,
                              reinterpret_cast<float4(*)[128]>(&lq[0][0]), red);
}

// Binary search of the simulation owning padded slot q (sims ordered by gbase).
__device__ __forceinline__ int find_gsim(const SimDesc* __restrict__ sims, int nsims, int q) {
    int lo = 0, hi = nsims - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sims[mid].gbase <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Padded layout of the mutual pass: per simulation, iron [0, gfe) then
// silicate [gfe, gn); slots past a segment's particles are ghosts.
__global__ void k_mutual_pad(const float4* __restrict__ pos, int gtot, const SimDesc* __restrict__ sims,
                             int nsims, float4* __restrict__ gpos) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= gtot) return;
    const SimDesc& sd = sims[find_gsim(sims, nsims, q)];
    const int u = q - sd.gbase;
    int src = -1;
    if (u < sd.gfe) {
        if (u < sd.fe_end - sd.start) src = sd.start + u;
    } else if (u - sd.gfe < sd.end - sd.fe_end) {
        src = sd.fe_end + (u - sd.gfe);
    }
    gpos[q] = src >= 0 ? pos[src] : make_float4(kGhostCoord, kGhostCoord, kGhostCoord, 0.0f);
}

// Per particle: the i-side slots of its row and the j-side entries of every
// earlier row (k_mutual), or of every earlier group and the earlier rows of
// its own group (k_mutual_grp, groups of W rows), added in compensated
// float sums (CompSum) in a fixed order.  A workgroup takes 64 padded
// places of one row (a SumBlock; the host lists the blocks holding a real
// particle, heaviest rows first), so the loop bounds are uniform.  Four
// wavefronts per block: wavefront w adds the w-th quarter of each particle's
// slots and slab entries (one particle per lane, so the loads stay
// coalesced), and the quarters are added in LDS in a fixed order -- four
// times the loads in flight of one thread per particle.
// With corr (the contact corrections of k_contact_corr), the correction plus
// the sum rounded to float goes to force.xyz -- k_finish's arithmetic --
// and, with mpos / mvel, the particle then takes the move's step (k_move's
// and k_move_collision's arithmetic) with the constants of mv, or with its
// simulation's own when mv.per_sim; without corr, the sum goes to gsum for
// k_finish / k_finish_potential.
constexpr int kSumSplit = 4;

// Move constants of a move fused into k_mutual_sum.
struct SumMove {
    float dt, dt_over_mass_fe, dt_over_mass_si;
    int per_sim;  // 1: take them from each SimDesc (the batched engine)
};

// Entries k0 .. k1 - 1 of one list (slots, or slab entries) added to
// (sx, sy, sz) in order.  at(k) gives the wavefront-uniform address of entry
// k's record for lane 0: lane l reads the (x, y, z) record at p + 3 l, so a
// wavefront reads 768 contiguous bytes per entry.  The loads of kSumBatch
// entries are issued before their additions, so a wavefront has that many
// round trips to memory in flight instead of one; the additions keep their
// order, so the sums are those of the one-at-a-time loop.  POT: x only.
constexpr int kSumBatch = 4;

template <bool POT, typename At>
__device__ __forceinline__ void add_entries(int k0, int k1, int lane, At at, CompSum& sx, CompSum& sy,
                                            CompSum& sz) {
    int k = k0;
    for (; k + kSumBatch <= k1; k += kSumBatch) {
        float vx[kSumBatch], vy[kSumBatch], vz[kSumBatch];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) {
            const float* p = at(k + u) + 3 * lane;
            vx[u] = p[0];
            if constexpr (!POT) {
                vy[u] = p[1];
                vz[u] = p[2];
            }
        }
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) {
            sx.add(vx[u]);
            if constexpr (!POT) {
                sy.add(vy[u]);
                sz.add(vz[u]);
            }
        }
    }
    for (; k < k1; ++k) {
        const float* p = at(k) + 3 * lane;
        sx.add(p[0]);
        if constexpr (!POT) {
            sy.add(p[1]);
            sz.add(p[2]);
        }
    }
}

template <bool POT>
__global__ __launch_bounds__(64 * kSumSplit) void k_mutual_sum(const SumBlock* __restrict__ blocks,
                                                              const SimDesc* __restrict__ sims,
                                                              const RowDesc* __restrict__ rows, int B, int W,
                                                              const float* __restrict__ iside,
                                                              const float* __restrict__ jside,
                                                              const float* __restrict__ jintra,
                                                              float4* __restrict__ gsum,
                                                              const float* __restrict__ corr, int ntot,
                                                              float4* __restrict__ force,
                                                              float4* __restrict__ mpos,
                                                              float4* __restrict__ mvel, SumMove mv) {
    __shared__ float red[3][2][kSumSplit][64];  // per component: sum and compensation of each quarter
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the addresses below stay scalar
    const SumBlock blk = blocks[blockIdx.x];
    const SimDesc& sd = sims[blk.sim];
    const int q = blk.q0 + lane;
    const int row = blk.q0 / B;
    const int o0 = blk.q0 - row * B;  // lane 0's place in its row
    // the particle at padded place q, or -1 for a ghost
    int i = -1;
    if (q < sd.gfe) {
        if (q < sd.fe_end - sd.start) i = sd.start + q;
    } else if (q - sd.gfe < sd.end - sd.fe_end) {
        i = sd.fe_end + (q - sd.gfe);
    }
    CompSum sx, sy, sz;
    sx.zero();
    sy.zero();
    sz.zero();
    if (i >= 0) {
        const RowDesc rd = rows[sd.row0 + row];
        add_entries<POT>(
            rd.nslot * w / kSumSplit, rd.nslot * (w + 1) / kSumSplit, lane,
            [&](int k) { return iside + (size_t)(rd.slot0 + k) * (3 * B) + 3 * o0; },
            sx, sy, sz);
        // the j-side entries of the earlier groups of W rows (W = 1: rows) ...
        const int group = row / W, WB = W * B;
        add_entries<POT>(
            group * w / kSumSplit, group * (w + 1) / kSumSplit, lane,
            [&](int r) {
                return jside + sd.jbase + 3 * (jside_row_offset(r, sd.gn, WB) + (blk.q0 - (long long)(r + 1) * WB));
            },
            sx, sy, sz);
        // ... and of the earlier rows of its own group
        const int in_group = row - group * W;
        const long long ilen = (long long)(W - 1) * B;
        add_entries<POT>(
            group * W + in_group * w / kSumSplit, group * W + in_group * (w + 1) / kSumSplit, lane,
            [&](int r) { return jintra + sd.jibase + 3 * (ilen * r + (blk.q0 - (long long)(r + 1) * B)); },
            sx, sy, sz);
    }
    red[0][0][w][lane] = sx.s;
    red[0][1][w][lane] = sx.c;
    red[1][0][w][lane] = sy.s;
    red[1][1][w][lane] = sy.c;
    red[2][0][w][lane] = sz.s;
    red[2][1][w][lane] = sz.c;
    __syncthreads();
    if (w == 0 && i >= 0) {
        // the quarters in order; their compensations join the total's
        CompSum ax, ay, az;
        ax.zero();
        ay.zero();
        az.zero();
#pragma unroll
        for (int k = 0; k < kSumSplit; ++k) {
            ax.add(red[0][0][k][lane]);
            ax.c += red[0][1][k][lane];
            ay.add(red[1][0][k][lane]);
            ay.c += red[1][1][k][lane];
            az.add(red[2][0][k][lane]);
            az.c += red[2][1][k][lane];
        }
        if (corr) {
            // k_finish's arithmetic: the correction's (sum, compensation)
            // pair plus the all-pairs sum rounded to float, rounded once
            const float2* rec = reinterpret_cast<const float2*>(corr + 6 * (size_t)i);
            const float2 rx = rec[0], ry = rec[1], rz = rec[2];
            CompSum cx{rx.x, rx.y}, cy{ry.x, ry.y}, cz{rz.x, rz.y};
            cx.add(ax.value());
            cy.add(ay.value());
            cz.add(az.value());
            const float fx = cx.value(), fy = cy.value(), fz = cz.value();
            force[i].x = fx;
            force[i].y = fy;
            force[i].z = fz;
            if (mpos) {
                const float dt = mv.per_sim ? sd.dt : mv.dt;
                const float dtm_fe = mv.per_sim ? sd.dt_over_mass_fe : mv.dt_over_mass_fe;
                const float dtm_si = mv.per_sim ? sd.dt_over_mass_si : mv.dt_over_mass_si;
                const float temp = i < sd.fe_end ? dtm_fe : dtm_si;
                float4 v = mvel[i];
                float4 p = mpos[i];
                v.x += fx * temp;
                v.y += fy * temp;
                v.z += fz * temp;
                p.x += v.x * dt;
                p.y += v.y * dt;
                p.z += v.z * dt;
                mvel[i] = v;
                mpos[i] = p;
            }
        } else {
            gsum[i] = make_float4(ax.value(), ay.value(), az.value(), 0.0f);
        }
    }
}

// ---------------------------------------------------------------------------
// Pass 2: spatial hash of unit cells (counting sort, deterministic buckets)
// ---------------------------------------------------------------------------

__global__ void k_fill_int(int* __restrict__ a, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_grid_count(const float4* __restrict__ pos, int ntot, const SimDesc* __restrict__ sims,
                             int nsims, int* __restrict__ counts, int* __restrict__ cellhash) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    const SimDesc& sd = sims[find_sim(sims, nsims, i)];
    float4 p = pos[i];
    unsigned h = cell_bucket(sd.hbase, sd.hmask, cell_coord(p.x), cell_coord(p.y), cell_coord(p.z));
    cellhash[i] = (int)h;
    atomicAdd(&counts[h], 1);
}

// Exclusive scan of counts[0..m) into starts[0..m], 1024 elements per block.
__global__ __launch_bounds__(256) void k_scan_blocks(const int* __restrict__ in, int m,
                                                     int* __restrict__ out,
                                                     int* __restrict__ block_sums) {
    __shared__ int sh[256];
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4];
    int local = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < m) ? in[base + k] : 0;
        local += v[k];
    }
    sh[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - local;  // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < m) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) block_sums[blockIdx.x] = sh[255];
}

// Exclusive scan of the block sums in one workgroup of kScanTopThreads
// threads, 16 sums each (nb <= 4096).  The workgroup is kept to one
// wavefront per SIMD: it runs on a side stream beside the pair kernel, and a
// 1,024-thread workgroup (four wavefronts on every SIMD of one CU) waited
// for a CU to drain for most of the pair kernel's run.
constexpr int kScanTopThreads = 256;
__global__ __launch_bounds__(kScanTopThreads) void k_scan_top(int* __restrict__ block_sums, int nb,
                                                              int* __restrict__ total_out) {
    constexpr int T = kScanTopThreads, E = 4096 / kScanTopThreads;
    __shared__ int sh[T];
    const int base = threadIdx.x * E;
    int v[E];
    int local = 0;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        v[k] = (base + k < nb) ? block_sums[base + k] : 0;
        local += v[k];
    }
    sh[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        int add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - local;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (base + k < nb) block_sums[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == T - 1) *total_out = sh[T - 1];
}

__global__ void k_scan_add(int* __restrict__ out, int m, const int* __restrict__ block_sums) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] += block_sums[i >> 10];
}

// Scatter particle ids into their buckets; counts[] is consumed back to zero.
__global__ void k_grid_scatter(int ntot, const int* __restrict__ cellhash,
                               const int* __restrict__ starts, int* __restrict__ counts,
                               int* __restrict__ sorted) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    int h = cellhash[i];
    int slot = starts[h] + atomicSub(&counts[h], 1) - 1;
    sorted[slot] = i;
}

// Sort each bucket by particle id so the contact sums run in a fixed order.
__global__ void k_grid_sort_buckets(int m, const int* __restrict__ starts, int* __restrict__ sorted) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= m) return;
    int s = starts[h], e = starts[h + 1];
    for (int a = s + 1; a < e; ++a) {
        int key = sorted[a];
        int b = a - 1;
        while (b >= s && sorted[b] > key) {
            sorted[b + 1] = sorted[b];
            --b;
        }
        sorted[b + 1] = key;
    }
}

// Positions and velocities in bucket order, so the contact pass reads the
// particles of a cell from consecutive memory.
__global__ void k_grid_gather(int ntot, const int* __restrict__ sorted, const float4* __restrict__ pos,
                              const float4* __restrict__ vel, float4* __restrict__ spos,
                              float4* __restrict__ svel) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    const int i = sorted[t];
    spos[t] = pos[i];
    if (svel) svel[t] = vel[i];
}

// ---------------------------------------------------------------------------
// Pass 3: chunk sums + exact contact correction
// ---------------------------------------------------------------------------

// The contact correction of a particle, for every pair with r < 1 found
// through the grid: the reference's exact contact term minus what the
// all-pairs pass added for that pair.  Contact terms are rounded exactly as
// the reference rounds them (float products of a float magnitude); each
// term and each gravity term subtracted goes into a compensated float sum,
// so cancellation between large contact forces adds no summation error.
//
// The 27-cell walk is a chain of dependent loads (bucket bounds, then each
// candidate).  It is cut into the nine rows of three cells (oz, oy) =
// (r / 3 - 1, r % 3 - 1), r = 0..8; each row is summed from zero (cells in x
// order, candidates in bucket order) and the nine row sums are added in row
// order.  LANES = 9 walks the rows of a particle on nine lanes at once (seven
// particles per wavefront, the row sums added through LDS): a chain a ninth
// as long, for plans too small to hide the walk behind the all-pairs pass.
// LANES = 1 walks them in turn on one lane.  The arithmetic is the same, so
// the two forms give bitwise the same corrections.  A simulation's buckets
// hold only its own particles, and the particle itself sits at position t,
// so a candidate's index is loaded only for a pair in contact.
constexpr int kCellRows = 9;
constexpr int kContactWaves = 4;  // wavefronts per workgroup

__device__ __forceinline__ void contact_row(const SimDesc& sd, int t, int i, int row,
                                            const int* __restrict__ starts, const int* __restrict__ sorted,
                                            const float4* __restrict__ spos, const float4* __restrict__ svel,
                                            int variant, CompSum& ex, CompSum& ey, CompSum& ez) {
    const float4 pi = spos[t];
    const float4 vi = svel[t];
    const int ccx = cell_coord(pi.x), ccy = cell_coord(pi.y), ccz = cell_coord(pi.z);
    const bool ife = i < sd.fe_end;
    const int qy = ccy + row % 3 - 1, qz = ccz + row / 3 - 1;
    int b0[3], b1[3];
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
        const unsigned h = cell_bucket(sd.hbase, sd.hmask, ccx + ox - 1, qy, qz);
        b0[ox] = starts[h];
        b1[ox] = starts[h + 1];
    }
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
        const int qx = ccx + ox - 1;
        for (int k = b0[ox]; k < b1[ox]; ++k) {
            const float4 pj = spos[k];
            if (k == t || cell_coord(pj.x) != qx || cell_coord(pj.y) != qy || cell_coord(pj.z) != qz)
                continue;  // i itself, or another cell sharing the bucket
            const float dpx = pj.x - pi.x;
            const float dpy = pj.y - pi.y;
            const float dpz = pj.z - pi.z;
            const float r2 = dpx * dpx + dpy * dpy + dpz * dpz;
            if (1.0f <= r2) continue;  // then sqrtf(r2) >= 1 too: the root only for candidates that may touch
            const float r = sqrtf(r2);
            if (1.0f <= r) continue;
            const float invr = 1.0f / r;
            const int j = sorted[k];
            const int test = (int)ife + (int)(j < sd.fe_end);
            const float4 vj = svel[k];
            const float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
            const float dot = dpx * dvx + dpy * dvy + dpz * dvz;
            const float fm = contact_pair_mag(sd.law, test, r, r2, dot);
            // The contact term minus what pass 1 added for this pair (for
            // r < 1 a bounded placeholder: |sf dp| < g), one addition each.
            const float g = gravity_product(sd.law, test);
            const float sf = g * inv_r3_variant(variant, fast_r2(dpx, dpy, dpz));
            ex.add(fm * dpx * invr - sf * dpx);
            ey.add(fm * dpy * invr - sf * dpy);
            ez.add(fm * dpz * invr - sf * dpz);
        }
    }
}

// Particles per wavefront of the contact walk.
template <int LANES>
__host__ __device__ constexpr int contact_per_wave() {
    return LANES == 1 ? 64 : 64 / LANES;
}

// Workgroups of k_finish / k_contact_corr for ntot particles.
template <int LANES>
__host__ __device__ constexpr int contact_blocks(int ntot) {
    return (ntot + kContactWaves * contact_per_wave<LANES>() - 1) / (kContactWaves * contact_per_wave<LANES>());
}

// The lane's particle (bucket position t, index i, simulation s; t = -1 for
// none) and, on the lane that owns the result (`lead`), its correction.
struct ContactLane {
    int t, i, s;
    bool lead;
    CompSum ex, ey, ez;  // the correction as an unrounded (sum, compensation) pair
};

template <int LANES>
__device__ __forceinline__ ContactLane contact_correction(int ntot, const SimDesc* __restrict__ sims, int nsims,
                                                          const int* __restrict__ starts,
                                                          const int* __restrict__ sorted,
                                                          const float4* __restrict__ spos,
                                                          const float4* __restrict__ svel, int variant) {
    static_assert(LANES == 1 || LANES == kCellRows, "one lane per particle or one per cell row");
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int slot = lane / LANES;  // LANES = 9: 0..6, and 7 for the idle lane 63
    const int row = LANES == 1 ? 0 : lane % LANES;
    ContactLane c;
    c.t = slot < contact_per_wave<LANES>()
              ? (blockIdx.x * kContactWaves + w) * contact_per_wave<LANES>() + slot
              : -1;
    if (c.t >= ntot) c.t = -1;
    c.i = c.t >= 0 ? sorted[c.t] : 0;
    c.s = c.t >= 0 ? find_sim(sims, nsims, c.i) : 0;
    const bool live = c.t >= 0 && sims[c.s].active;
    // each row summed from zero, then the nine rows' (sum, compensation)
    // pairs added in row order (the same arithmetic on one lane or nine)
    c.ex.zero();
    c.ey.zero();
    c.ez.zero();
    if constexpr (LANES == 1) {
        c.lead = live;
        if (live)
            for (int r = 0; r < kCellRows; ++r) {
                CompSum rx, ry, rz;
                rx.zero();
                ry.zero();
                rz.zero();
                contact_row(sims[c.s], c.t, c.i, r, starts, sorted, spos, svel, variant, rx, ry, rz);
                c.ex.add_sum(rx);
                c.ey.add_sum(ry);
                c.ez.add_sum(rz);
            }
    } else {
        __shared__ float part[kContactWaves][6][64];
        CompSum rx, ry, rz;
        rx.zero();
        ry.zero();
        rz.zero();
        if (live) contact_row(sims[c.s], c.t, c.i, row, starts, sorted, spos, svel, variant, rx, ry, rz);
        part[w][0][lane] = rx.s;
        part[w][1][lane] = rx.c;
        part[w][2][lane] = ry.s;
        part[w][3][lane] = ry.c;
        part[w][4][lane] = rz.s;
        part[w][5][lane] = rz.c;
        __syncthreads();
        c.lead = live && row == 0;
        if (c.lead) {
#pragma unroll
            for (int r = 0; r < kCellRows; ++r) {
                c.ex.add(part[w][0][lane + r]);
                c.ex.c += part[w][1][lane + r];
                c.ey.add(part[w][2][lane + r]);
                c.ey.c += part[w][3][lane + r];
                c.ez.add(part[w][4][lane + r]);
                c.ez.c += part[w][5][lane + r];
            }
        }
    }
    return c;
}

// Threads walk the particles in bucket (Morton) order: the lanes of a
// wavefront look up the same and adjacent cells.
template <int LANES>
__global__ __launch_bounds__(64 * kContactWaves) void k_finish(float4* __restrict__ force, int ntot,
                                                               const SimDesc* __restrict__ sims, int nsims,
                                                               const float4* __restrict__ partial,
                                                               const float4* __restrict__ gsum,
                                                               const int* __restrict__ starts,
                                                               const int* __restrict__ sorted,
                                                               const float4* __restrict__ spos,
                                                               const float4* __restrict__ svel, int variant) {
    const ContactLane c = contact_correction<LANES>(ntot, sims, nsims, starts, sorted, spos, svel, variant);
    if (!c.lead) return;
    const SimDesc& sd = sims[c.s];
    const int i = c.i;
    const int n = sd.end - sd.start;
    // the correction's pair plus the all-pairs sums, rounded once
    CompSum fx = c.ex, fy = c.ey, fz = c.ez;
    if (gsum && sd.gn > 0) {
        // mutual pass: already summed per particle
        const float4 gs = gsum[i];
        fx.add(gs.x);
        fy.add(gs.y);
        fz.add(gs.z);
    } else {
        for (int ch = 0; ch < sd.nchunk; ++ch) {
            float4 p = partial[sd.partial_base + ch * n + (i - sd.start)];
            fx.add(p.x);
            fy.add(p.y);
            fz.add(p.z);
        }
    }
    force[i].x = fx.value();
    force[i].y = fy.value();
    force[i].z = fz.value();
}

// The contact corrections alone, into corr (per particle, by index, one
// 24-byte record: the (sum, compensation) pairs of x, y and z -- the lanes
// walk in bucket order, so one record per lane keeps each store to one or
// two cache lines), for plans whose simulations all take the mutual pass:
// it needs only the grid, so it runs beside the all-pairs pass, and
// k_mutual_sum adds it to the sums -- the same arithmetic as k_finish.
template <int LANES>
__global__ __launch_bounds__(64 * kContactWaves) void k_contact_corr(int ntot, const SimDesc* __restrict__ sims,
                                                                     int nsims, const int* __restrict__ starts,
                                                                     const int* __restrict__ sorted,
                                                                     const float4* __restrict__ spos,
                                                                     const float4* __restrict__ svel, int variant,
                                                                     float* __restrict__ corr) {
    const ContactLane c = contact_correction<LANES>(ntot, sims, nsims, starts, sorted, spos, svel, variant);
    if (!c.lead) return;
    float2* rec = reinterpret_cast<float2*>(corr + 6 * (size_t)c.i);
    rec[0] = make_float2(c.ex.s, c.ex.c);
    rec[1] = make_float2(c.ey.s, c.ey.c);
    rec[2] = make_float2(c.ez.s, c.ez.c);
}

// ---------------------------------------------------------------------------
// Direct pass: the whole force of a small simulation in one launch
// ---------------------------------------------------------------------------
//
// For small N the three passes above are bound by their dozen launches, not by
// arithmetic.  k_direct computes the same force in one: a workgroup holds
// 64 / S i-particles, S lanes each, and its kDirectWaves * S lanes per
// particle ("j-streams") share the simulation's j-range: each segment (iron
// j, then silicate j, so the gravity product is one factor per segment) is
// cut into groups of four j dealt out in turn (stream k takes groups k,
// k + 16 S, ...).  Dealing keeps the contact work -- concentrated on the j
// near the block's own particles -- spread over all streams; S > 1 gives small
// simulations S times more wavefronts, each with a shorter chain (S = 8: 17%
// faster steps at 2,048 particles than S = 4, whose workgroups fill only half
// of the CUs there).  With S = 1
// the j come through wave-uniform scalar loads.  Every pair takes the
// all-pairs pass's branch-free gravity term (float tiles of kDirectTile j
// folded into a compensated float sum); when some lane of a group has a pair
// with r^2 < 1 the group also adds, for its pairs in contact, the exact
// contact term minus that gravity term, with the arithmetic of k_finish.  The
// streams' sums are added in a compensated float sum in a fixed order, so the
// result depends on the simulation alone.  (The potential variant sums in
// FP64: it feeds the energy reduction.)
// POT: the per-particle potential of k_finish_potential instead of the force.
constexpr int kDirectWaves = 16;
constexpr int kDirectTile = 256;  // j per float tile
// r^2 (with FMAs and kSelfEps) of every pair whose exactly rounded r is below
// 1 lies below this: the margin covers the few ulps between the two forms.
constexpr float kDirectContactR2 = 1.0001f;

// The move of a step fused into the direct pass (MOVE = 1:
// moveBodiesCollisionSingleGPU, impactAddTime.cu:379-394; MOVE = 2:
// moveBodiesDampedSeperate, generateTargetImpactor.cu:628-664), with the
// same arithmetic as k_move_collision / k_move_damped_separate.  The new
// state goes to other arrays (pout, vout): the pass still reads the old
// positions of every particle.  MOVE = 1 reads b1 as NFe.
struct FusedMove {
    float4* pout;
    float4* vout;
    float dt, dt_over_mass_fe, dt_over_mass_si;
    int b1, b2, b3;
    float damp1, damp2;
};

// FP64 sum with CompSum's interface: the potential pass feeds the energy
// reduction, which stays FP64.
struct F64Sum {
    double s;
    __device__ __forceinline__ void zero() { s = 0.0; }
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
};
struct PlainSum {
    float s;
    __device__ __forceinline__ void zero() { s = 0.0f; }
    __device__ __forceinline__ void add(float x) { s += x; }
    __device__ __forceinline__ float value() const { return s; }
};

template <bool POT, int S, int MOVE>
__global__ __launch_bounds__(64 * kDirectWaves) void k_direct(const float4* __restrict__ pos,
                                                              const float4* __restrict__ vel,
                                                              const DirectBlock* __restrict__ blocks,
                                                              const SimDesc* __restrict__ sims,
                                                              float4* __restrict__ force,
                                                              double* __restrict__ pe, FusedMove mv) {
    static_assert(MOVE == 0 || !POT, "the potential pass moves nothing");
    static_assert(S == 1 || S == 2 || S == 4 || S == 8, "lanes per particle");
    constexpr int IB = 64 / S;               // i-particles per workgroup
    constexpr int NS = kDirectWaves * S;     // j-streams per particle
    constexpr int TG = kDirectTile / 4;      // groups per stream per float tile
    using Sum = typename std::conditional<POT, F64Sum, CompSum>::type;
    using CSum = typename std::conditional<POT, F64Sum, PlainSum>::type;
    using Val = typename std::conditional<POT, double, float>::type;
    __shared__ Val red[3][kDirectWaves][64];
    __shared__ float redc[POT ? 1 : 3][kDirectWaves][64];  // the force's compensations
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int stream = wv * S + lane / IB;  // wave-uniform when S == 1
    const DirectBlock b = blocks[blockIdx.x];
    const int start = sims[b.sim].start, fe_end = sims[b.sim].fe_end, end = sims[b.sim].end;
    const PairLaw law = sims[b.sim].law;  // in registers: the contact branch reads it often
    const int i = b.i0 + lane % IB;
    const bool live = i < end;
    const int ic = live ? i : end - 1;
    const float4 pi = pos[ic];
    const bool ife = ic < fe_end;

    float4 vi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (!POT) vi = vel[ic];
    Sum sx, sy, sz;  // this wavefront's gravity plus contact corrections
    sx.zero();
    sy.zero();
    sz.zero();
    for (int seg = 0; seg < 2; ++seg) {
        const bool jfe = seg == 0;  // iron j first, then silicate j
        const int a = jfe ? start : fe_end;
        const int e = jfe ? fe_end : end;
        const int ngroups = (e - a + 3) / 4;
        if (stream >= ngroups) continue;
        const int test = (int)ife + (int)jfe;
        const float g = gravity_product(law, test);
        Sum ux, uy, uz;  // float tiles of kDirectTile j, folded (compensated; FP64 for the potential)
        // this stream's contact terms of the segment: a handful (the pairs in
        // contact are dealt over 16 S streams), so a plain float sum rounds
        // them to ~3 ulp of their magnitudes; FP64 for the potential
        CSum cx, cy, cz;
        ux.zero();
        uy.zero();
        uz.zero();
        cx.zero();
        cy.zero();
        cz.zero();
        // The exact contact term of pair (i, j) minus the gravity term s*dp
        // the tile sum took for it (k_finish's, resp. k_finish_potential's,
        // test and arithmetic).
        auto contact = [&](int j, const float4 pj, float s) {
            const float dx = pj.x - pi.x;
            const float dy = pj.y - pi.y;
            const float dz = pj.z - pi.z;
            if constexpr (!POT) {
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (!(r2 < 1.0f)) return;  // then sqrtf(r2) >= 1 too
                const float r = sqrtf(r2);
                if (r < 1.0f && j != i) {
                    const float invr = 1.0f / r;
                    const float4 vj = vel[j];
                    const float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
                    const float dot = dx * dvx + dy * dvy + dz * dvz;
                    const float fm = contact_pair_mag(law, test, r, r2, dot);
                    const float sf = g * s;
                    cx.add(fm * dx * invr - sf * dx);
                    cy.add(fm * dy * invr - sf * dy);
                    cz.add(fm * dz * invr - sf * dz);
                }
            } else {
                const double ddx = (double)pj.x - pi.x, ddy = (double)pj.y - pi.y, ddz = (double)pj.z - pi.z;
                const double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                if (r < 1.0 && j != i) {
                    const double G = gravity_product(law, test);
                    const double K = approach_k(law, test);
                    const double exact = G * (r - 2.0) + (K / 3.0) * (r - 1.0) * (r - 1.0) * (r + 2.0);
                    cx.add(exact + G * (double)s);
                }
            }
        };
        // this stream's groups, TG of them per float tile
        for (int g0 = stream; g0 < ngroups; g0 += NS * TG) {
            const int g1 = min(g0 + NS * TG, ngroups);
            float tx = 0.0f, ty = 0.0f, tz = 0.0f;
            auto pair = [&](const float4 q, float& s, bool& near) {
                const float dx = q.x - pi.x;
                const float dy = q.y - pi.y;
                const float dz = q.z - pi.z;
                if constexpr (!POT) {
                    const float r2f = fast_r2(dx, dy, dz);
                    s = fast_inv_r3(r2f);
                    tx = __builtin_fmaf(s, dx, tx);
                    ty = __builtin_fmaf(s, dy, ty);
                    tz = __builtin_fmaf(s, dz, tz);
                    near |= r2f < kDirectContactR2;
                } else {
                    const float r2f = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    s = r2f > 0.0f ? __builtin_amdgcn_rsqf(r2f) : 0.0f;
                    tx += s;
                    near |= r2f < kDirectContactR2;
                }
            };
            // The next group's loads are in flight while this one is computed
            // (indices clamped into the segment; a partial last group only
            // takes its first e - j pairs).
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = pos[min(a + 4 * g0 + u, e - 1)];
            for (int grp = g0; grp < g1; grp += NS) {
                const int j = a + 4 * grp;
                const int jn = grp + NS < g1 ? j + 4 * NS : j;
                float4 nq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) nq[u] = pos[min(jn + u, e - 1)];
                float s[4];
                bool near = false;
                if (j + 4 <= e) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) pair(q[u], s[u], near);
                    if (near) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) contact(j + u, q[u], s[u]);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (j + u < e) pair(q[u], s[u], near);
                    if (near) {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (j + u < e) contact(j + u, q[u], s[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) q[u] = nq[u];
            }
            ux.add(tx);
            uy.add(ty);
            uz.add(tz);
        }
        if constexpr (!POT) {
            sx.add_product(g, ux.s);
            sx.c += g * ux.c;
            sx.add(cx.s);
            sy.add_product(g, uy.s);
            sy.c += g * uy.c;
            sy.add(cy.s);
            sz.add_product(g, uz.s);
            sz.c += g * uz.c;
            sz.add(cz.s);
        } else {
            sx.add(-(double)g * ux.value() + cx.value());
        }
    }
    if constexpr (!POT) {
        red[0][wv][lane] = sx.s;
        red[1][wv][lane] = sy.s;
        red[2][wv][lane] = sz.s;
        redc[0][wv][lane] = sx.c;
        redc[1][wv][lane] = sy.c;
        redc[2][wv][lane] = sz.c;
    } else {
        red[0][wv][lane] = sx.value();
    }
    __syncthreads();
    // The force: wavefronts 0-2 take one component each; lane (q, p) adds the
    // 16 wavefronts' partials of particle p's q-th stream in wavefront order,
    // then the S streams' sums are added in order (a 16 + S chain per lane
    // instead of a 16 S chain on one wavefront).  The potential: one FP64
    // chain on wavefront 0.
    __shared__ float part[POT ? 1 : 3][2][64];
    if constexpr (!POT) {
        if (wv < 3) {
            CompSum a;
            a.zero();
#pragma unroll
            for (int w = 0; w < kDirectWaves; ++w) {
                a.add(red[wv][w][lane]);
                a.c += redc[wv][w][lane];
            }
            part[wv][0][lane] = a.s;
            part[wv][1][lane] = a.c;
        }
        __syncthreads();
    }
    if (wv == 0 && lane < IB && live) {
        if constexpr (!POT) {
            float f[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                CompSum t;
                t.zero();
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    t.add(part[c][0][q * IB + lane]);
                    t.c += part[c][1][q * IB + lane];
                }
                f[c] = t.value();
            }
            const float f0 = f[0], f1 = f[1], f2 = f[2];
            force[i].x = f0;
            force[i].y = f1;
            force[i].z = f2;
            if constexpr (MOVE != 0) {
                float temp, damp = 0.0f;
                if constexpr (MOVE == 1) {
                    temp = i < mv.b1 ? mv.dt_over_mass_fe : mv.dt_over_mass_si;
                } else if (mv.b3 <= i) {
                    temp = mv.dt_over_mass_si;
                    damp = mv.damp2;
                } else if (mv.b2 <= i) {
                    temp = mv.dt_over_mass_fe;
                    damp = mv.damp2;
                } else if (mv.b1 <= i) {
                    temp = mv.dt_over_mass_si;
                    damp = mv.damp1;
                } else {
                    temp = mv.dt_over_mass_fe;
                    damp = mv.damp1;
                }
                float4 v = vel[i];
                float4 p = pos[i];
                if constexpr (MOVE == 1) {
                    v.x += f0 * temp;
                    v.y += f1 * temp;
                    v.z += f2 * temp;
                } else {
                    v.x += (f0 - damp * v.x) * temp;
                    v.y += (f1 - damp * v.y) * temp;
                    v.z += (f2 - damp * v.z) * temp;
                }
                p.x += v.x * mv.dt;
                p.y += v.y * mv.dt;
                p.z += v.z * mv.dt;
                mv.vout[i] = v;
                mv.pout[i] = p;
            }
        } else {
            F64Sum fx;
            fx.zero();
#pragma unroll
            for (int w = 0; w < kDirectWaves; ++w)
#pragma unroll
                for (int q = 0; q < S; ++q) fx.add(red[0][w][q * IB + lane]);
            pe[i] = 0.5 * fx.value();
        }
    }
}

// ---------------------------------------------------------------------------
// Move kernels
// ---------------------------------------------------------------------------

// moveBodiesCollisionSingleGPU (impactAddTime.cu:379-394), generalised to
// packed simulations.  No contraction: v + f*t and x + v*dt round as written.
__global__ void k_move(float4* __restrict__ pos, float4* __restrict__ vel,
                       const float4* __restrict__ force, int ntot, const SimDesc* __restrict__ sims,
                       int nsims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    const SimDesc& sd = sims[nsims == 1 ? 0 : find_sim(sims, nsims, i)];
    if (!sd.active) return;
    const float temp = i < sd.fe_end ? sd.dt_over_mass_fe : sd.dt_over_mass_si;
    float4 v = vel[i];
    float4 p = pos[i];
    const float4 f = force[i];
    v.x += f.x * temp;
    v.y += f.y * temp;
    v.z += f.z * temp;
    p.x += v.x * sd.dt;
    p.y += v.y * sd.dt;
    p.z += v.z * sd.dt;
    vel[i] = v;
    pos[i] = p;
}

// moveBodiesCollisionSingleGPU with its constants by value (the drop-in path).
__global__ void k_move_collision(float4* __restrict__ pos, float4* __restrict__ vel,
                                 const float4* __restrict__ force, int n, float dt,
                                 float dt_over_mass_fe, float dt_over_mass_si, int nfe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float temp = i < nfe ? dt_over_mass_fe : dt_over_mass_si;
    float4 v = vel[i];
    float4 p = pos[i];
    const float4 f = force[i];
    v.x += f.x * temp;
    v.y += f.y * temp;
    v.z += f.z * temp;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    vel[i] = v;
    pos[i] = p;
}

// moveBodiesDampedSeperate (generateTargetImpactor.cu:628-664).
__global__ void k_move_damped_separate(float4* __restrict__ pos, float4* __restrict__ vel,
                                       const float4* __restrict__ force, int n, float dt,
                                       float dt_over_mass_fe, float dt_over_mass_si, int b1, int b2,
                                       int b3, float damp1, float damp2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float temp, damp;
    if (b3 <= i) {
        temp = dt_over_mass_si;
        damp = damp2;
    } else if (b2 <= i) {
        temp = dt_over_mass_fe;
        damp = damp2;
    } else if (b1 <= i) {
        temp = dt_over_mass_si;
        damp = damp1;
    } else {
        temp = dt_over_mass_fe;
        damp = damp1;
    }
    float4 v = vel[i];
    float4 p = pos[i];
    const float4 f = force[i];
    v.x += (f.x - damp * v.x) * temp;
    v.y += (f.y - damp * v.y) * temp;
    v.z += (f.z - damp * v.z) * temp;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    vel[i] = v;
    pos[i] = p;
}

// ---------------------------------------------------------------------------
// FP64 reductions: mass moments, momentum, angular momentum, energy
// ---------------------------------------------------------------------------

// Per-particle potential: -(1/2) * (sum of the chunk partials of 1/r scaled
// by G mi mj) plus, for pairs with r < 1, half the difference between the
// exact contact potential and the gravity term pass 1 added.  FP64 output.
__global__ __launch_bounds__(256) void k_finish_potential(int ntot, const SimDesc* __restrict__ sims,
                                                          int nsims,
                                                          const float4* __restrict__ partial,
                                                          const float4* __restrict__ gsum,
                                                          const int* __restrict__ starts,
                                                          const int* __restrict__ sorted,
                                                          const float4* __restrict__ spos,
                                                          double* __restrict__ pe) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    const int i = sorted[t];
    const int s = find_sim(sims, nsims, i);
    const SimDesc& sd = sims[s];
    const int n = sd.end - sd.start;
    double phi = 0.0;
    if (gsum && sd.gn > 0)
        phi = (double)gsum[i].x;
    else
        for (int c = 0; c < sd.nchunk; ++c) phi += (double)partial[sd.partial_base + c * n + (i - sd.start)].x;
    double u = -phi;
    const float4 pi = spos[t];
    const int ccx = cell_coord(pi.x), ccy = cell_coord(pi.y), ccz = cell_coord(pi.z);
    const bool ife = i < sd.fe_end;
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int qx = ccx + ox, qy = ccy + oy, qz = ccz + oz;
                const unsigned h = cell_bucket(sd.hbase, sd.hmask, qx, qy, qz);
                for (int k = starts[h]; k < starts[h + 1]; ++k) {
                    const float4 pj = spos[k];
                    if (cell_coord(pj.x) != qx || cell_coord(pj.y) != qy || cell_coord(pj.z) != qz)
                        continue;
                    const int j = sorted[k];
                    if (j == i || j < sd.start || j >= sd.end) continue;
                    const double dx = (double)pj.x - pi.x, dy = (double)pj.y - pi.y,
                                 dz = (double)pj.z - pi.z;
                    const double r = sqrt(dx * dx + dy * dy + dz * dz);
                    if (r >= 1.0) continue;
                    const int test = (int)ife + (int)(j < sd.fe_end);
                    const double G = gravity_product(sd.law, test);
                    const double K = approach_k(sd.law, test);
                    const double exact = G * (r - 2.0) + (K / 3.0) * (r - 1.0) * (r - 1.0) * (r + 2.0);
                    const float fx = pj.x - pi.x, fy = pj.y - pi.y, fz = pj.z - pi.z;
                    const float r2f = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
                    const double fast = -G * (double)__builtin_amdgcn_rsqf(r2f);
                    u += exact - fast;
                }
            }
    pe[i] = 0.5 * u;
}

// Block partial sums of the 14 FP64 moments of the particles of one
// simulation: M, M x (3), M v (3), sum m x cross v (3), kinetic, potential,
// sum m |x|^2 (unused by callers, kept for the rigid-body test), count.
constexpr int kMoments = 14;

__global__ __launch_bounds__(256) void k_moments(const float4* __restrict__ pos,
                                                 const float4* __restrict__ vel,
                                                 const double* __restrict__ pe,
                                                 const int* __restrict__ block_sim,
                                                 const int* __restrict__ block_first,
                                                 const SimDesc* __restrict__ sims,
                                                 double mass_fe_override, double mass_si_override,
                                                 double* __restrict__ out) {
    __shared__ double sh[kMoments][256];
    const int s = block_sim[blockIdx.x];
    const int i = block_first[blockIdx.x] + threadIdx.x;
    const SimDesc& sd = sims[s];
    double v[kMoments];
#pragma unroll
    for (int k = 0; k < kMoments; ++k) v[k] = 0.0;
    if (i < sd.end) {
        double m;
        if (mass_fe_override > 0.0)
            m = i < sd.fe_end ? mass_fe_override : mass_si_override;
        else
            m = i < sd.fe_end ? (double)sd.law.GMassFeSi : 1.0;
        const float4 p = pos[i];
        const float4 u = vel[i];
        const double x = p.x, y = p.y, z = p.z, a = u.x, b = u.y, c = u.z;
        v[0] = m;
        v[1] = m * x;
        v[2] = m * y;
        v[3] = m * z;
        v[4] = m * a;
        v[5] = m * b;
        v[6] = m * c;
        v[7] = m * (y * c - z * b);
        v[8] = m * (z * a - x * c);
        v[9] = m * (x * b - y * a);
        v[10] = 0.5 * m * (a * a + b * b + c * c);
        v[11] = pe ? pe[i] : 0.0;
        v[12] = m * (x * x + y * y + z * z);
        v[13] = 1.0;
    }
#pragma unroll
    for (int k = 0; k < kMoments; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < kMoments; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < kMoments) out[blockIdx.x * kMoments + threadIdx.x] = sh[threadIdx.x][0];
}

// Sum the block moments of each simulation in block order (one thread per
// simulation) and, for the recenter, turn them into the shift to apply.
__global__ void k_moments_final(const double* __restrict__ blocks, const int* __restrict__ sim_first_block,
                                int nsims, double* __restrict__ sums) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsims) return;
    double acc[kMoments];
#pragma unroll
    for (int k = 0; k < kMoments; ++k) acc[k] = 0.0;
    for (int b = sim_first_block[s]; b < sim_first_block[s + 1]; ++b)
#pragma unroll
        for (int k = 0; k < kMoments; ++k) acc[k] += blocks[b * kMoments + k];
#pragma unroll
    for (int k = 0; k < kMoments; ++k) sums[s * kMoments + k] = acc[k];
}

// FP64 mass moments of particles [0, n) (iron for i < nfe): per block
// {M, M x, M y, M z, M vx, M vy, M vz}, summed by k_moments_one_final.
constexpr int kMoments1 = 7;

__global__ __launch_bounds__(256) void k_moments_one(const float4* __restrict__ pos,
                                                     const float4* __restrict__ vel, int n, int nfe,
                                                     double mass_fe, double mass_si,
                                                     double* __restrict__ out) {
    __shared__ double sh[kMoments1][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v[kMoments1];
#pragma unroll
    for (int k = 0; k < kMoments1; ++k) v[k] = 0.0;
    if (i < n) {
        const double m = i < nfe ? mass_fe : mass_si;
        const float4 p = pos[i];
        const float4 u = vel[i];
        v[0] = m;
        v[1] = m * p.x;
        v[2] = m * p.y;
        v[3] = m * p.z;
        v[4] = m * u.x;
        v[5] = m * u.y;
        v[6] = m * u.z;
    }
#pragma unroll
    for (int k = 0; k < kMoments1; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < kMoments1; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < kMoments1) out[blockIdx.x * kMoments1 + threadIdx.x] = sh[threadIdx.x][0];
}

// One block: column k summed over the nb block rows in block order.
__global__ void k_moments_one_final(const double* __restrict__ blocks, int nb, double* __restrict__ sums) {
    const int k = threadIdx.x;
    if (k >= kMoments1) return;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += blocks[b * kMoments1 + k];
    sums[k] = acc;
}

// Recenter (impactAddTime.cu:433-455) with the means taken in FP64 on the
// device: positions += target_com - com, velocities += target_vel - mean_vel.
__global__ void k_recenter_apply(float4* __restrict__ pos, float4* __restrict__ vel, int n,
                                 const double* __restrict__ sums, double tcx, double tcy, double tcz,
                                 double tvx, double tvy, double tvz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double M = sums[0];
    const float sx = (float)(tcx - sums[1] / M);
    const float sy = (float)(tcy - sums[2] / M);
    const float sz = (float)(tcz - sums[3] / M);
    const float ux = (float)(tvx - sums[4] / M);
    const float uy = (float)(tvy - sums[5] / M);
    const float uz = (float)(tvz - sums[6] / M);
    float4 p = pos[i];
    float4 v = vel[i];
    p.x += sx;
    p.y += sy;
    p.z += sz;
    v.x += ux;
    v.y += uy;
    v.z += uz;
    pos[i] = p;
    vel[i] = v;
}

// The reference's own recenter (impactAddTime.cu:433-455), float for float.
// getMassCollision(0) / getCenterOfMassCollision(0) /
// getLinearVelocityCollision(0) (impactBasicFunctions.h:219-230,268-297,
// 388-417) accumulate in float on the host, particle by particle in index
// order: acc = float(double(acc) + double(x) * m) with m the double element
// mass (iron for i < nfe), and the mass the same way with x = 1.  Each of
// those seven sums is one sequential chain, so one workgroup stages the
// particles through LDS a tile at a time (coalesced loads by all its lanes)
// and seven lanes of its first wavefront run the chains side by side.  The
// result is written as the six float shifts target - mean that every
// particle then adds (k_recenter_ref_apply): bitwise the reference's host
// loop.
constexpr int kRecenterTile = 1024;

__global__ __launch_bounds__(256) void k_recenter_ref_sum(const float4* __restrict__ pos,
                                                          const float4* __restrict__ vel, int n, int nfe,
                                                          double mass_fe, double mass_si, float tcx, float tcy,
                                                          float tcz, float tvx, float tvy, float tvz,
                                                          float* __restrict__ shift) {
    __shared__ float tile[6][kRecenterTile];
    const int lane = threadIdx.x;
    float acc = 0.0f;  // lanes 0..5: sum m * (x, y, z, vx, vy, vz); lane 6: sum m
    for (int t0 = 0; t0 < n; t0 += kRecenterTile) {
        const int cnt = min(kRecenterTile, n - t0);
        __syncthreads();  // the previous tile's chains are done
        for (int k = lane; k < cnt; k += blockDim.x) {
            const float4 p = pos[t0 + k];
            const float4 v = vel[t0 + k];
            tile[0][k] = p.x;
            tile[1][k] = p.y;
            tile[2][k] = p.z;
            tile[3][k] = v.x;
            tile[4][k] = v.y;
            tile[5][k] = v.z;
        }
        __syncthreads();
        if (lane < 7) {
            for (int k = 0; k < cnt; ++k) {
                const double m = t0 + k < nfe ? mass_fe : mass_si;
                const double x = lane < 6 ? (double)tile[lane][k] : 1.0;
                acc = (float)((double)acc + x * m);
            }
        }
    }
    __shared__ float sums[7];
    if (lane < 7) sums[lane] = acc;
    __syncthreads();
    if (lane < 6) {
        // centerOfMass.x /= totalMass in float: the double quotient rounded
        // once to float is the correctly rounded float quotient (53 >= 2*24+2).
        const float mean = (float)((double)sums[lane] / (double)sums[6]);
        const float target = lane == 0 ? tcx : lane == 1 ? tcy : lane == 2 ? tcz : lane == 3 ? tvx : lane == 4 ? tvy : tvz;
        shift[lane] = target - mean;
    }
}

__global__ void k_recenter_ref_apply(float4* __restrict__ pos, float4* __restrict__ vel, int n,
                                     const float* __restrict__ shift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    float4 v = vel[i];
    p.x += shift[0];
    p.y += shift[1];
    p.z += shift[2];
    v.x += shift[3];
    v.y += shift[4];
    v.z += shift[5];
    pos[i] = p;
    vel[i] = v;
}

// Explicit instantiations used by the host code.
#define GI_GRAVITY_INST(R, POT, V) \
    template __global__ void k_gravity<R, 128, POT, V>(const float4*, const GravTask*, int, const SimDesc*, float4*);
GI_GRAVITY_INST(1, false, 0)
GI_GRAVITY_INST(1, false, 1)
GI_GRAVITY_INST(1, true, 0)
GI_GRAVITY_INST(2, false, 0)
GI_GRAVITY_INST(4, false, 0)
GI_GRAVITY_INST(8, false, 0)
GI_GRAVITY_INST(2, false, 1)
GI_GRAVITY_INST(4, false, 1)
GI_GRAVITY_INST(8, false, 1)
GI_GRAVITY_INST(2, true, 0)
GI_GRAVITY_INST(4, true, 0)
GI_GRAVITY_INST(8, true, 0)
#undef GI_GRAVITY_INST
#define GI_MUTUAL_INST(R, POT) \
    template __global__ void k_mutual<R, POT>(const float4*, const MutualTask*, int, const SimDesc*, float*, float*);
#define GI_DIRECT_INST(POT, S, MOVE)                                                                 \
    template __global__ void k_direct<POT, S, MOVE>(const float4*, const float4*, const DirectBlock*, \
                                                    const SimDesc*, float4*, double*, FusedMove);
#define GI_DIRECT_INST_S(S)       \
    GI_DIRECT_INST(false, S, 0)   \
    GI_DIRECT_INST(false, S, 1)   \
    GI_DIRECT_INST(false, S, 2)   \
    GI_DIRECT_INST(true, S, 0)
GI_DIRECT_INST_S(1)
GI_DIRECT_INST_S(2)
GI_DIRECT_INST_S(4)
GI_DIRECT_INST_S(8)
GI_MUTUAL_INST(1, false)
GI_MUTUAL_INST(2, false)
GI_MUTUAL_INST(4, false)
GI_MUTUAL_INST(8, false)
GI_MUTUAL_INST(1, true)
GI_MUTUAL_INST(2, true)
GI_MUTUAL_INST(4, true)
GI_MUTUAL_INST(8, true)
#undef GI_MUTUAL_INST
#define GI_MUTUAL_GRP_INST(R, POT)                                                                              \
    template __global__ void k_mutual_grp<R, POT>(const float4*, const GroupTask*, int, const MutualTask*,     \
                                                  const SimDesc*, const RowDesc*, float*, float*, float*);
GI_MUTUAL_GRP_INST(4, false)
GI_MUTUAL_GRP_INST(8, false)
GI_MUTUAL_GRP_INST(4, true)
GI_MUTUAL_GRP_INST(8, true)
#undef GI_MUTUAL_GRP_INST

}  // namespace gi

// gi_earth_moon.hip -- the Earth / Moon finder and the final statistics on the
// device (findEarthAndMoon and the get*Collision sums of
// Code/ImpactCode/impactBasicFunctions.h:109-217,219-652).
//
// The reference floods groups of elements closer than touch = 1.5 diameters
// on the host, O(N^2) per group, numbering each group by the index of its
// first (smallest) element; the Earth is the largest group and the Moon the
// next largest, ties going to the smaller number.  Here:
//   1. a hashed cell grid with cells a little larger than touch (cell
//      coordinates in FP64, so a touching pair always lies in adjacent cells);
//   2. union-find over the touching pairs found in the 27 neighbouring cells,
//      always hooking the larger root under the smaller (lock-free CAS), so
//      every root is the smallest index of its component -- the reference's
//      group number;
//   3. component sizes by integer atomics, the two largest by a 64-bit
//      atomicMax on (size, ~root);
//   4. FP64 moments (mass, mass-weighted position, momentum, position x
//      momentum) of the whole system, the Earth and the Moon, with element
//      counts by body and material, per block in a fixed order.
// Everything is deterministic: the components are unique, the counts are
// integers and the FP64 sums run in a fixed order.
#pragma once

#include "gi_device.h"

namespace gi {

// Cell coordinate of x for cells of size c, in FP64 (the float product x / c
// could misplace far-out elements by a cell), clamped like cell_coord.
__device__ __forceinline__ int cc_cell(float x, double inv_c) {
    double f = floor((double)x * inv_c);
    f = fmin(fmax(f, -1.0e9), 1.0e9);  // NaN -> -1e9
    return (int)f;
}

// The reference's touch test: mag = sqrt(dx*dx + dy*dy + dz*dz) < touch, all
// in float, products and sums rounded one by one
// (impactBasicFunctions.h:136-140).
__device__ __forceinline__ bool cc_touch(float4 a, float4 b, float touch) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    return sqrtf(r2) < touch;
}

__global__ void k_cc_count(const float4* __restrict__ pos, int n, double inv_c, unsigned hmask,
                           int* __restrict__ counts, int* __restrict__ cellhash) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    const unsigned h = cell_bucket(0, hmask, cc_cell(p.x, inv_c), cc_cell(p.y, inv_c), cc_cell(p.z, inv_c));
    cellhash[i] = (int)h;
    atomicAdd(&counts[h], 1);
}

__global__ void k_cc_init(int* __restrict__ parent, int* __restrict__ size, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    parent[i] = i;
    size[i] = 0;
}

__device__ __forceinline__ int cc_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x.  Parents only ever point to smaller indices, so the walk ends.
__device__ __forceinline__ int cc_find(int* parent, int x) {
    int p = cc_load(parent + x);
    while (p != x) {
        const int gp = cc_load(parent + p);
        if (gp != p) {
            // path halving; a racing update can only have made parent[x]
            // smaller, and a lost write is harmless
            __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void cc_union(int* parent, int a, int b) {
    while (true) {
        int ra = cc_find(parent, a);
        int rb = cc_find(parent, b);
        if (ra == rb) return;
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        // hook the larger root under the smaller one
        int expected = ra;
        if (__hip_atomic_compare_exchange_strong(parent + ra, &expected, rb, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        a = ra;
        b = rb;
    }
}

// Every touching pair (i, j > i) found in the 27 cells around i is united.
__global__ __launch_bounds__(256) void k_cc_union(const float4* __restrict__ pos, int n, double inv_c,
                                                  unsigned hmask, float touch, const int* __restrict__ starts,
                                                  const int* __restrict__ sorted, int* __restrict__ parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = pos[i];
    const int cx = cc_cell(pi.x, inv_c), cy = cc_cell(pi.y, inv_c), cz = cc_cell(pi.z, inv_c);
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int qx = cx + ox, qy = cy + oy, qz = cz + oz;
                const unsigned h = cell_bucket(0, hmask, qx, qy, qz);
                for (int k = starts[h]; k < starts[h + 1]; ++k) {
                    const int j = sorted[k];
                    if (j <= i) continue;
                    const float4 pj = pos[j];
                    if (cc_cell(pj.x, inv_c) != qx || cc_cell(pj.y, inv_c) != qy || cc_cell(pj.z, inv_c) != qz)
                        continue;  // another cell sharing the bucket
                    if (cc_touch(pi, pj, touch)) cc_union(parent, i, j);
                }
            }
}

// Root of x without path compression: after the unions the forest is fixed,
// and the only stores to parent are each thread's own final root.
__device__ __forceinline__ int cc_root(const int* parent, int x) {
    int p = cc_load(parent + x);
    while (p != x) {
        x = p;
        p = cc_load(parent + x);
    }
    return x;
}

// root[i] = the smallest index of i's component; size[root]++.  The walk
// must not compress paths: a path-splitting store from another thread's walk
// (a grandparent, not necessarily the root) could land after i's own root
// store and leave parent[i] pointing at an inner node.  Reading a parent
// that another thread has just replaced by its root is harmless: both are
// ancestors.
__global__ void k_cc_compress(int* __restrict__ parent, int* __restrict__ size, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = cc_root(parent, i);
    __hip_atomic_store(parent + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd(&size[r], 1);
}

// best = max over roots (other than `skip`) of (size << 32 | ~root): the
// largest component, ties to the smaller root.
__global__ void k_cc_top(const int* __restrict__ size, int n, int skip, unsigned long long* __restrict__ best) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i == skip) return;
    const int s = size[i];
    if (s == 0) return;
    const unsigned long long key = ((unsigned long long)(unsigned)s << 32) | (unsigned long long)(~(unsigned)i);
    atomicMax(best, key);
}

// Moments of one class: mass, sum m x (3), sum m v (3), sum m x cross v (3).
constexpr int kCCMoments = 10;
constexpr int kCCClasses = 3;  // 0 whole system, 1 Earth, 2 Moon
constexpr int kCCCounts = 8;   // Earth then Moon: Fe body 1, Fe body 2, Si body 1, Si body 2
constexpr int kCCBlock = 256;

// Per block: the moments of the system, the Earth and the Moon (FP64, a fixed
// grid-stride order and a fixed tree) and the element counts of the Earth and
// the Moon by body and material.  The impact layout is [Fe1, Fe2, Si1, Si2]
// (impactBasicFunctions.h:910-924).
__global__ __launch_bounds__(kCCBlock) void k_cc_moments(const float4* __restrict__ pos,
                                                         const float4* __restrict__ vel, int n,
                                                         const int* __restrict__ root, int earth, int moon,
                                                         int nfe1, int nfe2, int nsi1, double mass_fe,
                                                         double mass_si, double* __restrict__ out,
                                                         int* __restrict__ counts) {
    __shared__ double red[kCCBlock];
    __shared__ int cnt[kCCCounts];
    double m[kCCClasses][kCCMoments];
#pragma unroll
    for (int c = 0; c < kCCClasses; ++c)
#pragma unroll
        for (int k = 0; k < kCCMoments; ++k) m[c][k] = 0.0;
    if (threadIdx.x < kCCCounts) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int nfe = nfe1 + nfe2;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double mi = i < nfe ? mass_fe : mass_si;
        const float4 p = pos[i], v = vel[i];
        const double x = p.x, y = p.y, z = p.z, vx = v.x, vy = v.y, vz = v.z;
        const double t[kCCMoments] = {mi,          mi * x,      mi * y,      mi * z,      mi * vx,
                                      mi * vy,     mi * vz,     mi * (y * vz - z * vy),
                                      mi * (z * vx - x * vz),   mi * (x * vy - y * vx)};
        const int r = root[i];
        const int cls = r == earth ? 1 : (r == moon ? 2 : 0);
#pragma unroll
        for (int k = 0; k < kCCMoments; ++k) {
            m[0][k] += t[k];
            if (cls == 1) m[1][k] += t[k];
            if (cls == 2) m[2][k] += t[k];
        }
        if (cls != 0) {
            const int kind = i < nfe1 ? 0 : (i < nfe ? 1 : (i < nfe + nsi1 ? 2 : 3));
            atomicAdd(&cnt[(cls - 1) * 4 + kind], 1);
        }
    }
    for (int c = 0; c < kCCClasses; ++c)
        for (int k = 0; k < kCCMoments; ++k) {
            red[threadIdx.x] = m[c][k];
            __syncthreads();
            for (int s = kCCBlock / 2; s > 0; s >>= 1) {
                if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) out[(size_t)blockIdx.x * kCCClasses * kCCMoments + c * kCCMoments + k] = red[0];
            __syncthreads();
        }
    if (threadIdx.x < kCCCounts) counts[blockIdx.x * kCCCounts + threadIdx.x] = cnt[threadIdx.x];
}

}  // namespace gi

// gi_device.h -- device-side types and the exact pair law shared by the HIP
// kernels of the Giant-Impact step loop (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gi {

// Constants of one simulation: the pair law of forceCollisionKernalConstantsStruct
// (Code/ImpactCode/impactAddTime.cu:4-20) without the material split.
struct PairLaw {
    float GMassFeFe, GMassFeSi, KFeFe, KSiSi, KFeSi, KRFeFe, KRSiSi, KRFeSi, KRMix;
    float ShellBreakFe, ShellBreakSi, ShellBreakFeSi1, ShellBreakFeSi2;
};

// One simulation inside a packed particle array.  Particles [start, end);
// iron is [start, fe_end) (material by index, as the reference's id < NFe and
// boarder tests).  Partial sums of the all-pairs pass for particle i and
// j-chunk c live at partial[partial_base + c * (end - start) + (i - start)].
// The contact grid of the simulation owns buckets [hbase, hbase + hmask + 1).
struct SimDesc {
    int start, fe_end, end;
    int nchunk;
    int partial_base;
    int active;
    int hbase;
    unsigned hmask;
    PairLaw law;
    float dt, dt_over_mass_fe, dt_over_mass_si;
    float pad2;
    // Mutual all-pairs pass: the simulation's copy in the padded layout
    // (iron then silicate, each padded with ghosts to a whole number of
    // 64*R rows), its first row in the row table, its j-side slab and (the
    // grouped form) its within-group slab.
    int gbase, gfe, gn, row0;
    long long jbase;
    long long jibase;
};

// One wavefront of the mutual all-pairs pass: row `row` (64*R resident
// particles of the padded layout) against j in [j0, j1) of the same
// simulation, all of material jfe.  diag: the row against itself, one-sided;
// otherwise j lies beyond the row and each pair is evaluated once for both.
struct MutualTask {
    int sim, row, j0, j1, slot, jfe, diag, pad;
};

// One workgroup of the grouped mutual pass.  kind 0: the rows group*W ..
// group*W + W - 1 of simulation sim (wave w holds row group*W + w; waves past
// the simulation's last row only keep step) against j in [j0, j1) beyond the
// group, all of material jfe; `piece` numbers this j-range among the group's,
// so each row's i-side slot is rows[row].slot0 + piece.  kind 1: W
// independent single-wave row tasks rtask0 .. rtask0 + W - 1 of the row-task
// table (a row against itself, or against the rest of its group).
struct GroupTask {
    int kind, sim, group, j0, j1, piece, jfe, rtask0;
};

// One workgroup of the direct pass: 64 i-particles [i0, i0 + 64) of
// simulation sim (the last block of a simulation may be ragged).
struct DirectBlock {
    int sim, i0;
};

// The i-side slots of one row: slot0 .. slot0 + nslot - 1, in task order.
struct RowDesc {
    int slot0, nslot;
};

// One workgroup of k_mutual_sum: the 64 padded places [q0, q0 + 64) of
// simulation sim (inside one row, as 64 divides the row length).
struct SumBlock {
    int sim, q0;
};

// Ghost position of the padded layout: every pair with a real particle has
// r^-3 = 0 in float (r^2 ~ 3e36), and a ghost pair has dp = 0.
constexpr float kGhostCoord = 1e18f;
// Any r^2 at or above this is a ghost pair (real separations are far below).
constexpr float kGhostR2 = 1e30f;

// One wavefront of work for the all-pairs pass: 64*R i-particles starting at
// i0 of simulation sim, against j in [j0, j1) (all of one material: jfe).
struct GravTask {
    int i0, sim, j0, j1, chunk, jfe, pad0, pad1;
};

// Added to r^2 in the all-pairs pass so the self pair (dp = 0) contributes
// 0 * finite instead of 0 * inf.  For any pair with r^2 > 2^-40 it is below
// half an ulp of every partial sum of r^2 and changes nothing.
constexpr float kSelfEps = 1e-20f;

// Spreads the low 10 bits of v to every third bit (bit k -> bit 3k).
__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Bucket of an integer cell (cell size 1 = the element diameter) in the
// simulation's own range of the contact grid: the Morton code of the cell,
// so buckets that are close in memory hold cells that are close in space.
// Distinct cells may share a bucket (codes taken modulo the range); readers
// keep only the particles of the exact cell they look for.
__device__ __forceinline__ unsigned cell_bucket(int hbase, unsigned hmask, int cx, int cy, int cz) {
    const unsigned b = 1u << 20;  // bias: negative coordinates wrap like positive ones
    const unsigned m = spread3((unsigned)cx + b) | (spread3((unsigned)cy + b) << 1) |
                       (spread3((unsigned)cz + b) << 2);
    return (unsigned)hbase + (m & hmask);
}

__device__ __forceinline__ int cell_coord(float x) {
    float f = floorf(x);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);  // NaN -> -1e9 (fmaxf drops NaN)
    return (int)f;
}

// The reference's contact expression "G - K*(1.0 - r2)": the double literal
// promotes it to FP64 and the result is narrowed to float
// (impactAddTime.cu:306,314-315,326,334-335,343-344,355,363-364).  The
// library is built with -ffp-contract=off, so this is mul then sub.
__device__ __forceinline__ float contact_mag(float G, float K, float r2) {
    return (float)((double)G - (double)K * (1.0 - (double)r2));
}

// Pair force magnitude of getForcesCollisionSingleGPU (impactAddTime.cu:298-366)
// for a pair already known to be in contact (r < 1).  test = number of irons.
// The receding test (int)(dp.dv) > 0 holds exactly when dp.dv >= 1.0f
// (truncation toward zero; NaN converts to 0).
__device__ __forceinline__ float contact_pair_mag(const PairLaw& L, int test, float r, float r2,
                                                  float dot) {
    const bool receding = dot >= 1.0f;
    if (test == 0) {
        if (L.ShellBreakSi <= r) return contact_mag(1.0f, L.KSiSi, r2);
        return contact_mag(1.0f, receding ? L.KRSiSi : L.KSiSi, r2);
    } else if (test == 1) {
        if (L.ShellBreakFeSi1 <= r) return contact_mag(L.GMassFeSi, L.KFeSi, r2);
        if (L.ShellBreakFeSi2 <= r) return contact_mag(L.GMassFeSi, receding ? L.KRMix : L.KFeSi, r2);
        return contact_mag(L.GMassFeSi, receding ? L.KRFeSi : L.KFeSi, r2);
    } else {
        if (L.ShellBreakFe <= r) return contact_mag(L.GMassFeFe, L.KFeFe, r2);
        return contact_mag(L.GMassFeFe, receding ? L.KRFeFe : L.KFeFe, r2);
    }
}

// Gravity product G mi mj of a material pair as the force kernel uses it.
__device__ __forceinline__ float gravity_product(const PairLaw& L, int test) {
    return test == 0 ? 1.0f : (test == 1 ? L.GMassFeSi : L.GMassFeFe);
}
__device__ __forceinline__ float approach_k(const PairLaw& L, int test) {
    return test == 0 ? L.KSiSi : (test == 1 ? L.KFeSi : L.KFeFe);
}

// Compensated float sum (Neumaier's variant of Kahan summation): the
// rounding error of every addition, whichever operand is the larger, is
// carried in c, so a long sum of floats keeps float precision of its result
// instead of losing it with the number of terms.  The force path accumulates
// in it; FP64 is kept for the energy and angular-momentum reductions.
struct CompSum {
    float s, c;
    __device__ __forceinline__ void zero() { s = c = 0.0f; }
    __device__ __forceinline__ void add(float x) {
        const float t = s + x;
        c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    // another compensated sum, its compensation carried over
    __device__ __forceinline__ void add_sum(const CompSum& o) {
        add(o.s);
        c += o.c;
    }
    // the product a * b, its rounding error (exact, by FMA) carried over
    __device__ __forceinline__ void add_product(float a, float b) {
        const float p = a * b;
        add(p);
        c += __builtin_fmaf(a, b, -p);
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

// r^-3 as the all-pairs pass computes it (one hardware reciprocal square
// root, two multiplies) from r^2 + eps accumulated with FMAs.
//
// The reciprocal root is clamped to [0, 1] (the output clamp bit of
// v_rsq_f32 / v_rcp_f32; fmed3 against 0 and 1 folds into it: no extra
// instruction).  For r >= 1 that leaves the value as it is; for r < 1 -- a pair in contact,
// whose term the contact correction takes out again and replaces by the
// contact law -- the term becomes the bounded placeholder g dp (|dp| < 1)
// instead of g dp / r^3.  The float sums of the all-pairs pass therefore
// never carry a 1/r^2-sized term of a close pair, whose rounding would
// survive its exact removal; the self pair (r^2 = kSelfEps) gives 1 * 0 = 0.
// Every correction subtracts this same function's value for the pair.
__device__ __forceinline__ float fast_r2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, kSelfEps)));
}
__device__ __forceinline__ float clamp_unit(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }
// The three instructions are one asm block so that the multiplies follow the
// v_rsq_f32 directly (after the one wait state a transcendental result
// needs): left to itself the scheduler puts independent instructions between
// a v_rsq_f32 and its consumer for about a quarter of the pairs, and on
// gfx950 that costs more issue time than the wait (the mutual pass ran 2-5%
// slower; putting independent work between every v_rsq_f32 and its consumer
// made it 22-24% slower).  Same instructions, same operands, same value as
// clamp_unit(rsq(r2))^3.  Registers only.
__device__ __forceinline__ float fast_inv_r3(float r2) {
    float ri, sq;
    asm("v_rsq_f32_e64 %0, %2 clamp\n\t"
        "s_nop 0\n\t"
        "v_mul_f32_e32 %1, %0, %0\n\t"
        "v_mul_f32_e32 %0, %0, %1"
        : "=&v"(ri), "=&v"(sq)
        : "v"(r2));
    return ri;
}

// Variant 1: r^-3 = rsq(r^2) * rcp(r^2): one multiply and two transcendental
// ops, which issue beside the ordinary VALU stream (both clamped as above).
__device__ __forceinline__ float fast_inv_r3_rcp(float r2) {
    return clamp_unit(__builtin_amdgcn_rsqf(r2)) * clamp_unit(__builtin_amdgcn_rcpf(r2));
}

template <int V>
__device__ __forceinline__ float inv_r3(float r2) {
    if constexpr (V == 1)
        return fast_inv_r3_rcp(r2);
    else
        return fast_inv_r3(r2);
}

__device__ __forceinline__ float inv_r3_variant(int v, float r2) {
    return v == 1 ? fast_inv_r3_rcp(r2) : fast_inv_r3(r2);
}

}  // namespace gi

/*
 * gi_oracle.c -- CPU restatement of the Giant-Impact step loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the timed CPU baseline).  The
 * shipped library never links or calls this file.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  With
 * contraction off every float expression below rounds in the order written,
 * which is the canonical reading of the reference's source.
 */
#include "gi_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Physical constants of Code/CommonCode/commonSetup.h:29-33. */
static const double OR_PI = 3.141592654;
static const double OR_UNIVERSAL_GRAVITY = 6.67408e-20;
static const double OR_MASS_OF_EARTH = 5.97219e24;

/* (int) of a float as the reference's receding test does it
 * (impactAddTime.cu:313): truncation toward zero; NaN gives 0 and
 * out-of-range values saturate, as CUDA's cvt.rzi.s32.f32 does. */
int or_trunc_to_int(float v)
{
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
}

/* ------------------------------------------------------------------------ */
/* Units and constants                                                       */
/* ------------------------------------------------------------------------ */

/* setTargetImpactorParameters, generateTargetImpactor.cu:289-403. */
int or_units_compute(or_units *u)
{
    double massOfBody1 = OR_MASS_OF_EARTH * u->fraction_earth_mass_body1;
    double massOfBody2 = OR_MASS_OF_EARTH * u->fraction_earth_mass_body2;
    /* :295-304 */
    if (u->fraction_fe_body1 + u->fraction_si_body1 != 1.0) return -1;
    if (u->fraction_fe_body2 + u->fraction_si_body2 != 1.0) return -2;
    /* :307-312 */
    double totalMassOfFeBody1 = u->fraction_fe_body1 * massOfBody1;
    double totalMassOfSiBody1 = u->fraction_si_body1 * massOfBody1;
    double totalMassOfFeBody2 = u->fraction_fe_body2 * massOfBody2;
    double totalMassOfSiBody2 = u->fraction_si_body2 * massOfBody2;
    double totalMassOfFe = totalMassOfFeBody1 + totalMassOfFeBody2;
    double totalMassOfSi = totalMassOfSiBody1 + totalMassOfSiBody2;
    int N = u->total_number_of_elements;
    int NFe, NSi, NFe1, NFe2, NSi1, NSi2;
    /* :319-331 (assignments to int truncate) */
    if (totalMassOfFe != 0.0)
        NFe = (int)((double)N * (u->density_si / u->density_fe) /
                    (totalMassOfSi / totalMassOfFe + u->density_si / u->density_fe));
    else
        NFe = 0;
    NSi = N - NFe;
    if (totalMassOfFe != 0.0)
        NFe1 = (int)(NFe * totalMassOfFeBody1 / totalMassOfFe);
    else
        NFe1 = 0;
    NFe2 = NFe - NFe1;
    if (totalMassOfSi != 0.0)
        NSi1 = (int)(NSi * totalMassOfSiBody1 / totalMassOfSi);
    else
        NSi1 = 0;
    NSi2 = NSi - NSi1;
    /* :334-349 */
    double massFe = (NFe != 0) ? totalMassOfFe / NFe : 0.0;
    double massSi = (NSi != 0) ? totalMassOfSi / NSi : 0.0;
    double diameterOfElement;
    if (NSi != 0)
        diameterOfElement = pow((6.0 * massSi) / (OR_PI * u->density_si), (1.0 / 3.0));
    else
        diameterOfElement = pow((6.0 * massFe) / (OR_PI * u->density_fe), (1.0 / 3.0));
    double unitLength = diameterOfElement;
    double unitMass = (NSi != 0) ? massSi : massFe;
    double unitTime;
    /* :352-358 */
    if (NSi != 0)
        unitTime = sqrt((6.0 * massSi * (double)NSi) /
                        (OR_UNIVERSAL_GRAVITY * OR_PI * u->density_si * totalMassOfSi));
    else if (NFe != 0)
        unitTime = sqrt((6.0 * massFe * (double)NFe) /
                        (OR_UNIVERSAL_GRAVITY * OR_PI * u->density_fe * totalMassOfFe));
    else
        return -3;
    /* :362-378 */
    double MassSi = -1.0, MassFe = -1.0;
    if (NSi != 0) {
        MassSi = 1.0;
        MassFe = u->density_fe / u->density_si;
    } else {
        MassFe = 1.0;
    }
    u->nfe = NFe;
    u->nsi = NSi;
    u->nfe1 = NFe1;
    u->nsi1 = NSi1;
    u->nfe2 = NFe2;
    u->nsi2 = NSi2;
    u->unit_length = unitLength;
    u->unit_mass = unitMass;
    u->unit_time = unitTime;
    u->mass_fe = MassFe;
    u->mass_si = MassSi;
    u->diameter = 1.0;
    /* :381-382, :394-395 ("KFe *= a*b*c/d" evaluates the right side first) */
    u->mass_of_body1 = massOfBody1 / unitMass;
    u->mass_of_body2 = massOfBody2 / unitMass;
    u->k_fe_sim = u->k_fe * (unitTime * unitTime * unitLength / unitMass);
    u->k_si_sim = u->k_si * (unitTime * unitTime * unitLength / unitMass);
    return 0;
}

/* fprintf("%f") followed by an istream >> double read. */
double or_round_f(double v)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%f", v);
    return strtod(buf, NULL);
}

/* impactAddTime.cu:128-170 after the %f round trip of
 * generateTargetImpactorBasicFunctions.h:693-709 / impactBasicFunctions.h:39-75. */
void or_impact_constants(const or_units *u, or_force_collision_constants *fc,
                         or_move_collision_constants *mc)
{
    double MassFe = or_round_f(u->mass_fe);
    double MassSi = or_round_f(u->mass_si);
    double Diameter = or_round_f(u->diameter);
    double KFe = or_round_f(u->k_fe_sim);
    double KSi = or_round_f(u->k_si_sim);
    double KRFe = or_round_f(u->kr_fe);
    double KRSi = or_round_f(u->kr_si);
    double SDFe = or_round_f(u->sd_fe);
    double SDSi = or_round_f(u->sd_si);
    double Dt = u->dt; /* read from setupImpactAddTime, not rounded */
    fc->GMassFeFe = (float)(MassFe * MassFe);
    fc->GMassFeSi = (float)MassFe;
    fc->KFeFe = (float)(2.0 * KFe);
    fc->KSiSi = (float)(2.0 * KSi);
    fc->KFeSi = (float)(KFe + KSi);
    fc->KRFeFe = (float)(2.0 * KFe * KRFe);
    fc->KRSiSi = (float)(2.0 * KSi * KRSi);
    fc->KRFeSi = (float)(KFe * KRFe + KSi * KRSi);
    if (SDFe >= SDSi)
        fc->KRMix = (float)(KFe + KSi * KRSi);
    else
        fc->KRMix = (float)(KFe * KRFe + KSi);
    fc->ShellBreakFe = (float)(Diameter - Diameter * SDFe);
    fc->ShellBreakSi = (float)(Diameter - Diameter * SDSi);
    if (SDFe >= SDSi) {
        fc->ShellBreakFeSi1 = (float)(Diameter - Diameter * SDSi);
        fc->ShellBreakFeSi2 = (float)(Diameter - Diameter * SDFe);
    } else {
        fc->ShellBreakFeSi1 = (float)(Diameter - Diameter * SDFe);
        fc->ShellBreakFeSi2 = (float)(Diameter - Diameter * SDSi);
    }
    fc->NFe = u->nfe1 + u->nfe2; /* impactBasicFunctions.h:105 */
    mc->Dt = (float)Dt;
    mc->DtOverMassSi = (float)(Dt / MassSi);
    mc->DtOverMassFe = (float)(Dt / MassFe);
    mc->NFe = u->nfe1 + u->nfe2;
}

/* generateTargetImpactor.cu:406-451 (Gravity = 1, unrounded values). */
void or_separate_constants(const or_units *u, or_force_separate_constants *fc,
                           or_move_separate_constants *mc)
{
    double Gravity = 1.0, MassFe = u->mass_fe, MassSi = u->mass_si, Diameter = u->diameter;
    double KFe = u->k_fe_sim, KSi = u->k_si_sim, KRFe = u->kr_fe, KRSi = u->kr_si;
    double SDFe = u->sd_fe, SDSi = u->sd_si, Dt = u->dt;
    fc->GMassFeFe = (float)(Gravity * MassFe * MassFe);
    fc->GMassFeSi = (float)(Gravity * MassFe * MassSi);
    fc->KFeFe = (float)(2.0 * KFe);
    fc->KSiSi = (float)(2.0 * KSi);
    fc->KFeSi = (float)(KFe + KSi);
    fc->KRFeFe = (float)(2.0 * KFe * KRFe);
    fc->KRSiSi = (float)(2.0 * KSi * KRSi);
    fc->KRFeSi = (float)(KFe * KRFe + KSi * KRSi);
    if (SDFe >= SDSi)
        fc->KRMix = (float)(KFe + KSi * KRSi);
    else
        fc->KRMix = (float)(KFe * KRFe + KSi);
    fc->ShellBreakFe = (float)(Diameter - Diameter * SDFe);
    fc->ShellBreakSi = (float)(Diameter - Diameter * SDSi);
    if (SDFe >= SDSi) {
        fc->ShellBreakFeSi1 = (float)(Diameter - Diameter * SDSi);
        fc->ShellBreakFeSi2 = (float)(Diameter - Diameter * SDFe);
    } else {
        fc->ShellBreakFeSi1 = (float)(Diameter - Diameter * SDFe);
        fc->ShellBreakFeSi2 = (float)(Diameter - Diameter * SDSi);
    }
    fc->boarder1 = u->nfe1;
    fc->boarder2 = u->nfe1 + u->nsi1;
    fc->boarder3 = u->nfe1 + u->nsi1 + u->nfe2;
    mc->Dt = (float)Dt;
    mc->DtOverMassFe = (float)(Dt / MassFe);
    mc->DtOverMassSi = (float)(Dt / MassSi);
    mc->boarder1 = u->nfe1;
    mc->boarder2 = u->nsi1 + u->nfe1;
    mc->boarder3 = u->nfe1 + u->nsi1 + u->nfe2;
}

/* ------------------------------------------------------------------------ */
/* The pair force of impactAddTime.cu:298-366, float path                   */
/* ------------------------------------------------------------------------ */

/* "G - K*(1.0 - r2)" with the double literal promoting the expression
 * (impactAddTime.cu:306,314-315,326,334-335,343-344,355,363-364). */
static inline float contact_f32(float G, float K, float r2)
{
    return (float)((double)G - (double)K * (1.0 - (double)r2));
}

/* Pair force magnitude; test = number of irons in the pair (0 Si-Si, 1 Fe-Si,
 * 2 Fe-Fe); dvx.. are only read on the receding branches. */
static inline float pair_mag_f32(int test, float r, float r2, float dpx, float dpy, float dpz,
                                 float dvx, float dvy, float dvz, float GMassFeFe,
                                 float GMassFeSi, float KFeFe, float KSiSi, float KFeSi,
                                 float KRFeFe, float KRSiSi, float KRFeSi, float KRMix,
                                 float ShellBreakFe, float ShellBreakSi, float ShellBreakFeSi1,
                                 float ShellBreakFeSi2)
{
    float force_mag;
    int inout;
    if (test == 0) { /* :298-317 */
        if (1.0 <= r) {
            force_mag = (float)(1.0 / (double)r2);
        } else if (ShellBreakSi <= r) {
            force_mag = contact_f32(1.0f, KSiSi, r2);
        } else {
            inout = or_trunc_to_int(dpx * dvx + dpy * dvy + dpz * dvz);
            if (inout <= 0)
                force_mag = contact_f32(1.0f, KSiSi, r2);
            else
                force_mag = contact_f32(1.0f, KRSiSi, r2);
        }
    } else if (test == 1) { /* :318-346 */
        if (1.0 <= r) {
            force_mag = GMassFeSi / r2;
        } else if (ShellBreakFeSi1 <= r) {
            force_mag = contact_f32(GMassFeSi, KFeSi, r2);
        } else if (ShellBreakFeSi2 <= r) {
            inout = or_trunc_to_int(dpx * dvx + dpy * dvy + dpz * dvz);
            if (inout <= 0)
                force_mag = contact_f32(GMassFeSi, KFeSi, r2);
            else
                force_mag = contact_f32(GMassFeSi, KRMix, r2);
        } else {
            inout = or_trunc_to_int(dpx * dvx + dpy * dvy + dpz * dvz);
            if (inout <= 0)
                force_mag = contact_f32(GMassFeSi, KFeSi, r2);
            else
                force_mag = contact_f32(GMassFeSi, KRFeSi, r2);
        }
    } else { /* :347-366 */
        if (1.0 <= r) {
            force_mag = GMassFeFe / r2;
        } else if (ShellBreakFe <= r) {
            force_mag = contact_f32(GMassFeFe, KFeFe, r2);
        } else {
            inout = or_trunc_to_int(dpx * dvx + dpy * dvy + dpz * dvz);
            if (inout <= 0)
                force_mag = contact_f32(GMassFeFe, KFeFe, r2);
            else
                force_mag = contact_f32(GMassFeFe, KRFeFe, r2);
        }
    }
    return force_mag;
}

#define OR_OMP_THREADS(t) ((t) > 0 ? (t) : 1)

/* getForcesCollisionSingleGPU, impactAddTime.cu:245-377.  One "thread" per
 * id in [i_begin, i_end); j runs 0..n-1 in order, as the 256-wide LDS tiles
 * visit it (:277-283). */
void or_force_collision_f32(const or_float4 *pos, const or_float4 *vel, or_float4 *force, int n,
                            const or_force_collision_constants *c, int i_begin, int i_end,
                            int threads)
{
    const or_force_collision_constants k = *c;
    if (i_begin < 0) i_begin = 0;
    if (i_end > n) i_end = n;
#pragma omp parallel for schedule(dynamic, 16) num_threads(OR_OMP_THREADS(threads))
    for (int id = i_begin; id < i_end; id++) {
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        const float px = pos[id].x, py = pos[id].y, pz = pos[id].z;
        const float vx = vel[id].x, vy = vel[id].y, vz = vel[id].z;
        for (int ids = 0; ids < n; ids++) {
            float dpx = pos[ids].x - px;
            float dpy = pos[ids].y - py;
            float dpz = pos[ids].z - pz;
            float r2 = dpx * dpx + dpy * dpy + dpz * dpz;
            float r = sqrtf(r2);
            float invr = (id == ids) ? 0.0f : 1.0f / r;
            int test = 0;
            if (id < k.NFe) test = 1;
            if (ids < k.NFe) test++;
            float fm = pair_mag_f32(test, r, r2, dpx, dpy, dpz, vel[ids].x - vx, vel[ids].y - vy,
                                    vel[ids].z - vz, k.GMassFeFe, k.GMassFeSi, k.KFeFe, k.KSiSi,
                                    k.KFeSi, k.KRFeFe, k.KRSiSi, k.KRFeSi, k.KRMix,
                                    k.ShellBreakFe, k.ShellBreakSi, k.ShellBreakFeSi1,
                                    k.ShellBreakFeSi2);
            fx += fm * dpx * invr;
            fy += fm * dpy * invr;
            fz += fm * dpz * invr;
        }
        force[id].x = fx;
        force[id].y = fy;
        force[id].z = fz;
    }
}

/* moveBodiesCollisionSingleGPU, impactAddTime.cu:379-394. */
void or_move_collision_f32(or_float4 *pos, or_float4 *vel, const or_float4 *force, int n,
                           const or_move_collision_constants *c)
{
    for (int id = 0; id < n; id++) {
        float temp = (id < c->NFe) ? c->DtOverMassFe : c->DtOverMassSi;
        vel[id].x += (force[id].x) * temp;
        vel[id].y += (force[id].y) * temp;
        vel[id].z += (force[id].z) * temp;
        pos[id].x += vel[id].x * c->Dt;
        pos[id].y += vel[id].y * c->Dt;
        pos[id].z += vel[id].z * c->Dt;
    }
}

/* getForcesSeperate, generateTargetImpactor.cu:486-626.  Pairs across the
 * body border boarder2 are skipped (:529); the iron test uses boarder1 for
 * body 1 and boarder3 for body 2 (:531-544). */
void or_force_separate_f32(const or_float4 *pos, const or_float4 *vel, or_float4 *force, int n,
                           const or_force_separate_constants *c, int threads)
{
    const or_force_separate_constants k = *c;
#pragma omp parallel for schedule(dynamic, 16) num_threads(OR_OMP_THREADS(threads))
    for (int id = 0; id < n; id++) {
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        const float px = pos[id].x, py = pos[id].y, pz = pos[id].z;
        const float vx = vel[id].x, vy = vel[id].y, vz = vel[id].z;
        for (int ids = 0; ids < n; ids++) {
            if ((id < k.boarder2 && ids < k.boarder2) || (k.boarder2 <= id && k.boarder2 <= ids)) {
                int materialSwitch = 0;
                if ((id < k.boarder2) && (ids < k.boarder2)) materialSwitch = k.boarder1;
                if ((k.boarder2 <= id) && (k.boarder2 <= ids)) materialSwitch = k.boarder3;
                float dpx = pos[ids].x - px;
                float dpy = pos[ids].y - py;
                float dpz = pos[ids].z - pz;
                float r2 = dpx * dpx + dpy * dpy + dpz * dpz;
                float r = sqrtf(r2);
                float invr = (id == ids) ? 0.0f : 1.0f / r;
                int test = 0;
                if (id < materialSwitch) test = 1;
                if (ids < materialSwitch) test++;
                float fm = pair_mag_f32(test, r, r2, dpx, dpy, dpz, vel[ids].x - vx,
                                        vel[ids].y - vy, vel[ids].z - vz, k.GMassFeFe,
                                        k.GMassFeSi, k.KFeFe, k.KSiSi, k.KFeSi, k.KRFeFe,
                                        k.KRSiSi, k.KRFeSi, k.KRMix, k.ShellBreakFe,
                                        k.ShellBreakSi, k.ShellBreakFeSi1, k.ShellBreakFeSi2);
                fx += fm * dpx * invr;
                fy += fm * dpy * invr;
                fz += fm * dpz * invr;
            }
        }
        /* The reference stores after every tile (:621-623); the last store is the full sum. */
        force[id].x = fx;
        force[id].y = fy;
        force[id].z = fz;
    }
}

/* moveBodiesDampedSeperate, generateTargetImpactor.cu:628-664. */
void or_move_damped_separate_f32(or_float4 *pos, or_float4 *vel, const or_float4 *force, int n,
                                 const or_move_separate_constants *c, float damp1, float damp2)
{
    for (int id = 0; id < n; id++) {
        float temp, damp;
        if (c->boarder3 <= id) {
            temp = c->DtOverMassSi;
            damp = damp2;
        } else if (c->boarder2 <= id) {
            temp = c->DtOverMassFe;
            damp = damp2;
        } else if (c->boarder1 <= id) {
            temp = c->DtOverMassSi;
            damp = damp1;
        } else {
            temp = c->DtOverMassFe;
            damp = damp1;
        }
        vel[id].x += (force[id].x - damp * vel[id].x) * temp;
        vel[id].y += (force[id].y - damp * vel[id].y) * temp;
        vel[id].z += (force[id].z - damp * vel[id].z) * temp;
        pos[id].x += vel[id].x * c->Dt;
        pos[id].y += vel[id].y * c->Dt;
        pos[id].z += vel[id].z * c->Dt;
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 twins                                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    double G[3];  /* Si-Si, Fe-Si, Fe-Fe gravity products */
    double K[3];  /* approach stiffness */
    double KR[3]; /* receding stiffness in the deepest region */
    double KRMix;
    double sbSi, sbFe, sbFeSi1, sbFeSi2;
} pair_law64;

static pair_law64 law_from(float GMassFeFe, float GMassFeSi, float KFeFe, float KSiSi,
                           float KFeSi, float KRFeFe, float KRSiSi, float KRFeSi, float KRMix,
                           float sbFe, float sbSi, float sbFeSi1, float sbFeSi2)
{
    pair_law64 L;
    L.G[0] = 1.0;
    L.G[1] = GMassFeSi;
    L.G[2] = GMassFeFe;
    L.K[0] = KSiSi;
    L.K[1] = KFeSi;
    L.K[2] = KFeFe;
    L.KR[0] = KRSiSi;
    L.KR[1] = KRFeSi;
    L.KR[2] = KRFeFe;
    L.KRMix = KRMix;
    L.sbSi = sbSi;
    L.sbFe = sbFe;
    L.sbFeSi1 = sbFeSi1;
    L.sbFeSi2 = sbFeSi2;
    return L;
}

/* The same piecewise law in double.  *near is set when r or the receding dot
 * product lies within the margin of a threshold whose side matters. */
static inline double pair_mag_f64(const pair_law64 *L, int test, double r, double r2, double dot,
                                  double dotscale, double margin, int *near)
{
    double G = L->G[test], K = L->K[test];
    if (fabs(r - 1.0) < margin) *near = 1;
    if (1.0 <= r) return G / r2;
    double outer, inner = -1.0;
    double KRin;
    if (test == 0) {
        outer = L->sbSi;
        KRin = L->KR[0];
    } else if (test == 2) {
        outer = L->sbFe;
        KRin = L->KR[2];
    } else {
        outer = L->sbFeSi1;
        inner = L->sbFeSi2;
        KRin = L->KR[1];
    }
    if (fabs(r - outer) < margin) *near = 1;
    if (outer <= r) return G - K * (1.0 - r2);
    int receding = (dot >= 1.0); /* (int)dot > 0 */
    if (fabs(dot - 1.0) < margin * (1.0 + dotscale)) *near = 1;
    if (test == 1) {
        if (fabs(r - inner) < margin) *near = 1;
        if (inner <= r) return G - (receding ? L->KRMix : K) * (1.0 - r2);
    }
    return G - (receding ? KRin : K) * (1.0 - r2);
}

static void force_f64_generic(const or_float4 *pos, const or_float4 *vel, double *force3,
                              double *cond, int *near, double margin, int n, const pair_law64 *L,
                              int separate, int b1, int b2, int b3, int nfe, int i_begin,
                              int i_end, int threads)
{
    if (i_begin < 0) i_begin = 0;
    if (i_end > n) i_end = n;
#pragma omp parallel for schedule(dynamic, 16) num_threads(OR_OMP_THREADS(threads))
    for (int id = i_begin; id < i_end; id++) {
        double fx = 0, fy = 0, fz = 0, cs = 0;
        int nr = 0;
        int j0 = 0, j1 = n, sw = nfe;
        if (separate) {
            if (id < b2) {
                j1 = b2;
                sw = b1;
            } else {
                j0 = b2;
                sw = b3;
            }
        }
        for (int j = j0; j < j1; j++) {
            if (j == id) continue;
            double dx = (double)pos[j].x - pos[id].x;
            double dy = (double)pos[j].y - pos[id].y;
            double dz = (double)pos[j].z - pos[id].z;
            double r2 = dx * dx + dy * dy + dz * dz;
            double r = sqrt(r2);
            double ux = (double)vel[j].x - vel[id].x;
            double uy = (double)vel[j].y - vel[id].y;
            double uz = (double)vel[j].z - vel[id].z;
            double dot = dx * ux + dy * uy + dz * uz;
            double dots = r * sqrt(ux * ux + uy * uy + uz * uz);
            int test = (id < sw) + (j < sw);
            double fm = pair_mag_f64(L, test, r, r2, dot, dots, margin, &nr);
            fx += fm * dx / r;
            fy += fm * dy / r;
            fz += fm * dz / r;
            cs += fabs(fm);
        }
        force3[3 * id + 0] = fx;
        force3[3 * id + 1] = fy;
        force3[3 * id + 2] = fz;
        if (cond) cond[id] = cs;
        if (near) near[id] = nr;
    }
}

void or_force_collision_f64(const or_float4 *pos, const or_float4 *vel, double *force3,
                            double *cond, int *near, double rel_margin, int n,
                            const or_force_collision_constants *c, int i_begin, int i_end,
                            int threads)
{
    pair_law64 L = law_from(c->GMassFeFe, c->GMassFeSi, c->KFeFe, c->KSiSi, c->KFeSi, c->KRFeFe,
                            c->KRSiSi, c->KRFeSi, c->KRMix, c->ShellBreakFe, c->ShellBreakSi,
                            c->ShellBreakFeSi1, c->ShellBreakFeSi2);
    force_f64_generic(pos, vel, force3, cond, near, rel_margin, n, &L, 0, 0, 0, 0, c->NFe,
                      i_begin, i_end, threads);
}

void or_force_separate_f64(const or_float4 *pos, const or_float4 *vel, double *force3,
                           double *cond, int *near, double rel_margin, int n,
                           const or_force_separate_constants *c, int threads)
{
    pair_law64 L = law_from(c->GMassFeFe, c->GMassFeSi, c->KFeFe, c->KSiSi, c->KFeSi, c->KRFeFe,
                            c->KRSiSi, c->KRFeSi, c->KRMix, c->ShellBreakFe, c->ShellBreakSi,
                            c->ShellBreakFeSi1, c->ShellBreakFeSi2);
    force_f64_generic(pos, vel, force3, cond, near, rel_margin, n, &L, 1, c->boarder1,
                      c->boarder2, c->boarder3, 0, 0, n, threads);
}

/* Conserved quantities in FP64.  The pair potential integrates the approach
 * branch of the force law: U = -G/r (r >= 1), G (r - 2) + (K/3)(r-1)^2 (r+2)
 * (r < 1).  Masses: m_Si = 1, m_Fe = GMassFeSi. */
void or_conserved_f64(const or_float4 *pos, const or_float4 *vel, int n,
                      const or_force_collision_constants *c, double *out9, int threads)
{
    double G[3] = {1.0, c->GMassFeSi, c->GMassFeFe};
    double K[3] = {c->KSiSi, c->KFeSi, c->KFeFe};
    double mFe = c->GMassFeSi, mSi = 1.0;
    double ke = 0, px = 0, py = 0, pz = 0, mt = 0, cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < n; i++) {
        double m = (i < c->NFe) ? mFe : mSi;
        double vx = vel[i].x, vy = vel[i].y, vz = vel[i].z;
        ke += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        px += m * vx;
        py += m * vy;
        pz += m * vz;
        mt += m;
        cx += m * pos[i].x;
        cy += m * pos[i].y;
        cz += m * pos[i].z;
    }
    cx /= mt;
    cy /= mt;
    cz /= mt;
    double vcx = px / mt, vcy = py / mt, vcz = pz / mt;
    double lx = 0, ly = 0, lz = 0;
    for (int i = 0; i < n; i++) {
        double m = (i < c->NFe) ? mFe : mSi;
        double rx = pos[i].x - cx, ry = pos[i].y - cy, rz = pos[i].z - cz;
        double vx = vel[i].x - vcx, vy = vel[i].y - vcy, vz = vel[i].z - vcz;
        lx += m * (ry * vz - rz * vy);
        ly += m * (rz * vx - rx * vz);
        lz += m * (rx * vy - ry * vx);
    }
    double pe = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : pe) num_threads(OR_OMP_THREADS(threads))
    for (int i = 0; i < n; i++) {
        double acc = 0;
        for (int j = i + 1; j < n; j++) {
            double dx = (double)pos[j].x - pos[i].x;
            double dy = (double)pos[j].y - pos[i].y;
            double dz = (double)pos[j].z - pos[i].z;
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            int t = (i < c->NFe) + (j < c->NFe);
            if (r >= 1.0)
                acc += -G[t] / r;
            else
                acc += G[t] * (r - 2.0) + (K[t] / 3.0) * (r - 1.0) * (r - 1.0) * (r + 2.0);
        }
        pe += acc;
    }
    out9[0] = ke;
    out9[1] = pe;
    out9[2] = ke + pe;
    out9[3] = px;
    out9[4] = py;
    out9[5] = pz;
    out9[6] = lx;
    out9[7] = ly;
    out9[8] = lz;
}

/* ------------------------------------------------------------------------ */
/* Host helpers                                                              */
/* ------------------------------------------------------------------------ */

/* getMassCollision(0), getCenterOfMassCollision(0), getLinearVelocityCollision(0):
 * float accumulators, double MassFe/MassSi (impactBasicFunctions.h:219-230,268-297,388-417). */
void or_com_vel_collision_f32(const or_float4 *pos, const or_float4 *vel, int n, int nfe,
                              double mass_fe, double mass_si, float *out6)
{
    float mass = 0.0f;
    for (int i = 0; i < n; i++) {
        if (i < nfe)
            mass = (float)((double)mass + mass_fe);
        else
            mass = (float)((double)mass + mass_si);
    }
    float cx = 0, cy = 0, cz = 0, vx = 0, vy = 0, vz = 0;
    for (int i = 0; i < n; i++) {
        double m = (i < nfe) ? mass_fe : mass_si;
        cx = (float)((double)cx + (double)pos[i].x * m);
        cy = (float)((double)cy + (double)pos[i].y * m);
        cz = (float)((double)cz + (double)pos[i].z * m);
    }
    for (int i = 0; i < n; i++) {
        double m = (i < nfe) ? mass_fe : mass_si;
        vx = (float)((double)vx + (double)vel[i].x * m);
        vy = (float)((double)vy + (double)vel[i].y * m);
        vz = (float)((double)vz + (double)vel[i].z * m);
    }
    out6[0] = cx / mass;
    out6[1] = cy / mass;
    out6[2] = cz / mass;
    out6[3] = vx / mass;
    out6[4] = vy / mass;
    out6[5] = vz / mass;
}

/* impactAddTime.cu:433-455. */
void or_recenter_f32(or_float4 *pos, or_float4 *vel, int n, int nfe, double mass_fe,
                     double mass_si, const float *initial_com3, const float *initial_vel3)
{
    float cv[6];
    or_com_vel_collision_f32(pos, vel, n, nfe, mass_fe, mass_si, cv);
    for (int i = 0; i < n; i++) {
        pos[i].x += initial_com3[0] - cv[0];
        pos[i].y += initial_com3[1] - cv[1];
        pos[i].z += initial_com3[2] - cv[2];
        vel[i].x += initial_vel3[0] - cv[3];
        vel[i].y += initial_vel3[1] - cv[4];
        vel[i].z += initial_vel3[2] - cv[5];
    }
}

/* getCenterOfMass, generateTargetImpactorBasicFunctions.h:29-80. */
void or_body_com(const or_float4 *pos, int start_fe, int stop_fe, int stop_si, double mass_fe,
                 double mass_si, double total_mass, double *out4)
{
    double x = 0, y = 0, z = 0;
    out4[0] = out4[1] = out4[2] = out4[3] = 0.0;
    if (total_mass < 0.00000000001) return;
    for (int i = start_fe; i < stop_fe; i++) {
        x += pos[i].x * mass_fe;
        y += pos[i].y * mass_fe;
        z += pos[i].z * mass_fe;
    }
    for (int i = stop_fe; i < stop_si; i++) {
        x += pos[i].x * mass_si;
        y += pos[i].y * mass_si;
        z += pos[i].z * mass_si;
    }
    x /= total_mass;
    y /= total_mass;
    z /= total_mass;
    out4[0] = x;
    out4[1] = y;
    out4[2] = z;
    out4[3] = sqrt(x * x + y * y + z * z);
}

/* getAverageLinearVelocity, generateTargetImpactorBasicFunctions.h:119-171. */
void or_body_vel(const or_float4 *vel, int start_fe, int stop_fe, int stop_si, double mass_fe,
                 double mass_si, double total_mass, double *out4)
{
    or_body_com(vel, start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass, out4);
}

/* getAngularMomentum, generateTargetImpactorBasicFunctions.h:173-237: about
 * the body's centre of mass and relative to its mean velocity (both divided
 * by the nominal total_mass), iron then silicate.  out4 = x,y,z,|.| */
void or_body_angmom(const or_float4 *pos, const or_float4 *vel, int start_fe, int stop_fe, int stop_si,
                    double mass_fe, double mass_si, double total_mass, double *out4)
{
    double c[4], u[4];
    or_body_com(pos, start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass, c);
    or_body_vel(vel, start_fe, stop_fe, stop_si, mass_fe, mass_si, total_mass, u);
    double lx = 0.0, ly = 0.0, lz = 0.0;
    for (int i = start_fe; i < stop_si; i++) {
        double m = (i < stop_fe) ? mass_fe : mass_si;
        double rx = pos[i].x - c[0], ry = pos[i].y - c[1], rz = pos[i].z - c[2];
        double vx = vel[i].x - u[0], vy = vel[i].y - u[1], vz = vel[i].z - u[2];
        lx += (ry * vz - rz * vy) * m;
        ly += -(rx * vz - rz * vx) * m;
        lz += (rx * vy - ry * vx) * m;
    }
    out4[0] = lx;
    out4[1] = ly;
    out4[2] = lz;
    out4[3] = sqrt(lx * lx + ly * ly + lz * lz);
}

/* getRadiusOfBody, generateTargetImpactorBasicFunctions.h:82-117: the largest
 * distance of an element of [start, stop) from the body's centre of mass
 * (nominal total_mass).  The reference starts from an uninitialised radius
 * (:87); here from 0. */
double or_body_radius(const or_float4 *pos, int start, int stop_fe, int stop, double mass_fe, double mass_si,
                      double total_mass)
{
    double c[4];
    or_body_com(pos, start, stop_fe, stop, mass_fe, mass_si, total_mass, c);
    double radius = 0.0;
    for (int i = start; i < stop; i++) {
        double rx = (double)pos[i].x - c[0], ry = (double)pos[i].y - c[1], rz = (double)pos[i].z - c[2];
        double d = sqrt(rx * rx + ry * ry + rz * rz);
        if (d > radius) radius = d;
    }
    return radius;
}

/* impactInitialization.exe's set-up (impactInitialization.cu:235-249 units of
 * the offsets; :300-371 offsets added per body and the reorder
 * [Fe1, Si1, Fe2, Si2] -> [Fe1, Fe2, Si1, Si2]; :393-446 the whole system's
 * mass, centre of mass, mean velocity and angular momentum in double).
 * ip / iv: the setup file's offsets in km and km/s as read (float).
 * stats10 = {mass, com xyz, vel xyz, L xyz} in simulation units. */
void or_impact_initialize(const or_float4 *gpos, const or_float4 *gvel, int nfe1, int nsi1, int nfe2, int nsi2,
                          const float *ip1, const float *ip2, const float *iv1, const float *iv2,
                          double unit_length, double unit_time, double mass_fe, double mass_si, or_float4 *pos,
                          or_float4 *vel, double *stats10)
{
    const int n1 = nfe1 + nsi1, n = n1 + nfe2 + nsi2, nfe = nfe1 + nfe2;
    float p1[3], p2[3], v1[3], v2[3];
    for (int k = 0; k < 3; k++) {
        p1[k] = (float)(ip1[k] / unit_length); /* float /= double */
        p2[k] = (float)(ip2[k] / unit_length);
        v1[k] = (float)(iv1[k] * (unit_time / unit_length)); /* float *= double */
        v2[k] = (float)(iv2[k] * (unit_time / unit_length));
    }
    or_float4 *tp = (or_float4 *)malloc(sizeof(or_float4) * (size_t)(n > 0 ? n : 1));
    or_float4 *tv = (or_float4 *)malloc(sizeof(or_float4) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        const float *dp = i < n1 ? p1 : p2, *dv = i < n1 ? v1 : v2;
        tp[i] = gpos[i];
        tv[i] = gvel[i];
        tp[i].x += dp[0];
        tp[i].y += dp[1];
        tp[i].z += dp[2];
        tv[i].x += dv[0];
        tv[i].y += dv[1];
        tv[i].z += dv[2];
    }
    int k = 0;
    for (int i = 0; i < nfe1; i++, k++) { pos[k] = tp[i]; vel[k] = tv[i]; }
    for (int i = n1; i < n1 + nfe2; i++, k++) { pos[k] = tp[i]; vel[k] = tv[i]; }
    for (int i = nfe1; i < n1; i++, k++) { pos[k] = tp[i]; vel[k] = tv[i]; }
    for (int i = n1 + nfe2; i < n; i++, k++) { pos[k] = tp[i]; vel[k] = tv[i]; }
    free(tp);
    free(tv);
    double M = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
    for (int i = 0; i < n; i++) M += (i < nfe) ? mass_fe : mass_si;
    for (int i = 0; i < n; i++) {
        double m = (i < nfe) ? mass_fe : mass_si;
        cx += pos[i].x * m;
        cy += pos[i].y * m;
        cz += pos[i].z * m;
    }
    cx /= M;
    cy /= M;
    cz /= M;
    for (int i = 0; i < n; i++) {
        double m = (i < nfe) ? mass_fe : mass_si;
        ux += vel[i].x * m;
        uy += vel[i].y * m;
        uz += vel[i].z * m;
    }
    ux /= M;
    uy /= M;
    uz /= M;
    double lx = 0.0, ly = 0.0, lz = 0.0;
    for (int i = 0; i < n; i++) {
        double m = (i < nfe) ? mass_fe : mass_si;
        double rx = pos[i].x - cx, ry = pos[i].y - cy, rz = pos[i].z - cz;
        double vx = vel[i].x - ux, vy = vel[i].y - uy, vz = vel[i].z - uz;
        lx += (ry * vz - rz * vy) * m;
        ly += -(rx * vz - rz * vx) * m;
        lz += (rx * vy - ry * vx) * m;
    }
    stats10[0] = M;
    stats10[1] = cx;
    stats10[2] = cy;
    stats10[3] = cz;
    stats10[4] = ux;
    stats10[5] = uy;
    stats10[6] = uz;
    stats10[7] = lx;
    stats10[8] = ly;
    stats10[9] = lz;
}

/* spinBody, generateTargetImpactorBasicFunctions.h:407-470. */
int or_spin_body(const or_float4 *pos, or_float4 *vel, int start, int stop, int stop_fe,
                 double mass_fe, double mass_si, double total_mass, const float *spin4)
{
    float sx = spin4[0], sy = spin4[1], sz = spin4[2], sw = spin4[3];
    double mag = sqrtf(sx * sx + sy * sy + sz * sz); /* float expression, :428 */
    if (0.0000001 < mag) {
        sx = (float)(sx / mag);
        sy = (float)(sy / mag);
        sz = (float)(sz / mag);
    } else {
        return -1;
    }
    double com[4];
    or_body_com(pos, start, stop_fe, stop, mass_fe, mass_si, total_mass, com);
    for (int i = start; i < stop; i++) {
        double rx = pos[i].x - com[0], ry = pos[i].y - com[1], rz = pos[i].z - com[2];
        double magsquared = rx * rx + ry * ry + rz * rz;
        double spinDota = sx * rx + sy * ry + sz * rz;
        double perpendicularDistance = sqrt(magsquared - spinDota * spinDota);
        double perpendicularVelocity = sw * 2.0 * OR_PI * perpendicularDistance;
        double nx = (sy * rz - sz * ry);
        double ny = -(sx * rz - sz * rx);
        double nz = (sx * ry - sy * rx);
        double m = sqrt(nx * nx + ny * ny + nz * nz);
        if (m != 0.0) {
            nx /= m;
            ny /= m;
            nz /= m;
            vel[i].x = (float)(vel[i].x + perpendicularVelocity * nx);
            vel[i].y = (float)(vel[i].y + perpendicularVelocity * ny);
            vel[i].z = (float)(vel[i].z + perpendicularVelocity * nz);
        }
    }
    return 0;
}

/* nBodyCollisionSingleGPU (impactAddTime.cu:396-455) without draw and record. */
void or_impact_steps_f32(or_float4 *pos, or_float4 *vel, or_float4 *force, int n,
                         const or_force_collision_constants *fc,
                         const or_move_collision_constants *mc, int n_steps, int recenter_rate,
                         const float *initial_com3, const float *initial_vel3, double mass_fe,
                         double mass_si, int threads)
{
    int recenter_count = 0;
    for (int s = 0; s < n_steps; s++) {
        or_force_collision_f32(pos, vel, force, n, fc, 0, n, threads);
        or_move_collision_f32(pos, vel, force, n, mc);
        recenter_count++;
        if (recenter_rate > 0 && recenter_count == recenter_rate) {
            or_recenter_f32(pos, vel, n, fc->NFe, mass_fe, mass_si, initial_com3, initial_vel3);
            recenter_count = 0;
        }
    }
}

/* ---- final analysis --------------------------------------------------------- */

/* findEarthAndMoon, Code/ImpactCode/impactBasicFunctions.h:109-217.  Groups
 * are flooded from the first unassigned element; the group counter advances
 * for every i (:150), so a group's number is the index of its first element.
 * The touch test is float: mag = sqrt(dx*dx + dy*dy + dz*dz) < touch
 * (:136-140).  Earth = the largest group, Moon = the next, ties to the
 * smaller number (strict > in :170-187).  group[n] receives the numbers;
 * returns the number of non-empty groups. */
int or_find_earth_moon(const or_float4 *pos, int n, float touch, int *group, int *earth_id, int *moon_id,
                       int *earth_count, int *moon_count)
{
    int *used = (int *)calloc((size_t)n, sizeof(int));
    int *size = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; i++) group[i] = -1;
    int group_number = 0;
    for (int i = 0; i < n; i++) {
        if (group[i] == -1) {
            group[i] = group_number;
            int k = i;
            while (k < n) {
                if (group[k] == group_number && used[k] == 0) {
                    for (int j = i; j < n; j++) {
                        float dx = pos[k].x - pos[j].x;
                        float dy = pos[k].y - pos[j].y;
                        float dz = pos[k].z - pos[j].z;
                        float mag = sqrtf(dx * dx + dy * dy + dz * dz);
                        if (mag < touch) group[j] = group_number;
                    }
                    used[k] = 1;
                    k = i;
                } else {
                    k++;
                }
            }
        }
        group_number++;
    }
    for (int j = 0; j < n; j++) size[group[j]]++;
    int e = -1, ne = 0, m = -1, nm = 0, groups = 0;
    for (int g = 0; g < n; g++) {
        if (size[g] > 0) groups++;
        if (size[g] > ne) {
            ne = size[g];
            e = g;
        }
    }
    for (int g = 0; g < n; g++)
        if (size[g] > nm && g != e) {
            nm = size[g];
            m = g;
        }
    *earth_id = e;
    *moon_id = m;
    *earth_count = ne;
    *moon_count = nm;
    free(used);
    free(size);
    return groups;
}

/* Members of scope s in the reference's order: 0 every element, 1 the Earth's
 * then the Moon's (EarthIndex then MoonIndex, :198-212), 2 the Earth's, 3 the
 * Moon's. */
static int em_members(const int *group, int n, int earth, int moon, int scope, int *idx)
{
    int c = 0;
    if (scope == 0) {
        for (int i = 0; i < n; i++) idx[c++] = i;
        return c;
    }
    if (scope == 1 || scope == 2)
        for (int i = 0; i < n; i++)
            if (group[i] == earth) idx[c++] = i;
    if (scope == 1 || scope == 3)
        for (int i = 0; i < n; i++)
            if (moon >= 0 && group[i] == moon) idx[c++] = i;
    return c;
}

/* getMassCollision / getCenterOfMassCollision / getLinearVelocityCollision /
 * getAngularMomentumCollision for scopes 0..3 (impactBasicFunctions.h:219-652),
 * float accumulators with the double mass products as written there.
 * Outputs: mass4[4], com12[4][3], vel12[4][3], L12[4][3]. */
void or_final_stats_f32(const or_float4 *pos, const or_float4 *vel, int n, int nfe, const int *group,
                        int earth, int moon, double mass_fe, double mass_si, float *mass4, float *com12,
                        float *vel12, float *L12)
{
    int *idx = (int *)malloc((size_t)n * sizeof(int));
    for (int s = 0; s < 4; s++) {
        int c = em_members(group, n, earth, moon, s, idx);
        float mass = 0.0f;
        for (int a = 0; a < c; a++) mass = (float)((double)mass + (idx[a] < nfe ? mass_fe : mass_si));
        float cx = 0, cy = 0, cz = 0, vx = 0, vy = 0, vz = 0;
        for (int a = 0; a < c; a++) {
            int i = idx[a];
            double m = i < nfe ? mass_fe : mass_si;
            cx = (float)((double)cx + (double)pos[i].x * m);
            cy = (float)((double)cy + (double)pos[i].y * m);
            cz = (float)((double)cz + (double)pos[i].z * m);
        }
        cx /= mass;
        cy /= mass;
        cz /= mass;
        for (int a = 0; a < c; a++) {
            int i = idx[a];
            double m = i < nfe ? mass_fe : mass_si;
            vx = (float)((double)vx + (double)vel[i].x * m);
            vy = (float)((double)vy + (double)vel[i].y * m);
            vz = (float)((double)vz + (double)vel[i].z * m);
        }
        vx /= mass;
        vy /= mass;
        vz /= mass;
        float lx = 0, ly = 0, lz = 0;
        for (int a = 0; a < c; a++) {
            int i = idx[a];
            double m = i < nfe ? mass_fe : mass_si;
            float rx = pos[i].x - cx, ry = pos[i].y - cy, rz = pos[i].z - cz;
            float ux = vel[i].x - vx, uy = vel[i].y - vy, uz = vel[i].z - vz;
            lx = (float)((double)lx + (double)(ry * uz - rz * uy) * m);
            ly = (float)((double)ly + (double)(-(rx * uz - rz * ux)) * m);
            lz = (float)((double)lz + (double)(rx * uy - ry * ux) * m);
        }
        mass4[s] = mass;
        com12[3 * s + 0] = cx;
        com12[3 * s + 1] = cy;
        com12[3 * s + 2] = cz;
        vel12[3 * s + 0] = vx;
        vel12[3 * s + 1] = vy;
        vel12[3 * s + 2] = vz;
        L12[3 * s + 0] = lx;
        L12[3 * s + 1] = ly;
        L12[3 * s + 2] = lz;
    }
    free(idx);
}

/* The same sums in FP64 (truth for the product's FP64 reductions). */
void or_final_stats_f64(const or_float4 *pos, const or_float4 *vel, int n, int nfe, const int *group,
                        int earth, int moon, double mass_fe, double mass_si, double *mass4, double *com12,
                        double *vel12, double *L12)
{
    int *idx = (int *)malloc((size_t)n * sizeof(int));
    for (int s = 0; s < 4; s++) {
        int c = em_members(group, n, earth, moon, s, idx);
        double mass = 0, cx = 0, cy = 0, cz = 0, vx = 0, vy = 0, vz = 0, lx = 0, ly = 0, lz = 0;
        for (int a = 0; a < c; a++) {
            int i = idx[a];
            double m = i < nfe ? mass_fe : mass_si;
            mass += m;
            cx += m * pos[i].x;
            cy += m * pos[i].y;
            cz += m * pos[i].z;
            vx += m * vel[i].x;
            vy += m * vel[i].y;
            vz += m * vel[i].z;
        }
        cx /= mass;
        cy /= mass;
        cz /= mass;
        vx /= mass;
        vy /= mass;
        vz /= mass;
        for (int a = 0; a < c; a++) {
            int i = idx[a];
            double m = i < nfe ? mass_fe : mass_si;
            double rx = pos[i].x - cx, ry = pos[i].y - cy, rz = pos[i].z - cz;
            double ux = vel[i].x - vx, uy = vel[i].y - vy, uz = vel[i].z - vz;
            lx += m * (ry * uz - rz * uy);
            ly += m * (rz * ux - rx * uz);
            lz += m * (rx * uy - ry * ux);
        }
        mass4[s] = mass;
        com12[3 * s + 0] = cx;
        com12[3 * s + 1] = cy;
        com12[3 * s + 2] = cz;
        vel12[3 * s + 0] = vx;
        vel12[3 * s + 1] = vy;
        vel12[3 * s + 2] = vz;
        L12[3 * s + 0] = lx;
        L12[3 * s + 1] = ly;
        L12[3 * s + 2] = lz;
    }
    free(idx);
}

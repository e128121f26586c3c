/*
 * gi_oracle.h -- CPU restatement of the Giant-Impact step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing here is part of the shipped library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the checker (or as the timed CPU baseline).
 *
 * Every function cites the reference file:line it restates.  The float path
 * is compiled with -ffp-contract=off so each expression rounds exactly as
 * written (the reference's nvcc build may contract a*b+c into FMAs; that
 * choice is not observable without nvcc and is covered by the tolerance).
 */
#ifndef GI_ORACLE_H
#define GI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float x, y, z, w;
} or_float4;

typedef struct {
    float GMassFeFe, GMassFeSi, KFeFe, KSiSi, KFeSi, KRFeFe, KRSiSi, KRFeSi, KRMix, ShellBreakFe,
        ShellBreakSi, ShellBreakFeSi1, ShellBreakFeSi2;
    int NFe;
} or_force_collision_constants;

typedef struct {
    float Dt, DtOverMassFe, DtOverMassSi;
    int NFe;
} or_move_collision_constants;

typedef struct {
    float GMassFeFe, GMassFeSi, KFeFe, KSiSi, KFeSi, KRFeFe, KRSiSi, KRFeSi, KRMix, ShellBreakFe,
        ShellBreakSi, ShellBreakFeSi1, ShellBreakFeSi2;
    int boarder1, boarder2, boarder3;
} or_force_separate_constants;

typedef struct {
    float Dt, DtOverMassFe, DtOverMassSi;
    int boarder1, boarder2, boarder3;
} or_move_separate_constants;

/* Units of setTargetImpactorParameters (generateTargetImpactor.cu:289-403) and
 * the constants of loadKernalConstantStructures (impactAddTime.cu:128-170). */
typedef struct {
    /* inputs, in the setup file's units */
    double fraction_earth_mass_body1, fraction_earth_mass_body2;
    double fraction_fe_body1, fraction_si_body1, fraction_fe_body2, fraction_si_body2;
    int total_number_of_elements;
    double density_fe, density_si; /* kg / km^3 */
    double k_fe, k_si;             /* kg km^-1 s^-2 */
    double kr_fe, kr_si, sd_fe, sd_si;
    double dt;
    /* outputs */
    int nfe, nsi, nfe1, nsi1, nfe2, nsi2;
    double unit_length, unit_mass, unit_time; /* km, kg, s */
    double mass_fe, mass_si, diameter;        /* sim units */
    double k_fe_sim, k_si_sim;
    double mass_of_body1, mass_of_body2; /* sim units */
} or_units;

int or_units_compute(or_units *u);
/* The %f round trip of recordCarryForwardParameters -> readParametersFromGenerateBodies
 * (generateTargetImpactorBasicFunctions.h:693-709 -> impactBasicFunctions.h:39-75),
 * then loadKernalConstantStructures (impactAddTime.cu:128-170). */
void or_impact_constants(const or_units *u, or_force_collision_constants *fc,
                         or_move_collision_constants *mc);
/* loadKernalConstantStructures of the generator (generateTargetImpactor.cu:406-451),
 * from the unrounded values. */
void or_separate_constants(const or_units *u, or_force_separate_constants *fc,
                           or_move_separate_constants *mc);
double or_round_f(double v); /* printf("%f") then strtod */

/* ---- float restatements of the four kernels ---- */
void or_force_collision_f32(const or_float4 *pos, const or_float4 *vel, or_float4 *force, int n,
                            const or_force_collision_constants *c, int i_begin, int i_end,
                            int threads);
void or_move_collision_f32(or_float4 *pos, or_float4 *vel, const or_float4 *force, int n,
                           const or_move_collision_constants *c);
void or_force_separate_f32(const or_float4 *pos, const or_float4 *vel, or_float4 *force, int n,
                           const or_force_separate_constants *c, int threads);
void or_move_damped_separate_f32(or_float4 *pos, or_float4 *vel, const or_float4 *force, int n,
                                 const or_move_separate_constants *c, float damp1, float damp2);

/* ---- FP64 twins (truth) ---- */
/* Outputs are written for i in [i_begin, i_end) only (the separate variant:
 * all i).  force3: 3n doubles.  cond: n doubles = sum_j |f_ij| (the conditioning
 * denominator), may be NULL.  near: n ints, set to 1 when some pair's branch
 * decision of particle i lies within rel_margin of a threshold, may be NULL. */
void or_force_collision_f64(const or_float4 *pos, const or_float4 *vel, double *force3,
                            double *cond, int *near, double rel_margin, int n,
                            const or_force_collision_constants *c, int i_begin, int i_end,
                            int threads);
void or_force_separate_f64(const or_float4 *pos, const or_float4 *vel, double *force3,
                           double *cond, int *near, double rel_margin, int n,
                           const or_force_separate_constants *c, int threads);

/* out9: kinetic, potential, total, P(3), L about COM(3), in FP64. */
void or_conserved_f64(const or_float4 *pos, const or_float4 *vel, int n,
                      const or_force_collision_constants *c, double *out9, int threads);

/* ---- host-side helpers restated ---- */
/* getCenterOfMassCollision(0) / getLinearVelocityCollision(0) in float
 * (impactBasicFunctions.h:219-230,268-297,388-417). out6 = com, vel. */
void or_com_vel_collision_f32(const or_float4 *pos, const or_float4 *vel, int n, int nfe,
                              double mass_fe, double mass_si, float *out6);
/* Recenter block of nBodyCollisionSingleGPU (impactAddTime.cu:433-455). */
void or_recenter_f32(or_float4 *pos, or_float4 *vel, int n, int nfe, double mass_fe,
                     double mass_si, const float *initial_com3, const float *initial_vel3);
/* getCenterOfMass / getAverageLinearVelocity (generateTargetImpactorBasicFunctions.h:29-80,119-171)
 * for the body [start_fe, stop_si) with iron [start_fe, stop_fe). out4 = x,y,z,|.| */
void or_body_com(const or_float4 *pos, int start_fe, int stop_fe, int stop_si, double mass_fe,
                 double mass_si, double total_mass, double *out4);
void or_body_vel(const or_float4 *vel, int start_fe, int stop_fe, int stop_si, double mass_fe,
                 double mass_si, double total_mass, double *out4);
/* getAngularMomentum (generateTargetImpactorBasicFunctions.h:173-237). out4 = x,y,z,|.| */
void or_body_angmom(const or_float4 *pos, const or_float4 *vel, int start_fe, int stop_fe, int stop_si,
                    double mass_fe, double mass_si, double total_mass, double *out4);
/* getRadiusOfBody (generateTargetImpactorBasicFunctions.h:82-117), starting from 0. */
double or_body_radius(const or_float4 *pos, int start, int stop_fe, int stop, double mass_fe, double mass_si,
                      double total_mass);
/* impactInitialization.cu:235-249,300-371,393-446: offsets (km, km/s as read)
 * added per body, reorder to [Fe1, Fe2, Si1, Si2], and the system's FP64
 * stats10 = {mass, com xyz, vel xyz, L xyz}. */
void or_impact_initialize(const or_float4 *gpos, const or_float4 *gvel, int nfe1, int nsi1, int nfe2, int nsi2,
                          const float *ip1, const float *ip2, const float *iv1, const float *iv2,
                          double unit_length, double unit_time, double mass_fe, double mass_si, or_float4 *pos,
                          or_float4 *vel, double *stats10);
/* spinBody (generateTargetImpactorBasicFunctions.h:407-470). spin4 = direction xyz,
 * revolutions per time unit w. Returns 0, or -1 for a zero direction. */
int or_spin_body(const or_float4 *pos, or_float4 *vel, int start, int stop, int stop_fe,
                 double mass_fe, double mass_si, double total_mass, const float *spin4);
/* findEarthAndMoon (impactBasicFunctions.h:109-217); returns the number of groups. */
int or_find_earth_moon(const or_float4 *pos, int n, float touch, int *group, int *earth_id, int *moon_id,
                       int *earth_count, int *moon_count);
/* get{Mass,CenterOfMass,LinearVelocity,AngularMomentum}Collision(scope 0..3)
 * (impactBasicFunctions.h:219-652) in float, and an FP64 twin. */
void or_final_stats_f32(const or_float4 *pos, const or_float4 *vel, int n, int nfe, const int *group,
                        int earth, int moon, double mass_fe, double mass_si, float *mass4, float *com12,
                        float *vel12, float *L12);
void or_final_stats_f64(const or_float4 *pos, const or_float4 *vel, int n, int nfe, const int *group,
                        int earth, int moon, double mass_fe, double mass_si, double *mass4, double *com12,
                        double *vel12, double *L12);
/* One step of the nBodyCollisionSingleGPU loop (force, move) repeated n_steps times,
 * with the recenter every recenter_rate steps (0 = never). */
void or_impact_steps_f32(or_float4 *pos, or_float4 *vel, or_float4 *force, int n,
                         const or_force_collision_constants *fc,
                         const or_move_collision_constants *mc, int n_steps, int recenter_rate,
                         const float *initial_com3, const float *initial_vel3, double mass_fe,
                         double mass_si, int threads);

/* Conversion of the receding test (int)(dp.dv) > 0 with saturation as CUDA's
 * cvt.rzi does (exported for the threshold test). */
int or_trunc_to_int(float v);

#ifdef __cplusplus
}
#endif

#endif

/*
 * gi_host.h -- host side of the Giant-Impact step loop: units, constants and
 * seeded synthetic bodies.  Part of libgiant_impact.so (C ABI).
 */
#ifndef GI_HOST_H
#define GI_HOST_H

#include "giant_impact.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Setup of the generator (SetupTemplates/setupGeneratingTargetImpactor:1-67),
 * the values read by readTargetImpactorParameters
 * (Code/GenerateTargetImpactorCode/generateTargetImpactor.cu:169-248), in
 * file order. */
typedef struct gi_generator_setup {
    float initial_spin1[4]; /* direction xyz, revolutions per hour */
    float initial_spin2[4];
    double fraction_earth_mass_body1, fraction_earth_mass_body2;
    double fraction_fe_body1, fraction_si_body1, fraction_fe_body2, fraction_si_body2;
    float damp_rate_body1, damp_rate_body2;
    int total_number_of_elements;
    double damp_time, damp_rest_time, spin_rest_time; /* hours */
    double dt;
    double density_fe, density_si; /* kg km^-3 */
    double k_fe, k_si;             /* kg km^-1 s^-2 */
    double kr_fe, kr_si, sd_fe, sd_si;
    int draw_rate;
} gi_generator_setup;

/* The unit system and counts of setTargetImpactorParameters
 * (generateTargetImpactor.cu:289-403), before the %f round trip. */
typedef struct gi_units {
    int nfe, nsi, nfe1, nsi1, nfe2, nsi2;
    double unit_length, unit_mass, unit_time; /* km, kg, s */
    double mass_fe, mass_si, diameter, gravity;
    double k_fe, k_si; /* simulation units */
    double kr_fe, kr_si, sd_fe, sd_si;
    double mass_of_body1, mass_of_body2;
    double damp_time, damp_rest_time, spin_rest_time; /* simulation units */
    float spin1_w, spin2_w;                           /* revolutions per time unit */
} gi_units;

/* The template defaults of SetupTemplates/setupGeneratingTargetImpactor with
 * TotalNumberOfElements = n. */
void gi_generator_setup_defaults(gi_generator_setup *s, int n);
/* Parse a setupGeneratingTargetImpactor file the way the reference does:
 * the value after each '=' in order, key names ignored. */
int gi_generator_setup_read(const char *path, gi_generator_setup *s);
/* setTargetImpactorParameters. Returns GI_E_INVALID_ARG when the Fe and Si
 * fractions of a body do not add to 1 (generateTargetImpactor.cu:295-304). */
int gi_units_compute(const gi_generator_setup *s, gi_units *u);
/* loadKernalConstantStructures of the generator (generateTargetImpactor.cu:406-451). */
void gi_separate_constants(const gi_units *u, double dt, gi_force_separate_constants *fc,
                           gi_move_separate_constants *mc);
/* loadKernalConstantStructures of the impact (impactAddTime.cu:128-170), from
 * the values after the %f round trip of targetImpactorParameters. */
void gi_impact_constants(const gi_units *u, double dt, gi_force_collision_constants *fc,
                         gi_move_collision_constants *mc);
/* Write / read targetImpactorParameters (generateTargetImpactorBasicFunctions.h:688-711,
 * impactBasicFunctions.h:31-107). */
int gi_parameters_write(const char *path, const gi_units *u, int total_number_of_elements);
int gi_parameters_read(const char *path, gi_units *u, int *total_number_of_elements);

/* Seeded synthetic impact: two jittered simple-cubic-lattice bodies. */
typedef struct gi_impact_config {
    int n_target, n_impactor;
    double fe_fraction;       /* iron fraction by count, innermost */
    double jitter;            /* uniform per-coordinate jitter, in diameters */
    unsigned long long seed;  /* SplitMix64 seed */
    double angle_deg;         /* impact angle (0 = head on) */
    double speed_vesc;        /* relative speed in units of the mutual escape speed */
    double target_spin;       /* target spin in units of its breakup rate */
    double impactor_spin;     /* impactor spin in units of its breakup rate */
    double gap_in_target_radii; /* initial surface gap (default 1.0) */
} gi_impact_config;

typedef struct gi_impact_info {
    int n, nfe, nfe1, nfe2, nsi1, nsi2;
    double mass_fe, mass_si;
    double radius_target, radius_impactor;
    double mass_target, mass_impactor;
    double v_escape, v_relative, impact_parameter;
    double target_omega, impactor_omega;
    double com[3], vel[3];
} gi_impact_info;

/* Fill pos/vel (host arrays of n_target + n_impactor records) in the impact
 * layout [Fe1, Fe2, Si1, Si2] (impactInitialization.cu:326-371) with .w =
 * body/material tag of the generator (0 Fe1, 1 Si1, 2 Fe2, 3 Si2) in pos and
 * the element mass in vel; the centre of mass is at rest at the origin.
 * Constants come from the template defaults at n = n_target + n_impactor.
 * Any of fc, mc, info may be NULL. */
int gi_make_impact(const gi_impact_config *cfg, gi_float4 *pos, gi_float4 *vel,
                   gi_force_collision_constants *fc, gi_move_collision_constants *mc,
                   gi_impact_info *info);

/* ---- generation phase (generateTargetImpactor.cu:666-749) ---------------- */

/* spinBody (generateTargetImpactorBasicFunctions.h:407-470): add the rigid
 * spin of spin4 = {direction xyz, revolutions per time unit} about the
 * body's centre of mass to vel[start, stop); iron is [start, stop_fe).  The
 * centre of mass divides by total_mass (the body's nominal mass, as the
 * reference does).  Host arrays.  GI_E_INVALID_ARG for a zero direction. */
int gi_spin_body(const gi_float4 *pos, gi_float4 *vel, int start, int stop, int stop_fe,
                 double mass_fe, double mass_si, double total_mass, const float *spin4);
/* zeroOutTargetImpactorDrift (generateTargetImpactorBasicFunctions.h:472-506):
 * move each body's centre of mass to the origin and remove its mean velocity.
 * Generation layout [Fe1, Si1, Fe2, Si2] with counts from u.  Host arrays. */
int gi_zero_out_drift(gi_float4 *pos, gi_float4 *vel, const gi_units *u);

/* The raw bodies the generation starts from: seeded jittered lattices (the
 * reference's createRawTargetImpactor uses srand(time)), iron innermost,
 * both centred at the origin, layout [Fe1, Si1, Fe2, Si2], pos.w = tag 0..3,
 * vel.w = element mass, zero velocity.  Host arrays of n = sum of u's counts. */
int gi_make_raw_bodies(const gi_units *u, unsigned long long seed, double jitter, gi_float4 *pos,
                       gi_float4 *vel);

typedef struct gi_generate_info {
    int n;
    long long damp_steps, rest_steps, spin_rest_steps; /* steps per phase */
    double run_time;                                    /* simulation time at the end */
} gi_generate_info;

/* The generation run of generateTargetImpactor.exe without its display:
 * seeded raw bodies (jittered lattices, iron cores, layout [Fe1, Si1, Fe2,
 * Si2], both centred at the origin), then damped settling for DampTime, a
 * rest for DampRestTime, the spin of both bodies, a rest for SpinRestTime,
 * and the drift removal, with the reference's step bookkeeping.  pos/vel are
 * host arrays of setup->total_number_of_elements records (the result).  When
 * out_dir is not NULL the reference's files targetImpactorParameters and
 * targetImpactorInitialPosVel are written there. */
int gi_generate_bodies(const gi_generator_setup *setup, unsigned long long seed, double jitter,
                       gi_float4 *pos, gi_float4 *vel, gi_units *units, gi_generate_info *info,
                       const char *out_dir);

/* ---- impact initialisation and the impact run (impactInitialization.cu,
 *      impactAddTime.cu) ---------------------------------------------------- */

/* setupInitialConditions + recordInitialImpactStat's centring record
 * (impactInitialization.cu:260-391,393-533): read targetImpactorInitialPosVel
 * and targetImpactorParameters from in_dir and setupImpactInitialization from
 * setup_path; place and push the bodies; reorder to [Fe1, Fe2, Si1, Si2];
 * write frame 0 to out_dir/impactPosVel and out_dir/viewCenteringInfo. */
int gi_impact_initialize(const char *in_dir, const char *setup_path, const char *out_dir);

typedef struct gi_impact_run_info {
    int n;
    long long steps;
    int frames_written;
    double start_time, end_time; /* simulation units */
} gi_impact_run_info;

/* impactAddTime.exe without its display: read targetImpactorParameters from
 * param_dir, setupImpactAddTime from setup_path, viewCenteringInfo and the
 * last frame of impactPosVel from run_dir; advance AdditionalTime days,
 * appending a frame every RecordRate steps and recentring every RecenterRate
 * steps (FP64 on the device); rewrite viewCenteringInfo and write
 * runStatsFile (the Earth / Moon finder on the device) at the end. */
int gi_impact_run(const char *param_dir, const char *setup_path, const char *run_dir,
                  gi_impact_run_info *info);

/* Write runStatsFile as recordFinalImpactStat does
 * (Code/ImpactCode/impactBasicFunctions.h:883-1072): the Earth and Moon
 * masses, composition by body and material, angular momenta with their angle
 * off the ecliptic (xz) plane, centres and velocities, then the whole system,
 * in km, kg and s through the units of u.  time is the simulation time
 * (simulation units).  gi_impact_run writes it at the end of every run. */
int gi_write_run_stats(const char *path, const gi_earth_moon *em, const gi_units *u, double time);

/* ---- request loop (continuous batching of impact-sweep jobs) ------------- */

typedef struct gi_serve_request {
    gi_impact_config impact; /* the impact to simulate (seeded synthetic bodies) */
    int steps;               /* leapfrog steps to run */
    double arrival_s;        /* arrival time in seconds after the start of the run */
    gi_float4 *final_pos;    /* optional host buffers (n records) for the final state */
    gi_float4 *final_vel;
} gi_serve_request;

typedef struct gi_serve_result {
    int n;                  /* particles */
    int steps;              /* steps run */
    double arrival_s, admit_s, finish_s; /* seconds after the start of the run */
    double time_to_result_s;             /* finish_s - arrival_s */
    double energy0, energy1;             /* FP64 total energy at admission / completion */
    double angmom0[3], angmom1[3];       /* FP64 angular momentum about the centre of mass */
    double energy_rel_change;            /* |E1 - E0| / |E0| */
    double angmom_rel_change;            /* |L1 - L0| / |L0| */
} gi_serve_result;

typedef struct gi_serve_options {
    long long max_particles; /* resident particle capacity; 0 = from free HBM (capped at 2^21) */
    int max_sims;            /* resident simulation slots; 0 = 256 */
    int quantum_steps;       /* steps per scheduling round; 0 = 100 */
    double hbm_fraction;     /* share of free HBM the loop may use; 0 = 0.5 */
} gi_serve_options;

typedef struct gi_serve_summary {
    int completed;
    double wall_s;
    double total_sim_steps;       /* sum over requests of steps */
    double sim_steps_per_s;       /* total_sim_steps / wall_s */
    double pair_interactions_per_s;
    double p50_time_to_result_s, p99_time_to_result_s, max_time_to_result_s;
    long long capacity_particles; /* admission limit used */
    long long peak_resident_particles;
    int peak_resident_sims;
    int rounds;                   /* scheduling rounds (batched launches of quantum steps) */
    double max_energy_rel_change, max_angmom_rel_change;
} gi_serve_summary;

/* Serve `n_requests` requests on the current device: requests are admitted in
 * arrival order as soon as they have arrived and their particles fit in the
 * free capacity and a slot is free; every resident request advances in the
 * same batched launches (continuous batching); a request leaves as soon as
 * its steps are done, freeing room for the next.  results has n_requests
 * entries (request order).  Blocking; returns when every request is done. */
int gi_serve_run(const gi_serve_request *reqs, int n_requests, const gi_serve_options *opt,
                 gi_serve_result *results, gi_serve_summary *summary);

#ifdef __cplusplus
}
#endif

#endif

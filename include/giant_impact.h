/*
 * giant_impact.h -- C ABI of the MI355X (gfx950) step loop for Giant-Impact.
 *
 * The four kernels of the reference's step loop are replaced one for one:
 *
 *   getForcesCollisionSingleGPU   Code/ImpactCode/impactAddTime.cu:245-377   -> gi_force_collision
 *   moveBodiesCollisionSingleGPU  Code/ImpactCode/impactAddTime.cu:379-394   -> gi_move_collision
 *   getForcesSeperate             Code/GenerateTargetImpactorCode/generateTargetImpactor.cu:486-626
 *                                                                            -> gi_force_separate
 *   moveBodiesDampedSeperate      Code/GenerateTargetImpactorCode/generateTargetImpactor.cu:628-664
 *                                                                            -> gi_move_damped_separate
 *
 * Every pointer argument named pos / vel / force is a DEVICE pointer to n
 * float4 records (16 bytes each, .x .y .z used, .w passed through untouched),
 * exactly the Pos_DEV0 / Vel_DEV0 / Force_DEV0 arrays of the reference
 * (Code/CommonCode/commonSetup.h:38-39).  Material is given by index, as in
 * the reference: in the impact layout [Fe1, Fe2, Si1, Si2] particle id is iron
 * iff id < NFe; in the generation layout [Fe1, Si1, Fe2, Si2] the borders
 * boarder1..3 split the bodies and materials.
 *
 * Unlike the reference kernels (launched <<<N/256, 256>>> with no bounds
 * check, allocateMemoryAndSetupGPU at impactAddTime.cu:194-198 rejects
 * N % 256 != 0) these entry points accept any n >= 0.  gi_check_block_multiple
 * restates the reference's host-side N % 256 check for callers that want it.
 *
 * Return value of every int function: 0 on success, a positive hipError_t
 * code when the HIP runtime failed, or a negative GI_E_* code for an invalid
 * argument.  gi_last_error_string() describes the most recent failure of the
 * calling thread.  (The reference prints and calls exit(0) instead,
 * Code/CommonCode/commonFunctions.h:12-22.)
 *
 * Streams: a gi_stream_t is a hipStream_t; NULL means the default stream.
 * Work is enqueued asynchronously in stream order.  Scratch buffers are owned
 * by the library, one set per device, grown on demand and reused; calls that
 * overlap on one device from several host threads or streams are not
 * supported (the reference is single-threaded, one default stream).
 */
#ifndef GIANT_IMPACT_H
#define GIANT_IMPACT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GI_ABI_VERSION 1

/* 16-byte particle record, layout-identical to HIP/CUDA float4. */
typedef struct gi_float4 {
    float x, y, z, w;
} gi_float4;

typedef void *gi_stream_t;

/* Negative error codes (positive codes are hipError_t values). */
#define GI_OK 0
#define GI_E_INVALID_ARG (-1)
#define GI_E_NOT_MULTIPLE_OF_BLOCK (-2)
#define GI_E_NO_DEVICE (-3)
#define GI_E_TOO_LARGE (-4)
#define GI_E_IO (-5)
#define GI_E_PARSE (-6)
#define GI_E_STATE (-7)

/* ---- kernel constant structures (byte-identical to the reference) ------ */

/* forceCollisionKernalConstantsStruct, Code/ImpactCode/impactAddTime.cu:4-20 (56 bytes). */
typedef struct gi_force_collision_constants {
    float GMassFeFe;
    float GMassFeSi;
    float KFeFe;
    float KSiSi;
    float KFeSi;
    float KRFeFe;
    float KRSiSi;
    float KRFeSi;
    float KRMix;
    float ShellBreakFe;
    float ShellBreakSi;
    float ShellBreakFeSi1;
    float ShellBreakFeSi2;
    int NFe;
} gi_force_collision_constants;

/* moveCollisionKernalConstantsStruct, Code/ImpactCode/impactAddTime.cu:22-28 (16 bytes). */
typedef struct gi_move_collision_constants {
    float Dt;
    float DtOverMassFe;
    float DtOverMassSi;
    int NFe;
} gi_move_collision_constants;

/* forceSeperateKernalConstantsStruct,
 * Code/GenerateTargetImpactorCode/generateTargetImpactor.cu:3-21 (64 bytes). */
typedef struct gi_force_separate_constants {
    float GMassFeFe;
    float GMassFeSi;
    float KFeFe;
    float KSiSi;
    float KFeSi;
    float KRFeFe;
    float KRSiSi;
    float KRFeSi;
    float KRMix;
    float ShellBreakFe;
    float ShellBreakSi;
    float ShellBreakFeSi1;
    float ShellBreakFeSi2;
    int boarder1; /* NFe1:              end of body-1 iron     */
    int boarder2; /* NFe1+NSi1:         end of body 1          */
    int boarder3; /* NFe1+NSi1+NFe2:    end of body-2 iron     */
} gi_force_separate_constants;

/* moveSeperateKernalConstantsStruct,
 * Code/GenerateTargetImpactorCode/generateTargetImpactor.cu:23-31 (24 bytes). */
typedef struct gi_move_separate_constants {
    float Dt;
    float DtOverMassFe;
    float DtOverMassSi;
    int boarder1;
    int boarder2;
    int boarder3;
} gi_move_separate_constants;

/* ---- library / device -------------------------------------------------- */

int gi_abi_version(void);
const char *gi_last_error_string(void);
/* Code of the failure gi_last_error_string() describes (0 when none since the
 * last gi_clear_error on this thread); the reference's cudaErrorCheck
 * (Code/CommonCode/commonFunctions.h:12-22) would check it instead of
 * cudaGetLastError. */
int gi_last_error(void);
void gi_clear_error(void);
/* Number of visible HIP devices (0 when none); never fails. */
int gi_device_count(void);
/* Select the device used by subsequent calls from this thread. */
int gi_set_device(int device);
/* Restates the check of allocateMemoryAndSetupGPU (impactAddTime.cu:193-198):
 * returns GI_E_NOT_MULTIPLE_OF_BLOCK when n % 256 != 0, else 0. */
int gi_check_block_multiple(int n);

/* Device memory and copies (copyPosVelToGPU / copyPosVelFromGPU,
 * Code/CommonCode/commonFunctions.h:24-38).  gi_copy_* are blocking. */
int gi_malloc(void **dptr, size_t bytes);
int gi_free(void *dptr);
int gi_copy_h2d(void *dst, const void *src, size_t bytes);
int gi_copy_d2h(void *dst, const void *src, size_t bytes);
int gi_copy_d2d(void *dst, const void *src, size_t bytes, gi_stream_t stream);
int gi_stream_create(gi_stream_t *stream);
int gi_stream_destroy(gi_stream_t stream);
int gi_stream_sync(gi_stream_t stream);
int gi_device_sync(void);
/* Same as gi_device_sync: the cudaDeviceSynchronize the reference calls. */
int gi_sync(void);
/* Free and total device memory in bytes (hipMemGetInfo). */
int gi_mem_info(size_t *free_bytes, size_t *total_bytes);
/* The plan of the last drop-in force call on the current device: mutual = 1
 * when it uses the mutual all-pairs pass (k_mutual) for some simulation, 3
 * for its row-group form (k_mutual_grp), 2 when it uses the direct pass
 * (k_direct), 0 for the one-sided pass (k_gravity); r = i-particles per lane, pair_tasks = wavefront tasks of the
 * all-pairs kernel(s) (workgroups of the direct pass), chunk_len = j per task
 * of the mutual pass.  Zeros before the first call.  Any pointer may be NULL. */
int gi_plan_info(int *mutual, int *r, long long *pair_tasks, long long *chunk_len);

/* ---- the four step-loop kernels ----------------------------------------- */

/* getForcesCollisionSingleGPU (impactAddTime.cu:245-377): for every particle i,
 * force[i].xyz = sum over all j of force_mag(r, materials, shell, receding) * dp / r
 * with dp = pos[j] - pos[i]; force[i].w is not written. */
int gi_force_collision(const gi_float4 *pos, const gi_float4 *vel, gi_float4 *force, int n,
                       const gi_force_collision_constants *c, gi_stream_t stream);

/* No counterpart in the reference.  gi_force_collision for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same force, bitwise, where the
 * tuning table gives gi_force_collision the direct pass (n <= 12,288), with
 * the direct pass's scratch (the simulation descriptor and the block table,
 * ~13 KB) in a static device array that each device holds until the process
 * ends.  Allocates nothing, host or device.  n > 12,288 returns
 * GI_E_TOO_LARGE.  Every call on one device shares that array: calls that
 * overlap on one device (several host threads, or several streams without
 * ordering between them) race. */
int gi_force_collision_noalloc(const gi_float4 *pos, const gi_float4 *vel, gi_float4 *force, int n,
                               const gi_force_collision_constants *c, gi_stream_t stream);

/* moveBodiesCollisionSingleGPU (impactAddTime.cu:379-394):
 * vel += force * (Dt/m); pos += vel * Dt  (in that order; .w untouched). */
int gi_move_collision(gi_float4 *pos, gi_float4 *vel, const gi_float4 *force, int n,
                      const gi_move_collision_constants *c, gi_stream_t stream);

/* getForcesSeperate (generateTargetImpactor.cu:486-626): the same pair force,
 * but only between particles of the same body ([0,boarder2) and [boarder2,n)). */
int gi_force_separate(const gi_float4 *pos, const gi_float4 *vel, gi_float4 *force, int n,
                      const gi_force_separate_constants *c, gi_stream_t stream);

/* moveBodiesDampedSeperate (generateTargetImpactor.cu:628-664):
 * vel += (force - damp*vel) * (Dt/m); pos += vel * Dt, with damp = DampRateBody1
 * for body 1 and DampRateBody2 for body 2. */
int gi_move_damped_separate(gi_float4 *pos, gi_float4 *vel, const gi_float4 *force, int n,
                            const gi_move_separate_constants *c, float damp_rate_body1,
                            float damp_rate_body2, gi_stream_t stream);

/* ---- whole-system reductions (FP64) -------------------------------------- */

/* Mass-weighted centre of mass and mean velocity of particles [0,n) with
 * iron for id < nfe, in FP64 (the reference's getCenterOfMassCollision(0) /
 * getLinearVelocityCollision(0), impactBasicFunctions.h:268-297,388-417, sum
 * in float on the host).  out6 = {cmx, cmy, cmz, vx, vy, vz}; blocking. */
int gi_center_of_mass_velocity(const gi_float4 *pos, const gi_float4 *vel, int n, int nfe,
                               double mass_fe, double mass_si, double *out6, gi_stream_t stream);

/* Recenter (impactAddTime.cu:433-455): shift every position by
 * (target_com - com) and every velocity by (target_vel - mean_vel), where com
 * and mean_vel are the current mass-weighted means; entirely on the device. */
int gi_recenter(gi_float4 *pos, gi_float4 *vel, int n, int nfe, double mass_fe, double mass_si,
                const double *target_com3, const double *target_vel3, gi_stream_t stream);

/* The reference's own recenter, float for float (impactAddTime.cu:433-455
 * with getMassCollision(0) / getCenterOfMassCollision(0) /
 * getLinearVelocityCollision(0), impactBasicFunctions.h:219-230,268-297,
 * 388-417): the whole system's mass, centre of mass and mean velocity summed
 * in float particle by particle in index order (each term formed in double
 * with the double element masses mass_fe / mass_si, iron for id < nfe), the
 * means divided in float, then pos += initial_com3 - com and
 * vel += initial_vel3 - mean_vel in float.  Runs on the device (no host round
 * trip); the result is bitwise what the reference's host loop produces. */
int gi_recenter_float(gi_float4 *pos, gi_float4 *vel, int n, int nfe, double mass_fe, double mass_si,
                      const float *initial_com3, const float *initial_vel3, gi_stream_t stream);

/* FP64 conserved quantities of one simulation (particles [0,n), iron id < nfe):
 * out[0] kinetic, out[1] potential, out[2] total energy,
 * out[3..5] linear momentum, out[6..8] angular momentum about the centre of mass.
 * The potential of a pair is -G mi mj / r for r >= 1 and
 * G mi mj (r - 2) + (K/3)(r - 1)^2 (r + 2) for r < 1 (K = approach stiffness of
 * the pair), whose derivative is the force law.  Blocking. */
int gi_conserved_quantities(const gi_float4 *pos, const gi_float4 *vel, int n,
                            const gi_force_collision_constants *c, double *out9,
                            gi_stream_t stream);

/* ---- native single-simulation stepper ------------------------------------ */

/* nBodyCollisionSingleGPU (impactAddTime.cu:396-481) without the GL, file and
 * host round trips: n_steps of (force, move) on device-resident state, with an
 * optional device-side recenter every recenter_rate steps (0 = never) toward
 * target_com3 / target_vel3.  The step sequence is captured once in a HIP
 * graph and replayed. */
int gi_impact_steps(gi_float4 *pos, gi_float4 *vel, gi_float4 *force, int n,
                    const gi_force_collision_constants *fc, const gi_move_collision_constants *mc,
                    int n_steps, int recenter_rate, const double *target_com3,
                    const double *target_vel3, gi_stream_t stream);

/* nBodyCollisionSingleGPU's step and recenter bookkeeping exactly
 * (impactAddTime.cu:396-455): n_steps of (force, move), and after every
 * recenter_rate-th step (0 = never) the reference's float recenter
 * (gi_recenter_float) toward initial_com3 / initial_vel3 -- the
 * InitialCenterOfMass / InitialLinearVelocity floats of viewCenteringInfo --
 * with the %f-read element masses mass_fe / mass_si.  Graph-captured like
 * gi_impact_steps; the drop-in form of the reference's loop. */
int gi_impact_steps_float_recenter(gi_float4 *pos, gi_float4 *vel, gi_float4 *force, int n,
                                   const gi_force_collision_constants *fc, const gi_move_collision_constants *mc,
                                   int n_steps, int recenter_rate, const float *initial_com3,
                                   const float *initial_vel3, double mass_fe, double mass_si, gi_stream_t stream);

/* The generation phases' step (generateTargetImpactor.cu:670,684,696,711):
 * n_steps of (gi_force_separate, gi_move_damped_separate with damp1/damp2),
 * captured once in a HIP graph and replayed. */
int gi_separate_steps(gi_float4 *pos, gi_float4 *vel, gi_float4 *force, int n,
                      const gi_force_separate_constants *fc, const gi_move_separate_constants *mc,
                      float damp_rate_body1, float damp_rate_body2, int n_steps, gi_stream_t stream);

/* ---- batched ensemble (many impacts of uneven n in one launch per kernel) - */

typedef struct gi_ensemble gi_ensemble;

/* Create an ensemble of n_sims independent impacts.  Particle arrays are
 * packed back to back: simulation s owns particles [offset_s, offset_s + n[s])
 * with offset_s = n[0] + ... + n[s-1], each in the impact layout (iron first,
 * fc[s].NFe irons).  pos/vel are device pointers owned by the caller; force is
 * owned by the ensemble.  tile_config may be NULL (autotuned defaults). */
int gi_ensemble_create(gi_ensemble **out, int n_sims, const int *n,
                       const gi_force_collision_constants *fc,
                       const gi_move_collision_constants *mc, gi_float4 *pos, gi_float4 *vel,
                       gi_stream_t stream);
/* An ensemble over caller-owned arrays of `capacity` records with no
 * simulations yet: gi_ensemble_set_sims places them.  For a resident pool
 * whose simulations come and go (the request loop).  No reference
 * counterpart. */
int gi_ensemble_create_capacity(gi_ensemble **out, int capacity, gi_float4 *pos, gi_float4 *vel,
                                gi_stream_t stream);
/* Replace the simulation table: simulation s occupies records
 * [offset[s], offset[s] + n[s]) of the ensemble's arrays (offsets ascending,
 * no overlap, inside the capacity), iron first (fc[s].NFe irons).  Gaps
 * between simulations are left untouched.  Every simulation starts active.
 * Only the table changes: the arrays stay where they are, state in them is
 * not moved, the scratch is kept (grown when needed) and the captured step
 * graph is updated in place at the next gi_ensemble_steps.  Simulation
 * indices of the other calls (set_active, conserved) follow this table.
 * No reference counterpart. */
int gi_ensemble_set_sims(gi_ensemble *e, int n_sims, const int *offset, const int *n,
                         const gi_force_collision_constants *fc, const gi_move_collision_constants *mc);
/* Block until at most max_pending of this ensemble's gi_ensemble_steps calls
 * may still be running on the device (0: all done).  No reference
 * counterpart. */
int gi_ensemble_wait(gi_ensemble *e, int max_pending);
int gi_ensemble_destroy(gi_ensemble *e);
/* Advance every simulation whose active flag is set by n_steps steps. */
int gi_ensemble_steps(gi_ensemble *e, int n_steps);
/* Per-simulation active flags (host array of n_sims ints; 0 = frozen). */
int gi_ensemble_set_active(gi_ensemble *e, const int *active);
/* FP64 conserved quantities per simulation: out is n_sims x 9 doubles (see
 * gi_conserved_quantities).  Blocking. */
int gi_ensemble_conserved(gi_ensemble *e, double *out);
/* Total particles and total pair interactions of one step of the active set. */
int gi_ensemble_work(gi_ensemble *e, long long *particles, double *pairs);

/* ---- final analysis: the Earth / Moon finder ------------------------------ */

/* findEarthAndMoon and the final statistics' sums (recordFinalImpactStat,
 * Code/ImpactCode/impactBasicFunctions.h:109-217,219-652,883-1072).  Scopes:
 * 0 entire system, 1 Earth-Moon system, 2 Earth, 3 Moon, as the reference's
 * getMassCollision(scope) numbers them.  Sums are FP64 (the reference sums
 * in float); an empty scope has mass 0 and NaN centre and velocity, as the
 * reference's division by a zero mass gives. */
typedef struct gi_earth_moon {
    int n_groups;           /* groups of elements closer than touch (connected) */
    int earth_id, moon_id;  /* group numbers: the smallest element index of each; moon_id -1 when absent */
    int earth_count, moon_count;
    int earth_fe1, earth_fe2, earth_si1, earth_si2; /* elements by body and material */
    int moon_fe1, moon_fe2, moon_si1, moon_si2;
    double mass[4];
    double com[4][3];
    double vel[4][3];       /* mass-weighted mean velocity */
    double angmom[4][3];    /* about the scope's centre of mass, relative to its mean velocity */
} gi_earth_moon;

/* Groups of pos[n] (device, impact layout [Fe1, Fe2, Si1, Si2] with counts
 * nfe1, nfe2, nsi1; the rest is silicate of body 2) joined by pairs with the
 * reference's float test sqrt(dx*dx + dy*dy + dz*dz) < touch (the reference
 * uses touch = 1.5 * Diameter); the Earth is the largest group and the Moon
 * the next, ties to the smaller group number.  group (optional device int[n])
 * receives each element's group number.  Blocking. */
int gi_find_earth_moon(const gi_float4 *pos, const gi_float4 *vel, int n, int nfe1, int nfe2, int nsi1,
                       double mass_fe, double mass_si, float touch, int *group, gi_earth_moon *out,
                       gi_stream_t stream);

/* ---- tiling of the all-pairs pass (the autotuner's control) -------------- */

/* Override the tiling chosen from the tuning table (tuning/gfx950.txt beside
 * the library's lib/ directory, written by the autotuner): r = i-particles per
 * lane (1, 2, 4 or 8), variant = the all-pairs pass (0: one-sided, r^-3 as rsq
 * cubed; 1: one-sided, rsq*rcp; 2: mutual -- each pair evaluated once for both
 * particles, rows of 64*r particles; 3: direct -- small simulations, the whole
 * force, contacts included, in one launch, r = lanes per i-particle: 1, 2,
 * 4 or 8),
 * chunk_len = j-particles per task (0 =
 * derived from target_tasks),
 * target_tasks = wavefront tasks to aim for, min_chunk = lower bound of
 * chunk_len.  r = 0 restores the table.  Plans built with another tiling are
 * rebuilt on their next use. */
int gi_set_tiling(int r, int variant, long long chunk_len, long long target_tasks, long long min_chunk);

/* The tiling a single simulation of n particles gets (override, else the
 * tuning table's row, else the built-in default).  Host only: needs no
 * device.  Any output pointer may be NULL. */
int gi_tiling_for(long long n, int *r, int *variant, long long *chunk_len, long long *target_tasks,
                  long long *min_chunk);

/* Tiling of the batched engine (ensembles, request loop): every simulation's
 * j-range is cut into chunks of chunk_len whatever the batch, so a simulation
 * computes bitwise the same alone (an ensemble of one) and in any mix; r =
 * i-particles per lane of the launch, variant as for gi_set_tiling.  Default:
 * row "B chunk_len r [variant]" of the tuning file.  r = 0 restores the
 * default; takes effect at the next re-plan (a new ensemble or a change of the
 * active set). */
int gi_set_batched_tiling(int r, long long chunk_len, int variant);

/* Host only, no device needed: build the mutual pass's plan (variant 2, or
 * 4 with row groups) for simulations of n[s] particles (nfe[s] irons) with
 * r i-particles per lane and chunk_len j per task (batched = the batched
 * engine's planning), and check every index its kernels derive from it
 * against the buffers it sizes.  0 when all are in bounds, else GI_E_STATE
 * with the first violation in gi_last_error_string().  No reference
 * counterpart. */
int gi_plan_check(int n_sims, const int *n, const int *nfe, int r, int variant, long long chunk_len, int batched);

/* Wavefront tasks per all-pairs kernel dispatch: a longer task list is
 * launched as consecutive dispatches of at most max_tasks tasks on the same
 * stream (0, the default, = one dispatch).  Results are bitwise the same
 * either way: every task writes only its own partial sums.  For hardware
 * counter collection, which covers a bounded number of wavefronts per
 * dispatch, and for callers that share the GPU with latency-bound work.
 * Process-wide; applies to launches made (or graphs captured) after the call.
 * No reference counterpart (the reference launches one N-thread grid). */
int gi_set_dispatch_limit(long long max_tasks);

/* Contact overlap (enable = 1, the default): when every simulation of a
 * force evaluation takes the mutual all-pairs pass, its contact corrections
 * are computed beside that pass on an internal stream and added with the
 * pair sums (and a batched step also fuses its move there); with 0 they are
 * added by a finishing kernel after the pass, and a batched step moves in a
 * kernel of its own.  The two forms share their arithmetic: forces and
 * states are bitwise the same either way.  Process-wide; a plan built under
 * the other setting is rebuilt on its next use.  No reference counterpart. */
int gi_set_contact_overlap(int enable);

/* Contact walk (lanes = 0, the default: nine lanes per particle for force
 * evaluations of up to 49,152 particles, one lane per particle above): the
 * contact corrections' walk over a particle's 27 neighbour cells, forced to
 * one lane per particle (1) or one lane per row of three cells (9).  The two
 * forms share their arithmetic: the corrections are bitwise the same either
 * way.  Process-wide; a plan built under another setting is rebuilt on its
 * next use.  GI_E_INVALID_ARG for other values.  No reference counterpart. */
int gi_set_contact_walk(int lanes);

/* ---- kernel timing (HIP events around the all-pairs force kernel) ------- */

/* Enable/disable per-launch timing (HIP events on the launching stream) of the
 * all-pairs kernel(s) ("gravity"), the whole all-pairs pass ("gravity_pass"),
 * the whole force evaluation, contacts included ("force"), and of
 * gi_move_collision ("move"); when on, gi_kernel_timing returns the launch
 * count and the summed milliseconds since the last reset. */
int gi_set_kernel_timing(int enable);
int gi_kernel_timing(const char *kernel, long long *launches, double *total_ms);
int gi_reset_kernel_timing(void);

#ifdef __cplusplus
}
#endif

#endif /* GIANT_IMPACT_H */

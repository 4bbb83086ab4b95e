/*
 * yk.h -- C ABI of the MI355X Yukawa pair-force / velocity-Verlet library
 *         (libyukawa_hip.so, built from yukawamd_amd/csrc).
 *
 * The reference drives its hot path through two implicit-interface Fortran
 * subroutines that read and write module globals (src/md_globals.f:153-172):
 *
 *   accel(do_measurements)   src/int_control.f:52-81
 *       -> accel_ion_pure    src/int_ion_pure.f:98-254
 *       -> accel_ion_mix     src/int_ion_mix.f:105-295
 *       -> accel_nn          src/int_nn.f:130-345
 *   newton(do_measurements)  src/vel_verlet_2.f:18-120
 *
 * This ABI replaces both.  A handle owns one particle system resident in HBM
 * (structure-of-arrays); the host hands arrays over in the reference's
 * Fortran (3,n) column-major layout and gets them back the same way.  A
 * Fortran host binds every function below with iso_c_binding
 * (yukawamd_amd/fortran/yk_bind.f90).
 *
 * Conventions
 *   - Plain C types only; every function returns YK_OK (0) or a yk_status.
 *     The library never exits the process (the reference's md_exit,
 *     src/md_utils.f:158-183, is the caller's business); yk_last_error()
 *     returns a message for the calling thread's last failure.
 *   - Logicals are int32_t (0 = .false.).
 *   - Host arrays are borrowed for the duration of the call only.
 *   - One host thread per handle; each handle launches on its own HIP stream
 *     on the device that was current at yk_create.  Calls return after the
 *     work they describe is complete, except where noted.
 */
#ifndef YK_H
#define YK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YK_ABI_VERSION 1

typedef struct yk_system yk_system;

/* sim_type, src/int_control.f:60-79 ('B-field' sets a = 0, :71-72). */
enum yk_sim_type {
    YK_PURE_ION = 0,     /* 'pure-ion'    */
    YK_ION_MIXTURE = 1,  /* 'ion-mixture' */
    YK_NUCLEON = 2,      /* 'nucleon'     */
    YK_BFIELD_TEST = 3   /* 'B-field'     */
};

enum yk_precision {
    YK_FP64 = 0,   /* FP64 throughout: the drop-in form                        */
    YK_MIXED = 1   /* FP64 minimum image and accumulation, FP32 pair functions */
};

enum yk_flags {
    /* Keep the reference mixture kernel's last-particle behaviour: particle
     * n-1 never receives its zii factor (src/int_ion_mix.f:158, :201, :250). */
    YK_FLAG_REF_F1 = 1u << 0,
    /* Kick-drift-kick (src/vel_verlet_1.f:16-89) instead of the linked
     * predictor/corrector form (src/vel_verlet_2.f:33-79). */
    YK_FLAG_KDK = 1u << 1,
    /* Record HIP-event timings of every force launch (yk_get_timing). */
    YK_FLAG_TIMING = 1u << 2,
    /* Evaluate all tile pairs (AABB-culled) instead of the cluster-pair
     * lists; fixed at yk_create. */
    YK_FLAG_ALLPAIRS = 1u << 3,
    /* Spreiter-Walter velocity Verlet for the uniform field bfield along z
     * (src/vel_verlet_3.f:20-100) instead of the predictor/corrector form;
     * excludes YK_FLAG_KDK. */
    YK_FLAG_SW = 1u << 4,
    /* Pair lists per 64-particle tile only, without the 16-particle cluster
     * lists built from them (which systems up to 262,144 particles use by
     * default); fixed at yk_create. */
    YK_FLAG_TILE_LISTS = 1u << 5
};

enum yk_status {
    YK_OK = 0,
    YK_ERR_ARG = 1,         /* invalid argument                              */
    YK_ERR_HIP = 2,         /* HIP runtime error                             */
    YK_ERR_NOMEM = 3,       /* device or host allocation failed              */
    YK_ERR_STATE = 4,       /* call not valid in the handle's current state  */
    YK_ERR_NONFINITE = 5,   /* NaN/Inf in an input array                     */
    YK_ERR_UNSUPPORTED = 6, /* combination not implemented                   */
    YK_ERR_CAPACITY = 7     /* admission / capacity limit reached            */
};

/* Run parameters: the module globals accel/newton read
 * (src/md_globals.f:36-140). */
typedef struct yk_params {
    int32_t sim_type;    /* enum yk_sim_type                                  */
    int32_t precision;   /* enum yk_precision                                 */
    uint32_t flags;      /* enum yk_flags                                     */
    int32_t coulomb_on;  /* coulomb .eq. 'screened-coulomb' (int_nn.f:167)    */
    int32_t nuclear_on;  /* nuclear .eq. 'HPP'              (int_nn.f:168)    */
    int32_t resort_every;/* steps between Hilbert re-sorts of the device slots */
                         /* (0 = library default); results are identical up   */
                         /* to summation order whatever the value             */
    double xl[3];        /* box edge lengths (fm)                             */
    double xmuc;         /* inverse screening length (1/fm)                   */
    double rccut;        /* Coulomb cutoff radius (fm)                        */
    double frp;          /* ion form factor                                   */
    double vc;           /* Coulomb coupling ALPHA*HBARC (MeV fm)             */
    double xmass;        /* nucleon mass (MeV)                                */
    double rho;          /* number density (1/fm^3), pressure only            */
    double kT;           /* temperature (MeV), pressure only                  */
    double zi, ai;       /* pure-ion Z and A                                  */
    double aa, bb, cc;   /* HPP strengths (MeV)                               */
    double xpacket;      /* HPP Lambda (fm^2)                                 */
    double rncut;        /* nuclear cutoff radius (fm)                        */
    double bfield;       /* uniform B along z (Gauss), vel_verlet_2.f:61-69   */
} yk_params;

/* Scalars accel/newton leave in md_globals (ev, ek, px, pp). */
typedef struct yk_meas {
    double ev;     /* potential energy per particle (MeV)                    */
    double ek;     /* kinetic energy per particle (MeV); newton only         */
    double px;     /* pressure rho*(kT + vir/(3n)) (MeV/fm^3)                */
    double vir;    /* virial sum (MeV)                                       */
    double pp[6];  /* pressure tensor xx xy xz yy yz zz: accel returns the   */
                   /* virial part (MeV), newton the full tensor / volume     */
} yk_meas;

/* Accumulated device timings, filled when YK_FLAG_TIMING is set. */
typedef struct yk_timing {
    double force_ms;        /* sum of force-kernel durations (HIP events)  */
    int64_t force_launches; /* number of force-kernel launches timed       */
    double step_ms;         /* sum of whole-step durations (HIP events)    */
    int64_t steps;          /* number of steps timed                       */
    int64_t pairs_swept;    /* directed pairs evaluated by the timed force */
                            /* launches since the last yk_reset_timing     */
                            /* (masked lanes included)                     */
} yk_timing;

/* Library -------------------------------------------------------------- */
int32_t yk_abi_version(void);
const char *yk_last_error(void);

/* Create a system of n particles on the current HIP device.  zii/aii are the
 * per-particle charge and mass numbers (src/md_globals.f:158-159); they may be
 * NULL for YK_PURE_ION, which uses params->zi/ai.  Replaces the allocation
 * and parameter setup of input1/initconfig (src/input1.F:394-399,
 * src/initconfig.f:395-399) for the hot path. */
int yk_create(const yk_params *params, int32_t n, const double *zii,
              const double *aii, yk_system **out);
void yk_destroy(yk_system *sys);

/* Replace the run parameters (box strain, screening, cutoff, flags); n,
 * sim_type and precision stay fixed.  Replaces direct writes of md_globals
 * such as strain's xl update (src/strain.f:19-82). */
int yk_set_params(yk_system *sys, const yk_params *params);

/* Upload positions (and velocities; v may be NULL for zero) in (3,n)
 * layout.  Positions are used as given; the integrator wraps them into
 * [0, xl] like src/vel_verlet_2.f:39-40.  Returns YK_ERR_NONFINITE if any
 * coordinate is NaN/Inf. */
int yk_put_state(yk_system *sys, const double *x, const double *v);

/* Upload accelerations (3,n) belonging to the current positions -- e.g. the
 * a(3,0:n-1) a Fortran host kept across a host-side step -- and mark them
 * valid for the next yk_newton.  Replaces the host's own write of a
 * (src/md_globals.f:155). */
int yk_put_accel(yk_system *sys, const double *a);

/* Download positions, velocities, accelerations in (3,n) layout; any
 * pointer may be NULL.  Host readers of particle state (ttot, cancel_vcom,
 * checkpoint: src/gen_utils.f, src/md_utils.f) call this first. */
int yk_get_state(yk_system *sys, double *x, double *v, double *a);

/* accel(do_measurements): accelerations of the current positions, and when
 * do_measurements != 0 also ev, vir, px and the virial pressure tensor.
 * out may be NULL.  src/int_control.f:52-81. */
int yk_accel(yk_system *sys, int32_t do_measurements, yk_meas *out);

/* newton(do_measurements): one velocity-Verlet step of length dt.  Needs
 * accelerations of the current positions (a previous yk_accel/yk_newton).
 * src/vel_verlet_2.f:18-120 (src/vel_verlet_1.f:16-89 with YK_FLAG_KDK,
 * src/vel_verlet_3.f:20-100 with YK_FLAG_SW). */
int yk_newton(yk_system *sys, double dt, int32_t do_measurements, yk_meas *out);

/* nsteps newton steps with state kept on the device.  Step s (1-based)
 * measures when nind > 0 and s % nind == 0 -- the md.F cadence
 * do_measurements = mdmod(istep+1,nind).eq.0 (src/md.F:129) -- and writes
 * its record to records[s/nind - 1] (records may be NULL; it must hold
 * nsteps/nind entries otherwise).  One host synchronisation at the end. */
int yk_run(yk_system *sys, double dt, int32_t nsteps, int32_t nind,
           yk_meas *records);

/* Ttot: kinetic energy per particle of the current velocities,
 * (0.5*xmass*sum A v^2)/n, src/gen_utils.f:40-55. */
int yk_ttot(yk_system *sys, double *eka);

/* Velocity utilities of the step loop, on the device with fixed-order
 * reductions and no host synchronisation (the caller's next download or
 * yk_run sees their effect):
 *   yk_tnorm        tnorm, src/gen_utils.f:61-75: v /= sqrt(eka/(1.5*kT))
 *                   with eka the kinetic energy per particle and kT from the
 *                   handle's parameters;
 *   yk_cancel_vcom  cancel_vcom, src/gen_utils.f:82-107: subtract the
 *                   mass-weighted centre-of-mass velocity, then tnorm.
 * Both need kT > 0. */
int yk_tnorm(yk_system *sys);
int yk_cancel_vcom(yk_system *sys);

/* strain, src/strain.f:19-82: one step of normal strain of the box.
 * deps[k] = YK_XNOSTRAIN leaves axis k to the constant-volume rule (all three
 * unstrained: nothing to do); xl0: the gage dimensions.  The new box follows
 * the reference's formulas (host arithmetic; xl_out may be NULL); positions
 * beyond it wrap on the device; the pair lists are rebuilt at the next
 * evaluation; the accelerations stay valid (positions move by whole box
 * edges only). */
#define YK_XNOSTRAIN 1000.0
int yk_strain(yk_system *sys, double dt, const double deps[3], const double xl0[3],
              double xl_out[3]);

/* The measurement-phase step loop of src/md.F:124-133 on the device. */
typedef struct yk_loop {
    double dt;
    int32_t nsteps;    /* steps in this call                                   */
    int32_t istep0;    /* the step counter istep at entry (src/md.F:121-124)   */
    int32_t nind;      /* energy/pressure measurement period (0: never)        */
    int32_t nptensor;  /* pressure-tensor measurement period (0: never)        */
    int32_t ncom;      /* cancel_vcom period (0: never)                        */
    int32_t ntnorm;    /* tnorm period (0: never)                              */
    double deps[3];    /* strain rates, YK_XNOSTRAIN for none                  */
    double xl0[3];     /* gage dimensions (read when a rate is set)            */
} yk_loop;

/* Per step: cancel_vcom when mdmod(istep,ncom) == 0, tnorm when
 * mdmod(istep,ntnorm) == 0, strain, then newton measuring when
 * mdmod(istep+1,nind) == 0 or mdmod(istep+1,nptensor) == 0 (mdmod of a zero
 * period never is 0, src/md.F:248-256).  The r-th measuring step writes
 * records[r] and rec_steps[r] (its istep after the step); records must hold
 * every measuring step (max_records) when not NULL; *nrec gets their count and
 * xl_out the final box (any output may be NULL).  One host synchronisation
 * at the end. */
int yk_run_md(yk_system *sys, const yk_loop *loop, yk_meas *records, int32_t *rec_steps,
              int32_t max_records, int32_t *nrec, double xl_out[3]);

int yk_get_timing(yk_system *sys, yk_timing *out);
int yk_reset_timing(yk_system *sys);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: accel(do_measurements) (src/int_control.f:52-81) of n
 * particles whose arrays already live in device memory, without a handle and
 * without allocating.  x (3,n), zii (n), aii (n) in the reference's layout
 * (zii/aii may be NULL for YK_PURE_ION, which uses params->zi/ai); the
 * accelerations go to a (3,n) and, when meas is not NULL, the measurement
 * (zeros when do_measurements == 0) to the yk_meas meas points at; all five
 * pointers are device memory, each checked to lie in a device allocation of
 * its size (YK_ERR_ARG otherwise).  Forces come from the tiled all-pairs
 * kernel (the YK_FLAG_ALLPAIRS path; FP64 or mixed by params->precision);
 * results equal yk_accel's to summation-order rounding.  The kernels are
 * enqueued on `stream` (a hipStream_t; NULL: the null stream) and the call
 * returns without synchronising.  Work memory is one static device array per
 * GPU (about 24 MiB, held until the process ends) shared by every call on
 * that GPU: calls that overlap on one GPU race.  n > YK_NOALLOC_MAX_N:
 * YK_ERR_ARG. */
#define YK_NOALLOC_MAX_N 16384
int yk_accel_noalloc(const yk_params *params, int32_t n, const double *x, const double *zii,
                     const double *aii, double *a, int32_t do_measurements, yk_meas *meas,
                     void *stream);

/* Batched ensembles and continuous batching ------------------------------
 *
 * No counterpart in the reference, which runs one system per process.  A
 * yk_batch holds many independent ion systems (pure or mixture, FP64 or
 * mixed precision, each with its own N, box, screening, cutoff, timestep and
 * step budget) in one device arena and advances them together: one
 * persistent pair launch per step pulls (replica, target tile, source range)
 * work items from a device queue, and the step is replayed from a captured
 * hipGraph.  Replicas may be added and removed between runs; admission is by
 * arena footprint (yk_batch_footprint) and fails with YK_ERR_CAPACITY when
 * the arena or the replica table is full.  A replica's results do not depend
 * on the other replicas in the arena.  Integrator: vel_verlet_2
 * (src/vel_verlet_2.f:18-120) without a B field. */
typedef struct yk_batch yk_batch;

const char *yk_batch_last_error(void);

/* capacity: particle slots in the arena (rounded up to 256); max_records:
 * measurement records kept per replica. */
int yk_batch_create(int64_t capacity, int32_t max_replicas, int32_t max_records,
                    yk_batch **out);
void yk_batch_destroy(yk_batch *b);

/* Footprint of a replica of n particles under b's current tuning: the arena
 * slots it occupies and the device bytes it holds (its slots, the per-item
 * partial sums of its work items and, for a list replica, its list region of
 * n^2/4 bytes at most). */
int yk_batch_footprint(const yk_batch *b, int32_t n, int64_t *slots, int64_t *bytes);

/* Admit a replica: positions/velocities in (3,n) layout (v may be NULL),
 * zii/aii as in yk_create (NULL allowed for pure-ion), its timestep and its
 * step budget.  Accelerations of the initial positions are computed at the
 * next yk_batch_run (the initial record, yk_batch_get's init).  The device
 * memory the replica needs (its list region, the membership's work-item
 * buffers) is taken here: YK_ERR_NOMEM when the device is short of it,
 * YK_ERR_CAPACITY when the arena or the replica table is full; either leaves
 * the ensemble as it was. */
int yk_batch_add(yk_batch *b, const yk_params *params, int32_t n, const double *zii,
                 const double *aii, const double *x, const double *v, double dt,
                 int32_t steps, int32_t *id);
int yk_batch_remove(yk_batch *b, int32_t id);

/* Tuning per N bucket (first k with n <= n_upper[k]): list_items > 0 gives
 * the bucket's replicas (up to 262,144 particles) 16-slot cluster pair lists
 * with that many work items per 64-slot tile; list_items = 0 gives them tile
 * items of chunk_tiles source tiles each.  waves_per_block (1, 2, 4 or 8):
 * the persistent launch takes the waves of the bucket that holds the largest
 * share of the pair work (sum of n^2) among the active replicas.  A replica
 * keeps the list / tile choice it was admitted with.  Without a table:
 * lists with 16 items per tile up to 65,536 particles.  Returns YK_ERR_ARG
 * for chunk_tiles < 1, list_items outside [0, 1024], another waves_per_block
 * or n_upper not increasing. */
int yk_batch_set_tuning(yk_batch *b, int32_t nbuckets, const int32_t *n_upper,
                        const int32_t *chunk_tiles, const int32_t *waves_per_block,
                        const int32_t *list_items);

/* Re-sort a replica's slots into Hilbert order every `every` of its own steps
 * (0: never; default 10), keeping tiles compact as particles move.  The
 * re-sort points depend only on the replica's own step count. */
int yk_batch_set_resort(yk_batch *b, int32_t every);

/* Advance every replica by up to nsteps (a replica stops when its budget is
 * used up), measuring every nind-th step (src/md.F:129 cadence). */
int yk_batch_run(yk_batch *b, int32_t nsteps, int32_t nind);

int yk_batch_status(yk_batch *b, int32_t id, int32_t *steps_left, int32_t *nrec);

/* State of one replica in its caller order, the initial measurement and up
 * to rec_cap records; any pointer may be NULL. */
int yk_batch_get(yk_batch *b, int32_t id, double *x, double *v, double *a, yk_meas *init,
                 yk_meas *records, int32_t rec_cap, int32_t *nrec);

/* Device time (HIP events on the batch's stream) of the last yk_batch_run,
 * its step count and the number of work items per step. */
int yk_batch_last_run(yk_batch *b, double *ms, int64_t *steps, int64_t *items);

#ifdef __cplusplus
}
#endif
#endif /* YK_H */

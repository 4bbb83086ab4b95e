/*
 * yk_oracle.h -- CPU restatement of the yukawamd hot path (TEST INFRASTRUCTURE).
 *
 * This is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  The GPU library in
 * yukawamd_amd/csrc never links or calls it.
 *
 * Every routine restates one reference routine with the same loop nest and
 * the same floating-point evaluation order, so that with one thread it
 * reproduces the reference Fortran bit for bit (same libm exp/sqrt):
 *   yko_accel_ion_pure  <- accel_ion_pure   src/int_ion_pure.f:98-254
 *   yko_accel_ion_mix   <- accel_ion_mix    src/int_ion_mix.f:105-295
 *   yko_accel_nn        <- accel_nn         src/int_nn.f:130-345
 *   yko_accel           <- accel            src/int_control.f:52-81
 *   yko_newton          <- newton           src/vel_verlet_2.f:18-120
 *   yko_newton_kdk      <- newton           src/vel_verlet_1.f:16-89 (algorithm spec)
 *   yko_newton_sw       <- newton           src/vel_verlet_3.f:20-100 (algorithm spec)
 *   yko_ttot            <- ttot             src/gen_utils.f:40-55
 *   yko_tnorm           <- tnorm            src/gen_utils.f:61-75
 *   yko_cancel_vcom     <- cancel_vcom      src/gen_utils.f:82-107
 *   yko_strain          <- strain           src/strain.f:19-82
 *   yko_accel_truth     <- same physics, full N x N sum in long double
 *
 * Arrays are the reference's Fortran column-major (3,n) layout: element (k,i)
 * (k = 0..2) sits at x[3*i + k].
 */
#ifndef YK_ORACLE_H
#define YK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { YKO_PURE = 0, YKO_MIX = 1, YKO_NN = 2, YKO_BFIELD_TEST = 3 };

typedef struct yko_params {
    int kind;          /* sim_type: 'pure-ion', 'ion-mixture', 'nucleon', 'B-field' */
    int coulomb_on;    /* coulomb .eq. 'screened-coulomb' (nucleon path)          */
    int nuclear_on;    /* nuclear .eq. 'HPP' (nucleon path)                        */
    int ref_f1;        /* 1: keep the reference's missing zii(n-1) factor (mixture) */
    double xl[3];      /* box edge lengths (fm)                                     */
    double xmuc;       /* inverse screening length (1/fm)                           */
    double rccut;      /* Coulomb cutoff (fm)                                       */
    double frp;        /* ion form factor                                           */
    double vc;         /* Coulomb coupling ALPHA*HBARC (MeV fm)                     */
    double xmass;      /* nucleon mass (MeV)                                        */
    double rho;        /* number density (1/fm^3)                                   */
    double kT;         /* temperature (MeV)                                         */
    double zi, ai;     /* pure-ion charge and mass numbers                          */
    double aa, bb, cc; /* HPP strengths (MeV)                                       */
    double xpacket;    /* HPP Lambda (fm^2)                                         */
    double rncut;      /* nuclear cutoff (fm)                                       */
    double bfield;     /* uniform B along z (Gauss), vel_verlet_2 only              */
} yko_params;

/* meas[0]=ev (per particle) meas[1]=vir meas[2]=px
 * meas[3..8]=pp11 pp12 pp13 pp22 pp23 pp33, meas[9]=ek (newton only)          */
#define YKO_NMEAS 10

/* a must hold 3*(n+3) doubles (the reference's a(3,0:n+2)). Returns 0 or -1. */
int yko_accel(const yko_params *p, int n, const double *x, const double *zii,
              const double *aii, int do_meas, double *a, double *meas);

/* One vel_verlet_2 step on x, v, vold, a (a: 3*(n+3) doubles, holding the
 * accelerations of the current positions on entry). */
int yko_newton(const yko_params *p, int n, double dt, double *x, double *v,
               double *vold, double *a, const double *zii, const double *aii,
               int do_meas, double *meas);

/* One vel_verlet_1 (kick-drift-kick) step. */
int yko_newton_kdk(const yko_params *p, int n, double dt, double *x, double *v,
                   double *a, const double *zii, const double *aii, int do_meas,
                   double *meas);

/* One vel_verlet_3 (Spreiter-Walter, uniform B field along z) step. */
int yko_newton_sw(const yko_params *p, int n, double dt, double *x, double *v,
                  double *a, const double *zii, const double *aii, int do_meas,
                  double *meas);

double yko_ttot(const yko_params *p, int n, const double *v, const double *aii);
/* Scale v to the kinetic energy 1.5 kT per particle. */
void yko_tnorm(const yko_params *p, int n, double *v, const double *aii);
/* Remove the centre-of-mass velocity (mass weights aii), then tnorm. */
void yko_cancel_vcom(const yko_params *p, int n, double *v, const double *aii);
/* XNOSTRAIN in deps[k] means no strain rate on axis k (src/md_globals.f:22). */
#define YKO_XNOSTRAIN 1000.0
/* Box update for one step of strain rates deps from the gage dimensions
 * xl0 (p->xl is updated), then the wrap of positions beyond the new box. */
void yko_strain(yko_params *p, int n, double *x, const double *xl0, const double *deps,
                double dt);

/* Full-matrix force/energy/virial in long double: the "truth" that both the
 * reference and the GPU are compared with.  force[3*i+k] is the physical force
 * (MeV/fm) with no F1 defect; out: ev_total, vir, pp[6] (MeV). */
int yko_accel_truth(const yko_params *p, int n, const double *x,
                    const double *zii, double *force, double *out);

/* OpenMP-parallel restatement of the force-only branch (the common step),
 * used to time the CPU on the host cores.  Returns the thread count used. */
int yko_accel_omp(const yko_params *p, int n, const double *x, const double *zii,
                  const double *aii, double *a, int nthreads);

/* Conditioning of each particle's force: abs_sum[i] = sum_j |F_ij|, rms[i] =
 * sqrt(sum_j |F_ij|^2) (MeV/fm; test metrics, not a reference routine).
 * Returns the thread count used. */
int yko_pair_abs_sum(const yko_params *p, int n, const double *x, const double *zii,
                     double *abs_sum, double *rms, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

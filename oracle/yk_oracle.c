/*
 * yk_oracle.c -- CPU restatement of the yukawamd hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU library is
 * checked against; it is never linked into or called by the product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off -fopenmp (oracle/Makefile).
 * No -march and no fast-math: x86-64 baseline has no FMA, so every a*b+c is
 * two roundings, exactly as the reference Fortran compiles on this host.
 *
 * Evaluation order follows the reference line by line (cited per routine);
 * Fortran evaluates a*b*c as (a*b)*c, and real(4) literals that appear in
 * double expressions (0.5, 1., 2., 3.0) are exact.  Two real(4) subtleties
 * are kept: float(n) and 3.0*n are single-precision values.
 */
#include "yk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Minimum image exactly as src/int_ion_mix.f:163-165: two sequential ifs. */
static inline double min_image(double d, double half, double l)
{
    if (d > +half) d = d - l;
    if (d < -half) d = d + l;
    return d;
}

/* evcut switch, src/int_ion_mix.f:140-145 (rccut <= half box diagonal). */
static double half_diag(const double *xl)
{
    /* sqrt(xl(1)**2+xl(2)**2+xl(3)**2)/2.0 */
    return sqrt(xl[0] * xl[0] + xl[1] * xl[1] + xl[2] * xl[2]) / 2.0;
}

/* px = rho*(kT+vir/(3.0*n)); 3.0*n is a real(4) product. */
static double pressure(const yko_params *p, double vir, int n)
{
    float three_n = 3.0f * (float)n;
    return p->rho * (p->kT + vir / (double)three_n);
}

/* ------------------------------------------------------------------------ */
/* accel_ion_pure, src/int_ion_pure.f:98-254                                 */
static void accel_ion_pure(const yko_params *p, int n, const double *x,
                           int do_meas, double *a, double *meas)
{
    double *fj = (double *)calloc((size_t)3 * (n + 3), sizeof(double));
    double halfl[3], xx[3], fi[3], fcx[3];
    const double *xl = p->xl;
    double rccut2, evcut, evx, virx, pxx, pxy, pxz, pyy, pyz, pzz;
    for (int k = 0; k < 3; ++k) halfl[k] = 0.5 * xl[k];
    rccut2 = p->rccut * p->rccut;
    evcut = (p->rccut <= half_diag(xl)) ? exp(-(p->xmuc * p->rccut)) / p->rccut : 0.0;
    evx = virx = 0.0;
    fi[0] = fi[1] = fi[2] = 0.0;
    pxx = pxy = pxz = pyy = pyz = pzz = 0.0;

    for (int i = 0; i <= n - 2; ++i) {
        /* per-i partials of the omp reduction; combined into the shared
         * accumulators at the end of the j loop (:150-176) */
        double pevx = 0, pvirx = 0, pfi[3] = {0, 0, 0};
        double ppxx = 0, ppxy = 0, ppxz = 0, ppyy = 0, ppyz = 0, ppzz = 0;
        for (int j = i + 1; j <= n - 1; ++j) {
            double r2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                xx[k] = min_image(x[3 * i + k] - x[3 * j + k], halfl[k], xl[k]);
                r2 = r2 + xx[k] * xx[k];
            }
            if (r2 <= rccut2) {
                double r = sqrt(r2), fc;
                if (do_meas) {
                    fc = exp(-(p->xmuc * r)) / (r * r2);          /* :156 */
                    pevx = pevx + (fc * r2 - evcut);              /* :157 */
                    fc = (1.0 + p->xmuc * r) * fc;                /* :158 */
                    pvirx = pvirx + fc * r2;                      /* :159 */
                } else {
                    fc = (1.0 + p->xmuc * r) * exp(-(p->xmuc * r)) / (r * r2); /* :196 */
                }
                for (int k = 0; k < 3; ++k) {
                    fcx[k] = fc * xx[k];
                    pfi[k] = pfi[k] + fcx[k];
                    fj[3 * j + k] = fj[3 * j + k] - fcx[k];
                }
                if (do_meas) {
                    ppxx = ppxx + xx[0] * fcx[0];
                    ppxy = ppxy + xx[0] * fcx[1];
                    ppxz = ppxz + xx[0] * fcx[2];
                    ppyy = ppyy + xx[1] * fcx[1];
                    ppyz = ppyz + xx[1] * fcx[2];
                    ppzz = ppzz + xx[2] * fcx[2];
                }
            }
        }
        if (do_meas) {
            evx = evx + pevx; virx = virx + pvirx;
            pxx = pxx + ppxx; pxy = pxy + ppxy; pxz = pxz + ppxz;
            pyy = pyy + ppyy; pyz = pyz + ppyz; pzz = pzz + ppzz;
        }
        for (int k = 0; k < 3; ++k) fi[k] = fi[k] + pfi[k];
        for (int k = 0; k < 3; ++k) fj[3 * i + k] = fj[3 * i + k] + fi[k]; /* :172 */
        fi[0] = fi[1] = fi[2] = 0.0;
    }
    fj[3 * n + 0] = evx; fj[3 * n + 1] = virx; fj[3 * n + 2] = 0.0;
    fj[3 * n + 3] = pxx; fj[3 * n + 4] = pxy; fj[3 * n + 5] = pxz;
    fj[3 * n + 6] = pyy; fj[3 * n + 7] = pyz; fj[3 * n + 8] = pzz;
    memcpy(a, fj, sizeof(double) * 3 * (size_t)(n + 3));       /* allreduce stub */
    for (int i = 0; i < n; ++i)                                 /* :229 */
        for (int k = 0; k < 3; ++k)
            a[3 * i + k] = p->frp * p->zi * p->zi * p->vc * a[3 * i + k] / (p->ai * p->xmass);
    memset(meas, 0, sizeof(double) * YKO_NMEAS);
    if (do_meas) {
        double s = p->frp * p->zi * p->zi * p->vc;
        double ev = s * a[3 * n + 0], vir = s * a[3 * n + 1];
        meas[3] = s * a[3 * n + 3]; meas[4] = s * a[3 * n + 4]; meas[5] = s * a[3 * n + 5];
        meas[6] = s * a[3 * n + 6]; meas[7] = s * a[3 * n + 7]; meas[8] = s * a[3 * n + 8];
        meas[0] = ev / (double)(float)n;
        meas[1] = vir;
        meas[2] = pressure(p, vir, n);
    }
    free(fj);
}

/* ------------------------------------------------------------------------ */
/* accel_ion_mix, src/int_ion_mix.f:105-295 (including the F1 behaviour:    */
/* particle n-1 never receives its zii factor, :158/:201/:250)              */
static void accel_ion_mix(const yko_params *p, int n, const double *x,
                          const double *zii, const double *aii, int do_meas,
                          double *a, double *meas)
{
    double *fj = (double *)calloc((size_t)3 * (n + 3), sizeof(double));
    double halfl[3], xx[3];
    const double *xl = p->xl;
    double ev = 0, vir = 0, pp11 = 0, pp12 = 0, pp13 = 0, pp22 = 0, pp23 = 0, pp33 = 0;
    double rccut2, evcut;
    for (int k = 0; k < 3; ++k) halfl[k] = 0.5 * xl[k];
    rccut2 = p->rccut * p->rccut;
    evcut = (p->rccut <= half_diag(xl)) ? exp(-(p->xmuc * p->rccut)) / p->rccut : 0.0;

    for (int i = 0; i <= n - 2; ++i) {
        double evx = 0, virx = 0, fi[3] = {0, 0, 0};
        double pxx = 0, pxy = 0, pxz = 0, pyy = 0, pyz = 0, pzz = 0;
        const double zi = zii[i];
        for (int j = i + 1; j <= n - 1; ++j) {
            double r2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                xx[k] = min_image(x[3 * i + k] - x[3 * j + k], halfl[k], xl[k]);
                r2 = r2 + xx[k] * xx[k];
            }
            if (r2 <= rccut2) {
                const double zj = zii[j];
                double r = sqrt(r2), fc, f;
                if (do_meas) {
                    fc = exp(-(p->xmuc * r)) / (r * r2);          /* :170 */
                    evx = evx + zj * (fc * r2 - evcut);           /* :171 */
                    fc = (1.0 + p->xmuc * r) * fc;                /* :172 */
                    virx = virx + zj * fc * r2;                   /* :173 */
                } else {
                    fc = (1.0 + p->xmuc * r) * exp(-(p->xmuc * r)) / (r * r2); /* :234 */
                }
                f = fc * xx[0];                                   /* :178-181 */
                fi[0] = fi[0] + zj * f;
                fj[3 * j + 0] = fj[3 * j + 0] - zi * f;
                if (do_meas) pxx = pxx + xx[0] * (zj * f);
                f = fc * xx[1];                                   /* :183-187 */
                fi[1] = fi[1] + zj * f;
                fj[3 * j + 1] = fj[3 * j + 1] - zi * f;
                if (do_meas) { pxy = pxy + xx[0] * (zj * f); pyy = pyy + xx[1] * (zj * f); }
                f = fc * xx[2];                                   /* :189-194 */
                fi[2] = fi[2] + zj * f;
                fj[3 * j + 2] = fj[3 * j + 2] - zi * f;
                if (do_meas) {
                    pxz = pxz + xx[0] * (zj * f);
                    pyz = pyz + xx[1] * (zj * f);
                    pzz = pzz + xx[2] * (zj * f);
                }
            }
        }
        if (do_meas) {                                            /* :199-207 */
            ev = ev + zi * evx;
            vir = vir + zi * virx;
            pp11 = pp11 + zi * pxx; pp12 = pp12 + zi * pxy; pp13 = pp13 + zi * pxz;
            pp22 = pp22 + zi * pyy; pp23 = pp23 + zi * pyz; pp33 = pp33 + zi * pzz;
        }
        for (int k = 0; k < 3; ++k) fj[3 * i + k] = zi * (fj[3 * i + k] + fi[k]);
    }
    if (!p->ref_f1 && n >= 1)   /* corrected form: particle n-1 gets its factor too */
        for (int k = 0; k < 3; ++k) fj[3 * (n - 1) + k] = zii[n - 1] * fj[3 * (n - 1) + k];
    fj[3 * n + 0] = ev; fj[3 * n + 1] = vir; fj[3 * n + 2] = 0.0;
    fj[3 * n + 3] = pp11; fj[3 * n + 4] = pp12; fj[3 * n + 5] = pp13;
    fj[3 * n + 6] = pp22; fj[3 * n + 7] = pp23; fj[3 * n + 8] = pp33;
    memcpy(a, fj, sizeof(double) * 3 * (size_t)(n + 3));
    for (int i = 0; i < n; ++i)                                   /* :275 */
        for (int k = 0; k < 3; ++k)
            a[3 * i + k] = p->frp * p->vc * a[3 * i + k] / (aii[i] * p->xmass);
    memset(meas, 0, sizeof(double) * YKO_NMEAS);
    if (do_meas) {
        double s = p->frp * p->vc;
        double evs = s * a[3 * n + 0], virs = s * a[3 * n + 1];
        meas[3] = s * a[3 * n + 3]; meas[4] = s * a[3 * n + 4]; meas[5] = s * a[3 * n + 5];
        meas[6] = s * a[3 * n + 6]; meas[7] = s * a[3 * n + 7]; meas[8] = s * a[3 * n + 8];
        meas[0] = evs / (double)(float)n;
        meas[1] = virs;
        meas[2] = pressure(p, virs, n);
    }
    free(fj);
}

/* ------------------------------------------------------------------------ */
/* accel_nn, src/int_nn.f:130-345                                            */
static void accel_nn(const yko_params *p, int n, const double *x,
                     const double *zii, int do_meas, double *a, double *meas)
{
    double *fj = (double *)calloc((size_t)3 * (n + 3), sizeof(double));
    double halfl[3], xx[3], fi[3], fncx[3];
    const double *xl = p->xl;
    double xpacketi, xpacket2i, rncut2, rccut2, evcut;
    double evx = 0, virx = 0, pxx = 0, pxy = 0, pxz = 0, pyy = 0, pyz = 0, pzz = 0;
    for (int k = 0; k < 3; ++k) halfl[k] = 0.5 * xl[k];
    xpacketi = 1.0 / p->xpacket;                                  /* :172-173 */
    xpacket2i = 1.0 / (2.0 * p->xpacket);
    rncut2 = p->rncut * p->rncut;
    rccut2 = p->rccut * p->rccut;
    evcut = (p->rccut <= half_diag(xl)) ? p->vc * exp(-(p->xmuc * p->rccut)) / p->rccut : 0.0;
    fi[0] = fi[1] = fi[2] = 0.0;

    for (int i = 0; i <= n - 2; ++i) {
        double pevx = 0, pvirx = 0, pfi[3] = {0, 0, 0};
        double ppxx = 0, ppxy = 0, ppxz = 0, ppyy = 0, ppyz = 0, ppzz = 0;
        for (int j = i + 1; j <= n - 1; ++j) {
            double fn = 0.0, fc = 0.0, r2 = 0.0;
            int fzero = 1;
            for (int k = 0; k < 3; ++k) {
                xx[k] = min_image(x[3 * i + k] - x[3 * j + k], halfl[k], xl[k]);
                r2 = r2 + xx[k] * xx[k];
            }
            if (p->nuclear_on && r2 <= rncut2) {                  /* :208-219 */
                double expfac = exp(-(r2 * xpacket2i));
                double xbbcc = (zii[i] == zii[j]) ? p->bb + p->cc : p->bb - p->cc;
                if (do_meas) pevx = pevx + (p->aa * expfac + xbbcc) * expfac;
                fn = xpacketi * (2.0 * p->aa * expfac + xbbcc) * expfac;
                if (do_meas) pvirx = pvirx + fn * r2;
                fzero = 0;
            }
            if (p->coulomb_on && r2 <= rccut2) {                  /* :220-229 */
                if (zii[i] == 1.0 && zii[j] == 1.0) {
                    double r = sqrt(r2);
                    if (do_meas) {
                        fc = p->vc * exp(-(p->xmuc * r)) / (r * r2);
                        pevx = pevx + (fc * r2 - evcut);
                        fc = (1.0 + p->xmuc * r) * fc;
                        pvirx = pvirx + fc * r2;
                    } else {
                        fc = p->vc * (1.0 + p->xmuc * r) * exp(-(p->xmuc * r)) / (r * r2); /* :282 */
                    }
                    fzero = 0;
                }
            }
            if (!fzero) {
                for (int k = 0; k < 3; ++k) {
                    fncx[k] = (fn + fc) * xx[k];
                    pfi[k] = pfi[k] + fncx[k];
                    fj[3 * j + k] = fj[3 * j + k] - fncx[k];
                }
                if (do_meas) {
                    ppxx = ppxx + xx[0] * fncx[0];
                    ppxy = ppxy + xx[0] * fncx[1];
                    ppxz = ppxz + xx[0] * fncx[2];
                    ppyy = ppyy + xx[1] * fncx[1];
                    ppyz = ppyz + xx[1] * fncx[2];
                    ppzz = ppzz + xx[2] * fncx[2];
                }
            }
        }
        if (do_meas) {
            evx = evx + pevx; virx = virx + pvirx;
            pxx = pxx + ppxx; pxy = pxy + ppxy; pxz = pxz + ppxz;
            pyy = pyy + ppyy; pyz = pyz + ppyz; pzz = pzz + ppzz;
        }
        for (int k = 0; k < 3; ++k) fi[k] = fi[k] + pfi[k];
        for (int k = 0; k < 3; ++k) fj[3 * i + k] = fj[3 * i + k] + fi[k];
        fi[0] = fi[1] = fi[2] = 0.0;
    }
    fj[3 * n + 0] = evx; fj[3 * n + 1] = virx; fj[3 * n + 2] = 0.0;
    fj[3 * n + 3] = pxx; fj[3 * n + 4] = pxy; fj[3 * n + 5] = pxz;
    fj[3 * n + 6] = pyy; fj[3 * n + 7] = pyz; fj[3 * n + 8] = pzz;
    memcpy(a, fj, sizeof(double) * 3 * (size_t)(n + 3));
    for (int i = 0; i < n; ++i)                                   /* :318-322 */
        for (int k = 0; k < 3; ++k) a[3 * i + k] = a[3 * i + k] / p->xmass;
    memset(meas, 0, sizeof(double) * YKO_NMEAS);
    if (do_meas) {
        double ev = a[3 * n + 0], vir = a[3 * n + 1];
        meas[3] = a[3 * n + 3]; meas[4] = a[3 * n + 4]; meas[5] = a[3 * n + 5];
        meas[6] = a[3 * n + 6]; meas[7] = a[3 * n + 7]; meas[8] = a[3 * n + 8];
        meas[0] = ev / (double)(float)n;
        meas[1] = vir;
        meas[2] = pressure(p, vir, n);
    }
    free(fj);
}

/* accel, src/int_control.f:52-81 */
int yko_accel(const yko_params *p, int n, const double *x, const double *zii,
              const double *aii, int do_meas, double *a, double *meas)
{
    if (n < 1) return -1;
    switch (p->kind) {
    case YKO_PURE: accel_ion_pure(p, n, x, do_meas, a, meas); return 0;
    case YKO_MIX: accel_ion_mix(p, n, x, zii, aii, do_meas, a, meas); return 0;
    case YKO_NN: accel_nn(p, n, x, zii, do_meas, a, meas); return 0;
    case YKO_BFIELD_TEST:
        memset(a, 0, sizeof(double) * 3 * (size_t)(n + 3));
        memset(meas, 0, sizeof(double) * YKO_NMEAS);
        return 0;
    default: return -1;
    }
}

/* ------------------------------------------------------------------------ */
/* newton, src/vel_verlet_2.f:18-120                                         */
int yko_newton(const yko_params *p, int n, double dt, double *x, double *v,
               double *vold, double *a, const double *zii, const double *aii,
               int do_meas, double *meas)
{
    const double *xl = p->xl;
    double afac = 0.5 * (dt * dt);                                /* :33 */
    for (int i = 0; i < n; ++i)                                   /* :36-42 */
        for (int k = 0; k < 3; ++k) {
            double xk = x[3 * i + k] + v[3 * i + k] * dt + afac * a[3 * i + k];
            if (xk < 0.0) xk = xk + xl[k];
            if (xk > xl[k]) xk = xk - xl[k];
            x[3 * i + k] = xk;
        }
    for (int i = 0; i < n; ++i)                                   /* :49-52 */
        for (int k = 0; k < 3; ++k) {
            vold[3 * i + k] = v[3 * i + k];
            v[3 * i + k] = v[3 * i + k] + dt * a[3 * i + k];
        }
    if (yko_accel(p, n, x, zii, aii, do_meas, a, meas) != 0) return -1; /* :58 */
    if (p->bfield != 0.0) {                                       /* :61-69 */
        for (int i = 0; i < n; ++i) {
            double omega = (3.218404e-20) * zii[i] * p->bfield / aii[i];
            a[3 * i + 0] = a[3 * i + 0] + v[3 * i + 1] * omega;
            a[3 * i + 1] = a[3 * i + 1] - v[3 * i + 0] * omega;
        }
    }
    for (int i = 0; i < n; ++i)                                   /* :76-79 */
        for (int k = 0; k < 3; ++k) {
            double vp = vold[3 * i + k] + dt * a[3 * i + k];
            v[3 * i + k] = 0.5 * (v[3 * i + k] + vp);
        }
    if (do_meas) {                                                /* :84-106 */
        double vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0, vol;
        for (int i = 0; i < n; ++i) {
            const double *vi = v + 3 * i;
            vxx = vxx + aii[i] * vi[0] * vi[0];
            vxy = vxy + aii[i] * vi[0] * vi[1];
            vxz = vxz + aii[i] * vi[0] * vi[2];
            vyy = vyy + aii[i] * vi[1] * vi[1];
            vyz = vyz + aii[i] * vi[1] * vi[2];
            vzz = vzz + aii[i] * vi[2] * vi[2];
        }
        meas[9] = (0.5 * p->xmass) * (vxx + vyy + vzz) / (double)(float)n;
        vol = p->xl[0] * p->xl[1] * p->xl[2];
        meas[3] = (meas[3] + p->xmass * vxx) / vol;
        meas[4] = (meas[4] + p->xmass * vxy) / vol;
        meas[5] = (meas[5] + p->xmass * vxz) / vol;
        meas[6] = (meas[6] + p->xmass * vyy) / vol;
        meas[7] = (meas[7] + p->xmass * vyz) / vol;
        meas[8] = (meas[8] + p->xmass * vzz) / vol;
    } else {
        for (int k = 3; k < YKO_NMEAS; ++k) meas[k] = 0.0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* newton, src/vel_verlet_1.f:16-89: drift + wrap, half kick, force, half   */
/* kick.  (Not link-compatible in the reference; kept as algorithm spec.)    */
int yko_newton_kdk(const yko_params *p, int n, double dt, double *x, double *v,
                   double *a, const double *zii, const double *aii, int do_meas,
                   double *meas)
{
    const double *xl = p->xl;
    double afac = 0.5 * dt * dt;                                  /* :28 */
    double hdt = 0.5 * dt;                                        /* :43 */
    for (int i = 0; i < n; ++i)                                   /* :30-36 */
        for (int k = 0; k < 3; ++k) {
            double xk = x[3 * i + k] + v[3 * i + k] * dt + a[3 * i + k] * afac;
            if (xk < 0.0) xk = xk + xl[k];
            if (xk > xl[k]) xk = xk - xl[k];
            x[3 * i + k] = xk;
        }
    for (int i = 0; i < n; ++i)                                   /* :45-48 */
        for (int k = 0; k < 3; ++k) v[3 * i + k] = v[3 * i + k] + hdt * a[3 * i + k];
    if (yko_accel(p, n, x, zii, aii, do_meas, a, meas) != 0) return -1;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) v[3 * i + k] = v[3 * i + k] + hdt * a[3 * i + k];
    if (do_meas) {
        double vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0, vol;
        for (int i = 0; i < n; ++i) {
            const double *vi = v + 3 * i;
            vxx = vxx + aii[i] * vi[0] * vi[0];
            vxy = vxy + aii[i] * vi[0] * vi[1];
            vxz = vxz + aii[i] * vi[0] * vi[2];
            vyy = vyy + aii[i] * vi[1] * vi[1];
            vyz = vyz + aii[i] * vi[1] * vi[2];
            vzz = vzz + aii[i] * vi[2] * vi[2];
        }
        meas[9] = (0.5 * p->xmass) * (vxx + vyy + vzz) / (double)(float)n;
        vol = p->xl[0] * p->xl[1] * p->xl[2];
        meas[3] = (meas[3] + p->xmass * vxx) / vol;
        meas[4] = (meas[4] + p->xmass * vxy) / vol;
        meas[5] = (meas[5] + p->xmass * vxz) / vol;
        meas[6] = (meas[6] + p->xmass * vyy) / vol;
        meas[7] = (meas[7] + p->xmass * vyz) / vol;
        meas[8] = (meas[8] + p->xmass * vzz) / vol;
    } else {
        for (int k = 3; k < YKO_NMEAS; ++k) meas[k] = 0.0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* newton, src/vel_verlet_3.f:20-100: Spreiter-Walter velocity Verlet for a  */
/* uniform B field along z ("inversion of ez x v").  Drift + wrap (:31-40),  */
/* first velocity stage from v(t), a(t) (:45-54; the y update reads the x    */
/* component already updated on the line before, as the reference does),    */
/* force at t+dt (:58), second stage (:64-79).  The field constant is the    */
/* reference's single-precision literal 3.218404e-20 (:48, :67) widened to   */
/* double.  (Not link-compatible in the reference; kept as algorithm spec.)  */
int yko_newton_sw(const yko_params *p, int n, double dt, double *x, double *v,
                  double *a, const double *zii, const double *aii, int do_meas,
                  double *meas)
{
    const double *xl = p->xl;
    const double cb = (double)3.218404e-20f;
    double afac = 0.5 * (dt * dt);                                /* :31 */
    double halfdt = 0.5 * dt;                                     /* :45 */
    for (int i = 0; i < n; ++i)                                   /* :33-39 */
        for (int k = 0; k < 3; ++k) {
            double xk = x[3 * i + k] + v[3 * i + k] * dt + afac * a[3 * i + k];
            if (xk < 0.0) xk = xk + xl[k];
            if (xk > xl[k]) xk = xk - xl[k];
            x[3 * i + k] = xk;
        }
    for (int i = 0; i < n; ++i) {                                 /* :47-53 */
        double *vi = v + 3 * i;
        const double *ai = a + 3 * i;
        double omega = cb * zii[i] * p->bfield / aii[i];
        double omegap = halfdt * omega;
        vi[0] = (1.0 - omegap * omegap) * vi[0] + 2.0 * omegap * vi[1] +
                halfdt * (ai[0] + omegap * ai[1]);
        vi[1] = (1.0 - omegap * omegap) * vi[1] - 2.0 * omegap * vi[0] +
                halfdt * (ai[1] - omegap * ai[0]);
        vi[2] = vi[2] + halfdt * ai[2];
    }
    if (yko_accel(p, n, x, zii, aii, do_meas, a, meas) != 0) return -1;
    double vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
    for (int i = 0; i < n; ++i) {                                 /* :66-78 */
        double *vi = v + 3 * i;
        const double *ai = a + 3 * i;
        double omega = cb * zii[i] * p->bfield / aii[i];
        double omegap = halfdt * omega;
        vi[0] = (vi[0] + halfdt * (ai[0] + omegap * ai[1])) / (1.0 + omegap * omegap);
        vi[1] = (vi[1] + halfdt * (ai[1] - omegap * ai[0])) / (1.0 + omegap * omegap);
        vi[2] = vi[2] + halfdt * ai[2];
        vxx = vxx + aii[i] * vi[0] * vi[0];
        vxy = vxy + aii[i] * vi[0] * vi[1];
        vxz = vxz + aii[i] * vi[0] * vi[2];
        vyy = vyy + aii[i] * vi[1] * vi[1];
        vyz = vyz + aii[i] * vi[1] * vi[2];
        vzz = vzz + aii[i] * vi[2] * vi[2];
    }
    if (do_meas) {                                                /* :80-90 */
        double vol = p->xl[0] * p->xl[1] * p->xl[2];
        meas[9] = (0.5 * p->xmass) * (vxx + vyy + vzz) / (double)(float)n;
        meas[3] = (meas[3] + p->xmass * vxx) / vol;
        meas[4] = (meas[4] + p->xmass * vxy) / vol;
        meas[5] = (meas[5] + p->xmass * vxz) / vol;
        meas[6] = (meas[6] + p->xmass * vyy) / vol;
        meas[7] = (meas[7] + p->xmass * vyz) / vol;
        meas[8] = (meas[8] + p->xmass * vzz) / vol;
    } else {
        for (int k = 3; k < YKO_NMEAS; ++k) meas[k] = 0.0;
    }
    return 0;
}

/* ttot, src/gen_utils.f:40-55 */
double yko_ttot(const yko_params *p, int n, const double *v, const double *aii)
{
    double ekx = 0.0;
    for (int i = 0; i < n; ++i) {
        const double *vi = v + 3 * i;
        ekx = ekx + aii[i] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    }
    return (0.5 * p->xmass * ekx) / (double)(float)n;
}

/* tnorm, src/gen_utils.f:61-75: fac = sqrt(eka/(1.5*kT)), v = v/fac. */
void yko_tnorm(const yko_params *p, int n, double *v, const double *aii)
{
    const double eka = yko_ttot(p, n, v, aii);
    const double fac = sqrt(eka / (1.5 * p->kT));
    for (int i = 0; i < 3 * n; ++i) v[i] = v[i] / fac;
}

/* cancel_vcom, src/gen_utils.f:82-107: mass-weighted sums in particle order,
 * vcom = ptot/mtot subtracted, then tnorm. */
void yko_cancel_vcom(const yko_params *p, int n, double *v, const double *aii)
{
    double mtot = 0.0, ptot[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        mtot = mtot + aii[i];
        for (int k = 0; k < 3; ++k) ptot[k] = ptot[k] + aii[i] * v[3 * i + k];
    }
    double vcom[3];
    for (int k = 0; k < 3; ++k) vcom[k] = ptot[k] / mtot;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) v[3 * i + k] = v[3 * i + k] - vcom[k];
    yko_tnorm(p, n, v, aii);
}

/* strain, src/strain.f:19-82.  The strained axes grow by deps*dt*xl0; with one
 * or two strained axes the others follow from constant volume, with three
 * all follow their rates.  Then positions beyond the new box wrap once. */
void yko_strain(yko_params *p, int n, double *x, const double *xl0, const double *deps,
                double dt)
{
    double *xl = p->xl;
    const int c = (deps[0] != YKO_XNOSTRAIN ? 1 : 0) | (deps[1] != YKO_XNOSTRAIN ? 2 : 0) |
                  (deps[2] != YKO_XNOSTRAIN ? 4 : 0);
    switch (c) {
    case 1:
        xl[0] = xl[0] + deps[0] * dt * xl0[0];
        xl[1] = xl0[1] * sqrt(xl0[0] / xl[0]);
        xl[2] = xl0[2] * sqrt(xl0[0] / xl[0]);
        break;
    case 2:
        xl[1] = xl[1] + deps[1] * dt * xl0[1];
        xl[0] = xl0[0] * sqrt(xl0[1] / xl[1]);
        xl[2] = xl0[2] * sqrt(xl0[1] / xl[1]);
        break;
    case 3:
        xl[0] = xl[0] + deps[0] * dt * xl0[0];
        xl[1] = xl[1] + deps[1] * dt * xl0[1];
        xl[2] = xl0[2] * (xl0[0] / xl[0]) * (xl0[1] / xl[1]);
        break;
    case 4:
        xl[2] = xl[2] + deps[2] * dt * xl0[2];
        xl[0] = xl0[0] * sqrt(xl0[2] / xl[2]);
        xl[1] = xl0[1] * sqrt(xl0[2] / xl[2]);
        break;
    case 5:
        xl[0] = xl[0] + deps[0] * dt * xl0[0];
        xl[2] = xl[2] + deps[2] * dt * xl0[2];
        xl[1] = xl0[1] * (xl0[0] / xl[0]) * (xl0[2] / xl[2]);
        break;
    case 6:
        xl[1] = xl[1] + deps[1] * dt * xl0[1];
        xl[2] = xl[2] + deps[2] * dt * xl0[2];
        xl[0] = xl0[0] * (xl0[1] / xl[1]) * (xl0[2] / xl[2]);
        break;
    case 7:
        for (int k = 0; k < 3; ++k) xl[k] = xl[k] + deps[k] * dt * xl0[k];
        break;
    default:
        break;
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            if (x[3 * i + k] > xl[k]) x[3 * i + k] = x[3 * i + k] - xl[k];
}

/* ------------------------------------------------------------------------ */
/* Long-double full-matrix evaluation of the same physics (no F1 defect).    */
/* Pair energies are summed per i over all j != i and halved.                */
int yko_accel_truth(const yko_params *p, int n, const double *x,
                    const double *zii, double *force, double *out)
{
    typedef long double ld;
    const double *xl = p->xl;
    ld halfl[3], ev = 0, vir = 0, pp[6] = {0, 0, 0, 0, 0, 0};
    ld mu = p->xmuc, rc2 = (ld)p->rccut * p->rccut, evcut;
    ld rn2 = (ld)p->rncut * p->rncut;
    for (int k = 0; k < 3; ++k) halfl[k] = 0.5L * xl[k];
    if (p->kind == YKO_NN)
        evcut = (p->rccut <= half_diag(xl)) ? (ld)p->vc * expl(-mu * p->rccut) / p->rccut : 0;
    else
        evcut = (p->rccut <= half_diag(xl)) ? expl(-mu * p->rccut) / p->rccut : 0;
    for (int i = 0; i < n; ++i) {
        ld f[3] = {0, 0, 0}, e = 0, w = 0, q[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < n; ++j) {
            ld d[3], r2 = 0, s = 0, phi = 0;
            int hit = 0;
            if (j == i) continue;
            for (int k = 0; k < 3; ++k) {
                ld dk = (ld)x[3 * i + k] - (ld)x[3 * j + k];
                if (dk > halfl[k]) dk -= xl[k];
                if (dk < -halfl[k]) dk += xl[k];
                d[k] = dk;
                r2 += dk * dk;
            }
            if (p->kind == YKO_NN) {
                if (p->nuclear_on && r2 <= rn2) {
                    ld g = expl(-r2 / (2.0L * p->xpacket));
                    ld bc = (zii[i] == zii[j]) ? (ld)p->bb + p->cc : (ld)p->bb - p->cc;
                    phi += (p->aa * g + bc) * g;
                    s += (2.0L * p->aa * g + bc) * g / p->xpacket;
                    hit = 1;
                }
                if (p->coulomb_on && r2 <= rc2 && zii[i] == 1.0 && zii[j] == 1.0) {
                    ld r = sqrtl(r2), ex = expl(-mu * r);
                    phi += p->vc * ex / r - evcut;
                    s += p->vc * (1 + mu * r) * ex / (r * r2);
                    hit = 1;
                }
            } else if (r2 <= rc2) {
                ld zz = (p->kind == YKO_PURE) ? (ld)p->zi * p->zi : (ld)zii[i] * zii[j];
                ld r = sqrtl(r2), ex = expl(-mu * r);
                ld scale = (ld)p->frp * p->vc * zz;
                phi = scale * (ex / r - evcut);
                s = scale * (1 + mu * r) * ex / (r * r2);
                hit = 1;
            }
            if (!hit) continue;
            for (int k = 0; k < 3; ++k) f[k] += s * d[k];
            e += phi;
            w += s * r2;
            q[0] += d[0] * s * d[0]; q[1] += d[0] * s * d[1]; q[2] += d[0] * s * d[2];
            q[3] += d[1] * s * d[1]; q[4] += d[1] * s * d[2]; q[5] += d[2] * s * d[2];
        }
        for (int k = 0; k < 3; ++k) force[3 * i + k] = (double)f[k];
        ev += e; vir += w;
        for (int k = 0; k < 6; ++k) pp[k] += q[k];
    }
    out[0] = (double)(0.5L * ev);
    out[1] = (double)(0.5L * vir);
    for (int k = 0; k < 6; ++k) out[2 + k] = (double)(0.5L * pp[k]);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's OpenMP structure for the force-only mixture branch        */
/* (src/int_ion_mix.f:221-254): i serial, j work-shared with a reduction on  */
/* fi, one implicit barrier per i, reaction on fj(:,j) race-free because     */
/* each j is owned by one thread within an i.  Used only to time the CPU.    */
int yko_accel_omp(const yko_params *p, int n, const double *x, const double *zii,
                  const double *aii, double *a, int nthreads)
{
    double *fj = (double *)calloc((size_t)3 * (n + 3), sizeof(double));
    const double *xl = p->xl;
    double halfl[3];
    double rccut2 = p->rccut * p->rccut, mu = p->xmuc;
    int used = 1;
    int mix = (p->kind == YKO_MIX);
    double f0 = 0, f1 = 0, f2 = 0;   /* shared; the reduction targets */
    for (int k = 0; k < 3; ++k) halfl[k] = 0.5 * xl[k];
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
#ifdef _OPENMP
#pragma omp single
        used = omp_get_num_threads();
#endif
        for (int i = 0; i <= n - 2; ++i) {
            const double zi = mix ? zii[i] : 1.0;
#pragma omp single
            f0 = f1 = f2 = 0.0;
#pragma omp for reduction(+ : f0, f1, f2) schedule(static)
            for (int j = i + 1; j <= n - 1; ++j) {
                double xx[3], r2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    xx[k] = min_image(x[3 * i + k] - x[3 * j + k], halfl[k], xl[k]);
                    r2 = r2 + xx[k] * xx[k];
                }
                if (r2 <= rccut2) {
                    double r = sqrt(r2);
                    double fc = (1.0 + mu * r) * exp(-(mu * r)) / (r * r2);
                    double zj = mix ? zii[j] : 1.0;
                    double f;
                    f = fc * xx[0]; f0 += zj * f; fj[3 * j + 0] -= zi * f;
                    f = fc * xx[1]; f1 += zj * f; fj[3 * j + 1] -= zi * f;
                    f = fc * xx[2]; f2 += zj * f; fj[3 * j + 2] -= zi * f;
                }
            }
#pragma omp single
            {
                fj[3 * i + 0] = zi * (fj[3 * i + 0] + f0);
                fj[3 * i + 1] = zi * (fj[3 * i + 1] + f1);
                fj[3 * i + 2] = zi * (fj[3 * i + 2] + f2);
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        double s = mix ? p->frp * p->vc / (aii[i] * p->xmass)
                       : p->frp * p->zi * p->zi * p->vc / (p->ai * p->xmass);
        for (int k = 0; k < 3; ++k) a[3 * i + k] = s * fj[3 * i + k];
    }
    free(fj);
    return used;
}

/* ------------------------------------------------------------------------ */
/* Conditioning of each particle's force: abs_sum[i] = sum_j |F_ij| and      */
/* rms[i] = sqrt(sum_j |F_ij|^2) over the pairs inside the cutoffs (physical */
/* pair-force magnitudes in MeV/fm, the nucleon's nuclear and Coulomb terms  */
/* counted separately).  A force computed with per-pair relative error e has */
/* an error of at most e * abs_sum[i], and about e * rms[i] when the pair    */
/* errors are independent.  Full N x N loop over i in parallel.             */
int yko_pair_abs_sum(const yko_params *p, int n, const double *x, const double *zii,
                     double *abs_sum, double *rms, int nthreads)
{
    const double *xl = p->xl;
    double halfl[3];
    const double mu = p->xmuc, rc2 = p->rccut * p->rccut, rn2 = p->rncut * p->rncut;
    int used = 1;
    for (int k = 0; k < 3; ++k) halfl[k] = 0.5 * xl[k];
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
#ifdef _OPENMP
#pragma omp single
        used = omp_get_num_threads();
#endif
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            double sa = 0.0, s2 = 0.0;
            for (int j = 0; j < n; ++j) {
                double r2 = 0.0, f1 = 0.0, f2 = 0.0;
                if (j == i) continue;
                for (int k = 0; k < 3; ++k) {
                    const double dk = min_image(x[3 * i + k] - x[3 * j + k], halfl[k], xl[k]);
                    r2 += dk * dk;
                }
                const double r = sqrt(r2);
                if (p->kind == YKO_NN) {
                    if (p->nuclear_on && r2 <= rn2) {
                        const double g = exp(-r2 / (2.0 * p->xpacket));
                        const double bc = (zii[i] == zii[j]) ? p->bb + p->cc : p->bb - p->cc;
                        f1 = fabs((2.0 * p->aa * g + bc) * g / p->xpacket) * r;
                    }
                    if (p->coulomb_on && r2 <= rc2 && zii[i] == 1.0 && zii[j] == 1.0)
                        f2 = p->vc * (1.0 + mu * r) * exp(-mu * r) / r2;
                } else if (r2 <= rc2) {
                    const double zz = (p->kind == YKO_PURE) ? p->zi * p->zi : zii[i] * zii[j];
                    f1 = fabs(p->frp * p->vc * zz) * (1.0 + mu * r) * exp(-mu * r) / r2;
                }
                sa += f1 + f2;
                s2 += f1 * f1 + f2 * f2;
            }
            abs_sum[i] = sa;
            rms[i] = sqrt(s2);
        }
    }
    return used;
}

/*
 * qpad_oracle.c -- TEST INFRASTRUCTURE: serial FP64 restatement of QPAD's
 * slice step (source/simulation_class.f03:342-469), 3D step
 * (:294-512) and the kernels they call.  See qpad_oracle.h for the layout,
 * the role of this file and what is pinned by what.
 *
 * Arithmetic follows the reference's Fortran evaluation order (left to
 * right, as gfortran evaluates it without fast-math) so that the restatement
 * is faithful operation by operation.  Compile with -ffp-contract=off.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c11 */
#include "qpad_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define PI_DB (4.0 * atan(1.0)) /* source/param.f03:36 */

/* Timers by QPAD's own event names (source/sysutil_module.f03:307-330 and
   the start_tprof calls of the routines restated here), for the CPU
   baseline's breakdown.  Off unless ora_timing(1); they do not touch the
   arithmetic. */
static int g_timing = 0;
static double g_timer[ORA_NTIMER];
static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}
#define TSTART() double t_ = g_timing ? now_s() : 0.0
#define TSTOP(k) do { if (g_timing) { double e_ = now_s(); g_timer[k] += e_ - t_; t_ = e_; } } while (0)

void ora_timing(int enable)
{
    g_timing = enable;
    for (int k = 0; k < ORA_NTIMER; k++) g_timer[k] = 0.0;
}

void ora_timers(double *out)
{
    for (int k = 0; k < ORA_NTIMER; k++) out[k] = g_timer[k];
}

static inline int kre(int m) { return m == 0 ? 0 : 2 * m - 1; }
static inline int kim(int m) { return 2 * m; }

/* element (component c, mode array k, node i) of a slice array */
#define FA(arr, c, k, i) ((arr)[((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)])

/* ------------------------------------------------------------------------ */
/* Tridiagonal matrices: source/fields/field_solver_class.f03:256-561        */
/* Row r = 0..nr-1 is QPAD grid point r+1 (noff = 0, one radial partition). */
/* ------------------------------------------------------------------------ */
void ora_matrix(int kind, int m, int nr, double dr, double relax, int bnd,
                double *a, double *b, double *c)
{
    double dr2 = dr * dr;
    double m2 = (double)(m * m);
    /* interior rows: field_solver_class.f03:295-300 (psi/bt/ez/bz),
       :336-343 (B+), :371-378 (B-) */
    for (int r = 1; r < nr; r++) {
        double j = (double)r;
        a[r] = 1.0 - 0.5 / j;
        c[r] = 1.0 + 0.5 / j;
        if (kind == ORA_FK_BPLUS) {
            double t = (double)(m + 1) / j;
            b[r] = -2.0 - t * t - relax;
        } else if (kind == ORA_FK_BMINUS) {
            double t = (double)(m - 1) / j;
            b[r] = -2.0 - t * t - relax;
        } else {
            b[r] = -2.0 - m2 / (j * j);
        }
    }
    /* first grid point: :302-330 (psi group), :345-357 (B+), :380-400 (B-) */
    int identity = 0;
    if (kind == ORA_FK_BPLUS) {
        identity = 1;
    } else if (kind == ORA_FK_BMINUS) {
        if (m == 1) { a[0] = 0.0; b[0] = -4.0 - relax; c[0] = 4.0; }
        else identity = 1;
    } else {
        if (m == 0) { a[0] = 0.0; b[0] = -4.0; c[0] = 4.0; }
        else identity = 1;
    }
    if (identity) {
        a[0] = 0.0; b[0] = 1.0; c[0] = 0.0;
        if (nr > 1) a[1] = 0.0;   /* HYPRE_BUF(4) = 0: decouples node 2 */
    }
    /* upper boundary: :491-548 */
    int rl = nr - 1;
    if (bnd == ORA_BND_ZERO) {
        c[rl] = 0.0;
    } else {
        double jmax = (double)nr;
        double coef = 0.0;
        int fold = 1;
        switch (kind) {
        case ORA_FK_PSI:
        case ORA_FK_EZ:
        case ORA_FK_BZ:
            if (m == 0) fold = 0;
            else coef = 1.0 - (double)m / jmax;
            break;
        case ORA_FK_BT:
            if (m == 0) coef = 1.0 + 1.0 / (jmax * log(jmax * dr));
            else coef = 1.0 - (double)m / jmax;
            break;
        default: /* B+ and B- both use m+1 (field_solver_class.f03:532-540) */
            coef = 1.0 - (double)(m + 1) / jmax;
            break;
        }
        if (fold) b[rl] = b[rl] + coef * c[rl];
        c[rl] = 0.0;
    }
    /* HYPRE_BUF = HYPRE_BUF / dr2 (:551) */
    for (int r = 0; r < nr; r++) {
        a[r] = a[r] / dr2;
        b[r] = b[r] / dr2;
        c[r] = c[r] / dr2;
    }
}

/* Thomas factorisation (textbook; source/math_module.f03:270-292 restates
   the same elimination).  lo = sub-diagonal, inv = 1/pivot, cp = c/pivot. */
void ora_thomas_factor(int n, const double *a, const double *b,
                       const double *c, double *lo, double *inv, double *cp)
{
    inv[0] = 1.0 / b[0];
    cp[0] = c[0] * inv[0];
    lo[0] = 0.0;
    for (int r = 1; r < n; r++) {
        double den = b[r] - a[r] * cp[r - 1];
        inv[r] = 1.0 / den;
        cp[r] = c[r] * inv[r];
        lo[r] = a[r];
    }
}

void ora_thomas_solve(int n, const double *lo, const double *inv,
                      const double *cp, double *d)
{
    d[0] = d[0] * inv[0];
    for (int r = 1; r < n; r++) d[r] = (d[r] - lo[r] * d[r - 1]) * inv[r];
    for (int r = n - 2; r >= 0; r--) d[r] = d[r] - cp[r] * d[r + 1];
}

static void solve_sys(const ora_sim *s, int kind, int m, double *buf)
{
    const ora_factor *f = &s->fac[kind][m];
    ora_thomas_solve(s->g.nr, f->lo, f->inv, f->cp, buf);
}

/* ------------------------------------------------------------------------ */
/* psi: source/fields/field_psi_class.f03:131-254                            */
/* ------------------------------------------------------------------------ */
void ora_solve_psi(const ora_sim *s, const double *q, double *psi)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double *bre = malloc(sizeof(double) * nr), *bim = malloc(sizeof(double) * nr);
    for (int m = 0; m <= s->g.max_mode; m++) {
        for (int i = 1; i <= nr; i++) bre[i - 1] = -1.0 * FA(q, 0, kre(m), i);
        solve_sys(s, ORA_FK_PSI, m, bre);
        for (int i = 1; i <= nr; i++) FA(psi, 0, kre(m), i) = bre[i - 1];
        if (m == 0) continue;
        for (int i = 1; i <= nr; i++) bim[i - 1] = -1.0 * FA(q, 0, kim(m), i);
        solve_sys(s, ORA_FK_PSI, m, bim);
        for (int i = 1; i <= nr; i++) FA(psi, 0, kim(m), i) = bim[i - 1];
        FA(psi, 0, kre(m), 1) = 0.0; /* psi(m>0) vanishes on axis (:207-210) */
        FA(psi, 0, kim(m), 1) = 0.0;
    }
    free(bre); free(bim);
}

/* ------------------------------------------------------------------------ */
/* Ez: source/fields/field_e_class.f03:148-334                               */
/* ------------------------------------------------------------------------ */
void ora_solve_ez(const ora_sim *s, const double *jay, double *e)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    const int ldj = ld; (void)ldj;
    double idr = 1.0 / s->g.dr, idrh = 0.5 * idr;
    double *bre = calloc(nr + 1, sizeof(double)), *bim = calloc(nr + 1, sizeof(double));
    /* 1-based buffers: buf[i], i = 1..nr (buf[0] unused) */
#define JRE(c, i) FA(jay, (c) - 1, kre(m), (i))
#define JIM(c, i) FA(jay, (c) - 1, kim(m), (i))
    for (int m = 0; m <= s->g.max_mode; m++) {
        memset(bre, 0, sizeof(double) * (nr + 1));
        memset(bim, 0, sizeof(double) * (nr + 1));
        if (m == 0) {
            int i1 = 2, i2 = nr - 2;
            double div = 0.0, ir;
            for (int i = i1; i <= i2; i++) {
                ir = idr / (double)(i - 1);
                bre[i] = idrh * (JRE(1, i + 1) - JRE(1, i - 1)) + ir * JRE(1, i);
                div = div + bre[i] * (double)(i - 1);
            }
            ir = idr / (double)(nr - 2);
            bre[nr - 1] = idrh * (JRE(1, nr) - JRE(1, nr - 2)) + ir * JRE(1, nr - 1);
            ir = idr / (double)(nr - 1);
            bre[nr] = idr * (JRE(1, nr) - JRE(1, nr - 1)) + ir * JRE(1, nr);
            div = div - idrh * (JRE(1, nr - 2) + JRE(1, nr - 1)) * ((double)nr - 2.5);
            bre[1] = div;               /* MPI_REDUCE over one rank */
            bre[1] = -8.0 * bre[1];
        } else {
            double dm = (double)m, ir;
            for (int i = 2; i <= nr; i++) {
                ir = idr / (double)(i - 1);
                bre[i] = idrh * (JRE(1, i + 1) - JRE(1, i - 1)) + ir * JRE(1, i) - dm * ir * JIM(2, i);
                bim[i] = idrh * (JIM(1, i + 1) - JIM(1, i - 1)) + ir * JIM(1, i) + dm * ir * JRE(2, i);
            }
            if (m % 2 == 0) {
                bre[1] = 2.0 * idr * JRE(1, 2) - dm * idr * JIM(2, 2);
                bim[1] = 2.0 * idr * JIM(1, 2) + dm * idr * JRE(2, 2);
            } else {
                bre[1] = 0.0;
                bim[1] = 0.0;
                if (m == 1) {
                    ir = idr;
                    bre[2] = idr * (JRE(1, 3) - JRE(1, 2)) + ir * JRE(1, 2) - dm * ir * JIM(2, 2);
                    bim[2] = idr * (JIM(1, 3) - JIM(1, 2)) + ir * JIM(1, 2) + dm * ir * JRE(2, 2);
                }
            }
            ir = idr / (double)(nr - 1);
            bre[nr] = idrh * (3.0 * JRE(1, nr) - 4.0 * JRE(1, nr - 1) + JRE(1, nr - 2)) + ir * JRE(1, nr)
                      - dm * ir * JIM(2, nr);
            bim[nr] = idrh * (3.0 * JIM(1, nr) - 4.0 * JIM(1, nr - 1) + JIM(1, nr - 2)) + ir * JIM(1, nr)
                      + dm * ir * JRE(2, nr);
        }
        solve_sys(s, ORA_FK_EZ, m, bre + 1);
        for (int i = 1; i <= nr; i++) FA(e, 2, kre(m), i) = bre[i];
        if (m > 0) {
            solve_sys(s, ORA_FK_EZ, m, bim + 1);
            for (int i = 1; i <= nr; i++) FA(e, 2, kim(m), i) = bim[i];
            FA(e, 2, kre(m), 1) = 0.0;  /* Ez(m>0) vanishes on axis (:288-290) */
            FA(e, 2, kim(m), 1) = 0.0;
        }
    }
#undef JRE
#undef JIM
    free(bre); free(bim);
}

/* ------------------------------------------------------------------------ */
/* E_perp from B and psi: source/fields/field_e_class.f03:412-514            */
/* ------------------------------------------------------------------------ */
void ora_solve_et(const ora_sim *s, const double *b, const double *psi, double *e)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double idr = 1.0 / s->g.dr, idrh = idr * 0.5;
    /* m = 0 */
    {
        int k = 0;
        for (int i = 1; i <= nr; i++) {
            FA(e, 0, k, i) = FA(b, 1, k, i) - idrh * (FA(psi, 0, k, i + 1) - FA(psi, 0, k, i - 1));
            FA(e, 1, k, i) = -FA(b, 0, k, i);
        }
        FA(e, 0, k, 1) = 0.0;
        FA(e, 1, k, 1) = 0.0;
        FA(e, 0, k, nr) = FA(b, 1, k, nr) + idrh * (4.0 * FA(psi, 0, k, nr - 1) - FA(psi, 0, k, nr - 2)
                                                    - 3.0 * FA(psi, 0, k, nr));
    }
    for (int m = 1; m <= s->g.max_mode; m++) {
        int r = kre(m), q = kim(m);
        double dm = (double)m, ir;
        for (int i = 2; i <= nr; i++) {
            ir = idr / (double)(i - 1);
            FA(e, 0, r, i) = FA(b, 1, r, i) - idrh * (FA(psi, 0, r, i + 1) - FA(psi, 0, r, i - 1));
            FA(e, 0, q, i) = FA(b, 1, q, i) - idrh * (FA(psi, 0, q, i + 1) - FA(psi, 0, q, i - 1));
            FA(e, 1, r, i) = -FA(b, 0, r, i) + ir * dm * FA(psi, 0, q, i);
            FA(e, 1, q, i) = -FA(b, 0, q, i) - ir * dm * FA(psi, 0, r, i);
        }
        if (m == 1) {
            FA(e, 0, r, 1) = FA(b, 1, r, 1) - idr * FA(psi, 0, r, 2);
            FA(e, 0, q, 1) = FA(b, 1, q, 1) - idr * FA(psi, 0, q, 2);
            FA(e, 1, r, 1) = -FA(b, 0, r, 1) + idr * FA(psi, 0, q, 2);
            FA(e, 1, q, 1) = -FA(b, 0, q, 1) - idr * FA(psi, 0, r, 2);
        } else {
            FA(e, 0, r, 1) = 0.0; FA(e, 0, q, 1) = 0.0;
            FA(e, 1, r, 1) = 0.0; FA(e, 1, q, 1) = 0.0;
        }
        FA(e, 0, r, nr) = FA(b, 1, r, nr) + idrh * (4.0 * FA(psi, 0, r, nr - 1) - FA(psi, 0, r, nr - 2)
                                                    - 3.0 * FA(psi, 0, r, nr));
        FA(e, 0, q, nr) = FA(b, 1, q, nr) + idrh * (4.0 * FA(psi, 0, q, nr - 1) - FA(psi, 0, q, nr - 2)
                                                    - 3.0 * FA(psi, 0, q, nr));
    }
}

/* ------------------------------------------------------------------------ */
/* Bz: source/fields/field_b_class.f03:205-307 (source), :508-543 (solution),*/
/* :760-796 (driver)                                                         */
/* ------------------------------------------------------------------------ */
void ora_solve_bz(const ora_sim *s, const double *jay, double *bf)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double idr = 1.0 / s->g.dr, idrh = 0.5 * idr;
    double *bre = calloc(nr + 1, sizeof(double)), *bim = calloc(nr + 1, sizeof(double));
#define JRE(c, i) FA(jay, (c) - 1, kre(m), (i))
#define JIM(c, i) FA(jay, (c) - 1, kim(m), (i))
    for (int m = 0; m <= s->g.max_mode; m++) {
        memset(bre, 0, sizeof(double) * (nr + 1));
        memset(bim, 0, sizeof(double) * (nr + 1));
        double ir;
        if (m == 0) {
            for (int i = 2; i <= nr; i++) {
                ir = idr / (double)(i - 1);
                bre[i] = -idrh * (JRE(2, i + 1) - JRE(2, i - 1)) - ir * JRE(2, i);
            }
            bre[1] = -2.0 * idr * JRE(2, 2);
            ir = idr / (double)(nr - 1);
            bre[nr] = -idrh * (3.0 * JRE(2, nr) - 4.0 * JRE(2, nr - 1) + JRE(2, nr - 2)) - ir * JRE(2, nr);
        } else {
            double dm = (double)m;
            for (int i = 2; i <= nr; i++) {
                ir = idr / (double)(i - 1);
                bre[i] = -idrh * (JRE(2, i + 1) - JRE(2, i - 1)) - ir * JRE(2, i) - dm * ir * JIM(1, i);
                bim[i] = -idrh * (JIM(2, i + 1) - JIM(2, i - 1)) - ir * JIM(2, i) + dm * ir * JRE(1, i);
            }
            if (m % 2 == 0) {
                bre[1] = -2.0 * idr * JRE(2, 2) - dm * idr * JIM(1, 2);
                bim[1] = -2.0 * idr * JIM(2, 2) + dm * idr * JRE(1, 2);
            } else {
                bre[1] = 0.0;
                bim[1] = 0.0;
                if (m == 1) {
                    ir = idr;
                    bre[2] = -idr * (JRE(2, 3) - JRE(2, 2)) - ir * JRE(2, 2) - dm * ir * JIM(1, 2);
                    bim[2] = -idr * (JIM(2, 3) - JIM(2, 2)) - ir * JIM(2, 2) + dm * ir * JRE(1, 2);
                }
            }
            ir = idr / (double)(nr - 1);
            bre[nr] = -idrh * (3.0 * JRE(2, nr) - 4.0 * JRE(2, nr - 1) + JRE(2, nr - 2)) - ir * JRE(2, nr)
                      - dm * ir * JIM(1, nr);
            bim[nr] = -idrh * (3.0 * JIM(2, nr) - 4.0 * JIM(2, nr - 1) + JIM(2, nr - 2)) - ir * JIM(2, nr)
                      + dm * ir * JRE(1, nr);
        }
        solve_sys(s, ORA_FK_BZ, m, bre + 1);
        for (int i = 1; i <= nr; i++) FA(bf, 2, kre(m), i) = bre[i];
        if (m > 0) {
            solve_sys(s, ORA_FK_BZ, m, bim + 1);
            for (int i = 1; i <= nr; i++) FA(bf, 2, kim(m), i) = bim[i];
            FA(bf, 2, kre(m), 1) = 0.0;   /* Bz(m>0) vanishes on axis */
            FA(bf, 2, kim(m), 1) = 0.0;
        }
    }
#undef JRE
#undef JIM
    free(bre); free(bim);
}

/* ------------------------------------------------------------------------ */
/* Beam B_perp: source/fields/field_b_class.f03:309-358 (source),            */
/* :545-701 (solution), :798-834 (driver)                                    */
/* ------------------------------------------------------------------------ */
void ora_solve_bt_beam(const ora_sim *s, const double *q, double *bf)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double idr = 1.0 / s->g.dr, idrh = 0.5 * idr;
    double *bre = calloc(nr + 1, sizeof(double)), *bim = calloc(nr + 1, sizeof(double));
    for (int m = 0; m <= s->g.max_mode; m++) {
        for (int i = 1; i <= nr; i++) bre[i] = -1.0 * FA(q, 0, kre(m), i);
        solve_sys(s, ORA_FK_BT, m, bre + 1);
        if (m == 0) {
            for (int i = 2; i <= nr - 1; i++) {
                FA(bf, 0, 0, i) = 0.0;
                FA(bf, 1, 0, i) = -idrh * (bre[i + 1] - bre[i - 1]);
            }
            FA(bf, 0, 0, 1) = 0.0;
            FA(bf, 1, 0, 1) = 0.0;
            FA(bf, 0, 0, nr) = 0.0;
            FA(bf, 1, 0, nr) = -idrh * (3.0 * bre[nr] - 4.0 * bre[nr - 1] + bre[nr - 2]);
            continue;
        }
        for (int i = 1; i <= nr; i++) bim[i] = -1.0 * FA(q, 0, kim(m), i);
        solve_sys(s, ORA_FK_BT, m, bim + 1);
        int r = kre(m), k = kim(m);
        double dm = (double)m, ir;
        for (int i = 2; i <= nr - 1; i++) {
            ir = idr / (double)(i - 1);
            FA(bf, 0, r, i) = -ir * dm * bim[i];
            FA(bf, 0, k, i) = ir * dm * bre[i];
            FA(bf, 1, r, i) = -idrh * (bre[i + 1] - bre[i - 1]);
            FA(bf, 1, k, i) = -idrh * (bim[i + 1] - bim[i - 1]);
        }
        if (m == 1) {
            FA(bf, 0, r, 1) = -idr * dm * bim[2];
            FA(bf, 0, k, 1) = idr * dm * bre[2];
            FA(bf, 1, r, 1) = -idr * bre[2];
            FA(bf, 1, k, 1) = -idr * bim[2];
        } else {
            FA(bf, 0, r, 1) = 0.0; FA(bf, 0, k, 1) = 0.0;
            FA(bf, 1, r, 1) = 0.0; FA(bf, 1, k, 1) = 0.0;
        }
        ir = idr / (double)(nr - 1);
        FA(bf, 0, r, nr) = -ir * dm * bim[nr];
        FA(bf, 0, k, nr) = ir * dm * bre[nr];
        FA(bf, 1, r, nr) = -idrh * (3.0 * bre[nr] - 4.0 * bre[nr - 1] + bre[nr - 2]);
        FA(bf, 1, k, nr) = -idrh * (3.0 * bim[nr] - 4.0 * bim[nr - 1] + bim[nr - 2]);
    }
    free(bre); free(bim);
}

/* ------------------------------------------------------------------------ */
/* Plasma B_perp predictor-corrector: source/fields/field_b_class.f03:360-506*/
/* (source, relaxation uses the current b), :703-758 (B+/B- recombination),  */
/* :836-879 (driver)                                                         */
/* ------------------------------------------------------------------------ */
void ora_solve_bt_iter(const ora_sim *s, const double *djdxi, const double *jay, double *bf)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double idr = 1.0 / s->g.dr, idrh = 0.5 * idr;
    double rx = s->g.relax_fac * (idr * idr);   /* relax_fac * idr**2 */
    double *b1r = calloc(nr + 1, sizeof(double)), *b1i = calloc(nr + 1, sizeof(double));
    double *b2r = calloc(nr + 1, sizeof(double)), *b2i = calloc(nr + 1, sizeof(double));
    for (int m = 0; m <= s->g.max_mode; m++) {
        int kr = kre(m), ki = kim(m);
        memset(b1r, 0, sizeof(double) * (nr + 1)); memset(b2r, 0, sizeof(double) * (nr + 1));
        memset(b1i, 0, sizeof(double) * (nr + 1)); memset(b2i, 0, sizeof(double) * (nr + 1));
#define F1R(c, i) FA(djdxi, (c) - 1, kr, (i))
#define F1I(c, i) FA(djdxi, (c) - 1, ki, (i))
#define F2R(c, i) FA(jay, (c) - 1, kr, (i))
#define F2I(c, i) FA(jay, (c) - 1, ki, (i))
#define F3R(c, i) FA(bf, (c) - 1, kr, (i))
#define F3I(c, i) FA(bf, (c) - 1, ki, (i))
        if (m == 0) {
            for (int i = 2; i <= nr; i++) {
                b1r[i] = -F1R(2, i) - F3R(1, i) * rx;
                b2r[i] = F1R(1, i) + idrh * (F2R(3, i + 1) - F2R(3, i - 1)) - F3R(2, i) * rx;
            }
            b1r[1] = 0.0;
            b2r[1] = 0.0;
            b2r[2] = F1R(1, 2) + idr * (F2R(3, 3) - F2R(3, 2)) - F3R(2, 2) * rx;
            b1r[nr] = -F1R(2, nr) - F3R(1, nr) * rx;
            b2r[nr] = F1R(1, nr) + idrh * (3.0 * F2R(3, nr) - 4.0 * F2R(3, nr - 1) + F2R(3, nr - 2))
                      - F3R(2, nr) * rx;
            solve_sys(s, ORA_FK_BPLUS, 0, b1r + 1);
            solve_sys(s, ORA_FK_BMINUS, 0, b2r + 1);
            for (int i = 1; i <= nr; i++) {
                FA(bf, 0, 0, i) = b1r[i];
                FA(bf, 1, 0, i) = b2r[i];
            }
            FA(bf, 0, 0, 1) = 0.0;
            FA(bf, 1, 0, 1) = 0.0;
            continue;
        }
        double dm = (double)m, ir, s1r, s1i, s2r, s2i;
        for (int i = 2; i <= nr; i++) {
            ir = idr / (double)(i - 1);
            s1r = -F1R(2, i) + dm * F2I(3, i) * ir;
            s1i = -F1I(2, i) - dm * F2R(3, i) * ir;
            s2r = F1R(1, i) + idrh * (F2R(3, i + 1) - F2R(3, i - 1));
            s2i = F1I(1, i) + idrh * (F2I(3, i + 1) - F2I(3, i - 1));
            b1r[i] = s1r - s2i - (F3R(1, i) - F3I(2, i)) * rx;
            b1i[i] = s1i + s2r - (F3I(1, i) + F3R(2, i)) * rx;
            b2r[i] = s1r + s2i - (F3R(1, i) + F3I(2, i)) * rx;
            b2i[i] = s1i - s2r - (F3I(1, i) - F3R(2, i)) * rx;
        }
        if (m == 1) {
            s1r = -F1R(2, 1) + idr * dm * F2I(3, 2);
            s1i = -F1I(2, 1) - idr * dm * F2R(3, 2);
            s2r = F1R(1, 1) + idr * F2R(3, 2);
            s2i = F1I(1, 1) + idr * F2I(3, 2);
        } else if (m % 2 == 0) {
            s1r = 0.0; s1i = 0.0; s2r = 0.0; s2i = 0.0;
        } else {
            s1r = idr * dm * F2I(3, 2);
            s1i = idr * dm * F2R(3, 2);
            s2r = idr * F2R(3, 2);
            s2i = idr * F2I(3, 2);
        }
        b1r[1] = s1r - s2i - (F3R(1, 1) - F3I(2, 1)) * rx;
        b1i[1] = s1i + s2r - (F3I(1, 1) + F3R(2, 1)) * rx;
        b2r[1] = s1r + s2i - (F3R(1, 1) + F3I(2, 1)) * rx;
        b2i[1] = s1i - s2r - (F3I(1, 1) - F3R(2, 1)) * rx;
        ir = idr / (double)(nr - 1);
        s1r = -F1R(2, nr) + dm * F2I(3, nr) * ir;
        s1i = -F1I(2, nr) - dm * F2R(3, nr) * ir;
        s2r = F1R(1, nr) + idrh * (3.0 * F2R(3, nr) - 4.0 * F2R(3, nr - 1) + F2R(3, nr - 2));
        s2i = F1I(1, nr) + idrh * (3.0 * F2I(3, nr) - 4.0 * F2I(3, nr - 1) + F2I(3, nr - 2));
        b1r[nr] = s1r - s2i - (F3R(1, nr) - F3I(2, nr)) * rx;
        b1i[nr] = s1i + s2r - (F3I(1, nr) + F3R(2, nr)) * rx;
        b2r[nr] = s1r + s2i - (F3R(1, nr) + F3I(2, nr)) * rx;
        b2i[nr] = s1i - s2r - (F3I(1, nr) - F3R(2, nr)) * rx;
        solve_sys(s, ORA_FK_BPLUS, m, b1r + 1);
        solve_sys(s, ORA_FK_BPLUS, m, b1i + 1);
        solve_sys(s, ORA_FK_BMINUS, m, b2r + 1);
        solve_sys(s, ORA_FK_BMINUS, m, b2i + 1);
        for (int i = 1; i <= nr; i++) {
            FA(bf, 0, kr, i) = 0.5 * (b1r[i] + b2r[i]);
            FA(bf, 0, ki, i) = 0.5 * (b1i[i] + b2i[i]);
            FA(bf, 1, kr, i) = 0.5 * (b1i[i] - b2i[i]);
            FA(bf, 1, ki, i) = 0.5 * (-b1r[i] + b2r[i]);
        }
        if (m != 1) {
            FA(bf, 0, kr, 1) = 0.0; FA(bf, 0, ki, 1) = 0.0;
            FA(bf, 1, kr, 1) = 0.0; FA(bf, 1, ki, 1) = 0.0;
        }
#undef F1R
#undef F1I
#undef F2R
#undef F2I
#undef F3R
#undef F3I
    }
    free(b1r); free(b1i); free(b2r); free(b2i);
}

/* ------------------------------------------------------------------------ */
/* dJ_perp/dxi: source/fields/field_src_class.f03:273-405                    */
/* ------------------------------------------------------------------------ */
void ora_solve_djdxi(const ora_sim *s, const double *acu, const double *amu, double *dcu)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double idr = 1.0 / s->g.dr, idrh = idr * 0.5;
    for (int m = 0; m <= s->g.max_mode; m++) {
        int kr = kre(m), ki = kim(m);
#define AR(c, i) FA(acu, (c) - 1, kr, (i))
#define AI(c, i) FA(acu, (c) - 1, ki, (i))
#define MR(c, i) FA(amu, (c) - 1, kr, (i))
#define MI(c, i) FA(amu, (c) - 1, ki, (i))
#define DR(c, i) FA(dcu, (c) - 1, kr, (i))
#define DI(c, i) FA(dcu, (c) - 1, ki, (i))
        double ir;
        if (m == 0) {
            for (int i = 2; i <= nr; i++) {
                ir = idr / (double)(i - 1);
                DR(1, i) = AR(1, i) - idrh * (MR(1, i + 1) - MR(1, i - 1)) - ir * MR(1, i);
                DR(2, i) = AR(2, i) - idrh * (MR(2, i + 1) - MR(2, i - 1)) - ir * MR(2, i);
            }
            DR(1, 1) = 0.0;
            DR(2, 1) = 0.0;
            ir = idr / (double)(nr - 1);
            DR(1, nr) = AR(1, nr) + idrh * (4.0 * MR(1, nr - 1) - MR(1, nr - 2) - 3.0 * MR(1, nr)) - ir * MR(1, nr);
            DR(2, nr) = AR(2, nr) + idrh * (4.0 * MR(2, nr - 1) - MR(2, nr - 2) - 3.0 * MR(2, nr)) - ir * MR(2, nr);
            continue;
        }
        double dm = (double)m;
        for (int i = 2; i <= nr; i++) {
            ir = idr / (double)(i - 1);
            DR(1, i) = AR(1, i) - idrh * (MR(1, i + 1) - MR(1, i - 1)) - ir * MR(1, i) + dm * ir * MI(2, i);
            DR(2, i) = AR(2, i) - idrh * (MR(2, i + 1) - MR(2, i - 1)) - ir * MR(2, i) + dm * ir * MI(3, i);
            DI(1, i) = AI(1, i) - idrh * (MI(1, i + 1) - MI(1, i - 1)) - ir * MI(1, i) - dm * ir * MR(2, i);
            DI(2, i) = AI(2, i) - idrh * (MI(2, i + 1) - MI(2, i - 1)) - ir * MI(2, i) - dm * ir * MR(3, i);
        }
        if (m == 1) {
            DR(1, 1) = AR(1, 1) - 2.0 * idr * MR(1, 2) + dm * idr * MI(2, 2);
            DR(2, 1) = AR(2, 1) - 2.0 * idr * MR(2, 2) + dm * idr * MI(3, 2);
            DI(1, 1) = AI(1, 1) - 2.0 * idr * MI(1, 2) - dm * idr * MR(2, 2);
            DI(2, 1) = AI(2, 1) - 2.0 * idr * MI(2, 2) - dm * idr * MR(3, 2);
        } else {
            DR(1, 1) = 0.0; DR(2, 1) = 0.0; DI(1, 1) = 0.0; DI(2, 1) = 0.0;
            if (m == 2) {
                ir = idr;
                DR(1, 2) = AR(1, 2) - idr * (MR(1, 3) - MR(1, 2)) - ir * MR(1, 2) + dm * ir * MI(2, 2);
                DR(2, 2) = AR(2, 2) - idr * (MR(2, 3) - MR(2, 2)) - ir * MR(2, 2) + dm * ir * MI(3, 2);
                DI(1, 2) = AI(1, 2) - idr * (MI(1, 3) - MI(1, 2)) - ir * MI(1, 2) - dm * ir * MR(2, 2);
                DI(2, 2) = AI(2, 2) - idr * (MI(2, 3) - MI(2, 2)) - ir * MI(2, 2) - dm * ir * MR(3, 2);
            }
        }
        ir = idr / (double)(nr - 1);
        DR(1, nr) = AR(1, nr) + idrh * (4.0 * MR(1, nr - 1) - MR(1, nr - 2) - 3.0 * MR(1, nr)) - ir * MR(1, nr)
                    + dm * ir * MI(2, nr);
        DR(2, nr) = AR(2, nr) + idrh * (4.0 * MR(2, nr - 1) - MR(2, nr - 2) - 3.0 * MR(2, nr)) - ir * MR(2, nr)
                    + dm * ir * MI(3, nr);
        DI(1, nr) = AI(1, nr) + idrh * (4.0 * MI(1, nr - 1) - MI(1, nr - 2) - 3.0 * MI(1, nr)) - ir * MI(1, nr)
                    - dm * ir * MR(2, nr);
        DI(2, nr) = AI(2, nr) + idrh * (4.0 * MI(2, nr - 1) - MI(2, nr - 2) - 3.0 * MI(2, nr)) - ir * MI(2, nr)
                    - dm * ir * MR(3, nr);
#undef AR
#undef AI
#undef MR
#undef MI
#undef DR
#undef DI
    }
}

/* ------------------------------------------------------------------------ */
/* convergence_tester: source/simulation_class.f03:522-606 (dim = 2, B_phi)  */
/* ------------------------------------------------------------------------ */
void ora_conv_record(ora_sim *s, const double *b)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    for (int i = 1; i <= nr; i++) { s->conv_re[i] = 0.0; s->conv_im[i] = 0.0; }
    for (int m = 0; m <= s->g.max_mode; m++) {
        for (int i = 1; i <= nr; i++) s->conv_re[i] = s->conv_re[i] + fabs(FA(b, 1, kre(m), i));
        if (m == 0) continue;
        for (int i = 1; i <= nr; i++) s->conv_im[i] = s->conv_im[i] + fabs(FA(b, 1, kim(m), i));
    }
}

void ora_conv_compare(ora_sim *s, const double *b, double *rel, double *abs_)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld;
    double mx = -INFINITY;
    for (int i = 1; i <= nr; i++) {
        double v = s->conv_re[i] * s->conv_re[i] + s->conv_im[i] * s->conv_im[i];
        if (v > mx) mx = v;
    }
    double old_norm = sqrt(mx);
    for (int m = 0; m <= s->g.max_mode; m++) {
        for (int i = 1; i <= nr; i++) s->conv_re[i] = s->conv_re[i] - fabs(FA(b, 1, kre(m), i));
        if (m == 0) continue;
        for (int i = 1; i <= nr; i++) s->conv_im[i] = s->conv_im[i] - fabs(FA(b, 1, kim(m), i));
    }
    mx = -INFINITY;
    for (int i = 1; i <= nr; i++) {
        double v = s->conv_re[i] * s->conv_re[i] + s->conv_im[i] * s->conv_im[i];
        if (v > mx) mx = v;
    }
    *abs_ = sqrt(mx);
    if (old_norm > 2.220446049250313e-16) *rel = *abs_ / old_norm;   /* epsilon(1.0d0) */
    else *rel = 1.7976931348623157e308;                              /* huge(1.0)     */
}

/* ------------------------------------------------------------------------ */
/* Plasma particles                                                          */
/* ------------------------------------------------------------------------ */

/* gen_interp_info: source/species/interp_part2d.f03:28-64 */
static inline void interp_info(double x1, double x2, double idr, int *idx,
                               double *w0, double *w1, double *pc, double *ps)
{
    double pos = sqrt(x1 * x1 + x2 * x2);
    *pc = x1 / pos;
    *ps = x2 / pos;
    pos = pos * idr;
    int ip = (int)pos;
    *idx = ip + 1;
    pos = pos - (double)ip;
    *w0 = 1.0 - pos;   /* spline_linear, source/interpolation.f03:10-20 */
    *w1 = pos;
}

/* interp_field (vector form): source/species/interp_part2d.f03:66-110 */
static inline void interp_vec(const ora_sim *s, const double *f, int ncomp,
                              int idx, double w0, double w1, double pc, double ps,
                              double *fp)
{
    int nm = s->nm, ld = s->ld;
    for (int c = 0; c < ncomp; c++) fp[c] = 0.0;
    for (int c = 0; c < ncomp; c++) {
        fp[c] = fp[c] + FA(f, c, 0, idx) * w0;
        fp[c] = fp[c] + FA(f, c, 0, idx + 1) * w1;
    }
    double phr = 1.0, phi = 0.0;
    for (int m = 1; m <= s->g.max_mode; m++) {
        double nr_ = phr * pc - phi * ps, ni_ = phr * ps + phi * pc;
        phr = nr_; phi = ni_;
        double ph_re = 2.0 * phr, ph_im = 2.0 * phi;
        for (int c = 0; c < ncomp; c++) {
            fp[c] = fp[c] + (FA(f, c, kre(m), idx) * ph_re - FA(f, c, kim(m), idx) * ph_im) * w0;
            fp[c] = fp[c] + (FA(f, c, kre(m), idx + 1) * ph_re - FA(f, c, kim(m), idx + 1) * ph_im) * w1;
        }
    }
}

void ora_interp_point(const ora_sim *s, const double *f, int ncomp, double x1,
                      double x2, double *out)
{
    int idx; double w0, w1, pc, ps;
    interp_info(x1, x2, 1.0 / s->g.dr, &idx, &w0, &w1, &pc, &ps);
    interp_vec(s, f, ncomp, idx, w0, w1, pc, ps, out);
}

/* qdeposit_part2d: source/species/part2d_class.f03:231-359 */
void ora_qdeposit(const ora_sim *s, const ora_species *sp, double *q)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld, M = s->g.max_mode;
    double idr = 1.0 / s->g.dr;
    for (int64_t p = 0; p < sp->npp; p++) {
        double pos = sqrt(sp->x1[p] * sp->x1[p] + sp->x2[p] * sp->x2[p]) * idr;
        double ph0r = sp->x1[p] / pos * idr, ph0i = -sp->x2[p] / pos * idr;
        int nn = (int)floor(pos);
        pos = pos - (double)nn;
        nn = nn + 1;
        double wt[2] = {1.0 - pos, pos};
        double phr = 1.0 * sp->q[p], phi = 0.0 * sp->q[p];
        for (int j = 0; j < 2; j++) FA(q, 0, 0, nn + j) = FA(q, 0, 0, nn + j) + wt[j] * phr;
        for (int m = 1; m <= M; m++) {
            double a = phr * ph0r - phi * ph0i, b = phr * ph0i + phi * ph0r;
            phr = a; phi = b;
            for (int j = 0; j < 2; j++) {
                FA(q, 0, kre(m), nn + j) = FA(q, 0, kre(m), nn + j) + wt[j] * phr;
                FA(q, 0, kim(m), nn + j) = FA(q, 0, kim(m), nn + j) + wt[j] * phi;
            }
        }
    }
    /* axis and 1/r (:313-336) */
    FA(q, 0, 0, 0) = 0.0;
    FA(q, 0, 0, 1) = 8.0 * FA(q, 0, 0, 1);
    for (int j = 2; j <= nr + 1; j++) {
        double ir = 1.0 / (double)(j - 1);
        FA(q, 0, 0, j) = FA(q, 0, 0, j) * ir;
    }
    for (int m = 1; m <= M; m++) {
        FA(q, 0, kre(m), 0) = 0.0; FA(q, 0, kim(m), 0) = 0.0;
        FA(q, 0, kre(m), 1) = 0.0; FA(q, 0, kim(m), 1) = 0.0;
        for (int j = 2; j <= nr + 1; j++) {
            double ir = 1.0 / (double)(j - 1);
            FA(q, 0, kre(m), j) = FA(q, 0, kre(m), j) * ir;
            FA(q, 0, kim(m), j) = FA(q, 0, kim(m), j) * ir;
        }
    }
}

/* axis rules (part2d_class.f03:914-960) and division by the node index
   (:962-1004) of the current deposits; amjdeposit_std_part2d has the same
   (:683-742) */
static void deposit_axis_and_radius(const ora_sim *s, double *cu, double *amu, double *dcu)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld, M = s->g.max_mode;
    for (int c = 0; c < 3; c++) { FA(cu, c, 0, 0) = 0.0; FA(amu, c, 0, 0) = 0.0; }
    for (int c = 0; c < 2; c++) FA(dcu, c, 0, 0) = 0.0;
    FA(cu, 0, 0, 1) = 0.0; FA(cu, 1, 0, 1) = 0.0; FA(cu, 2, 0, 1) = 8.0 * FA(cu, 2, 0, 1);
    FA(dcu, 0, 0, 1) = 0.0; FA(dcu, 1, 0, 1) = 0.0;
    for (int c = 0; c < 3; c++) FA(amu, c, 0, 1) = 0.0;
    for (int m = 1; m <= M; m++) {
        int ks[2] = {kre(m), kim(m)};
        for (int t = 0; t < 2; t++) {
            int k = ks[t];
            for (int c = 0; c < 3; c++) { FA(cu, c, k, 0) = 0.0; FA(amu, c, k, 0) = 0.0; }
            for (int c = 0; c < 2; c++) FA(dcu, c, k, 0) = 0.0;
            if (m == 1) {
                FA(cu, 0, k, 1) = 8.0 * FA(cu, 0, k, 1); FA(cu, 1, k, 1) = 8.0 * FA(cu, 1, k, 1);
                FA(cu, 2, k, 1) = 0.0;
                FA(dcu, 0, k, 1) = 8.0 * FA(dcu, 0, k, 1); FA(dcu, 1, k, 1) = 8.0 * FA(dcu, 1, k, 1);
                for (int c = 0; c < 3; c++) FA(amu, c, k, 1) = 0.0;
            } else if (m == 2) {
                for (int c = 0; c < 3; c++) FA(cu, c, k, 1) = 0.0;
                for (int c = 0; c < 2; c++) FA(dcu, c, k, 1) = 0.0;
                for (int c = 0; c < 3; c++) FA(amu, c, k, 1) = 8.0 * FA(amu, c, k, 1);
            } else {
                for (int c = 0; c < 3; c++) FA(cu, c, k, 1) = 0.0;
                for (int c = 0; c < 2; c++) FA(dcu, c, k, 1) = 0.0;
                for (int c = 0; c < 3; c++) FA(amu, c, k, 1) = 0.0;
            }
        }
    }
    /* divide by the node index (:962-1004) */
    for (int k = 0; k < nm; k++) {
        for (int j = 2; j <= nr + 1; j++) {
            double ir = 1.0 / (double)(j - 1);
            for (int c = 0; c < 3; c++) FA(cu, c, k, j) = FA(cu, c, k, j) * ir;
            for (int c = 0; c < 2; c++) FA(dcu, c, k, j) = FA(dcu, c, k, j) * ir;
            for (int c = 0; c < 3; c++) FA(amu, c, k, j) = FA(amu, c, k, j) * ir;
        }
    }
}

/* amjdeposit_robust_part2d: source/species/part2d_class.f03:746-1010 */
void ora_amjdeposit_robust(const ora_sim *s, ora_species *sp, const double *ef,
                           const double *bf, double *cu, double *amu,
                           double *dcu, double dt)
{
    int nm = s->nm, ld = s->ld, M = s->g.max_mode;
    double idr = 1.0 / s->g.dr;
    double idt = 1.0 / dt;
    double qtmh = 0.5 * sp->qbm * dt;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        double ep[3], bp[3], wp[3], u0[3], u[3], ut[3], du[3], u2[3];
        interp_vec(s, ef, 3, ix, w0, w1, pc, ps, ep);
        interp_vec(s, bf, 3, ix, w0, w1, pc, ps, bp);
        wp[0] = ep[0] - bp[1];
        wp[1] = ep[1] + bp[0];
        wp[2] = ep[2];
        u0[0] = sp->p1[p] * pc + sp->p2[p] * ps;
        u0[1] = sp->p2[p] * pc - sp->p1[p] * ps;
        u0[2] = sp->p3[p];
        double gam = sqrt(1.0 + u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
        double qtmh1 = qtmh * gam / (gam - u0[2]);
        for (int c = 0; c < 3; c++) { ep[c] = ep[c] * qtmh1; ut[c] = u0[c] + ep[c]; }
        gam = sqrt(1.0 + ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
        double qtmh2 = qtmh / (gam - ut[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * qtmh2;
        u[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
        u[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
        u[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
        double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * ostq;
        ut[0] = ut[0] + u[1] * bp[2] - u[2] * bp[1];
        ut[1] = ut[1] + u[2] * bp[0] - u[0] * bp[2];
        ut[2] = ut[2] + u[0] * bp[1] - u[1] * bp[0];
        for (int c = 0; c < 3; c++) u[c] = ut[c] + ep[c];

        du[0] = idt * (u[0] - u0[0]);
        du[1] = idt * (u[1] - u0[1]);
        for (int c = 0; c < 3; c++) u[c] = 0.5 * (u[c] + u0[c]);
        sp->gamma[p] = sqrt(1.0 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double ipsi = 1.0 / (sp->gamma[p] - u[2]);
        sp->psi[p] = (1.0 - 1.0 / ipsi) / sp->qbm;
        double dpsi = sp->qbm * (wp[2] - (wp[0] * u[0] + wp[1] * u[1]) * ipsi);
        du[0] = du[0] + u[0] * dpsi * ipsi;
        du[1] = du[1] + u[1] * dpsi * ipsi;
        u2[0] = u[0] * u[0] * ipsi;
        u2[1] = u[0] * u[1] * ipsi;
        u2[2] = u[1] * u[1] * ipsi;
        double ph0r = pc, ph0i = -ps;
        double phr = 1.0 * sp->q[p] * ipsi, phi = 0.0 * sp->q[p] * ipsi;
        double wt[2] = {w0, w1};
        for (int j = 0; j < 2; j++) {
            double w = wt[j] * phr;
            for (int c = 0; c < 3; c++) FA(cu, c, 0, ix + j) = FA(cu, c, 0, ix + j) + w * u[c];
            for (int c = 0; c < 2; c++) FA(dcu, c, 0, ix + j) = FA(dcu, c, 0, ix + j) + w * du[c];
            for (int c = 0; c < 3; c++) FA(amu, c, 0, ix + j) = FA(amu, c, 0, ix + j) + w * u2[c];
        }
        for (int m = 1; m <= M; m++) {
            double a = phr * ph0r - phi * ph0i, b = phr * ph0i + phi * ph0r;
            phr = a; phi = b;
            int kr = kre(m), ki = kim(m);
            for (int j = 0; j < 2; j++) {
                double w = wt[j] * phr;
                for (int c = 0; c < 3; c++) FA(cu, c, kr, ix + j) = FA(cu, c, kr, ix + j) + w * u[c];
                for (int c = 0; c < 2; c++) FA(dcu, c, kr, ix + j) = FA(dcu, c, kr, ix + j) + w * du[c];
                for (int c = 0; c < 3; c++) FA(amu, c, kr, ix + j) = FA(amu, c, kr, ix + j) + w * u2[c];
                w = wt[j] * phi;
                for (int c = 0; c < 3; c++) FA(cu, c, ki, ix + j) = FA(cu, c, ki, ix + j) + w * u[c];
                for (int c = 0; c < 2; c++) FA(dcu, c, ki, ix + j) = FA(dcu, c, ki, ix + j) + w * du[c];
                for (int c = 0; c < 3; c++) FA(amu, c, ki, ix + j) = FA(amu, c, ki, ix + j) + w * u2[c];
            }
        }
    }
    deposit_axis_and_radius(s, cu, amu, dcu);
}

/* push_u_robust_part2d: source/species/part2d_class.f03:1879-1965 */
void ora_push_u_robust(const ora_sim *s, ora_species *sp, const double *ef,
                       const double *bf, double dt)
{
    double idr = 1.0 / s->g.dr;
    double qtmh = sp->qbm * dt * 0.5;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        double ep[3], bp[3], ut[3], pp[3];
        interp_vec(s, ef, 3, ix, w0, w1, pc, ps, ep);
        interp_vec(s, bf, 3, ix, w0, w1, pc, ps, bp);
        /* transform_to_cartesian (interp_part2d.f03:155-171) */
        double t = ep[0] * pc - ep[1] * ps; ep[1] = ep[0] * ps + ep[1] * pc; ep[0] = t;
        t = bp[0] * pc - bp[1] * ps; bp[1] = bp[0] * ps + bp[1] * pc; bp[0] = t;
        pp[0] = sp->p1[p]; pp[1] = sp->p2[p]; pp[2] = sp->p3[p];
        double gam = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
        double qtmh1 = qtmh / (gam - pp[2]);
        double qtmh2 = qtmh1 * gam;
        for (int c = 0; c < 3; c++) { ep[c] = ep[c] * qtmh2; bp[c] = bp[c] * qtmh1; }
        for (int c = 0; c < 3; c++) ut[c] = pp[c] + ep[c];
        pp[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
        pp[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
        pp[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
        double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * ostq;
        ut[0] = ut[0] + pp[1] * bp[2] - pp[2] * bp[1];
        ut[1] = ut[1] + pp[2] * bp[0] - pp[0] * bp[2];
        ut[2] = ut[2] + pp[0] * bp[1] - pp[1] * bp[0];
        for (int c = 0; c < 3; c++) pp[c] = ut[c] + ep[c];
        sp->p1[p] = pp[0]; sp->p2[p] = pp[1]; sp->p3[p] = pp[2];
        sp->gamma[p] = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
    }
}

/* interp_psi_part2d: source/species/part2d_class.f03:2264-2305 -- the
   field psi at each particle, which the standard pusher uses in place of
   the robust pusher's gamma - u_z */
void ora_interp_psi(const ora_sim *s, ora_species *sp, const double *psi)
{
    double idr = 1.0 / s->g.dr;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        interp_vec(s, psi, 1, ix, w0, w1, pc, ps, &sp->psi[p]);
    }
}

/* amjdeposit_std_part2d: source/species/part2d_class.f03:478-744.  The
   predictor uses the interpolated psi of each particle (interp_psi) and
   stores the time-centred gamma; the axis rules and the division by the
   node index are the robust deposit's (:683-742 = :914-1004) */
void ora_amjdeposit_std(const ora_sim *s, ora_species *sp, const double *ef,
                        const double *bf, double *cu, double *amu,
                        double *dcu, double dt)
{
    int nm = s->nm, ld = s->ld, M = s->g.max_mode;
    double idr = 1.0 / s->g.dr;
    double idt = 1.0 / dt;
    double qtmh = 0.5 * sp->qbm * dt;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        double ep[3], bp[3], wp[3], u0[3], u[3], ut[3], du[3], u2[3];
        interp_vec(s, ef, 3, ix, w0, w1, pc, ps, ep);
        interp_vec(s, bf, 3, ix, w0, w1, pc, ps, bp);
        wp[0] = ep[0] - bp[1];
        wp[1] = ep[1] + bp[0];
        wp[2] = ep[2];
        u0[0] = sp->p1[p] * pc + sp->p2[p] * ps;
        u0[1] = sp->p2[p] * pc - sp->p1[p] * ps;
        u0[2] = sp->p3[p];
        double gam = sqrt(1.0 + u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
        double qtmh2 = qtmh / (1.0 - sp->qbm * sp->psi[p]);
        double qtmh1 = qtmh2 * gam;
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * qtmh2;
        for (int c = 0; c < 3; c++) ut[c] = u0[c] + ep[c] * qtmh1;
        u[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
        u[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
        u[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
        double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * ostq;
        ut[0] = ut[0] + u[1] * bp[2] - u[2] * bp[1];
        ut[1] = ut[1] + u[2] * bp[0] - u[0] * bp[2];
        ut[2] = ut[2] + u[0] * bp[1] - u[1] * bp[0];
        gam = sqrt(1.0 + ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
        qtmh1 = qtmh2 * gam;
        for (int c = 0; c < 3; c++) u[c] = ut[c] + ep[c] * qtmh1;

        du[0] = idt * (u[0] - u0[0]);
        du[1] = idt * (u[1] - u0[1]);
        for (int c = 0; c < 3; c++) u[c] = 0.5 * (u[c] + u0[c]);
        sp->gamma[p] = sqrt(1.0 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double ipsi = 1.0 / (1.0 - sp->qbm * sp->psi[p]);
        double dpsi = sp->qbm * (wp[2] - (wp[0] * u[0] + wp[1] * u[1]) * ipsi);
        du[0] = du[0] + u[0] * dpsi * ipsi;
        du[1] = du[1] + u[1] * dpsi * ipsi;
        u2[0] = u[0] * u[0] * ipsi;
        u2[1] = u[0] * u[1] * ipsi;
        u2[2] = u[1] * u[1] * ipsi;
        double ph0r = pc, ph0i = -ps;
        double phr = 1.0 * sp->q[p] * ipsi, phi = 0.0 * sp->q[p] * ipsi;
        double wt[2] = {w0, w1};
        for (int j = 0; j < 2; j++) {
            double w = wt[j] * phr;
            for (int c = 0; c < 3; c++) FA(cu, c, 0, ix + j) = FA(cu, c, 0, ix + j) + w * u[c];
            for (int c = 0; c < 2; c++) FA(dcu, c, 0, ix + j) = FA(dcu, c, 0, ix + j) + w * du[c];
            for (int c = 0; c < 3; c++) FA(amu, c, 0, ix + j) = FA(amu, c, 0, ix + j) + w * u2[c];
        }
        for (int m = 1; m <= M; m++) {
            double a = phr * ph0r - phi * ph0i, b = phr * ph0i + phi * ph0r;
            phr = a; phi = b;
            int kr = kre(m), ki = kim(m);
            for (int j = 0; j < 2; j++) {
                double w = wt[j] * phr;
                for (int c = 0; c < 3; c++) FA(cu, c, kr, ix + j) = FA(cu, c, kr, ix + j) + w * u[c];
                for (int c = 0; c < 2; c++) FA(dcu, c, kr, ix + j) = FA(dcu, c, kr, ix + j) + w * du[c];
                for (int c = 0; c < 3; c++) FA(amu, c, kr, ix + j) = FA(amu, c, kr, ix + j) + w * u2[c];
                w = wt[j] * phi;
                for (int c = 0; c < 3; c++) FA(cu, c, ki, ix + j) = FA(cu, c, ki, ix + j) + w * u[c];
                for (int c = 0; c < 2; c++) FA(dcu, c, ki, ix + j) = FA(dcu, c, ki, ix + j) + w * du[c];
                for (int c = 0; c < 3; c++) FA(amu, c, ki, ix + j) = FA(amu, c, ki, ix + j) + w * u2[c];
            }
        }
    }
    deposit_axis_and_radius(s, cu, amu, dcu);
}

/* push_u_std_part2d: source/species/part2d_class.f03:1790-1877 -- the
   interpolated psi and the gamma stored by the last current deposit in
   place of the robust pusher's own gamma - p_z and gamma */
void ora_push_u_std(const ora_sim *s, ora_species *sp, const double *ef,
                    const double *bf, double dt)
{
    double idr = 1.0 / s->g.dr;
    double qtmh = sp->qbm * dt * 0.5;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        double ep[3], bp[3], ut[3], pp[3];
        interp_vec(s, ef, 3, ix, w0, w1, pc, ps, ep);
        interp_vec(s, bf, 3, ix, w0, w1, pc, ps, bp);
        double t = ep[0] * pc - ep[1] * ps; ep[1] = ep[0] * ps + ep[1] * pc; ep[0] = t;
        t = bp[0] * pc - bp[1] * ps; bp[1] = bp[0] * ps + bp[1] * pc; bp[0] = t;
        pp[0] = sp->p1[p]; pp[1] = sp->p2[p]; pp[2] = sp->p3[p];
        double qtmh1 = qtmh / (1.0 - sp->qbm * sp->psi[p]);
        double qtmh2 = qtmh1 * sp->gamma[p];
        for (int c = 0; c < 3; c++) { ep[c] = ep[c] * qtmh2; bp[c] = bp[c] * qtmh1; }
        for (int c = 0; c < 3; c++) ut[c] = pp[c] + ep[c];
        pp[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
        pp[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
        pp[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
        double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * ostq;
        ut[0] = ut[0] + pp[1] * bp[2] - pp[2] * bp[1];
        ut[1] = ut[1] + pp[2] * bp[0] - pp[0] * bp[2];
        ut[2] = ut[2] + pp[0] * bp[1] - pp[1] * bp[0];
        for (int c = 0; c < 3; c++) pp[c] = ut[c] + ep[c];
        sp->p1[p] = pp[0]; sp->p2[p] = pp[1]; sp->p3[p] = pp[2];
        sp->gamma[p] = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
    }
}

/* ------------------------------------------------------------------------ */
/* Laser (PGC) path                                                          */
/* ------------------------------------------------------------------------ */

/* the envelope history of the laser (field_laser's cfr, cfi f2; with
   several lasers, their sum), slices -1 .. nxi+1: field_laser keeps two
   guard slices below the box (sim_lasers_class.f03:105-110) */
static int laser_alloc(ora_sim *s)
{
    const size_t n1 = (size_t)s->nm * s->ld, nh = (size_t)(s->g.nxi + 3) * n1;
    if (!s->laser_on) {
        s->lz_ar = calloc(nh, sizeof(double));
        s->lz_ai = calloc(nh, sizeof(double));
        s->ar = calloc(n1, sizeof(double));
        s->ai = calloc(n1, sizeof(double));
        s->ar_grad = calloc(3 * n1, sizeof(double));
        s->ai_grad = calloc(3 * n1, sizeof(double));
        s->chi = calloc(n1, sizeof(double));
        s->h_chi = calloc((size_t)(s->g.nxi + 2) * n1, sizeof(double));
        if (!s->lz_ar || !s->lz_ai || !s->ar || !s->ai || !s->ar_grad || !s->ai_grad || !s->chi || !s->h_chi)
            return -1;
        s->laser_on = 1;
    }
    return 0;
}

int ora_set_laser(ora_sim *s, const double *ar_hist, const double *ai_hist)
{
    const size_t nh = (size_t)(s->g.nxi + 3) * s->nm * s->ld;
    if (laser_alloc(s)) return -1;
    memcpy(s->lz_ar, ar_hist, nh * sizeof(double));
    memcpy(s->lz_ai, ai_hist, nh * sizeof(double));
    return 0;
}

/* laser_all of slice j with one laser: zero_field_laser(only_f1)
   (field_laser_class.f03:952-1012), copy_slice of the envelope, the
   gradients of set_grad_field_laser (:637-750) and gather (:929-946).
   Gradient components are (r, phi, z): central differences in r, the
   (1/r) d/dphi of the mode (-m Im, m Re over r) and a three-point backward
   difference in xi over slices j, j-1, j-2.  For m > 0 the axis values of
   the first rank are written after the radial loop with its index, that is
   at node nr + 1 (:710-724), and node 1 keeps zero: restated as is. */
void ora_laser_slice(ora_sim *s, int j)
{
    const int nm = s->nm, ld = s->ld, nr = s->g.nr, M = s->g.max_mode;
    const size_t n1 = (size_t)nm * ld;
    const double idrh = 0.5 / s->g.dr, idzh = 0.5 / s->g.dxi;
#define LZ(h, sl, k, i) (h)[((size_t)((sl) + 1) * nm + (k)) * ld + (i)]
    memset(s->ar_grad, 0, 3 * n1 * sizeof(double));
    memset(s->ai_grad, 0, 3 * n1 * sizeof(double));
    for (int k = 0; k < nm; k++)
        for (int i = 0; i <= nr + 1; i++) {
            FA(s->ar, 0, k, i) = LZ(s->lz_ar, j, k, i);
            FA(s->ai, 0, k, i) = LZ(s->lz_ai, j, k, i);
        }
    const double *hs[2] = {s->lz_ar, s->lz_ai};
    double *gs[2] = {s->ar_grad, s->ai_grad};
    for (int t = 0; t < 2; t++) {
        const double *h = hs[t];
        double *gr = gs[t];
        for (int i = 1; i <= nr; i++)
            FA(gr, 2, 0, i) = idzh * (3.0 * LZ(h, j, 0, i) - 4.0 * LZ(h, j - 1, 0, i) + LZ(h, j - 2, 0, i));
        for (int i = 2; i <= nr; i++) {
            FA(gr, 0, 0, i) = idrh * (LZ(h, j, 0, i + 1) - LZ(h, j, 0, i - 1));
            FA(gr, 1, 0, i) = 0.0;
        }
        FA(gr, 0, 0, 1) = 0.0;
        FA(gr, 1, 0, 1) = 0.0;
        for (int m = 1; m <= M; m++) {
            const int kr = kre(m), ki = kim(m);
            for (int i = 1; i <= nr; i++) {
                FA(gr, 2, kr, i) = idzh * (3.0 * LZ(h, j, kr, i) - 4.0 * LZ(h, j - 1, kr, i) + LZ(h, j - 2, kr, i));
                FA(gr, 2, ki, i) = idzh * (3.0 * LZ(h, j, ki, i) - 4.0 * LZ(h, j - 1, ki, i) + LZ(h, j - 2, ki, i));
            }
            for (int i = 2; i <= nr; i++) {
                const double ir = 1.0 / ((double)(i - 1) * s->g.dr);
                FA(gr, 0, kr, i) = idrh * (LZ(h, j, kr, i + 1) - LZ(h, j, kr, i - 1));
                FA(gr, 0, ki, i) = idrh * (LZ(h, j, ki, i + 1) - LZ(h, j, ki, i - 1));
                FA(gr, 1, kr, i) = -ir * m * LZ(h, j, ki, i);
                FA(gr, 1, ki, i) = ir * m * LZ(h, j, kr, i);
            }
            const int ie = nr + 1;   /* the loop index after "do i = 2, nrp" */
            if (m % 2 == 1) {
                FA(gr, 0, kr, ie) = 2.0 * idrh * LZ(h, j, kr, 2);
                FA(gr, 0, ki, ie) = 2.0 * idrh * LZ(h, j, ki, 2);
                FA(gr, 1, kr, ie) = -m * FA(gr, 0, ki, ie);
                FA(gr, 1, ki, ie) = m * FA(gr, 0, kr, ie);
            } else {
                FA(gr, 0, kr, ie) = 0.0;
                FA(gr, 0, ki, ie) = 0.0;
                FA(gr, 1, kr, ie) = 0.0;
                FA(gr, 1, ki, ie) = 0.0;
            }
        }
    }
#undef LZ
}

/* interp_field of the laser envelope and its gradients at a particle
   (part2d_class.f03:1406-1409, interp_part2d.f03:66-110) */
static void laser_gather(const ora_sim *s, int ix, double w0, double w1, double pc, double ps,
                         double *apr, double *api, double *aprg, double *apig)
{
    interp_vec(s, s->ar, 1, ix, w0, w1, pc, ps, apr);
    interp_vec(s, s->ai, 1, ix, w0, w1, pc, ps, api);
    interp_vec(s, s->ar_grad, 3, ix, w0, w1, pc, ps, aprg);
    interp_vec(s, s->ai_grad, 3, ix, w0, w1, pc, ps, apig);
}

/* amjdeposit_robust_pgc_part2d (part2d_class.f03:1310-1602) and, std_,
   amjdeposit_std_pgc_part2d (:1012-1308): the ponderomotive correction
   gam_corr = (q/m)^2 |a|^2 / 2 in every gamma, and the effective field
   E -+ (q/m)/(2 gamma) grad |a|^2 / 2 (r, phi components minus, z plus)
   with the initial gamma; the robust form normalises each half kick by its
   own gamma - u_z, the standard one by the interpolated psi */
void ora_amjdeposit_pgc(const ora_sim *s, ora_species *sp, const double *ef, const double *bf,
                        double *cu, double *amu, double *dcu, double dt, int std_)
{
    int nm = s->nm, ld = s->ld, M = s->g.max_mode;
    double idr = 1.0 / s->g.dr, idt = 1.0 / dt, qtmh = 0.5 * sp->qbm * dt;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        double ep[3], bp[3], wp[3], u0[3], u[3], ut[3], du[3], u2[3];
        double apr, api, aprg[3], apig[3];
        interp_vec(s, ef, 3, ix, w0, w1, pc, ps, ep);
        interp_vec(s, bf, 3, ix, w0, w1, pc, ps, bp);
        laser_gather(s, ix, w0, w1, pc, ps, &apr, &api, aprg, apig);
        double gam_corr = 0.5 * sp->qbm * sp->qbm * (apr * apr + api * api);
        u0[0] = sp->p1[p] * pc + sp->p2[p] * ps;
        u0[1] = sp->p2[p] * pc - sp->p1[p] * ps;
        u0[2] = sp->p3[p];
        double gam = sqrt(1.0 + u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2] + gam_corr);
        wp[0] = ep[0] - bp[1];
        wp[1] = ep[1] + bp[0];
        wp[2] = ep[2];
        double tmp = 0.5 * sp->qbm / gam;
        ep[0] = ep[0] - tmp * (apr * aprg[0] + api * apig[0]);
        ep[1] = ep[1] - tmp * (apr * aprg[1] + api * apig[1]);
        ep[2] = ep[2] + tmp * (apr * aprg[2] + api * apig[2]);
        double qtmh_e, qtmh_b = 0.0;
        if (std_) {
            qtmh_b = qtmh / (1.0 - sp->qbm * sp->psi[p]);
            qtmh_e = qtmh_b * gam;
        } else {
            qtmh_e = qtmh * gam / (gam - u0[2]);
        }
        for (int c = 0; c < 3; c++) ut[c] = u0[c] + ep[c] * qtmh_e;
        if (!std_) {
            gam = sqrt(1.0 + ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2] + gam_corr);
            qtmh_b = qtmh / (gam - ut[2]);
        }
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * qtmh_b;
        u[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
        u[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
        u[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
        double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * ostq;
        ut[0] = ut[0] + u[1] * bp[2] - u[2] * bp[1];
        ut[1] = ut[1] + u[2] * bp[0] - u[0] * bp[2];
        ut[2] = ut[2] + u[0] * bp[1] - u[1] * bp[0];
        gam = sqrt(1.0 + ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2] + gam_corr);
        qtmh_e = std_ ? qtmh_b * gam : qtmh * gam / (gam - ut[2]);
        for (int c = 0; c < 3; c++) u[c] = ut[c] + ep[c] * qtmh_e;
        du[0] = idt * (u[0] - u0[0]);
        du[1] = idt * (u[1] - u0[1]);
        for (int c = 0; c < 3; c++) u[c] = 0.5 * (u[c] + u0[c]);
        sp->gamma[p] = sqrt(1.0 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + gam_corr);
        double ipsi;
        if (std_) {
            ipsi = 1.0 / (1.0 - sp->qbm * sp->psi[p]);
        } else {
            ipsi = 1.0 / (sp->gamma[p] - u[2]);
            sp->psi[p] = (1.0 - 1.0 / ipsi) / sp->qbm;
        }
        double dpsi = sp->qbm * (wp[2] - (wp[0] * u[0] + wp[1] * u[1]) * ipsi);
        du[0] = du[0] + u[0] * dpsi * ipsi;
        du[1] = du[1] + u[1] * dpsi * ipsi;
        u2[0] = u[0] * u[0] * ipsi;
        u2[1] = u[0] * u[1] * ipsi;
        u2[2] = u[1] * u[1] * ipsi;
        double ph0r = pc, ph0i = -ps;
        double phr = 1.0 * sp->q[p] * ipsi, phi = 0.0 * sp->q[p] * ipsi;
        double wt[2] = {w0, w1};
        for (int jj = 0; jj < 2; jj++) {
            double w = wt[jj] * phr;
            for (int c = 0; c < 3; c++) FA(cu, c, 0, ix + jj) = FA(cu, c, 0, ix + jj) + w * u[c];
            for (int c = 0; c < 2; c++) FA(dcu, c, 0, ix + jj) = FA(dcu, c, 0, ix + jj) + w * du[c];
            for (int c = 0; c < 3; c++) FA(amu, c, 0, ix + jj) = FA(amu, c, 0, ix + jj) + w * u2[c];
        }
        for (int m = 1; m <= M; m++) {
            double a = phr * ph0r - phi * ph0i, b = phr * ph0i + phi * ph0r;
            phr = a; phi = b;
            int kr = kre(m), ki = kim(m);
            for (int jj = 0; jj < 2; jj++) {
                double w = wt[jj] * phr;
                for (int c = 0; c < 3; c++) FA(cu, c, kr, ix + jj) = FA(cu, c, kr, ix + jj) + w * u[c];
                for (int c = 0; c < 2; c++) FA(dcu, c, kr, ix + jj) = FA(dcu, c, kr, ix + jj) + w * du[c];
                for (int c = 0; c < 3; c++) FA(amu, c, kr, ix + jj) = FA(amu, c, kr, ix + jj) + w * u2[c];
                w = wt[jj] * phi;
                for (int c = 0; c < 3; c++) FA(cu, c, ki, ix + jj) = FA(cu, c, ki, ix + jj) + w * u[c];
                for (int c = 0; c < 2; c++) FA(dcu, c, ki, ix + jj) = FA(dcu, c, ki, ix + jj) + w * du[c];
                for (int c = 0; c < 3; c++) FA(amu, c, ki, ix + jj) = FA(amu, c, ki, ix + jj) + w * u2[c];
            }
        }
    }
    deposit_axis_and_radius(s, cu, amu, dcu);
}

/* push_u_robust_pgc_part2d (part2d_class.f03:1967-2092) and
   push_u_std_pgc_part2d (:2094-2219), the same steps: both normalise by
   the time-centred gamma and psi the last current deposit stored; the new
   gamma carries gam_corr extrapolated over dxi with the xi gradient of the
   envelope */
void ora_push_u_pgc(const ora_sim *s, ora_species *sp, const double *ef, const double *bf, double dt)
{
    double idr = 1.0 / s->g.dr;
    double qtmh = sp->qbm * dt * 0.5, qbm2_hf = sp->qbm * sp->qbm * 0.5;
    for (int64_t p = 0; p < sp->npp; p++) {
        int ix; double w0, w1, pc, ps;
        interp_info(sp->x1[p], sp->x2[p], idr, &ix, &w0, &w1, &pc, &ps);
        double ep[3], bp[3], ut[3], pp[3], apr, api, aprg[3], apig[3];
        interp_vec(s, ef, 3, ix, w0, w1, pc, ps, ep);
        interp_vec(s, bf, 3, ix, w0, w1, pc, ps, bp);
        laser_gather(s, ix, w0, w1, pc, ps, &apr, &api, aprg, apig);
        double t = ep[0] * pc - ep[1] * ps; ep[1] = ep[0] * ps + ep[1] * pc; ep[0] = t;
        t = bp[0] * pc - bp[1] * ps; bp[1] = bp[0] * ps + bp[1] * pc; bp[0] = t;
        t = aprg[0] * pc - aprg[1] * ps; aprg[1] = aprg[0] * ps + aprg[1] * pc; aprg[0] = t;
        t = apig[0] * pc - apig[1] * ps; apig[1] = apig[0] * ps + apig[1] * pc; apig[0] = t;
        double gam_corr = qbm2_hf * (apr * apr + api * api);
        double tmp = 0.5 * sp->qbm / sp->gamma[p];
        ep[0] = ep[0] - tmp * (apr * aprg[0] + api * apig[0]);
        ep[1] = ep[1] - tmp * (apr * aprg[1] + api * apig[1]);
        ep[2] = ep[2] + tmp * (apr * aprg[2] + api * apig[2]);
        double qtmh_b = qtmh / (1.0 - sp->qbm * sp->psi[p]);
        double qtmh_e = qtmh_b * sp->gamma[p];
        for (int c = 0; c < 3; c++) { ep[c] = ep[c] * qtmh_e; bp[c] = bp[c] * qtmh_b; }
        pp[0] = sp->p1[p]; pp[1] = sp->p2[p]; pp[2] = sp->p3[p];
        for (int c = 0; c < 3; c++) ut[c] = pp[c] + ep[c];
        pp[0] = ut[0] + ut[1] * bp[2] - ut[2] * bp[1];
        pp[1] = ut[1] + ut[2] * bp[0] - ut[0] * bp[2];
        pp[2] = ut[2] + ut[0] * bp[1] - ut[1] * bp[0];
        double ostq = 2.0 / (1.0 + bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int c = 0; c < 3; c++) bp[c] = bp[c] * ostq;
        ut[0] = ut[0] + pp[1] * bp[2] - pp[2] * bp[1];
        ut[1] = ut[1] + pp[2] * bp[0] - pp[0] * bp[2];
        ut[2] = ut[2] + pp[0] * bp[1] - pp[1] * bp[0];
        for (int c = 0; c < 3; c++) pp[c] = ut[c] + ep[c];
        sp->p1[p] = pp[0]; sp->p2[p] = pp[1]; sp->p3[p] = pp[2];
        tmp = aprg[2] * dt; gam_corr = gam_corr + qbm2_hf * (apr + 0.25 * tmp) * tmp;
        tmp = apig[2] * dt; gam_corr = gam_corr + qbm2_hf * (api + 0.25 * tmp) * tmp;
        sp->gamma[p] = sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2] + gam_corr);
    }
}

/* deposit_chi_part2d (part2d_class.f03:361-476): the susceptibility
   -(q/m) q / (1 - (q/m) psi) of each particle with linear weights and
   e^{-im phi}; then, on the whole array, node 0 zero, the m = 0 axis node
   times get_deposit_ax_corr(ppc_r) = 12 ppc^2 / (1 + 2 ppc^2) (:2581-2586),
   m > 0 axis zero, node j divided by j - 1.  The rescaling acts on
   whatever chi holds, so a second species rescales the first one's
   deposit again (sim_lasers_class.f03:176-200 calls it per species into
   one array): restated as is. */
void ora_deposit_chi(const ora_sim *s, const ora_species *sp, double *chi)
{
    int nr = s->g.nr, nm = s->nm, ld = s->ld, M = s->g.max_mode;
    double idr = 1.0 / s->g.dr;
    for (int64_t p = 0; p < sp->npp; p++) {
        double pos = sqrt(sp->x1[p] * sp->x1[p] + sp->x2[p] * sp->x2[p]) * idr;
        double ph0r = sp->x1[p] / pos * idr, ph0i = -sp->x2[p] / pos * idr;
        int nn = (int)floor(pos);
        pos = pos - (double)nn;
        nn = nn + 1;
        double wt[2] = {1.0 - pos, pos};
        double phr = -1.0 * sp->qbm * sp->q[p] / (1.0 - sp->qbm * sp->psi[p]), phi = 0.0;
        for (int jj = 0; jj < 2; jj++) FA(chi, 0, 0, nn + jj) = FA(chi, 0, 0, nn + jj) + wt[jj] * phr;
        for (int m = 1; m <= M; m++) {
            double a = phr * ph0r - phi * ph0i, b = phr * ph0i + phi * ph0r;
            phr = a; phi = b;
            for (int jj = 0; jj < 2; jj++) {
                FA(chi, 0, kre(m), nn + jj) = FA(chi, 0, kre(m), nn + jj) + wt[jj] * phr;
                FA(chi, 0, kim(m), nn + jj) = FA(chi, 0, kim(m), nn + jj) + wt[jj] * phi;
            }
        }
    }
    const double ppc = (double)sp->cfg.ppc1;
    const double corr = (12.0 * ppc * ppc) / (1.0 + 2.0 * ppc * ppc);
    FA(chi, 0, 0, 0) = 0.0;
    FA(chi, 0, 0, 1) = FA(chi, 0, 0, 1) * corr;
    for (int j = 2; j <= nr + 1; j++) FA(chi, 0, 0, j) = FA(chi, 0, 0, j) * (1.0 / (double)(j - 1));
    for (int m = 1; m <= M; m++) {
        FA(chi, 0, kre(m), 0) = 0.0; FA(chi, 0, kim(m), 0) = 0.0;
        FA(chi, 0, kre(m), 1) = 0.0; FA(chi, 0, kim(m), 1) = 0.0;
        for (int j = 2; j <= nr + 1; j++) {
            double ir = 1.0 / (double)(j - 1);
            FA(chi, 0, kre(m), j) = FA(chi, 0, kre(m), j) * ir;
            FA(chi, 0, kim(m), j) = FA(chi, 0, kim(m), j) * ir;
        }
    }
}

/* push_x_part2d: source/species/part2d_class.f03:2221-2262 */
void ora_push_x(ora_species *sp, double dt)
{
    for (int64_t p = 0; p < sp->npp; p++) {
        double dtc = dt / (sp->gamma[p] - sp->p3[p]);
        sp->x1[p] = sp->x1[p] + sp->p1[p] * dtc;
        sp->x2[p] = sp->x2[p] + sp->p2[p] * dtc;
    }
}

/* update_bound_part2d: source/species/part2d_class.f03:2307-2353
   (swap-with-last removal of r >= edge) */
void ora_update_bound(const ora_sim *s, ora_species *sp)
{
    if (sp->npp == 0) return;
    double edge = (double)s->g.nr * s->g.dr;
    int64_t i = 0;
#define MOVE(dst, src) do { sp->x1[dst] = sp->x1[src]; sp->x2[dst] = sp->x2[src]; \
        sp->p1[dst] = sp->p1[src]; sp->p2[dst] = sp->p2[src]; sp->p3[dst] = sp->p3[src]; \
        sp->gamma[dst] = sp->gamma[src]; sp->psi[dst] = sp->psi[src]; sp->q[dst] = sp->q[src]; \
        sp->id[dst] = sp->id[src]; } while (0)
    /* !(pos < edge): a particle whose position is no longer a number leaves
       too (QPAD's pos >= edge keeps it, and its next interpolation indexes
       with int(NaN)) */
    while (i < sp->npp - 1) {
        double pos = sqrt(sp->x1[i] * sp->x1[i] + sp->x2[i] * sp->x2[i]);
        if (!(pos < edge)) {
            MOVE(i, sp->npp - 1);
            sp->npp--;
            continue;
        }
        i++;
    }
    int64_t l = sp->npp - 1;
    double pos = sqrt(sp->x1[l] * sp->x1[l] + sp->x2[l] * sp->x2[l]);
    if (!(pos < edge)) sp->npp--;
#undef MOVE
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,        */
/* "Parallel random numbers: as easy as 1, 2, 3", SC'11)                     */
/* ------------------------------------------------------------------------ */
void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Two uniforms (53-bit) -> Box-Muller pair; the stand-in for rand_norm
   (source/math_module.f03:77-102).  out[0..1] = pair from sub, out[2..3] =
   pair from sub+1. */
void ora_normal4(uint64_t seed, uint64_t stream, uint32_t sub, double out[4])
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int h = 0; h < 2; h++) {
        uint32_t ctr[4] = {(uint32_t)stream, (uint32_t)(stream >> 32), sub + (uint32_t)h, 0x51ED2D0Bu};
        uint32_t o[4];
        ora_philox4x32(ctr, key, o);
        double u1 = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
        double u2 = ((double)(o[2] >> 5) * 67108864.0 + (double)(o[3] >> 6)) * (1.0 / 9007199254740992.0);
        double theta = 2.0 * PI_DB * u1;
        double amp = sqrt(-2.0 * log(1.0 - u2));
        out[2 * h] = amp * cos(theta);
        out[2 * h + 1] = amp * sin(theta);
    }
}

/* inject_fdist2d: source/species/fdist2d_class.f03:289-357 (uniform
   perpendicular and longitudinal profiles, deterministic theta) */
void ora_species_inject(ora_sim *s, int isp)
{
    ora_species *sp = &s->sp[isp];
    const ora_species_cfg *c = &sp->cfg;
    int nr = s->g.nr;
    double dr = s->g.dr;
    double dtheta = 2.0 * PI_DB / (double)c->num_theta;
    int ppc_tot = c->ppc1 * c->ppc2;
    const double den_lon = sp->den_perp ? sp->den_lon : 1.0;
    double coef = copysign(1.0, c->qm) / ((double)ppc_tot * (double)c->num_theta);
    int64_t ip = 0, l = 0;
    for (int j = 1; j <= c->num_theta; j++) {
        for (int i = 1; i <= nr; i++) {
            for (int i1 = 1; i1 <= c->ppc1; i1++) {
                double rn = ((double)i1 - 0.5) / (double)c->ppc1 + (double)(i - 1);
                for (int i2 = 1; i2 <= c->ppc2; i2++, l++) {
                    double theta = (((double)i2 - 0.5) / (double)c->ppc2 + (double)j - 1.0) * dtheta;
                    double xt1 = rn * dr * cos(theta), xt2 = rn * dr * sin(theta);
                    const double den_perp = sp->den_perp ? sp->den_perp[l] : 1.0;
                    if (den_lon * den_perp * c->density < c->den_min) continue;
                    sp->x1[ip] = xt1;
                    sp->x2[ip] = xt2;
                    sp->q[ip] = rn * den_perp * den_lon * c->density * coef;
                    double th[4] = {0.0, 0.0, 0.0, 0.0};
                    if (c->uth[0] != 0.0 || c->uth[1] != 0.0 || c->uth[2] != 0.0)
                        ora_normal4(c->seed, (uint64_t)ip, 0u, th);
                    sp->p1[ip] = c->uth[0] * th[0];
                    sp->p2[ip] = c->uth[1] * th[1];
                    sp->p3[ip] = c->uth[2] * th[2];
                    sp->gamma[ip] = sqrt(1.0 + sp->p1[ip] * sp->p1[ip] + sp->p2[ip] * sp->p2[ip]
                                         + sp->p3[ip] * sp->p3[ip]);
                    sp->psi[ip] = (1.0 - sp->gamma[ip] + sp->p3[ip]) / c->qm;
                    sp->id[ip] = (int32_t)ip;
                    ip++;
                }
            }
        }
    }
    sp->npp = ip;
}

int ora_species_set_profile(ora_sim *s, int isp, const double *den_perp, double den_lon)
{
    ora_species *sp = &s->sp[isp];
    const size_t n = (size_t)sp->cfg.num_theta * s->g.nr * sp->cfg.ppc1 * sp->cfg.ppc2;
    free(sp->den_perp);
    sp->den_perp = NULL;
    sp->den_lon = den_lon;
    if (!den_perp) return 0;
    sp->den_perp = malloc(n * sizeof(double));
    if (!sp->den_perp) return -1;
    memcpy(sp->den_perp, den_perp, n * sizeof(double));
    return 0;
}

/* species2d%new / renew charge set-up: source/species/species2d_class.f03:
   108-130 and :154-184 (q deposit, slice-1 copy, neutralising background) */
void ora_species_init_q(ora_sim *s, int isp)
{
    ora_species *sp = &s->sp[isp];
    int nm = s->nm, ld = s->ld;
    size_t n1 = (size_t)nm * ld;
    memset(sp->q_own, 0, sizeof(double) * n1);
    ora_qdeposit(s, sp, sp->q_own);
    memcpy(sp->q_hist + 1 * n1, sp->q_own, sizeof(double) * n1);
    if (sp->cfg.neutralized) {
        for (size_t t = 0; t < n1; t++) sp->qn[t] = -1.0 * sp->q_own[t];
    }
}

/* ------------------------------------------------------------------------ */
/* Beam                                                                      */
/* ------------------------------------------------------------------------ */

/* inject_fdist3d_rnd: source/beam/fdist3d_rnd_class.f03:360-513 (Cartesian
   geometry, Gaussian profiles, no Twiss tilt, no xi-dependent offset).  The
   draws come from Philox (counter = particle index, attempt) instead of
   gfortran's random_number. */
void ora_beam_inject(ora_sim *s, int ib, const ora_beam_cfg *c)
{
    ora_beam *bm = &s->bm[ib];
    double dr = s->g.dr, dz = s->g.dxi;
    double edge_r = (double)s->g.nr * dr, edge_z = (double)s->g.nxi * dz;
    double q_per_part = copysign(1.0, c->qm) * c->total_charge
                        / (2.0 * PI_DB * (double)c->total_num * dr * dr * dz);
    int64_t n = 0;
    for (int64_t i = 0; i < c->total_num; i++) {
        double xt[3];
        for (uint32_t a = 0;; a++) {
            double g[4];
            ora_normal4(c->seed, (uint64_t)i, 2u * a, g);
            int ok = 1;
            for (int d = 0; d < 3 && ok; d++) {
                xt[d] = g[d] * c->sigma[d] + c->center[d];
                if (xt[d] < c->range_lo[d] || xt[d] >= c->range_hi[d]) ok = 0;
            }
            if (ok) break;
        }
        double r_tmp = sqrt(xt[0] * xt[0] + xt[1] * xt[1]);
        if (r_tmp < 0.0 || r_tmp >= edge_r) continue;
        if (xt[2] < 0.0 || xt[2] >= edge_z) continue;
        bm->x1[n] = xt[0]; bm->x2[n] = xt[1]; bm->x3[n] = xt[2];
        bm->q[n] = q_per_part;
        double g[4];
        ora_normal4(c->seed, (uint64_t)i, 0x80000000u, g);
        bm->p1[n] = c->uth[0] * g[0];
        bm->p2[n] = c->uth[1] * g[1];
        bm->p3[n] = c->uth[2] * g[2] + c->gamma;
        bm->p3[n] = sqrt(bm->p3[n] * bm->p3[n] - bm->p1[n] * bm->p1[n] - bm->p2[n] * bm->p2[n] - 1.0);
        bm->id[n] = (int32_t)i;
        n++;
    }
    if (c->quiet) {
        for (int64_t i = 0; i < n; i++) {
            bm->x1[n + i] = -bm->x1[i]; bm->x2[n + i] = -bm->x2[i]; bm->x3[n + i] = bm->x3[i];
            bm->p1[n + i] = -bm->p1[i]; bm->p2[n + i] = -bm->p2[i]; bm->p3[n + i] = bm->p3[i];
            bm->q[i] = 0.5 * bm->q[i];
            bm->q[n + i] = bm->q[i];
            bm->id[n + i] = bm->id[i] + (int32_t)c->total_num;
        }
        n *= 2;
    }
    bm->npp = n;
}

/* density profiles of fdist3d_std_lib.f03: get_den_uniform (:33-39),
   get_den_gaussian (:114-127), get_den_pw_linear (:332-363) */
static double std_profile(const ora_beam_std_cfg *c, int d, double x)
{
    if (c->profile[d] == 1) {
        double mu = c->gauss_center[d], sigma = c->gauss_sigma[d];
        double t = (x - mu) / sigma;
        return exp(-0.5 * (t * t));
    }
    if (c->profile[d] == 2) {
        int n = c->pw_n[d];
        const double *xa = c->pw_x[d], *fa = c->pw_fx[d];
        if (x < xa[0] || x > xa[n - 1]) return 0.0;
        for (int i = 1; i < n; i++)
            if (x <= xa[i]) return fa[i - 1] + (fa[i] - fa[i - 1]) / (xa[i] - xa[i - 1]) * (x - xa[i - 1]);
        return 0.0;
    }
    return 1.0;
}

/* inject_fdist3d_std: source/beam/fdist3d_std_class.f03:439-613 (one radial
   and longitudinal partition).  Pass 0 counts the kept lattice points, pass
   1 stores them; momenta draw from Philox (counter = kept index) instead of
   gfortran's rand_norm. */
/* a tabulated lattice density (ora_add_beam_std_table): the box xi cells
   k0 .. k0 + nk - 1, r cells 1 .. ni, in the lattice order */
static const double *g_tab;
static int g_tab_k0, g_tab_nk, g_tab_ni;

static int64_t std_lattice(ora_sim *s, const ora_beam_std_cfg *c, ora_beam *bm, int store)
{
    int nrp = s->g.nr, nzp = s->g.nxi, nth = c->num_theta;
    double dr = s->g.dr, dz = s->g.dxi;
    double dtheta = 2.0 * PI_DB / (double)nth;
    int ppc_tot = c->ppc[0] * c->ppc[1] * c->ppc[2];
    double coef = copysign(1.0, c->qm) / ((double)ppc_tot * (double)nth);
    int64_t ip = 0;
    for (int k = 1; k <= nzp; k++)
        for (int j = 1; j <= nth; j++)
            for (int i = 1; i <= nrp; i++)
                for (int i3 = 1; i3 <= c->ppc[2]; i3++) {
                    double zn = ((double)i3 - 0.5) / (double)c->ppc[2] + (double)(k - 1);
                    for (int i2 = 1; i2 <= c->ppc[1]; i2++) {
                        double theta = (((double)i2 - 0.5) / (double)c->ppc[1] + (double)j - 1.0) * dtheta;
                        for (int i1 = 1; i1 <= c->ppc[0]; i1++) {
                            double rn = ((double)i1 - 0.5) / (double)c->ppc[0] + (double)(i - 1);
                            double xt[3];
                            if (c->geometry == 1) {
                                xt[0] = rn * dr;
                                xt[1] = theta;
                            } else {
                                xt[0] = rn * dr * cos(theta);
                                xt[1] = rn * dr * sin(theta);
                            }
                            xt[2] = zn * dz;
                            int out = 0;
                            for (int d = 0; d < 3; d++)
                                if (xt[d] < c->range_lo[d] || xt[d] > c->range_hi[d]) out = 1;
                            if (out) continue;
                            double den_loc;
                            if (g_tab) {
                                if (k < g_tab_k0 || k >= g_tab_k0 + g_tab_nk || i > g_tab_ni) abort();
                                const int64_t L = (((((int64_t)(k - g_tab_k0) * nth + (j - 1)) * g_tab_ni + (i - 1))
                                                    * c->ppc[2] + (i3 - 1)) * c->ppc[1] + (i2 - 1)) * c->ppc[0]
                                                  + (i1 - 1);
                                den_loc = g_tab[L] * c->density;
                            } else {
                                double dv[3];
                                for (int d = 0; d < 3; d++) dv[d] = std_profile(c, d, xt[d]);
                                den_loc = dv[0] * dv[1] * dv[2] * c->density;
                            }
                            if (den_loc < c->den_min) continue;
                            if (store) {
                                if (c->geometry == 1) {
                                    bm->x1[ip] = xt[0] * cos(xt[1]);
                                    bm->x2[ip] = xt[0] * sin(xt[1]);
                                } else {
                                    bm->x1[ip] = xt[0];
                                    bm->x2[ip] = xt[1];
                                }
                                bm->x3[ip] = xt[2];
                                double g[4];
                                ora_normal4(c->seed, (uint64_t)ip, 0x80000000u, g);
                                bm->p1[ip] = c->uth[0] * g[0];
                                bm->p2[ip] = c->uth[1] * g[1];
                                bm->p3[ip] = c->uth[2] * g[2] + c->gamma;
                                bm->p3[ip] = sqrt(bm->p3[ip] * bm->p3[ip] - bm->p1[ip] * bm->p1[ip]
                                                  - bm->p2[ip] * bm->p2[ip] - 1.0);
                                bm->q[ip] = rn * den_loc * coef;
                                bm->id[ip] = (int32_t)ip;
                            }
                            ip++;
                        }
                    }
                }
    return ip;
}

int ora_add_beam_std_table(ora_sim *s, const ora_beam_std_cfg *c, const double *den, int k0, int nk, int ni)
{
    g_tab = den;
    g_tab_k0 = k0;
    g_tab_nk = nk;
    g_tab_ni = ni;
    const int r = ora_add_beam_std(s, c);
    g_tab = NULL;
    return r;
}

int ora_add_beam_std(ora_sim *s, const ora_beam_std_cfg *c)
{
    if (s->nbeams >= 8) return -1;
    int64_t n = std_lattice(s, c, NULL, 0);
    int ib = s->nbeams++;
    ora_beam *bm = &s->bm[ib];
    memset(bm, 0, sizeof(*bm));
    int64_t npmax = n * (c->quiet ? 2 : 1);
    if (npmax < 1) npmax = 1;
    bm->npmax = npmax;
    bm->x1 = calloc(npmax, sizeof(double)); bm->x2 = calloc(npmax, sizeof(double));
    bm->x3 = calloc(npmax, sizeof(double)); bm->p1 = calloc(npmax, sizeof(double));
    bm->p2 = calloc(npmax, sizeof(double)); bm->p3 = calloc(npmax, sizeof(double));
    bm->q = calloc(npmax, sizeof(double)); bm->id = calloc(npmax, sizeof(int32_t));
    bm->qbm = c->qm / c->mass;
    bm->dt = s->g.dt;
    bm->evol = c->evol;
    bm->q_own_hist = calloc(((size_t)s->g.nxi + 2) * s->nm * s->ld, sizeof(double));
    std_lattice(s, c, bm, 1);
    /* quiet start (:554-579) */
    if (c->quiet) {
        for (int64_t i = 0; i < n; i++) {
            bm->x1[n + i] = -bm->x1[i]; bm->x2[n + i] = -bm->x2[i]; bm->x3[n + i] = bm->x3[i];
            bm->p1[n + i] = -bm->p1[i]; bm->p2[n + i] = -bm->p2[i]; bm->p3[n + i] = bm->p3[i];
            bm->q[i] = 0.5 * bm->q[i];
            bm->q[n + i] = bm->q[i];
            bm->id[n + i] = bm->id[i] + (int32_t)n;
        }
        n *= 2;
    }
    /* Twiss tilt (:581-592) */
    for (int d = 0; d < 2; d++) {
        if (c->alpha[d] == 0.0) continue;
        double cf = c->gamma * c->alpha[d] / c->twiss_beta[d];
        double *x = d == 0 ? bm->x1 : bm->x2, *p = d == 0 ? bm->p1 : bm->p2;
        for (int64_t i = 0; i < n; i++) p[i] = p[i] - cf * (x[i] - c->twiss_ctr[d]);
    }
    /* xi-dependent transverse offset, eval_polynomial (math_module.f03:
       134-149) (:594-608) */
    for (int d = 0; d < 2; d++) {
        if (c->offset_n[d] <= 0) continue;
        double *x = d == 0 ? bm->x1 : bm->x2;
        for (int64_t i = 0; i < n; i++) {
            double t = bm->x3[i] - c->offset_xi0[d], y = 0.0, tp = 1.0;
            for (int k = 0; k < c->offset_n[d]; k++) {
                y = y + c->offset_c[d][k] * tp;
                tp = tp * t;
            }
            x[i] = x[i] + y;
        }
    }
    bm->npp = n;
    return ib;
}

/* qdeposit_part3d: source/beam/part3d_class.f03:221-356 into the beam's f2 */
void ora_beam_qdeposit(ora_sim *s, int ib, double *h)
{
    ora_beam *bm = &s->bm[ib];
    int nr = s->g.nr, nm = s->nm, ld = s->ld, M = s->g.max_mode, nzp = s->g.nxi;
    double idr = 1.0 / s->g.dr, idz = 1.0 / s->g.dxi;
    size_t sl = (size_t)nm * ld;   /* one slice of a 1-component field */
#define HQ(k, i, z) h[(size_t)(z) * sl + (size_t)(k) * ld + (size_t)(i)]
    for (int64_t p = 0; p < bm->npp; p++) {
        double pr = sqrt(bm->x1[p] * bm->x1[p] + bm->x2[p] * bm->x2[p]) * idr;
        double pz = bm->x3[p] * idz;
        double ph0r = bm->x1[p] / pr * idr, ph0i = -bm->x2[p] / pr * idr;
        int nn = (int)floor(pr), mm = (int)floor(pz);
        pr = pr - (double)nn;
        pz = pz - (double)mm;
        nn = nn + 1; mm = mm + 1;
        double wr[2] = {1.0 - pr, pr}, wz[2] = {1.0 - pz, pz};
        double phr = 1.0 * bm->q[p], phi = 0.0 * bm->q[p];
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 2; j++)
                HQ(0, nn + j, mm + k) = HQ(0, nn + j, mm + k) + wr[j] * wz[k] * phr;
        for (int m = 1; m <= M; m++) {
            double a = phr * ph0r - phi * ph0i, b = phr * ph0i + phi * ph0r;
            phr = a; phi = b;
            for (int k = 0; k < 2; k++)
                for (int j = 0; j < 2; j++) {
                    HQ(kre(m), nn + j, mm + k) = HQ(kre(m), nn + j, mm + k) + wr[j] * wz[k] * phr;
                    HQ(kim(m), nn + j, mm + k) = HQ(kim(m), nn + j, mm + k) + wr[j] * wz[k] * phi;
                }
        }
    }
    for (int z = 1; z <= nzp; z++) {
        HQ(0, 0, z) = 0.0;
        HQ(0, 1, z) = 8.0 * HQ(0, 1, z);
        for (int m = 1; m <= M; m++) {
            HQ(kre(m), 0, z) = 0.0; HQ(kim(m), 0, z) = 0.0;
            HQ(kre(m), 1, z) = 0.0; HQ(kim(m), 1, z) = 0.0;
        }
    }
    for (int j = 2; j <= nr + 1; j++) {
        double ir = 1.0 / (double)(j - 1);
        for (int z = 1; z <= nzp; z++) HQ(0, j, z) = HQ(0, j, z) * ir;
    }
    for (int m = 1; m <= M; m++)
        for (int j = 2; j <= nr + 1; j++) {
            double ir = 1.0 / (double)(j - 1);
            for (int z = 1; z <= nzp; z++) {
                HQ(kre(m), j, z) = HQ(kre(m), j, z) * ir;
                HQ(kim(m), j, z) = HQ(kim(m), j, z) * ir;
            }
        }
#undef HQ
}

/* push_reduced_part3d + interp_emf_part3d + update_bound_part3d:
   source/beam/part3d_class.f03:477-576, :691-790, :640-689 */
void ora_beam_push(ora_sim *s, int ib)
{
    ora_beam *bm = &s->bm[ib];
    if (!bm->evol) return;
    int nm = s->nm, ld = s->ld, M = s->g.max_mode;
    double idr = 1.0 / s->g.dr, idz = 1.0 / s->g.dxi;
    double qtmh = bm->qbm * bm->dt * 0.5;
    size_t sl = (size_t)3 * nm * ld;
    const double *he = s->h_e, *hb = s->h_b;
#define H3(h, c, k, i, z) (h)[(size_t)(z) * sl + ((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)]
    for (int64_t p = 0; p < bm->npp; p++) {
        double pr = sqrt(bm->x1[p] * bm->x1[p] + bm->x2[p] * bm->x2[p]) * idr;
        double pz = bm->x3[p] * idz;
        double cc = bm->x1[p] / pr * idr, ss = bm->x2[p] / pr * idr;
        int nn = (int)pr, mm = (int)pz;
        pr = pr - (double)nn;
        pz = pz - (double)mm;
        nn = nn + 1; mm = mm + 1;
        double wr[2] = {1.0 - pr, pr}, wz[2] = {1.0 - pz, pz};
        double ep[3] = {0, 0, 0}, bp[3] = {0, 0, 0};
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 2; j++) {
                double wt = wr[j] * wz[k];
                for (int c = 0; c < 3; c++) {
                    ep[c] = ep[c] + H3(he, c, 0, nn + j, mm + k) * wt;
                    bp[c] = bp[c] + H3(hb, c, 0, nn + j, mm + k) * wt;
                }
            }
        double phr = 1.0, phi = 0.0;
        for (int m = 1; m <= M; m++) {
            double a = phr * cc - phi * ss, b = phr * ss + phi * cc;
            phr = a; phi = b;
            double ph_r = 2.0 * phr, ph_i = 2.0 * phi;
            for (int k = 0; k < 2; k++)
                for (int j = 0; j < 2; j++) {
                    double wt = wr[j] * wz[k];
                    for (int c = 0; c < 3; c++) {
                        ep[c] = ep[c] + (H3(he, c, kre(m), nn + j, mm + k) * ph_r
                                         - H3(he, c, kim(m), nn + j, mm + k) * ph_i) * wt;
                        bp[c] = bp[c] + (H3(hb, c, kre(m), nn + j, mm + k) * ph_r
                                         - H3(hb, c, kim(m), nn + j, mm + k) * ph_i) * wt;
                    }
                }
        }
        double t = ep[0] * cc - ep[1] * ss; double u = ep[0] * ss + ep[1] * cc; ep[0] = t; ep[1] = u;
        t = bp[0] * cc - bp[1] * ss; u = bp[0] * ss + bp[1] * cc; bp[0] = t; bp[1] = u;
        double wp[3];
        for (int c = 0; c < 3; c++) { ep[c] = ep[c] * qtmh; bp[c] = bp[c] * qtmh; }
        wp[0] = ep[0] - bp[1];
        wp[1] = ep[1] + bp[0];
        wp[2] = ep[2];
        bm->p1[p] = bm->p1[p] + wp[0]; bm->p2[p] = bm->p2[p] + wp[1]; bm->p3[p] = bm->p3[p] + wp[2];
        bm->p1[p] = bm->p1[p] + wp[0]; bm->p2[p] = bm->p2[p] + wp[1]; bm->p3[p] = bm->p3[p] + wp[2];
        double dt_gam = bm->dt / sqrt(1.0 + bm->p1[p] * bm->p1[p] + bm->p2[p] * bm->p2[p]
                                      + bm->p3[p] * bm->p3[p]);
        bm->x1[p] = bm->x1[p] + bm->p1[p] * dt_gam;
        bm->x2[p] = bm->x2[p] + bm->p2[p] * dt_gam;
        bm->x3[p] = bm->x3[p] - bm->p3[p] * dt_gam + bm->dt;
    }
#undef H3
    /* update_bound_part3d (swap with last) */
    if (bm->npp == 0) return;
    double e1 = (double)s->g.nr * s->g.dr, e2 = (double)s->g.nxi * s->g.dxi;
#define BMOVE(dst, src) do { bm->x1[dst] = bm->x1[src]; bm->x2[dst] = bm->x2[src]; \
        bm->x3[dst] = bm->x3[src]; bm->p1[dst] = bm->p1[src]; bm->p2[dst] = bm->p2[src]; \
        bm->p3[dst] = bm->p3[src]; bm->q[dst] = bm->q[src]; bm->id[dst] = bm->id[src]; } while (0)
    int64_t i = 0;
    while (i < bm->npp - 1) {
        double pr = sqrt(bm->x1[i] * bm->x1[i] + bm->x2[i] * bm->x2[i]);
        if (pr >= e1 || bm->x3[i] >= e2) { BMOVE(i, bm->npp - 1); bm->npp--; continue; }
        i++;
    }
    int64_t l = bm->npp - 1;
    double pr = sqrt(bm->x1[l] * bm->x1[l] + bm->x2[l] * bm->x2[l]);
    if (pr >= e1 || bm->x3[l] >= e2) bm->npp--;
#undef BMOVE
}

/* ------------------------------------------------------------------------ */
/* Simulation object                                                         */
/* ------------------------------------------------------------------------ */
static const int field_ncomp[ORA_NFIELD] = {1, 3, 3, 3, 3, 3, 3, 2, 3, 2, 1, 1};

ora_sim *ora_create(const ora_grid *g)
{
    ora_sim *s = calloc(1, sizeof(ora_sim));
    s->g = *g;
    s->nm = 2 * g->max_mode + 1;
    s->ld = g->nr + 2;
    size_t n1 = (size_t)s->nm * s->ld;
    for (int f = 0; f < ORA_NFIELD; f++) {
        s->ncomp[f] = field_ncomp[f];
        s->f[f] = calloc(n1 * field_ncomp[f], sizeof(double));
    }
    size_t ns = (size_t)g->nxi + 2;
    s->h_psi = calloc(ns * n1, sizeof(double));
    s->h_e = calloc(ns * n1 * 3, sizeof(double));
    s->h_b = calloc(ns * n1 * 3, sizeof(double));
    s->h_b_spe = calloc(ns * n1 * 3, sizeof(double));
    s->h_e_spe = calloc(ns * n1 * 3, sizeof(double));
    s->h_cu = calloc(ns * n1 * 3, sizeof(double));
    s->h_q_spe = calloc(ns * n1, sizeof(double));
    s->h_q_beam = calloc(ns * n1, sizeof(double));
    s->conv_re = calloc(g->nr + 2, sizeof(double));
    s->conv_im = calloc(g->nr + 2, sizeof(double));
    int kinds[6] = {ORA_FK_PSI, ORA_FK_EZ, ORA_FK_BZ, ORA_FK_BT, ORA_FK_BPLUS, ORA_FK_BMINUS};
    double *a = malloc(sizeof(double) * g->nr), *b = malloc(sizeof(double) * g->nr),
           *c = malloc(sizeof(double) * g->nr);
    for (int t = 0; t < 6; t++)
        for (int m = 0; m <= g->max_mode; m++) {
            ora_factor *f = &s->fac[kinds[t]][m];
            f->lo = malloc(sizeof(double) * g->nr);
            f->inv = malloc(sizeof(double) * g->nr);
            f->cp = malloc(sizeof(double) * g->nr);
            ora_matrix(kinds[t], m, g->nr, g->dr, g->relax_fac, g->bnd, a, b, c);
            ora_thomas_factor(g->nr, a, b, c, f->lo, f->inv, f->cp);
        }
    free(a); free(b); free(c);
    return s;
}

void ora_destroy(ora_sim *s)
{
    if (!s) return;
    for (int f = 0; f < ORA_NFIELD; f++) free(s->f[f]);
    free(s->h_psi); free(s->h_e); free(s->h_b); free(s->h_b_spe); free(s->h_e_spe);
    free(s->h_cu); free(s->h_q_spe); free(s->h_q_beam);
    free(s->conv_re); free(s->conv_im);
    free(s->lz_ar); free(s->lz_ai); free(s->ar); free(s->ai); free(s->ar_grad); free(s->ai_grad);
    free(s->chi); free(s->h_chi);
    for (int t = 0; t < 6; t++)
        for (int m = 0; m <= s->g.max_mode; m++) {
            free(s->fac[t][m].lo); free(s->fac[t][m].inv); free(s->fac[t][m].cp);
        }
    for (int i = 0; i < s->nspecies; i++) {
        ora_species *sp = &s->sp[i];
        free(sp->x1); free(sp->x2); free(sp->p1); free(sp->p2); free(sp->p3);
        free(sp->q); free(sp->gamma); free(sp->psi); free(sp->id);
        free(sp->q_own); free(sp->qn); free(sp->cu_own); free(sp->acu_own); free(sp->amu_own);
        free(sp->q_hist);
        free(sp->multi_ion); free(sp->ion_old); free(sp->ppc_add); free(sp->den_perp);
        free(sp->add_x1); free(sp->add_x2); free(sp->add_q); free(sp->rho_hist);
    }
    for (int i = 0; i < s->nbeams; i++) {
        ora_beam *bm = &s->bm[i];
        free(bm->x1); free(bm->x2); free(bm->x3); free(bm->p1); free(bm->p2); free(bm->p3);
        free(bm->q); free(bm->id); free(bm->q_own_hist);
    }
    free(s);
}

int ora_add_species(ora_sim *s, const ora_species_cfg *cfg)
{
    if (s->nspecies >= 8) return -1;
    int isp = s->nspecies++;
    ora_species *sp = &s->sp[isp];
    memset(sp, 0, sizeof(*sp));
    sp->cfg = *cfg;
    /* a PGC species with no laser pushes in a zero envelope (laser_all of
       sim_lasers exists with nlasers = 0) */
    if ((cfg->push_type == ORA_PUSH2_STD_PGC || cfg->push_type == ORA_PUSH2_ROBUST_PGC) && laser_alloc(s))
        return -1;
    /* npmax: fdist2d_class.f03:259-262 (xtra = 2) */
    sp->npmax = (int64_t)s->g.nr * cfg->ppc1 * cfg->ppc2 * cfg->num_theta * 2;
    int64_t n = sp->npmax;
    sp->x1 = calloc(n, sizeof(double)); sp->x2 = calloc(n, sizeof(double));
    sp->p1 = calloc(n, sizeof(double)); sp->p2 = calloc(n, sizeof(double));
    sp->p3 = calloc(n, sizeof(double)); sp->q = calloc(n, sizeof(double));
    sp->gamma = calloc(n, sizeof(double)); sp->psi = calloc(n, sizeof(double));
    sp->id = calloc(n, sizeof(int32_t));
    sp->qm = cfg->qm;
    sp->qbm = cfg->qm / cfg->mass;           /* sim_plasma_class.f03:120-122 */
    sp->dt = s->g.dxi;
    size_t n1 = (size_t)s->nm * s->ld;
    sp->q_own = calloc(n1, sizeof(double));
    sp->qn = calloc(n1, sizeof(double));
    sp->cu_own = calloc(3 * n1, sizeof(double));
    sp->acu_own = calloc(2 * n1, sizeof(double));
    sp->amu_own = calloc(3 * n1, sizeof(double));
    sp->q_hist = calloc(((size_t)s->g.nxi + 2) * n1, sizeof(double));
    ora_species_inject(s, isp);
    ora_species_init_q(s, isp);
    return isp;
}

/* ------------------------------------------------------------------------ */
/* Neutral gas and field ionisation (neutral_class.f03)                      */
/* ------------------------------------------------------------------------ */

/* ADK coefficients per level, w = C1 E^-C3 exp(-C2 / E) with E in GV/m and
   w in 1/s (neutral_class.f03:23-299, the elements of atomic number 1, 2, 3,
   6, 7, 8, 10, 18).  Each row is the ADK rate (l = 0, m = 0) of charge
   state i with its ionisation energy U: n* = i sqrt(U_H / U), C3 = 2n* - 1,
   C2 = (2/3) (U/U_H)^(3/2) E_a, C1 = w_a 2^(2n*) / (n* Gamma(2n*)) (U/2U_H)
   (2 (U/U_H)^(3/2) E_a)^(2n* - 1). */
static const double ADK_H[1][3] = {{8.522542995398661e19, 342.53947239007687, 1.0005337056631487}};
static const double ADK_He[2][3] = {{7.2207661763501e18, 832.809878216992, 0.48776427204592254},
                                    {2.7226733893691e21, 2742.1316798375965, 1.0000920088118899}};
static const double ADK_Li[3][3] = {{3.460272990838495e21, 85.51998980232813, 2.1770706138013733},
                                    {3.6365138642921554e20, 4493.713340713575, 0.6964625952167312},
                                    {2.0659396971422902e22, 9256.32561931876, 0.9999745128918196}};
static const double ADK_C[6][3] = {{1.8809310466441196e20, 258.1083104918942, 1.1984440550739555},
                                   {5.509352426405289e23, 822.523017116882, 1.987881643259267},
                                   {4.572825176241475e25, 2263.6763526780164, 2.198152910400111},
                                   {9.835564484116034e27, 3537.9525592618193, 2.674447703635764},
                                   {7.536123777733801e22, 53034.28979013096, 0.8628040191262976},
                                   {6.587972799911127e23, 74090.46603762524, 0.9996157827449552}};
static const double ADK_N[7][3] = {{6.441568477283914e19, 378.4956025896395, 0.9350660865589433},
                                   {1.4701035287257558e23, 1100.1258100528623, 1.711847679118431},
                                   {4.963404971976876e25, 2232.374603536868, 2.2130314611564854},
                                   {1.429510866064162e27, 4658.081063546513, 2.3525381223431663},
                                   {1.3012849307544791e29, 6615.849778853221, 2.7281285044063783},
                                   {2.1409091975134166e23, 88606.4475724123, 0.8838466900060211},
                                   {1.4212916247536375e24, 117682.27130021283, 0.9994495038042972}};
static const double ADK_O[8][3] = {{8.471015773053202e19, 343.2810702373722, 0.9990920676510977},
                                   {4.659951179950614e22, 1421.770921017062, 1.4896379815554006},
                                   {1.375943467286961e25, 2781.3610873184325, 1.985966133301126},
                                   {1.4410401568574448e27, 4652.670876621719, 2.353837077475246},
                                   {2.1812887948673846e28, 8303.41171218499, 2.4562134926799617},
                                   {1.0859737168339012e30, 11088.09514588445, 2.766300924086769},
                                   {5.135265231597041e23, 137319.4144041063, 0.8991975448040144},
                                   {2.7649244731504574e24, 175715.8649323391, 0.999259077262769}};
static const double ADK_Ne[10][3] = {{1.2380498982127483e19, 684.0495717312216, 0.5886207199036224},
                                     {1.6861743227006676e22, 1790.87087951275, 1.3052851455522987},
                                     {4.012462791033353e24, 3450.2505445538854, 1.7789909006599105},
                                     {1.4251452589514637e26, 6544.999998027671, 1.9932260907792414},
                                     {6.670617838240815e27, 9689.667572133209, 2.282840577038655},
                                     {1.9391656057538128e29, 13557.847041347119, 2.522116807457813},
                                     {1.3881927947789378e30, 20383.797283570835, 2.586898525643846},
                                     {3.1075134088626144e31, 25254.466329809107, 2.8167466774677434},
                                     {2.1020658703689154e24, 282468.0585535283, 0.9200040384519763},
                                     {8.391062111505151e24, 343429.9456106982, 0.9988031600227598}};
static const double ADK_Ar[18][3] = {{4.583155091154178e19, 427.3616288942586, 0.858307468844768},
                                     {2.347069522730368e23, 992.0665907024354, 1.8069357289301329},
                                     {1.931637188507353e26, 1775.9423286421918, 2.4675897958480384},
                                     {2.3004409750295724e28, 3141.43462246603, 2.8229627269683206},
                                     {3.468592152886133e30, 4422.602894394151, 3.263766733070865},
                                     {2.965675081118506e32, 5958.158993507227, 3.6326551045811497},
                                     {2.1249901067576396e33, 9479.178143644607, 3.629746649292696},
                                     {8.991061313075681e34, 11737.03857874398, 3.92742350307885},
                                     {7.611312705906995e29, 59342.85460567607, 2.229751529947494},
                                     {6.602179339057654e30, 71781.6325368281, 2.3680482459080237},
                                     {4.815600903407338e31, 85819.59318777094, 2.490706079578982},
                                     {2.055113377686043e32, 105186.65111959413, 2.558310025830369},
                                     {1.1945429165015238e33, 122591.38628416909, 2.663021333332675},
                                     {6.212435859768618e33, 141745.69057711778, 2.7584389081812857},
                                     {1.6060852787486324e34, 170916.68217921766, 2.783373249974532},
                                     {9.327897444360857e34, 190110.7378894357, 2.894937691775023},
                                     {6.6878740700780965e25, 1.8068758775178685e6, 0.9536895901987263},
                                     {1.5247916822996254e26, 2.0115330349032478e6, 0.9959340925415332}};

const double *ora_adk(int element, int level)
{
    const double(*t)[3] = NULL;
    int n = 0;
    switch (element) {
    case 1: t = ADK_H; n = 1; break;
    case 2: t = ADK_He; n = 2; break;
    case 3: t = ADK_Li; n = 3; break;
    case 6: t = ADK_C; n = 6; break;
    case 7: t = ADK_N; n = 7; break;
    case 8: t = ADK_O; n = 8; break;
    case 10: t = ADK_Ne; n = 10; break;
    case 18: t = ADK_Ar; n = 18; break;
    default: return NULL;
    }
    return (level >= 1 && level <= n) ? t[level - 1] : NULL;
}

/* renew_neutral (neutral_class.f03:839-878): no electrons, every cell back
   to the neutral gas, no ion charge */
void ora_neutral_renew(const ora_sim *s, ora_species *sp)
{
    const int L = sp->multi_max, nc = sp->cfg.num_theta * s->g.nr;
    sp->npp = 0;
    sp->n_add = 0;
    sp->n_added = 0;
    memset(sp->multi_ion, 0, (size_t)nc * (L + 2) * sizeof(double));
    for (int c = 0; c < nc; c++) sp->multi_ion[(size_t)c * (L + 2) + L] = 1.0;
    memset(sp->ion_old, 0, (size_t)nc * sizeof(double));
    memset(sp->ppc_add, 0, (size_t)nc * sizeof(int));
    memset(sp->qn, 0, (size_t)s->nm * s->ld * sizeof(double));
}

/* init_neutral (neutral_class.f03:408-574, sim_plasma_class.f03:148-176):
   an electron species with no particles yet, qbm = q/m, the ADK rows of
   the first multi_max = ion_max levels, wp = sqrt(n0) 5.641460231180626e4.
   QPAD sets the level indices from ion_max before clipping it to the
   element's levels; ion_max beyond them is refused here. */
int ora_add_neutral(ora_sim *s, const ora_neutral_cfg *c)
{
    if (s->nspecies >= 8 || c->ion_max < 1 || !ora_adk(c->element, c->ion_max)) return -1;
    if (c->density * 1.0 * 1.0 < c->den_min) return -1;
    ora_species_cfg sc;
    memset(&sc, 0, sizeof(sc));
    sc.ppc1 = c->ppc1;
    sc.ppc2 = c->ppc2;
    sc.num_theta = c->num_theta;
    sc.qm = c->qm;
    sc.mass = c->mass;
    sc.density = c->density;
    sc.neutralized = 0;
    sc.den_min = c->den_min;
    sc.push_type = c->push_type;
    const int isp = s->nspecies++;
    ora_species *sp = &s->sp[isp];
    memset(sp, 0, sizeof(*sp));
    sp->cfg = sc;
    sp->is_neutral = 1;
    sp->multi_max = c->ion_max;
    sp->rate = ora_adk(c->element, 1);
    sp->wp = sqrt(c->n0) * 5.641460231180626e4;
    sp->npmax = (int64_t)s->g.nr * c->ppc1 * c->ppc2 * c->num_theta * 2;
    const int64_t n = sp->npmax;
    sp->x1 = calloc(n, sizeof(double)); sp->x2 = calloc(n, sizeof(double));
    sp->p1 = calloc(n, sizeof(double)); sp->p2 = calloc(n, sizeof(double));
    sp->p3 = calloc(n, sizeof(double)); sp->q = calloc(n, sizeof(double));
    sp->gamma = calloc(n, sizeof(double)); sp->psi = calloc(n, sizeof(double));
    sp->id = calloc(n, sizeof(int32_t));
    sp->add_x1 = calloc(n, sizeof(double)); sp->add_x2 = calloc(n, sizeof(double));
    sp->add_q = calloc(n, sizeof(double));
    sp->qm = c->qm;
    sp->qbm = c->qm / c->mass;
    sp->dt = s->g.dxi;
    const size_t n1 = (size_t)s->nm * s->ld, nc = (size_t)c->num_theta * s->g.nr;
    sp->q_own = calloc(n1, sizeof(double));
    sp->qn = calloc(n1, sizeof(double));
    sp->cu_own = calloc(3 * n1, sizeof(double));
    sp->acu_own = calloc(2 * n1, sizeof(double));
    sp->amu_own = calloc(3 * n1, sizeof(double));
    sp->q_hist = calloc(((size_t)s->g.nxi + 2) * n1, sizeof(double));
    sp->rho_hist = calloc(((size_t)s->g.nxi + 2) * n1, sizeof(double));
    sp->multi_ion = calloc(nc * (c->ion_max + 2), sizeof(double));
    sp->ion_old = calloc(nc, sizeof(double));
    sp->ppc_add = calloc(nc, sizeof(int));
    ora_neutral_renew(s, sp);
    return isp;
}

/* ionize_neutral (neutral_class.f03:600-753): per theta bin k and radial
   cell j, |E| at the cell centre (the mean of nodes j, j+1, modes summed
   at theta_k; for m > 0 QPAD reads the real part where it means the
   imaginary one, e_im => e%rf_re(m): restated as is), in GV/m via
   wp 1.708e-12; where it exceeds 1e-6 and the cell is not fully ionised,
   the ADK rate of every level, the second-order (as QPAD notes, not quite
   right) update of the level fractions with the overshoot rules, and the
   total ion charge rounded to the particles per cell it stands for */
void ora_neutral_ionize(const ora_sim *s, ora_species *sp, const double *ef)
{
    const int nr = s->g.nr, nm = s->nm, ld = s->ld, M = s->g.max_mode;
    const int nth = sp->cfg.num_theta, L = sp->multi_max, ppc_tot = sp->cfg.ppc1 * sp->cfg.ppc2;
    const int in = L, ii = L + 1;   /* neutral residue, total ion charge */
    const double pi2_ntheta = 2.0 * PI_DB / (double)nth, dt = sp->dt, wp = sp->wp;
    double w_ion[20];
    for (int k = 1; k <= nth; k++) {
        const double theta = pi2_ntheta * (double)(k - 1);
        const double cr = cos(theta), ci = sin(theta);
        for (int j = 1; j <= nr; j++) {
            double e1 = 0.5 * (FA(ef, 0, 0, j) + FA(ef, 0, 0, j + 1));
            double e2 = 0.5 * (FA(ef, 1, 0, j) + FA(ef, 1, 0, j + 1));
            double e3 = 0.5 * (FA(ef, 2, 0, j) + FA(ef, 2, 0, j + 1));
            double phr = 1.0, phi = 0.0;
            for (int m = 1; m <= M; m++) {
                const double a = phr * cr - phi * ci, b = phr * ci + phi * cr;
                phr = a;
                phi = b;
                const int kr = kre(m), kq = kre(m);   /* the quirk: the real part twice */
                e1 = e1 + (FA(ef, 0, kr, j) + FA(ef, 0, kr, j + 1)) * phr - (FA(ef, 0, kq, j) + FA(ef, 0, kq, j + 1)) * phi;
                e2 = e2 + (FA(ef, 1, kr, j) + FA(ef, 1, kr, j + 1)) * phr - (FA(ef, 1, kq, j) + FA(ef, 1, kq, j + 1)) * phi;
                e3 = e3 + (FA(ef, 2, kr, j) + FA(ef, 2, kr, j + 1)) * phr - (FA(ef, 2, kq, j) + FA(ef, 2, kq, j + 1)) * phi;
            }
            (void)nm; (void)ld;
            const double eij = sqrt(e1 * e1 + e2 * e2 + e3 * e3) * wp * 1.708e-12;
            double *mi = sp->multi_ion + ((size_t)(k - 1) * nr + (j - 1)) * (L + 2);
            if (!(eij > 1.0e-6)) continue;
            if (!(mi[ii] < (double)L)) continue;
            int shoot = 0;
            for (int i = 0; i < L; i++) {
                const double *r = sp->rate + 3 * i;
                w_ion[i] = r[0] * pow(eij, -r[2]) * exp(-r[1] / eij) / wp;
            }
            double cons = mi[in] * w_ion[0] * dt * (1.0 + 0.5 * w_ion[0] * dt);
            if (cons > mi[in]) {
                shoot = 1;
                cons = mi[in];
            }
            mi[in] = mi[in] - cons;
            for (int i = 1; i <= L - 1; i++) {
                const double inj = cons;
                double dens_temp = 0.0;
                if (shoot) {
                    dens_temp = inj * 0.5;
                    shoot = 0;
                }
                cons = (mi[i - 1] + dens_temp) * w_ion[i] * dt * (1.0 + 0.5 * w_ion[i] * dt);
                if (cons > mi[i - 1] + dens_temp) {
                    shoot = 1;
                    cons = mi[i - 1] + dens_temp;
                }
                mi[i - 1] = fmin(mi[i - 1] - cons + inj, 1.0);
            }
            mi[L - 1] = fmin(mi[L - 1] + cons, 1.0);
            mi[ii] = 0.0;
            for (int i = 1; i <= L; i++) mi[ii] = mi[ii] + (double)i * mi[i - 1];
            mi[ii] = (double)L / (double)ppc_tot * (double)(int)(mi[ii] * (double)ppc_tot / (double)L + 0.5);
        }
    }
}

/* add_particles (neutral_class.f03:755-837): per theta bin k and radial
   cell j, as many electrons as the rounded ion charge grew by, at rest on
   an even radial lattice of the cell at theta_k, charge rn n coef,
   coef = multi_max sign(q) / (ppc nth); part_add keeps their positions
   with the opposite charge for the ion deposit.  Ids count the electrons
   released since the renew. */
void ora_neutral_add(const ora_sim *s, ora_species *sp)
{
    const int nr = s->g.nr, nth = sp->cfg.num_theta, L = sp->multi_max;
    const int ppc_tot = sp->cfg.ppc1 * sp->cfg.ppc2;
    const double dr = s->g.dr, dtheta = 2.0 * PI_DB / (double)nth;
    const double den_lon = 1.0, den_perp = 1.0;
    const double coef = (double)L * copysign(1.0, sp->cfg.qm) / ((double)ppc_tot * (double)nth);
    sp->n_add = 0;
    for (int k = 1; k <= nth; k++)
        for (int j = 1; j <= nr; j++) {
            const size_t c = (size_t)(k - 1) * nr + (j - 1);
            const double *mi = sp->multi_ion + c * (L + 2);
            const int ppc_add = (int)((mi[L + 1] - sp->ion_old[c]) / (double)L * (double)ppc_tot + 0.5);
            sp->ppc_add[c] = ppc_add;
            for (int i = 1; i <= ppc_add; i++) {
                const double rn = (double)(j - 1) + ((double)i - 0.5) / (double)ppc_add;
                const double theta = (double)(k - 1) * dtheta;
                const double x1 = rn * dr * cos(theta), x2 = rn * dr * sin(theta);
                const int64_t p = sp->npp++;
                sp->x1[p] = x1;
                sp->x2[p] = x2;
                sp->q[p] = rn * den_lon * den_perp * sp->cfg.density * coef;
                sp->p1[p] = 0.0;
                sp->p2[p] = 0.0;
                sp->p3[p] = 0.0;
                sp->gamma[p] = 1.0;
                sp->psi[p] = 0.0;
                sp->id[p] = (int32_t)(sp->n_added++);
                sp->add_x1[sp->n_add] = x1;
                sp->add_x2[sp->n_add] = x2;
                sp->add_q[sp->n_add] = -sp->q[p];
                sp->n_add++;
            }
        }
}

/* ion_deposit_neutral (neutral_class.f03:904-930): the charge deposit of
   part_add into rho_ion_add, added to rho_ion (kept in qn, which the slice
   adds to q_spe after the electrons' charge); part_add emptied */
void ora_neutral_ion_deposit(const ora_sim *s, ora_species *sp)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    double *add = calloc(n1, sizeof(double));
    ora_species buf;
    memset(&buf, 0, sizeof(buf));
    buf.npp = sp->n_add;
    buf.x1 = sp->add_x1;
    buf.x2 = sp->add_x2;
    buf.q = sp->add_q;
    ora_qdeposit(s, &buf, add);
    for (size_t t = 0; t < n1; t++) sp->qn[t] = sp->qn[t] + add[t];
    free(add);
    sp->n_add = 0;
}

int ora_add_beam(ora_sim *s, const ora_beam_cfg *cfg, int64_t npmax)
{
    if (s->nbeams >= 8) return -1;
    int ib = s->nbeams++;
    ora_beam *bm = &s->bm[ib];
    memset(bm, 0, sizeof(*bm));
    if (npmax <= 0) npmax = cfg->total_num * (cfg->quiet ? 2 : 1);
    bm->npmax = npmax;
    int64_t n = npmax;
    bm->x1 = calloc(n, sizeof(double)); bm->x2 = calloc(n, sizeof(double));
    bm->x3 = calloc(n, sizeof(double)); bm->p1 = calloc(n, sizeof(double));
    bm->p2 = calloc(n, sizeof(double)); bm->p3 = calloc(n, sizeof(double));
    bm->q = calloc(n, sizeof(double)); bm->id = calloc(n, sizeof(int32_t));
    bm->qbm = cfg->qm / cfg->mass;
    bm->dt = s->g.dt;
    bm->evol = cfg->evol;
    bm->q_own_hist = calloc(((size_t)s->g.nxi + 2) * s->nm * s->ld, sizeof(double));
    if (cfg->total_num > 0) ora_beam_inject(s, ib, cfg);
    return ib;
}

int ora_beam_set(ora_sim *s, int ib, int64_t n, const double *x1, const double *x2,
                 const double *x3, const double *p1, const double *p2,
                 const double *p3, const double *q)
{
    ora_beam *bm = &s->bm[ib];
    if (n > bm->npmax) return -1;
    memcpy(bm->x1, x1, n * sizeof(double)); memcpy(bm->x2, x2, n * sizeof(double));
    memcpy(bm->x3, x3, n * sizeof(double)); memcpy(bm->p1, p1, n * sizeof(double));
    memcpy(bm->p2, p2, n * sizeof(double)); memcpy(bm->p3, p3, n * sizeof(double));
    memcpy(bm->q, q, n * sizeof(double));
    for (int64_t i = 0; i < n; i++) bm->id[i] = (int32_t)i;
    bm->npp = n;
    return 0;
}

/* 3D step prologue: source/simulation_class.f03:297-340 */
void ora_step3d_begin(ora_sim *s)
{
    size_t n1 = (size_t)s->nm * s->ld, ns = (size_t)s->g.nxi + 2;
    memset(s->h_q_beam, 0, ns * n1 * sizeof(double));   /* q_beam%as(0.0) */
    memset(s->h_q_spe, 0, ns * n1 * sizeof(double));    /* q_spe%as(0.0)  */
    TSTART();
    for (int ib = 0; ib < s->nbeams; ib++) {            /* beam3d%qdp     */
        ora_beam *bm = &s->bm[ib];
        memset(bm->q_own_hist, 0, ns * n1 * sizeof(double));
        ora_beam_qdeposit(s, ib, bm->q_own_hist);
        for (size_t t = 0; t < ns * n1; t++) s->h_q_beam[t] = s->h_q_beam[t] + bm->q_own_hist[t];
    }
    TSTOP(ORA_T_DEPOSIT_3D);
    int zero[] = {ORA_B, ORA_E, ORA_B_SPE, ORA_E_SPE, ORA_PSI, ORA_CU, ORA_ACU, ORA_AMU};
    for (int t = 0; t < 8; t++) memset(s->f[zero[t]], 0, n1 * s->ncomp[zero[t]] * sizeof(double));
}

static void add_into(double *dst, const double *src, size_t n)
{
    for (size_t t = 0; t < n; t++) dst[t] = dst[t] + src[t];
}

/* The pieces of a slice (source/simulation_class.f03:342-469), shared by
   the standard slice and the sub-cycled one
   (proj_subcyc/simulation_subcyc_class.f03:230-390). */

/* q_spe of every species and neutral (qdp, ion_deposit), psi, interp_psi
   of the plasma species (QPAD's loop leaves out the neutrals, whose
   electrons keep the psi = 0 of their release), the laser of slice j
   (laser != 0), Bz from the J predictor: :346-367 */
static void slice_charge(ora_sim *s, int j, int laser)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    double *psi = s->f[ORA_PSI], *cu = s->f[ORA_CU], *b_spe = s->f[ORA_B_SPE], *q_spe = s->f[ORA_Q_SPE];
    TSTART();
    memset(q_spe, 0, n1 * sizeof(double));
    for (int k = 0; k < s->nspecies; k++) {                              /* species2d%qdp */
        ora_species *sp = &s->sp[k];
        memset(sp->q_own, 0, n1 * sizeof(double));
        ora_qdeposit(s, sp, sp->q_own);
        add_into(q_spe, sp->q_own, n1);
        if (sp->is_neutral) {                                            /* neutral%ion_deposit */
            ora_neutral_ion_deposit(s, sp);
            add_into(q_spe, sp->qn, n1);
        } else if (sp->cfg.neutralized) {
            add_into(q_spe, sp->qn, n1);
        }
    }
    TSTOP(ORA_T_DEPOSIT_2D);
    ora_solve_psi(s, q_spe, psi);
    TSTOP(ORA_T_SOLVE_PSI);
    for (int k = 0; k < s->nspecies; k++) {                              /* species2d%interp_psi */
        const int pt = s->sp[k].cfg.push_type;
        if (s->sp[k].is_neutral) continue;
        if (pt == ORA_PUSH2_STD || pt == ORA_PUSH2_STD_PGC) ora_interp_psi(s, &s->sp[k], psi);
    }
    if (laser && s->laser_on) ora_laser_slice(s, j);                     /* laser_all of slice j */
    ora_solve_bz(s, cu, b_spe);
    TSTOP(ORA_T_SOLVE_BZ);
}

/* the predictor-corrector iteration (:370-398) with the 2D step dt;
   returns the iterations executed */
static int slice_iterate(ora_sim *s, double dt)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    double *psi = s->f[ORA_PSI], *e = s->f[ORA_E];
    double *b = s->f[ORA_B], *b_spe = s->f[ORA_B_SPE], *b_beam = s->f[ORA_B_BEAM];
    double *cu = s->f[ORA_CU], *acu = s->f[ORA_ACU], *amu = s->f[ORA_AMU], *dcu = s->f[ORA_DCU];
    TSTART();
    int l, nit = 0;
    double rel = 0.0, ab = 0.0;
    for (l = 1; l <= s->g.iter_max; l++) {
        nit = l;
        ora_conv_record(s, b_spe);
        for (size_t t = 0; t < 3 * n1; t++) b[t] = b_spe[t] + b_beam[t];
        TSTOP(ORA_T_ARITHMETICS);
        ora_solve_ez(s, cu, e);
        TSTOP(ORA_T_SOLVE_EZ);
        ora_solve_et(s, b, psi, e);
        TSTOP(ORA_T_SOLVE_PLASMA_ET);
        memset(cu, 0, 3 * n1 * sizeof(double));
        memset(acu, 0, 2 * n1 * sizeof(double));
        memset(amu, 0, 3 * n1 * sizeof(double));
        for (int k = 0; k < s->nspecies; k++) {                          /* species2d%amjdp */
            ora_species *sp = &s->sp[k];
            memset(sp->cu_own, 0, 3 * n1 * sizeof(double));
            memset(sp->acu_own, 0, 2 * n1 * sizeof(double));
            memset(sp->amu_own, 0, 3 * n1 * sizeof(double));
            const int pt = sp->cfg.push_type;
            if (pt == ORA_PUSH2_STD)
                ora_amjdeposit_std(s, sp, e, b, sp->cu_own, sp->amu_own, sp->acu_own, dt);
            else if (pt == ORA_PUSH2_STD_PGC || pt == ORA_PUSH2_ROBUST_PGC)
                ora_amjdeposit_pgc(s, sp, e, b, sp->cu_own, sp->amu_own, sp->acu_own, dt,
                                   pt == ORA_PUSH2_STD_PGC);
            else
                ora_amjdeposit_robust(s, sp, e, b, sp->cu_own, sp->amu_own, sp->acu_own, dt);
            add_into(cu, sp->cu_own, 3 * n1);
            add_into(acu, sp->acu_own, 2 * n1);
            add_into(amu, sp->amu_own, 3 * n1);
        }
        TSTOP(ORA_T_DEPOSIT_2D);
        ora_solve_djdxi(s, acu, amu, dcu);
        TSTOP(ORA_T_SET_SOURCE);
        ora_solve_bt_iter(s, dcu, cu, b_spe);
        TSTOP(ORA_T_SOLVE_PLASMA_BT);
        ora_solve_bz(s, cu, b_spe);
        TSTOP(ORA_T_SOLVE_BZ);
        ora_conv_compare(s, b_spe, &rel, &ab);
        TSTOP(ORA_T_ARITHMETICS);
        if (rel < s->g.iter_reltol || ab < s->g.iter_abstol) break;
    }
    s->last_rel = rel;
    s->last_abs = ab;
    return nit;
}

/* chi of the plasma species, cbq, the cu and q_spe histories (:400-412) */
static void slice_chi_cbq(ora_sim *s, int j)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    double *cu = s->f[ORA_CU], *q_spe = s->f[ORA_Q_SPE];
    TSTART();
    if (s->laser_on) {                                                   /* sim_lasers%deposit_chi */
        memset(s->chi, 0, n1 * sizeof(double));
        for (int k = 0; k < s->nspecies; k++)
            if (!s->sp[k].is_neutral) ora_deposit_chi(s, &s->sp[k], s->chi);
        memcpy(s->h_chi + (size_t)j * n1, s->chi, n1 * sizeof(double));
    }
    for (int k = 0; k < s->nspecies; k++) {                              /* species2d%cbq */
        ora_species *sp = &s->sp[k];
        for (size_t t = 0; t < n1; t++) sp->q_own[t] = sp->q_own[t] + sp->cu_own[2 * n1 + t];
        memcpy(sp->q_hist + (size_t)j * n1, sp->q_own, n1 * sizeof(double));
        if (sp->is_neutral) memcpy(sp->rho_hist + (size_t)j * n1, sp->qn, n1 * sizeof(double));
    }
    memcpy(s->h_cu + (size_t)j * 3 * n1, cu, 3 * n1 * sizeof(double));
    for (size_t t = 0; t < n1; t++) q_spe[t] = q_spe[t] + cu[2 * n1 + t];
    memcpy(s->h_q_spe + (size_t)j * n1, q_spe, n1 * sizeof(double));
    TSTOP(ORA_T_COPY_SLICES);
}

/* b = b_spe + b_beam, E of the plasma alone, E of the converged currents
   (:413-416) */
static void slice_final_fields(ora_sim *s)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    double *psi = s->f[ORA_PSI], *e = s->f[ORA_E], *e_spe = s->f[ORA_E_SPE];
    double *b = s->f[ORA_B], *b_spe = s->f[ORA_B_SPE], *b_beam = s->f[ORA_B_BEAM], *cu = s->f[ORA_CU];
    TSTART();
    for (size_t t = 0; t < 3 * n1; t++) b[t] = b_spe[t] + b_beam[t];
    TSTOP(ORA_T_COPY_SLICES);
    ora_solve_et(s, b_spe, psi, e_spe);
    TSTOP(ORA_T_SOLVE_PLASMA_ET);
    ora_solve_ez(s, cu, e);
    TSTOP(ORA_T_SOLVE_EZ);
    ora_solve_et(s, b, psi, e);
    TSTOP(ORA_T_SOLVE_PLASMA_ET);
}

/* the J_perp predictor of the next slice: dcu *= dxi, cu += dcu (:425-426) */
static void slice_predictor(ora_sim *s)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    double *cu = s->f[ORA_CU], *dcu = s->f[ORA_DCU];
    TSTART();
    for (size_t t = 0; t < 2 * n1; t++) dcu[t] = s->g.dxi * dcu[t];      /* dot_f1 */
    for (size_t t = 0; t < 2 * n1; t++) cu[t] = cu[t] + dcu[t];          /* J_perp predictor */
    TSTOP(ORA_T_ARITHMETICS);
}

/* the neutrals' update, then push_u, (clamp > 0) clamp_exp_fac, push_x and
   update_bound of every species with the 2D step dt (:436-449; the
   sub-cycled form :337-364) */
static void slice_push(ora_sim *s, double dt, double clamp)
{
    double *e = s->f[ORA_E], *b = s->f[ORA_B];
    TSTART();
    for (int k = 0; k < s->nspecies; k++) {
        ora_species *sp = &s->sp[k];
        if (sp->is_neutral) {                                            /* neutral%update */
            const int nc = sp->cfg.num_theta * s->g.nr, L = sp->multi_max;
            for (int c = 0; c < nc; c++) sp->ion_old[c] = sp->multi_ion[(size_t)c * (L + 2) + L + 1];
            ora_neutral_ionize(s, sp, e);
            ora_neutral_add(s, sp);
        }
        const int pt = sp->cfg.push_type;
        if (pt == ORA_PUSH2_STD) ora_push_u_std(s, sp, e, b, dt);
        else if (pt == ORA_PUSH2_STD_PGC || pt == ORA_PUSH2_ROBUST_PGC) ora_push_u_pgc(s, sp, e, b, dt);
        else ora_push_u_robust(s, sp, e, b, dt);
        if (clamp > 0.0) ora_clamp_exp_fac(sp, clamp);
        ora_push_x(sp, dt);
        TSTOP(ORA_T_PUSH_2D);
        ora_update_bound(s, sp);
        TSTOP(ORA_T_MOVE_2D);
    }
}

/* the e, b, psi, b_spe, e_spe histories (:452-456) */
static void slice_copies(ora_sim *s, int j)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    TSTART();
    memcpy(s->h_e + (size_t)j * 3 * n1, s->f[ORA_E], 3 * n1 * sizeof(double));
    memcpy(s->h_b + (size_t)j * 3 * n1, s->f[ORA_B], 3 * n1 * sizeof(double));
    memcpy(s->h_psi + (size_t)j * n1, s->f[ORA_PSI], n1 * sizeof(double));
    memcpy(s->h_b_spe + (size_t)j * 3 * n1, s->f[ORA_B_SPE], 3 * n1 * sizeof(double));
    memcpy(s->h_e_spe + (size_t)j * 3 * n1, s->f[ORA_E_SPE], 3 * n1 * sizeof(double));
    TSTOP(ORA_T_COPY_SLICES);
}

/* get_exp_fac_max of every species and neutral
   (proj_subcyc/sim_plasma_subcyc_class.f03:58-86, part2d_subcyc_class.f03:
   27-44): the largest gamma / (gamma - p_z) of any particle, at least 1 */
double ora_exp_fac_max(const ora_sim *s)
{
    double m = 1.0;
    for (int k = 0; k < s->nspecies; k++) {
        const ora_species *sp = &s->sp[k];
        for (int64_t p = 0; p < sp->npp; p++) {
            const double f = sp->gamma[p] / (sp->gamma[p] - sp->p3[p]);
            if (f > m) m = f;
        }
    }
    return m;
}

/* get_subcyc_step (proj_subcyc/simulation_subcyc_class.f03:431-449): the
   sub-cycles of a slice and their 2D step */
int ora_subcyc_step(double exp_fac, double exp_fac_max, double dt, double dt_min, double *dt_sub)
{
    int n = 1;
    double d = dt;
    if (exp_fac > exp_fac_max) {
        n = (int)ceil(exp_fac / exp_fac_max);
        d = dt / (double)n;
        if (d < dt_min) {
            n = (int)floor(dt / dt_min);
            d = dt / (double)n;
        }
    }
    *dt_sub = d;
    return n;
}

/* clamp_exp_fac (proj_subcyc/part2d_subcyc_class.f03:46-66): a particle
   whose gamma / (gamma - p_z) exceeds f has its momentum scaled so that the
   ratio becomes f, and its gamma recomputed */
void ora_clamp_exp_fac(ora_species *sp, double f)
{
    for (int64_t p = 0; p < sp->npp; p++) {
        const double g = sp->gamma[p], pz = sp->p3[p];
        const double ef = g / (g - pz);
        if (ef > f) {
            double sc = (f - 1.0) * (f - 1.0);
            sc = sc / (f * f * pz * pz - sc * (g * g - 1.0));
            sc = sqrt(sc);
            sp->p1[p] = sp->p1[p] * sc;
            sp->p2[p] = sp->p2[p] * sc;
            sp->p3[p] = sp->p3[p] * sc;
            sp->gamma[p] = sqrt(1.0 + sp->p1[p] * sp->p1[p] + sp->p2[p] * sp->p2[p] + sp->p3[p] * sp->p3[p]);
        }
    }
}

void ora_set_subcycling(ora_sim *s, int on, double exp_fac_max, double exp_fac_clamped, double dt_2d_min)
{
    s->subcyc = on;
    s->sc_fac_max = exp_fac_max;
    s->sc_fac_clamped = exp_fac_clamped;
    s->sc_dt_min = dt_2d_min;
}

void ora_last_subcycles(const ora_sim *s, int *n, double *dt_sub, double *exp_fac)
{
    *n = s->last_nsub;
    *dt_sub = s->last_dt_sub;
    *exp_fac = s->last_exp_fac;
}

/* One xi-slice: source/simulation_class.f03:342-469, or with sub-cycling
   on, proj_subcyc/simulation_subcyc_class.f03:230-390: the slice's plasma
   response and push repeated n times with the 2D step dxi / n (n from the
   largest gamma / (gamma - p_z) at the slice start), each push followed by
   clamp_exp_fac, the neutrals ionised in every sub-cycle with their own
   step dxi (neutral%dt); chi, cbq, the histories and the J_perp predictor
   (with dxi) once after the last.  n_iter: the iterations summed over the
   sub-cycles. */
void ora_slice(ora_sim *s, int j, int *n_iter)
{
    const size_t n1 = (size_t)s->nm * s->ld;
    TSTART();
    memcpy(s->f[ORA_Q_BEAM], s->h_q_beam + (size_t)j * n1, n1 * sizeof(double));  /* copy_slice 2to1 */
    TSTOP(ORA_T_COPY_SLICES);
    ora_solve_bt_beam(s, s->f[ORA_Q_BEAM], s->f[ORA_B_BEAM]);
    TSTOP(ORA_T_SOLVE_BEAM_BT);
    int nit = 0;
    if (!s->subcyc) {
        slice_charge(s, j, 1);
        nit = slice_iterate(s, s->g.dxi);
        slice_chi_cbq(s, j);
        slice_final_fields(s);
        slice_predictor(s);
        slice_push(s, s->g.dxi, 0.0);
        s->last_nsub = 1;
        s->last_dt_sub = s->g.dxi;
        s->last_exp_fac = 0.0;
    } else {
        if (s->laser_on) ora_laser_slice(s, j);
        const double ef = ora_exp_fac_max(s);
        double dsub;
        const int n = ora_subcyc_step(ef, s->sc_fac_max, s->g.dxi, s->sc_dt_min, &dsub);
        for (int i = 1; i <= n; i++) {
            slice_charge(s, j, 0);
            nit += slice_iterate(s, dsub);
            slice_final_fields(s);
            slice_push(s, dsub, s->sc_fac_clamped);
        }
        slice_chi_cbq(s, j);
        slice_predictor(s);
        s->last_nsub = n;
        s->last_dt_sub = dsub;
        s->last_exp_fac = ef;
    }
    s->last_iters = nit;
    if (n_iter) *n_iter = nit;
    slice_copies(s, j);
}

/* 3D step epilogue: source/simulation_class.f03:485-501 (beam push, species
   renew) */
void ora_step3d_end(ora_sim *s, int istep)
{
    (void)istep;
    TSTART();
    for (int ib = 0; ib < s->nbeams; ib++) ora_beam_push(s, ib);
    TSTOP(ORA_T_PUSH_3D);
    for (int k = 0; k < s->nspecies; k++) {
        if (s->sp[k].is_neutral) {
            ora_neutral_renew(s, &s->sp[k]);
            continue;
        }
        ora_species_inject(s, k);
        ora_species_init_q(s, k);
    }
}

void ora_run(ora_sim *s, int nstep3d, int nslices)
{
    if (nslices <= 0 || nslices > s->g.nxi) nslices = s->g.nxi;
    for (int i = 1; i <= nstep3d; i++) {
        ora_step3d_begin(s);
        for (int j = 1; j <= nslices; j++) ora_slice(s, j, NULL);
        ora_step3d_end(s, i);
    }
}

double *ora_field(ora_sim *s, int fid) { return (fid >= 0 && fid < ORA_NFIELD) ? s->f[fid] : NULL; }

double *ora_laser_array(ora_sim *s, int which)
{
    switch (which) {
    case 0: return s->ar;
    case 1: return s->ai;
    case 2: return s->ar_grad;
    case 3: return s->ai_grad;
    case 4: return s->chi;
    case 5: return s->h_chi;
    case 6: return s->lz_ar;
    case 7: return s->lz_ai;
    default: return NULL;
    }
}

double *ora_hist(ora_sim *s, int fid)
{
    switch (fid) {
    case ORA_PSI: return s->h_psi;
    case ORA_E: return s->h_e;
    case ORA_B: return s->h_b;
    case ORA_B_SPE: return s->h_b_spe;
    case ORA_E_SPE: return s->h_e_spe;
    case ORA_CU: return s->h_cu;
    case ORA_Q_SPE: return s->h_q_spe;
    case ORA_Q_BEAM: return s->h_q_beam;
    default: return NULL;
    }
}

int ora_nm(const ora_sim *s) { return s->nm; }
int ora_ld(const ora_sim *s) { return s->ld; }
ora_species *ora_get_species(ora_sim *s, int isp) { return &s->sp[isp]; }
ora_beam *ora_get_beam(ora_sim *s, int ib) { return &s->bm[ib]; }
int ora_last_iters(const ora_sim *s) { return s->last_iters; }
void ora_last_residuals(const ora_sim *s, double *rel, double *abs_)
{
    *rel = s->last_rel;
    *abs_ = s->last_abs;
}

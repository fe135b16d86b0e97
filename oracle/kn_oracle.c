/* kn_oracle.c -- CPU restatement of the kspace-neutrinos hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function follows the reference file:line given in its comment.  GSL
 * (absent from this image) is restated from its published algorithms:
 * QUADPACK QAG (Piessens et al. 1983) with the 61-point Gauss-Kronrod rule
 * whose constants oracle/gen_quad_tables.py derives with mpmath, GSL's
 * natural cubic spline (tridiagonal L D L^T solve) and linear interpolation,
 * and the small-argument Taylor branch of the spherical Bessel j0.
 *
 * Parity pins (tests/test_oracle_*.py): the reference's own known answers
 * (powerspectrum_test.c, delta_pow_test.c, delta_tot_table_test.c,
 * omega_nu_single_test.c, transfer_init_test.c), closed forms, brute force
 * and mpmath.  Build: make -C oracle   (strict IEEE: -ffp-contract=off).
 */
#include "kn_oracle.h"
#include "quad_tables.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* kspace_neutrino_const.h:9-19, gadget_defines.h:8, omega_nu_single.c:10-14 */
#define ORC_LIGHTCGS 2.99792458e10
#define ORC_BOLEVK 8.61734e-5
#define ORC_FLOAT_ACC 1e-6
#define ORC_GSL_VAL 200
#define ORC_HUBBLE 3.24077929e-18
#define ORC_HBAR 6.582119e-16
#define ORC_STEFAN_BOLTZMANN 5.670373e-5
#define ORC_GRAVITY 6.67408e-8
#define ORC_NRHOTAB 200
#define ORC_NU_SW 100
#define ORC_ZETA3 1.202056903159594

static double orc_tnucmb(void) { return pow(4 / 11., 1 / 3.) * 1.00328; } /* omega_nu_single.h:17 */

/* ======================= QAG / qk61 (GSL qag.c, qk.c semantics) ======================= */

static double rescale_error(double err, double result_abs, double result_asc)
{
    err = fabs(err);
    if (result_asc != 0 && err != 0) {
        double scale = pow((200 * err / result_asc), 1.5);
        err = (scale < 1) ? result_asc * scale : result_asc;
    }
    if (result_abs > DBL_MIN / (50 * DBL_EPSILON)) {
        double min_err = 50 * DBL_EPSILON * result_abs;
        if (min_err > err) err = min_err;
    }
    return err;
}

void orc_qk61(orc_fn f, void *p, double a, double b, double *result, double *abserr,
              double *resabs, double *resasc)
{
    const int n = 31;
    double fv1[31], fv2[31];
    const double center = 0.5 * (a + b);
    const double half_length = 0.5 * (b - a);
    const double abs_half_length = fabs(half_length);
    const double f_center = f(center, p);
    double result_gauss = 0;
    double result_kronrod = f_center * orc_wgk61[n - 1];
    double result_abs = fabs(result_kronrod);
    double result_asc, mean, err;
    int j;
    for (j = 0; j < (n - 1) / 2; j++) {        /* Gauss nodes: odd indices */
        const int jtw = 2 * j + 1;
        const double absc = half_length * orc_xgk61[jtw];
        const double f1 = f(center - absc, p), f2 = f(center + absc, p);
        const double fsum = f1 + f2;
        fv1[jtw] = f1;
        fv2[jtw] = f2;
        result_gauss += orc_wg30[j] * fsum;
        result_kronrod += orc_wgk61[jtw] * fsum;
        result_abs += orc_wgk61[jtw] * (fabs(f1) + fabs(f2));
    }
    for (j = 0; j < n / 2; j++) {              /* Kronrod-only nodes: even indices */
        const int jtwm1 = 2 * j;
        const double absc = half_length * orc_xgk61[jtwm1];
        const double f1 = f(center - absc, p), f2 = f(center + absc, p);
        fv1[jtwm1] = f1;
        fv2[jtwm1] = f2;
        result_kronrod += orc_wgk61[jtwm1] * (f1 + f2);
        result_abs += orc_wgk61[jtwm1] * (fabs(f1) + fabs(f2));
    }
    mean = result_kronrod * 0.5;
    result_asc = orc_wgk61[n - 1] * fabs(f_center - mean);
    for (j = 0; j < n - 1; j++)
        result_asc += orc_wgk61[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    err = (result_kronrod - result_gauss) * half_length;
    result_kronrod *= half_length;
    result_abs *= abs_half_length;
    result_asc *= abs_half_length;
    *result = result_kronrod;
    *resabs = result_abs;
    *resasc = result_asc;
    *abserr = rescale_error(err, result_abs, result_asc);
}

/* Workspace of the adaptive bisection: interval list kept partially sorted by
 * error (QUADPACK dqpsrt). */
typedef struct {
    int limit, size, nrmax, i, maximum_level;
    double alist[ORC_GSL_VAL], blist[ORC_GSL_VAL], rlist[ORC_GSL_VAL], elist[ORC_GSL_VAL];
    int order[ORC_GSL_VAL], level[ORC_GSL_VAL];
} orc_ws;

static void ws_qpsrt(orc_ws *w)
{
    const int last = w->size - 1, limit = w->limit;
    int i_nrmax = w->nrmax, i_maxerr = w->order[i_nrmax], i, k, top;
    double errmax, errmin;
    if (last < 2) {
        w->order[0] = 0;
        w->order[1] = 1;
        w->i = i_maxerr;
        return;
    }
    errmax = w->elist[i_maxerr];
    while (i_nrmax > 0 && errmax > w->elist[w->order[i_nrmax - 1]]) {
        w->order[i_nrmax] = w->order[i_nrmax - 1];
        i_nrmax--;
    }
    top = (last < (limit / 2 + 2)) ? last : limit - last + 1;
    i = i_nrmax + 1;
    while (i < top && errmax < w->elist[w->order[i]]) {
        w->order[i - 1] = w->order[i];
        i++;
    }
    w->order[i - 1] = i_maxerr;
    errmin = w->elist[last];
    k = top - 1;
    while (k > i - 2 && errmin >= w->elist[w->order[k]]) {
        w->order[k + 1] = w->order[k];
        k--;
    }
    w->order[k + 1] = last;
    w->i = w->order[i_nrmax];
    w->nrmax = i_nrmax;
}

static void ws_update(orc_ws *w, double a1, double b1, double r1, double e1,
                      double a2, double b2, double r2, double e2)
{
    const int i_max = w->i, i_new = w->size, lvl = w->level[i_max] + 1;
    if (e2 > e1) {
        w->alist[i_max] = a2; w->rlist[i_max] = r2; w->elist[i_max] = e2; w->level[i_max] = lvl;
        w->alist[i_new] = a1; w->blist[i_new] = b1; w->rlist[i_new] = r1; w->elist[i_new] = e1;
        w->level[i_new] = lvl;
    } else {
        w->blist[i_max] = b1; w->rlist[i_max] = r1; w->elist[i_max] = e1; w->level[i_max] = lvl;
        w->alist[i_new] = a2; w->blist[i_new] = b2; w->rlist[i_new] = r2; w->elist[i_new] = e2;
        w->level[i_new] = lvl;
    }
    w->size++;
    if (lvl > w->maximum_level) w->maximum_level = lvl;
    ws_qpsrt(w);
}

static int subinterval_too_small(double a1, double a2, double b2)
{
    const double e = DBL_EPSILON, u = DBL_MIN;
    double tmp = (1 + 100 * e) * (fabs(a2) + 1000 * u);
    return fabs(a1) <= tmp && fabs(b2) <= tmp;
}

/* Return codes follow GSL: 0 success, 14 EMAXITER-like failures are reported as
 * the GSL errno value (GSL_EMAXITER=11, GSL_EROUND=18, GSL_ESING=21). */
static int qag61_impl(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
                      int limit, double *result, double *abserr, int *nintervals, double *ivl);

int orc_qag61(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
              int limit, double *result, double *abserr, int *nintervals)
{
    return qag61_impl(f, p, a, b, epsabs, epsrel, limit, result, abserr, nintervals, NULL);
}

int orc_qag61_record(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
                     int limit, double *result, double *abserr, int *nintervals, double *ivl)
{
    return qag61_impl(f, p, a, b, epsabs, epsrel, limit, result, abserr, nintervals, ivl);
}

static void record_ivl(const orc_ws *w, double *ivl)
{
    int k;
    if (!ivl) return;
    for (k = 0; k < w->size; k++) {
        ivl[2 * k] = w->alist[k];
        ivl[2 * k + 1] = w->blist[k];
    }
}

static int qag61_impl(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
                      int limit, double *result, double *abserr, int *nintervals, double *ivl)
{
    orc_ws w;
    double area, errsum, result0, abserr0, resabs0, resasc0, tolerance, round_off;
    int iteration, roundoff_type1 = 0, roundoff_type2 = 0, error_type = 0, k;
    w.limit = limit; w.size = 0; w.nrmax = 0; w.i = 0; w.maximum_level = 0;
    w.alist[0] = a; w.blist[0] = b; w.rlist[0] = 0; w.elist[0] = 0; w.order[0] = 0; w.level[0] = 0;
    *result = 0; *abserr = 0;
    if (nintervals) *nintervals = 1;
    orc_qk61(f, p, a, b, &result0, &abserr0, &resabs0, &resasc0);
    w.size = 1; w.rlist[0] = result0; w.elist[0] = abserr0;
    record_ivl(&w, ivl);
    tolerance = fmax(epsabs, epsrel * fabs(result0));
    round_off = 50 * DBL_EPSILON * resabs0;
    if (abserr0 <= round_off && abserr0 > tolerance) {
        *result = result0; *abserr = abserr0;
        return 18;
    } else if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) {
        *result = result0; *abserr = abserr0;
        return 0;
    } else if (limit == 1) {
        *result = result0; *abserr = abserr0;
        return 11;
    }
    area = result0;
    errsum = abserr0;
    iteration = 1;
    do {
        double a1, b1, a2, b2, a_i, b_i, r_i, e_i;
        double area1 = 0, area2 = 0, area12, error1 = 0, error2 = 0, error12;
        double resasc1, resasc2, resabs1, resabs2;
        a_i = w.alist[w.i]; b_i = w.blist[w.i]; r_i = w.rlist[w.i]; e_i = w.elist[w.i];
        a1 = a_i; b1 = 0.5 * (a_i + b_i); a2 = b1; b2 = b_i;
        orc_qk61(f, p, a1, b1, &area1, &error1, &resabs1, &resasc1);
        orc_qk61(f, p, a2, b2, &area2, &error2, &resabs2, &resasc2);
        area12 = area1 + area2;
        error12 = error1 + error2;
        errsum += (error12 - e_i);
        area += area12 - r_i;
        if (resasc1 != error1 && resasc2 != error2) {
            double delta = r_i - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i) roundoff_type1++;
            if (iteration >= 10 && error12 > e_i) roundoff_type2++;
        }
        tolerance = fmax(epsabs, epsrel * fabs(area));
        if (errsum > tolerance) {
            if (roundoff_type1 >= 6 || roundoff_type2 >= 20) error_type = 2;
            if (subinterval_too_small(a1, a2, b2)) error_type = 3;
        }
        ws_update(&w, a1, b1, area1, error1, a2, b2, area2, error2);
        iteration++;
    } while (iteration < limit && !error_type && errsum > tolerance);
    {
        double s = 0;
        for (k = 0; k < w.size; k++) s += w.rlist[k];
        *result = s;
    }
    *abserr = errsum;
    if (nintervals) *nintervals = w.size;
    record_ivl(&w, ivl);
    if (errsum <= tolerance) return 0;
    if (error_type == 2) return 18;
    if (error_type == 3) return 21;
    if (iteration == limit) return 11;
    return 5;
}

/* ======================= splines (GSL cspline natural, linear) ======================= */

size_t orc_bsearch(const double *xa, double x, size_t lo, size_t hi)
{
    while (hi > lo + 1) {
        size_t i = (hi + lo) / 2;
        if (xa[i] > x) hi = i; else lo = i;
    }
    return lo;
}

int orc_spline_init(orc_spline *s, const double *x, const double *y, int n, int linear)
{
    s->n = n; s->linear = linear; s->x = x; s->y = y; s->c = NULL;
    if (linear) return 0;
    {
        const int max_index = n - 1, sys = max_index - 1;
        double *c = calloc((size_t)n, sizeof(double));
        double *diag = malloc(sizeof(double) * (size_t)(sys > 0 ? sys : 1));
        double *off = malloc(sizeof(double) * (size_t)(sys > 0 ? sys : 1));
        double *g = malloc(sizeof(double) * (size_t)(sys > 0 ? sys : 1));
        int i;
        s->c = c;
        for (i = 0; i < sys; i++) {
            const double h_i = x[i + 1] - x[i], h_ip1 = x[i + 2] - x[i + 1];
            const double yd_i = y[i + 1] - y[i], yd_ip1 = y[i + 2] - y[i + 1];
            const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
            const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
            off[i] = h_ip1;
            diag[i] = 2.0 * (h_ip1 + h_i);
            g[i] = 3.0 * (yd_ip1 * g_ip1 - yd_i * g_i);
        }
        if (sys == 1) {
            c[1] = g[0] / diag[0];
        } else if (sys > 1) {
            /* symmetric tridiagonal solve, L D L^T (GSL linalg solve_symm_tridiag) */
            double *gam = malloc(sizeof(double) * (size_t)sys), *alp = malloc(sizeof(double) * (size_t)sys);
            double *z = malloc(sizeof(double) * (size_t)sys);
            const int N = sys;
            alp[0] = diag[0];
            gam[0] = off[0] / alp[0];
            for (i = 1; i < N - 1; i++) {
                alp[i] = diag[i] - off[i - 1] * gam[i - 1];
                gam[i] = off[i] / alp[i];
            }
            alp[N - 1] = diag[N - 1] - off[N - 2] * gam[N - 2];
            z[0] = g[0];
            for (i = 1; i < N; i++) z[i] = g[i] - gam[i - 1] * z[i - 1];
            for (i = 0; i < N; i++) z[i] = z[i] / alp[i];
            c[1 + N - 1] = z[N - 1];
            for (i = N - 2; i >= 0; i--) c[1 + i] = z[i] - gam[i] * c[1 + i + 1];
            free(gam); free(alp); free(z);
        }
        free(diag); free(off); free(g);
    }
    return 0;
}

double orc_spline_eval(const orc_spline *s, double xv, int *err)
{
    const double *xa = s->x, *ya = s->y;
    size_t idx;
    double x_lo, x_hi, dx, y_lo, y_hi, dy;
    if (xv < xa[0] || xv > xa[s->n - 1]) {   /* GSL_EDOM -> handler -> terminate(2001) */
        if (err) *err = 2001;
        return NAN;
    }
    idx = orc_bsearch(xa, xv, 0, (size_t)s->n - 1);
    x_lo = xa[idx]; x_hi = xa[idx + 1]; dx = x_hi - x_lo;
    y_lo = ya[idx]; y_hi = ya[idx + 1];
    if (s->linear) {
        if (dx > 0.0) return y_lo + (xv - x_lo) / dx * (y_hi - y_lo);
        return y_lo;
    }
    dy = y_hi - y_lo;
    {
        const double delx = xv - x_lo;
        const double c_i = s->c[idx], c_ip1 = s->c[idx + 1];
        const double b_i = (dy / dx) - dx * (c_ip1 + 2.0 * c_i) / 3.0;
        const double d_i = (c_ip1 - c_i) / (3.0 * dx);
        return y_lo + delx * (b_i + delx * (c_i + delx * d_i));
    }
}

void orc_spline_free(orc_spline *s) { free(s->c); s->c = NULL; }

/* ======================= Omega_nu (omega_nu_single.c) ======================= */

/* omega_nu_single.c:100-114 */
static double rho_nu_conversion(void)
{
    double convert = 4 * M_PI * 2;
    const double chbar = 1. / (2 * M_PI * ORC_LIGHTCGS * ORC_HBAR);
    convert *= (chbar * chbar * chbar);
    convert *= (1.60217646e-12 / ORC_LIGHTCGS / ORC_LIGHTCGS);
    return convert;
}

/* omega_nu_single.c:89-96 */
static double rho_nu_int(double q, void *params)
{
    const double *pp = (const double *)params;
    const double amnu = pp[0], kT = pp[1];
    double epsilon = sqrt(q * q + amnu * amnu);
    double f0 = 1. / (exp(q / kT) + 1);
    return q * q * epsilon * f0;
}

double orc_rho_nu_exact(double a, double mnu, double kBtnu)
{
    double param[2] = {mnu * a, kBtnu}, r, e;
    orc_qag61(rho_nu_int, param, 0, 500 * kBtnu, 0, 1e-9, ORC_GSL_VAL, &r, &e, NULL);
    return r * rho_nu_conversion() / pow(a, 4);
}

/* omega_nu_single.c:117-153 */
static void rho_nu_init(orc_rho_nu *t, double a0, double mnu, double kBtnu)
{
    const double logA0 = log(a0) - log(1.2);
    const double logaf = log(ORC_NU_SW * kBtnu / mnu) + log(1.2);
    int i;
    t->mnu = mnu;
    t->have_table = 0;
    if (mnu < 1e-6 * kBtnu || logaf < logA0) return;
    for (i = 0; i < ORC_NRHOTAB; i++) {
        double param[2], r, e;
        t->loga[i] = logA0 + i * (logaf - logA0) / (ORC_NRHOTAB - 1);
        param[0] = mnu * exp(t->loga[i]);
        param[1] = kBtnu;
        orc_qag61(rho_nu_int, param, 0, 500 * kBtnu, 0, 1e-9, ORC_GSL_VAL, &r, &e, NULL);
        t->rhonu[i] = r / pow(exp(t->loga[i]), 4) * rho_nu_conversion();
    }
    {
        orc_spline s;
        orc_spline_init(&s, t->loga, t->rhonu, ORC_NRHOTAB, 0);
        memcpy(t->c2, s.c, sizeof(t->c2));
        orc_spline_free(&s);
    }
    t->have_table = 1;
}

/* omega_nu_single.c:156-167 */
static double non_rel_rho_nu(double a, double kT, double amnu, double kTamnu2)
{
    return amnu * (kT * kT * kT) / (a * a * a * a) *
           (1.5 * 1.202056903159594 + kTamnu2 * 45. / 4. * 1.0369277551433704 +
            2835. / 32. * kTamnu2 * kTamnu2 * 1.0083492773819229 +
            80325 / 32. * kTamnu2 * kTamnu2 * kTamnu2 * 1.0020083928260826) *
           rho_nu_conversion();
}

static double rel_rho_nu(double a, double kT)
{
    return 7 * pow(M_PI * kT / a, 4) / 120. * rho_nu_conversion();
}

/* omega_nu_single.c:171-202 */
static double rho_nu(const orc_rho_nu *t, double a, double kT)
{
    const double amnu = a * t->mnu;
    const double kTamnu2 = (kT * kT / amnu / amnu);
    if (ORC_NU_SW * ORC_NU_SW * kTamnu2 < 1) return non_rel_rho_nu(a, kT, amnu, kTamnu2);
    if (amnu < 1e-6 * kT) return rel_rho_nu(a, kT);
    {
        const double loga = log(a);
        if (!t->have_table || loga < t->loga[0]) {
            if (amnu < 1e-4 * kT) return rel_rho_nu(a, kT);
            return non_rel_rho_nu(a, kT, amnu, kTamnu2);
        }
        {
            orc_spline s = {ORC_NRHOTAB, 0, t->loga, t->rhonu, (double *)t->c2};
            return orc_spline_eval(&s, loga, NULL);
        }
    }
}

/* omega_nu_single.c:16-51 */
void orc_init_omega_nu(orc_omega_nu *om, const double MNu[3], double a0, double h, double tcmb0)
{
    int mi;
    memset(om, 0, sizeof(*om));
    om->hybnu.enabled = 0;
    om->tcmb0 = tcmb0;
    om->kBtnu = ORC_BOLEVK * orc_tnucmb() * tcmb0;
    om->rhocrit = (3 * ORC_HUBBLE * h * ORC_HUBBLE * h) / (8 * M_PI * ORC_GRAVITY);
    for (mi = 0; mi < ORC_NUSPECIES; mi++) {
        int mmi;
        om->deg[mi] = 0;
        for (mmi = 0; mmi < mi; mmi++) {
            if (fabs(MNu[mi] - MNu[mmi]) < ORC_FLOAT_ACC) {
                om->deg[mmi] += 1;
                break;
            }
        }
        if (mmi == mi) om->deg[mi] = 1;
    }
    for (mi = 0; mi < ORC_NUSPECIES; mi++)
        if (om->deg[mi]) rho_nu_init(&om->tab[mi], a0, MNu[mi], om->kBtnu);
}

/* omega_nu_single.c:55-65 */
double orc_get_omega_nu(const orc_omega_nu *om, double a)
{
    double rhonu = 0;
    int mi;
    for (mi = 0; mi < ORC_NUSPECIES; mi++)
        if (om->deg[mi] > 0) rhonu += om->deg[mi] * rho_nu(&om->tab[mi], a, om->kBtnu);
    return rhonu / om->rhocrit;
}

/* omega_nu_single.c:246-257 */
double orc_particle_nu_fraction(const orc_hybrid_nu *h, double a, int i)
{
    if (!h->enabled) return 0;
    if (a > h->nu_crit_time) return h->nufrac_low[i];
    return 0;
}

/* omega_nu_single.c:69-74 */
double orc_get_omega_nu_nopart(const orc_omega_nu *om, double a)
{
    double omega_nu = orc_get_omega_nu(om, a);
    double part_nu = orc_get_omega_nu(om, 1) * orc_particle_nu_fraction(&om->hybnu, a, 0) / (a * a * a);
    return omega_nu - part_nu;
}

/* omega_nu_single.c:77-81 */
double orc_get_omegag(const orc_omega_nu *om, double a)
{
    const double omegag = 4 * ORC_STEFAN_BOLTZMANN / (ORC_LIGHTCGS * ORC_LIGHTCGS * ORC_LIGHTCGS) *
                          pow(om->tcmb0, 4) / om->rhocrit;
    return omegag / pow(a, 4);
}

/* omega_nu_single.c:262-279 */
double orc_omega_nu_single(const orc_omega_nu *om, double a, int i)
{
    double omega_nu, omega_part;
    if (om->deg[i] == 0) {
        int j;
        for (j = i; j >= 0; j--)
            if (om->deg[j]) { i = j; break; }
    }
    omega_nu = rho_nu(&om->tab[i], a, om->kBtnu) / om->rhocrit;
    omega_part = rho_nu(&om->tab[i], 1, om->kBtnu) / om->rhocrit;
    omega_part *= orc_particle_nu_fraction(&om->hybnu, a, i) / (a * a * a);
    return omega_nu - omega_part;
}

/* omega_nu_single.c:207-228 */
static double fermi_dirac_kernel(double x, void *p)
{
    (void)p;
    return x * x / (exp(x) + 1);
}

double orc_nufrac_low(double qc)
{
    double total, err;
    orc_qag61(fermi_dirac_kernel, NULL, 0, qc, 0, 1e-6, 100, &total, &err, NULL);
    return total / (1.5 * ORC_ZETA3);
}

/* omega_nu_single.c:230-240 */
void orc_init_hybrid_nu(orc_hybrid_nu *h, const double mnu[3], double vcrit, double light, double nu_crit_time, double kBtnu)
{
    int i;
    h->enabled = 1;
    h->nu_crit_time = nu_crit_time;
    h->vcrit = vcrit / light;
    for (i = 0; i < ORC_NUSPECIES; i++) {
        const double qc = mnu[i] * vcrit / light / kBtnu;
        h->nufrac_low[i] = orc_nufrac_low(qc);
    }
}

/* ======================= Hubble function of the tests ======================= */
/* delta_tot_table_test.c:25-45 */
static const orc_omega_nu *h_omnu;
static double h_Omega_nonu, h_OmegaLambda, h_Hubble;

void orc_init_hubble(const orc_omega_nu *om, double Omega0, double UnitTime_in_s)
{
    h_omnu = om;
    h_Omega_nonu = Omega0 - orc_get_omega_nu(om, 1);
    h_OmegaLambda = 1 - Omega0;
    h_Hubble = ORC_HUBBLE * UnitTime_in_s;
}

double orc_hubble(double a)
{
    double omega_tot = h_Omega_nonu / pow(a, 3) + h_OmegaLambda;
    omega_tot += orc_get_omega_nu(h_omnu, a);
    omega_tot += orc_get_omegag(h_omnu, a);
    return h_Hubble * sqrt(omega_tot);
}

/* ======================= specialJ / fslength (delta_tot_table.c) ======================= */

double orc_bessel_j0(double x)
{
    /* sin(x)/x with its Taylor series for |x| < 0.5 (GSL sf_bessel_j0 branches) */
    const double ax = fabs(x);
    if (ax < 0.5) {
        const double y = x * x;
        const double c1 = -1.0 / 6.0, c2 = 1.0 / 120.0, c3 = -1.0 / 5040.0, c4 = 1.0 / 362880.0,
                     c5 = -1.0 / 39916800.0, c6 = 1.0 / 6227020800.0;
        return 1.0 + y * (c1 + y * (c2 + y * (c3 + y * (c4 + y * (c5 + y * c6)))));
    }
    return sin(x) / x;
}

/* delta_tot_table.c:417-428 */
static double specialJ_fit(double x)
{
    double x2, x4, x8;
    if (x <= 0.) return 1.;
    x2 = x * x;
    x4 = x2 * x2;
    x8 = x4 * x4;
    return (1. + 0.0168 * x2 + 0.0407 * x4) / (1. + 2.1734 * x2 + 1.6787 * exp(4.1811 * log(x)) + 0.1467 * x8);
}

/* delta_tot_table.c:431-434 */
static double II(double x, double qc, int n)
{
    return (n * n + n * n * n * qc + n * qc * x * x - x * x) * qc * orc_bessel_j0(qc * x) +
           (2 * n + n * n * qc + qc * x * x) * cos(qc * x);
}

/* delta_tot_table.c:442-454 */
static double Jfrac_high(double x, double qc, double nufrac_low)
{
    double integ = 0;
    int n;
    for (n = 1; n < 20; n++)
        integ += -1 * pow((-1), n) * exp(-n * qc) / (n * n + x * x) / (n * n + x * x) * II(x, qc, n);
    integ /= 1.5 * 1.202056903159594 * (1 - nufrac_low);
    return integ;
}

/* delta_tot_table.c:457-463 */
double orc_specialJ(double x, double qc, double nufrac_low)
{
    if (qc > 0) return Jfrac_high(x, qc, nufrac_low);
    return specialJ_fit(x);
}

/* delta_tot_table.c:378-383 */
static double fslength_int(double loga, void *p)
{
    const double a = exp(loga);
    (void)p;
    return 1. / a / (a * orc_hubble(a));
}

static int g_qag_error;

/* delta_tot_table.c:394-407 */
double orc_fslength(double logai, double logaf, double light)
{
    double v, e;
    if (logai >= logaf) return 0;
    if (orc_qag61(fslength_int, NULL, logai, logaf, 0, 1e-6, ORC_GSL_VAL, &v, &e, NULL)) g_qag_error = 2001;
    return light * v;
}

/* delta_tot_table.c:613-617 */
double orc_get_delta_tot(double dnu, double dcdm, double OmegaNua3, double Omeganonu, double Omeganu1, double partnu)
{
    const double fcdm = 1 - OmegaNua3 / (Omeganonu + Omeganu1);
    return fcdm * (dcdm + dnu * OmegaNua3 / (Omeganonu + Omeganu1 * partnu));
}

/* ======================= transfer table (transfer_init.c:9-83) ======================= */

int orc_transfer_load(orc_transfer *t, double BoxSize, double UnitLength_in_cm, double InputSpectrum_UnitLength_in_cm, const char *file)
{
    char line[1000];
    const double scale = InputSpectrum_UnitLength_in_cm / UnitLength_in_cm;
    const double kmin = M_PI / BoxSize * scale;
    int count = 0;
    FILE *fd = fopen(file, "r");
    t->NPowerTable = 0;
    t->logk = t->T_nu = NULL;
    if (!fd) return 2019;
    while (fgets(line, sizeof line, fd)) {
        double k, c1, c2, c3, c4, c5, c6, c7;
        if (line[0] == '#') continue;
        if (sscanf(line, " %lg %lg %lg %lg %lg %lg %lg %lg", &k, &c1, &c2, &c3, &c4, &c5, &c6, &c7) == 8) {
            if (k > kmin) t->NPowerTable++;
        } else
            break;
    }
    fclose(fd);
    t->logk = malloc(2 * (size_t)t->NPowerTable * sizeof(double));
    t->T_nu = t->logk + t->NPowerTable;
    fd = fopen(file, "r");
    if (!fd) return 2020;
    while (count < t->NPowerTable && fgets(line, sizeof line, fd)) {
        double k, T_cdm, T_b, T_g, T_rnu, T_nu, T_tot, T_nonu;
        if (line[0] == '#') continue;
        if (sscanf(line, " %lg %lg %lg %lg %lg %lg %lg %lg", &k, &T_cdm, &T_b, &T_g, &T_rnu, &T_nu, &T_tot, &T_nonu) == 8) {
            if (k > kmin) {
                t->T_nu[count] = T_nu / T_nonu;
                k /= scale;
                t->logk[count] = log(k);
                count++;
            }
        } else
            break;
    }
    fclose(fd);
    if (count < t->NPowerTable) return 2021;
    return 0;
}

void orc_transfer_free(orc_transfer *t) { free(t->logk); t->logk = t->T_nu = NULL; }

/* ======================= integrator state (delta_tot_table.c) ======================= */

#define DT(d, k, a) ((d)->block_delta_tot[(size_t)(k) * (size_t)(d)->namax + (size_t)(a)])

/* delta_tot_table.c:25-60 */
void orc_dtt_allocate(orc_dtt *d, int nk_in, double TimeTransfer, double TimeMax, double Omega0,
                      const orc_omega_nu *om, double UnitTime_in_s, double UnitLength_in_cm)
{
    memset(d, 0, sizeof(*d));
    d->nk_allocated = nk_in;
    d->nk = nk_in;
    d->TimeTransfer = TimeTransfer;
    d->namax = (int)ceil(100 * (TimeMax - TimeTransfer)) + 2;
    d->ia = 0;
    d->scalefact = calloc((size_t)d->namax * (size_t)(nk_in + 1), sizeof(double));
    d->block_delta_tot = d->scalefact + d->namax;
    d->delta_nu_init = calloc(3 * (size_t)nk_in, sizeof(double));
    d->delta_nu_last = d->delta_nu_init + nk_in;
    d->wavenum = d->delta_nu_init + 2 * nk_in;
    d->omnu = om;
    d->light = ORC_LIGHTCGS * UnitTime_in_s / UnitLength_in_cm;
    d->delta_nu_prefac = 1.5 * Omega0 * ORC_HUBBLE * ORC_HUBBLE * pow(UnitTime_in_s, 2) / d->light;
    d->Omeganonu = Omega0 - orc_get_omega_nu(om, 1);
}

void orc_dtt_free(orc_dtt *d)
{
    free(d->scalefact);
    free(d->delta_nu_init);
    d->scalefact = d->delta_nu_init = NULL;
}

/* Integrand of delta_tot_table.c:492-500 with its two interpolants. */
typedef struct {
    double k, mnubykT, qc, nufrac_low, light, loga;
    orc_spline fs, dt;
    int exact_fsl;
    int err;
} dnu_params;

static double dnu_integrand_at(double logai, double fsl_aia, dnu_params *p)
{
    double delta_tot_at_a = orc_spline_eval(&p->dt, logai, &p->err);
    double specJ = orc_specialJ(p->k * fsl_aia / p->mnubykT, p->qc, p->nufrac_low);
    double ai = exp(logai);
    return fsl_aia / (ai * orc_hubble(ai)) * specJ * delta_tot_at_a;
}

static double dnu_integrand(double logai, void *vp)
{
    dnu_params *p = (dnu_params *)vp;
    double fsl_aia = orc_spline_eval(&p->fs, logai, &p->err);
    return dnu_integrand_at(logai, fsl_aia, p);
}

/* Exact free-streaming length light * int_x^{xa} e^{-2u}/H(e^u) du by a
 * 16-point Gauss-Legendre rule on [x, r] plus the precomputed tail from r. */
static double fsl_piece(double x, double r, double tail, double light)
{
    const double c = 0.5 * (x + r), h = 0.5 * (r - x);
    double s = 0;
    int q;
    for (q = 0; q < 16; q++) s += orc_gl16_w[q] * fslength_int(c + h * orc_gl16_x[q], NULL);
    return light * (tail + h * s);
}

/* Converged integral of the interpolated integrand: split [x0, xa] at every
 * knot of both interpolants and apply 8-point Gauss-Legendre on each piece. */
static double dnu_converged(dnu_params *p, const double *fsx, int nfs, const double *sx, int ns,
                            double x0, double xa)
{
    double total = 0, tail = 0;
    int i = nfs - 1, j = ns - 1;
    double r = xa;
    /* walk pieces from the right so the exact fslength tail accumulates */
    while (r > x0) {
        double l = x0, s = 0, c, h;
        int q;
        while (i >= 0 && fsx[i] >= r) i--;
        while (j >= 0 && sx[j] >= r) j--;
        if (i >= 0 && fsx[i] > l) l = fsx[i];
        if (j >= 0 && sx[j] > l) l = sx[j];
        c = 0.5 * (l + r);
        h = 0.5 * (r - l);
        for (q = 0; q < 8; q++) {
            const double xq = c + h * orc_gl8_x[q];
            double fsl;
            if (p->exact_fsl) fsl = fsl_piece(xq, r, tail, p->light);
            else fsl = orc_spline_eval(&p->fs, xq, &p->err);
            s += orc_gl8_w[q] * dnu_integrand_at(xq, fsl, p);
        }
        total += h * s;
        if (p->exact_fsl) {
            double t = 0;
            for (q = 0; q < 16; q++) t += orc_gl16_w[q] * fslength_int(c + h * orc_gl16_x[q], NULL);
            tail += h * t;
        }
        r = l;
    }
    return total;
}

/* delta_tot_table.c:507-611 */
int orc_get_delta_nu(orc_dtt *d, double a, const double wavenum[], double out[], double mnu)
{
    double fsl_A0a, deriv_prefac, qc = 0, relerr = 1e-6;
    const int Na = d->ia;
    const double mnubykT = mnu / d->omnu->kBtnu;
    double partnu;
    int ik;
    g_qag_error = 0;
    fsl_A0a = orc_fslength(log(d->TimeTransfer), log(a), d->light);
    deriv_prefac = d->TimeTransfer * (orc_hubble(d->TimeTransfer) / d->light) * d->TimeTransfer;
    for (ik = 0; ik < d->nk; ik++) {
        const double specJ = orc_specialJ(wavenum[ik] * fsl_A0a / (mnubykT > 0 ? mnubykT : 1), qc,
                                          d->omnu->hybnu.nufrac_low[0]);
        out[ik] = specJ * d->delta_nu_init[ik] * (1. + deriv_prefac * fsl_A0a);
    }
    partnu = orc_particle_nu_fraction(&d->omnu->hybnu, a, 0);
    if (partnu > 0) {
        if (1 - partnu < 1e-3) return g_qag_error;
        qc = d->omnu->hybnu.vcrit * mnubykT;
        relerr /= (1. + 1e-5 - orc_particle_nu_fraction(&d->omnu->hybnu, a, 0));
    }
    d->qag_max_intervals = 0;
    if (Na > 1 && mnubykT > 0) {
        const int Nfs = Na * 16;
        double *fslengths = malloc(sizeof(double) * (size_t)Nfs);
        double *fsscales = malloc(sizeof(double) * (size_t)Nfs);
        double *row = malloc(sizeof(double) * (size_t)Na);
        dnu_params P;
        memset(&P, 0, sizeof P);
        P.mnubykT = mnubykT;
        P.qc = qc;
        P.nufrac_low = d->omnu->hybnu.nufrac_low[0];
        P.light = d->light;
        P.exact_fsl = (d->mode == 2);
        for (ik = 0; ik < Nfs; ik++) {
            fsscales[ik] = log(d->TimeTransfer) + ik * (log(a) - log(d->TimeTransfer)) / (Nfs - 1.);
            fslengths[ik] = orc_fslength(fsscales[ik], log(a), d->light);
        }
        orc_spline_init(&P.fs, fsscales, fslengths, Nfs, 0);
        for (ik = 0; ik < d->nk; ik++) {
            double res = 0, abserr;
            int nint = 0, j;
            for (j = 0; j < Na; j++) row[j] = DT(d, ik, j);
            P.k = wavenum[ik];
            orc_spline_init(&P.dt, d->scalefact, row, Na, Na <= 2);
            if (d->mode == 0) {
                if (orc_qag61(dnu_integrand, &P, log(d->TimeTransfer), log(a), 0, relerr, ORC_GSL_VAL, &res, &abserr, &nint))
                    g_qag_error = 2001;
                if (nint > d->qag_max_intervals) d->qag_max_intervals = nint;
            } else {
                res = dnu_converged(&P, fsscales, Nfs, d->scalefact, Na, log(d->TimeTransfer), log(a));
            }
            out[ik] += d->delta_nu_prefac * res;
            orc_spline_free(&P.dt);
        }
        orc_spline_free(&P.fs);
        free(fslengths);
        free(fsscales);
        free(row);
        if (P.err) g_qag_error = P.err;
    }
    return g_qag_error;
}

/* delta_tot_table.c:153-172 */
int orc_get_delta_nu_combined(orc_dtt *d, double a, const double wavenum[], double out[])
{
    const double Omega_nu_tot = orc_get_omega_nu_nopart(d->omnu, a);
    double *single = malloc(sizeof(double) * (size_t)d->nk);
    int mi, ik, err = 0;
    memset(out, 0, sizeof(double) * (size_t)d->nk);
    for (mi = 0; mi < ORC_NUSPECIES; mi++) {
        if (d->omnu->deg[mi] > 0) {
            const double omeganu = d->omnu->deg[mi] * orc_omega_nu_single(d->omnu, a, mi);
            int e = orc_get_delta_nu(d, a, wavenum, single, d->omnu->tab[mi].mnu);
            if (e) err = e;
            for (ik = 0; ik < d->nk; ik++) out[ik] += single[ik] * omeganu / Omega_nu_tot;
        }
    }
    free(single);
    return err;
}

/* delta_tot_table.c:79-149 */
int orc_delta_tot_init(orc_dtt *d, int nk_in, const double wavenum[], const double delta_cdm[],
                       const orc_transfer *t, double Time)
{
    int ik;
    for (ik = 0; ik < d->ia; ik++) {
        if (log(Time) <= d->scalefact[ik]) {
            d->ia = ik;
            break;
        }
    }
    if (Time > d->TimeTransfer + 0.01 && d->ia == 0) return d->last_error = 2023;
    if (Time < d->TimeTransfer - 1e-5) return d->last_error = 2024;
    if (nk_in > d->nk_allocated) return d->last_error = 2011;
    d->nk = nk_in;
    {
        const double OmegaNua3 = orc_get_omega_nu_nopart(d->omnu, d->TimeTransfer) * pow(d->TimeTransfer, 3);
        const double OmegaNu1 = orc_get_omega_nu(d->omnu, 1);
        orc_spline s;
        orc_spline_init(&s, t->logk, t->T_nu, t->NPowerTable, 0);
        if (log(wavenum[d->nk - 1]) > t->logk[t->NPowerTable - 1]) {
            orc_spline_free(&s);
            return d->last_error = 2;
        }
        for (ik = 0; ik < d->nk; ik++) {
            int err = 0;
            const double T_nubyT_notnu = orc_spline_eval(&s, log(wavenum[ik]), &err);
            const double OmegaMa = (d->Omeganonu + OmegaNua3);
            if (err) { orc_spline_free(&s); return d->last_error = err; }
            if (d->ia == 0) {
                const double partnu = orc_particle_nu_fraction(&d->omnu->hybnu, d->TimeTransfer, 0);
                DT(d, ik, 0) = orc_get_delta_tot(delta_cdm[ik] * T_nubyT_notnu, delta_cdm[ik], OmegaNua3,
                                                 d->Omeganonu, OmegaNu1, partnu);
            }
            d->delta_nu_init[ik] = DT(d, ik, 0) * OmegaMa / (OmegaMa - OmegaNua3 + T_nubyT_notnu * OmegaNua3) *
                                   fabs(T_nubyT_notnu);
            d->wavenum[ik] = wavenum[ik];
        }
        orc_spline_free(&s);
    }
    if (d->ia == 0) {
        d->scalefact[0] = log(d->TimeTransfer);
        d->ia = 1;
    }
    {
        int e = orc_get_delta_nu_combined(d, exp(d->scalefact[d->ia - 1]), wavenum, d->delta_nu_last);
        d->delta_tot_init_done = 1;
        if (e) return d->last_error = e;
    }
    return 0;
}

/* delta_tot_table.c:177-191 */
void orc_update_delta_tot(orc_dtt *d, double a, const double dcdm[], const double dnu[], int overwrite)
{
    const double OmegaNua3 = orc_get_omega_nu_nopart(d->omnu, a) * pow(a, 3);
    const double OmegaNu1 = orc_get_omega_nu(d->omnu, 1);
    const double partnu = orc_particle_nu_fraction(&d->omnu->hybnu, a, 0);
    int ik;
    if (!overwrite) d->ia++;
    d->scalefact[d->ia - 1] = log(a);
    for (ik = 0; ik < d->nk; ik++)
        DT(d, ik, d->ia - 1) = orc_get_delta_tot(dnu[ik], dcdm[ik], OmegaNua3, d->Omeganonu, OmegaNu1, partnu);
}

/* delta_tot_table.c:193-250 */
int orc_get_delta_nu_update(orc_dtt *d, double a, int nk_in, const double keff[], const double dcdm[],
                            double out[], const orc_transfer *t)
{
    int ik, e;
    if (!d->delta_tot_init_done) {
        e = orc_delta_tot_init(d, nk_in, keff, dcdm, t, a);
        if (e) return e;
    }
    if (nk_in != d->nk) return d->last_error = 2002;
    if (d->nk < 2) return d->last_error = 2003;
    if (log(a) - d->scalefact[d->ia - 1] < ORC_FLOAT_ACC) {
        for (ik = 0; ik < d->nk; ik++) out[ik] = d->delta_nu_last[ik];
        return 0;
    }
    orc_update_delta_tot(d, a, dcdm, d->delta_nu_last, 0);
    e = orc_get_delta_nu_combined(d, a, keff, out);
    if (e) return d->last_error = e;
    for (ik = 0; ik < d->nk; ik++) d->delta_nu_last[ik] = out[ik];
    if (a >= exp(d->scalefact[d->ia - 2]) + 0.009)
        orc_update_delta_tot(d, a, dcdm, out, 1);
    else
        d->ia--;
    for (ik = 0; ik < d->nk; ik++) {
        if (isnan(out[ik])) return d->last_error = 2004;
        if (out[ik] < 0) out[ik] = 0;
    }
    return 0;
}

/* interface_common.c:139-145 */
double orc_kspace_prefac(const orc_dtt *d, double Time)
{
    const double OmegaNu_nop = orc_get_omega_nu_nopart(d->omnu, Time);
    const double omega_hybrid = orc_get_omega_nu(d->omnu, Time) - OmegaNu_nop;
    return OmegaNu_nop / (d->Omeganonu / pow(Time, 3) + omega_hybrid);
}

/* ======================= k-shell binning (powerspectrum.c) ======================= */

/* powerspectrum.c:8-12 */
static double onedinvwindow(int kx, int n)
{
    return kx ? M_PI * kx / (n * sin(M_PI * kx / (double)n)) : 1.0;
}

/* powerspectrum.c:16-24 */
static double invwindow(int kx, int ky, int kz, int n)
{
    if (n == 0) return 0;
    {
        double iwx = onedinvwindow(kx, n), iwy = onedinvwindow(ky, n), iwz = onedinvwindow(kz, n);
        return pow(iwx * iwy * iwz, 2);
    }
}

int orc_psindex(long long k2, int dims, int nrbins)
{
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    return (int)floor(binsperunit * log(sqrt((double)k2)));
}

/* The accumulation loop of powerspectrum.c:56-89 on one rank's slabs: per bin
 * the sums of c |F|^2 W^-4, c and c |k| (c the multiplicity), before the
 * allreduce and normalisation. */
void orc_bin_sums(int dims, const double *field, int nrbins, int startslab, int nslab, double *power,
                  long long *count, double *keffs)
{
#define KVAL(n) ((n) <= dims / 2 ? (n) : ((n)-dims))
#define RE(ix) field[2 * (size_t)(ix)]
#define IM(ix) field[2 * (size_t)(ix) + 1]
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    int i;
    memset(power, 0, sizeof(double) * (size_t)nrbins);
    memset(count, 0, sizeof(long long) * (size_t)nrbins);
    memset(keffs, 0, sizeof(double) * (size_t)nrbins);
    for (i = startslab; i < startslab + nslab; i++) {
        const size_t indx = (size_t)(i - startslab) * (size_t)dims * (size_t)(dims / 2 + 1);
        int j;
        for (j = 0; j < dims; j++) {
            const size_t indy = (size_t)j * (size_t)(dims / 2 + 1);
            size_t index = indx + indy;
            double kk = sqrt(pow(KVAL(i), 2) + pow(KVAL(j), 2));
            int k, psindex;
            if (kk > 0) {
                psindex = (int)floor(binsperunit * log(kk));
                power[psindex] += (RE(index) * RE(index) + IM(index) * IM(index)) *
                                  pow(invwindow(KVAL(i), KVAL(j), 0, dims), 2);
                keffs[psindex] += kk;
                count[psindex]++;
            }
            index = indx + indy + dims / 2;
            kk = sqrt(pow(KVAL(i), 2) + pow(KVAL(j), 2) + pow(KVAL(dims / 2), 2));
            psindex = (int)floor(binsperunit * log(kk));
            power[psindex] += (RE(index) * RE(index) + IM(index) * IM(index)) *
                              pow(invwindow(KVAL(i), KVAL(j), KVAL(dims / 2), dims), 2);
            keffs[psindex] += kk;
            count[psindex]++;
            for (k = 1; k < dims / 2; k++) {
                index = indx + indy + k;
                kk = sqrt(pow(KVAL(i), 2) + pow(KVAL(j), 2) + pow(KVAL(k), 2));
                psindex = (int)floor(binsperunit * log(kk));
                power[psindex] += 2 * (RE(index) * RE(index) + IM(index) * IM(index)) *
                                  pow(invwindow(KVAL(i), KVAL(j), KVAL(k), dims), 2);
                count[psindex] += 2;
                keffs[psindex] += 2 * kk;
            }
        }
    }
#undef RE
#undef IM
#undef KVAL
}

/* powerspectrum.c:33-117, single rank (the allreduces are identities) */
int orc_total_powerspectrum(int dims, const double *field, int nrbins, int startslab, int nslab,
                            double *power, long long *count, double *keffs)
{
    double total_mass2 = 0;
    int i, nonzero;
    if (startslab == 0) total_mass2 = field[0] * field[0] + field[1] * field[1];  /* powerspectrum.c:45-47 */
    orc_bin_sums(dims, field, nrbins, startslab, nslab, power, count, keffs);
    for (i = 0; i < nrbins; i++) {
        power[i] /= total_mass2;
        if (count[i]) {
            keffs[i] /= count[i];
            power[i] /= count[i];
        }
    }
    for (i = 0, nonzero = 0; i < nrbins; i++) {
        if (count[i]) {
            if (nonzero < i) {
                power[nonzero] = power[i];
                keffs[nonzero] = keffs[i];
                count[nonzero] = count[i];
            }
            nonzero++;
        }
    }
    return nonzero;
}

/* The Green's-function multiply of a PM potential pass, the step that follows
 * the neutrino rescale in Gadget-2's pmforce_periodic (the hook of
 * 0002-Add-hooks-to-kspace-neutrino-code.patch:116-125 sits right before it):
 * for k2 > 0, F *= -exp(-k2 asmth2)/k2 * ff^4 with ff = 1/(fx fy fz),
 * f = sin(pi k/N)/(pi k/N) per axis (1 at k = 0); the k = 0 mode is zeroed on
 * the rank holding slab 0.  asmth2 = (2 pi asmth / BoxSize)^2.  Restated from
 * the published Gadget-2 source (not part of the reference tree). */
void orc_green(int dims, double *field, int startslab, int nslab, double asmth2)
{
    int i, j, k;
    for (i = startslab; i < startslab + nslab; i++)
        for (j = 0; j < dims; j++)
            for (k = 0; k < dims / 2 + 1; k++) {
                const int ky = i > dims / 2 ? i - dims : i, kx = j > dims / 2 ? j - dims : j, kz = k;
                const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
                const size_t ip = ((size_t)(i - startslab) * dims + j) * (size_t)(dims / 2 + 1) + k;
                double smth, fx = 1, fy = 1, fz = 1, ff;
                if (k2 <= 0) continue;
                smth = -exp(-k2 * asmth2) / k2;
                if (kx != 0) {
                    fx = (M_PI * kx) / dims;
                    fx = sin(fx) / fx;
                }
                if (ky != 0) {
                    fy = (M_PI * ky) / dims;
                    fy = sin(fy) / fy;
                }
                if (kz != 0) {
                    fz = (M_PI * kz) / dims;
                    fz = sin(fz) / fz;
                }
                ff = 1 / (fx * fy * fz);
                smth *= ff * ff * ff * ff;
                field[2 * ip] *= smth;
                field[2 * ip + 1] *= smth;
            }
    if (startslab == 0) field[0] = field[1] = 0.0;
}

/* ======================= delta_pow + rescale ======================= */

/* delta_pow.c:19-37 */
double orc_get_dnudcdm(const orc_delta_pow *p, double kk)
{
    orc_spline s;
    if (kk < p->logkk[0] && kk > p->logkk[0] - log(2)) kk = p->logkk[0];
    if (kk < p->logkk[0] - log(2)) kk = p->logkk[0];
    if (kk > p->logkk[p->nbins - 1]) kk = p->logkk[p->nbins - 1];
    s.n = p->nbins; s.linear = 1; s.x = p->logkk; s.y = p->ratio; s.c = NULL;
    return p->norm * orc_spline_eval(&s, kk, NULL);
}

/* interface_gadget.c:158-194 (the mode loop; y is the slab index) */
int orc_rescale(int dims, double *field, int startslab, int nslab, double BoxSize, const orc_delta_pow *p)
{
    int x, y, z;
    for (y = startslab; y < startslab + nslab; y++)
        for (x = 0; x < dims; x++)
            for (z = 0; z < dims / 2 + 1; z++) {
                double kx, ky, kz, k2, smth;
                size_t ip;
                kx = x > dims / 2 ? x - dims : x;
                ky = y > dims / 2 ? y - dims : y;
                kz = z > dims / 2 ? z - dims : z;
                k2 = kx * kx + ky * ky + kz * kz;
                if (k2 <= 0) continue;
                k2 = log(sqrt(k2) * 2 * M_PI / BoxSize);
                smth = (1 + orc_get_dnudcdm(p, k2));
                if (isnan(smth)) return 5;
                ip = (size_t)dims * (size_t)(dims / 2 + 1) * (size_t)(y - startslab) + (size_t)(dims / 2 + 1) * x + z;
                field[2 * ip] *= smth;
                field[2 * ip + 1] *= smth;
            }
    return 0;
}

/* ======================= adjoint of get_delta_nu (test infrastructure) =======================
 *
 * The reference has no adjoint; these define the derivative the GPU adjoint
 * must reproduce: the derivative of the computed quadrature, i.e. of
 *   IC(k) + prefac * sum over the final QAG subintervals of GK61(I_k),
 * with the subdivision of the forward evaluation held fixed.  They are built
 * from independent means so that they pin the GPU's hand-derived formulas:
 *  - d/d delta_tot[k][j]: evaluation of the natural spline of the unit vector
 *    e_j (one spline per j, the oracle's own spline routine) at every node;
 *  - d/d mnu: complex-step differentiation (Im f(m + i h) / h) of a complex
 *    restatement of specialJ inside the integrand.
 * Their own pins (tests/test_oracle_adjoint.py): central differences of
 * orc_dnu_frozen, which is linear in delta_tot and smooth in mnu. */
#include <complex.h>

static double complex cj0(double complex z)
{
    if (cabs(z) < 0.5) {
        const double complex y = z * z;
        return 1.0 + y * (-1.0 / 6.0 + y * (1.0 / 120.0 + y * (-1.0 / 5040.0 + y * (1.0 / 362880.0 +
                     y * (-1.0 / 39916800.0 + y * (1.0 / 6227020800.0))))));
    }
    return csin(z) / z;
}

/* specialJ (delta_tot_table.c:417-463) on complex x and qc; hybrid when creal(qc) > 0. */
static double complex specialJ_c(double complex x, double complex qc, double nufrac_low)
{
    if (creal(qc) > 0) {
        double complex integ = 0;
        int n;
        for (n = 1; n < 20; n++) {
            const double complex ii = (n * n + n * n * n * qc + n * qc * x * x - x * x) * qc * cj0(qc * x) +
                                      (2 * n + n * n * qc + qc * x * x) * ccos(qc * x);
            integ += -1 * pow(-1, n) * cexp(-n * qc) / (n * n + x * x) / (n * n + x * x) * ii;
        }
        return integ / (1.5 * 1.202056903159594 * (1 - nufrac_low));
    }
    if (creal(x) <= 0.) return 1.;
    {
        const double complex x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
        return (1. + 0.0168 * x2 + 0.0407 * x4) / (1. + 2.1734 * x2 + 1.6787 * cexp(4.1811 * clog(x)) + 0.1467 * x8);
    }
}

/* d specialJ(k fsl / mu, vcrit mu [hybrid]) / d mu by complex step. */
static double dJ_dmu(double kfsl, double mu, double vcrit, int hybrid, double nufrac_low)
{
    const double h = 1e-30 * mu;
    const double complex m = mu + I * h;
    const double complex J = specialJ_c(kfsl / m, hybrid ? vcrit * m : 0.0, nufrac_low);
    return cimag(J) / h;
}

double orc_specialJ_dmu(double kfsl, double mu, double vcrit, int hybrid, double nufrac_low)
{
    return dJ_dmu(kfsl, mu, vcrit, hybrid, nufrac_low);
}

typedef struct {
    double fsl_A0a, deriv_prefac, qc, relerr, mnubykT;
    int integrate, hybrid, Na, Nfs;
    double *fsscales, *fslengths;
    orc_spline fs;
} dnu_setup;

/* The steps of delta_tot_table.c:507-575 that precede the per-k integrals. */
static void setup_dnu(orc_dtt *d, double a, double mnu, dnu_setup *S)
{
    double partnu;
    int i;
    memset(S, 0, sizeof *S);
    S->Na = d->ia;
    S->mnubykT = mnu / d->omnu->kBtnu;
    S->relerr = 1e-6;
    S->fsl_A0a = orc_fslength(log(d->TimeTransfer), log(a), d->light);
    S->deriv_prefac = d->TimeTransfer * (orc_hubble(d->TimeTransfer) / d->light) * d->TimeTransfer;
    partnu = orc_particle_nu_fraction(&d->omnu->hybnu, a, 0);
    S->integrate = 1;
    if (partnu > 0) {
        if (1 - partnu < 1e-3) S->integrate = 0;
        S->qc = d->omnu->hybnu.vcrit * S->mnubykT;
        S->hybrid = 1;
        S->relerr /= (1. + 1e-5 - orc_particle_nu_fraction(&d->omnu->hybnu, a, 0));
    }
    if (!(S->Na > 1 && S->mnubykT > 0)) S->integrate = 0;
    if (!S->integrate) return;
    S->Nfs = S->Na * 16;
    S->fslengths = malloc(sizeof(double) * (size_t)S->Nfs);
    S->fsscales = malloc(sizeof(double) * (size_t)S->Nfs);
    for (i = 0; i < S->Nfs; i++) {
        S->fsscales[i] = log(d->TimeTransfer) + i * (log(a) - log(d->TimeTransfer)) / (S->Nfs - 1.);
        S->fslengths[i] = orc_fslength(S->fsscales[i], log(a), d->light);
    }
    orc_spline_init(&S->fs, S->fsscales, S->fslengths, S->Nfs, 0);
}

static void free_setup(dnu_setup *S)
{
    if (!S->integrate) return;
    orc_spline_free(&S->fs);
    free(S->fsscales);
    free(S->fslengths);
}

static double ic_value(const orc_dtt *d, const dnu_setup *S, double k, double init)
{
    const double specJ = orc_specialJ(k * S->fsl_A0a / (S->mnubykT > 0 ? S->mnubykT : 1), 0,
                                      d->omnu->hybnu.nufrac_low[0]);
    return specJ * init * (1. + S->deriv_prefac * S->fsl_A0a);
}

/* Forward of orc_get_delta_nu (mode 0) that also records each k's final
 * subdivision: ivl[(k * ORC_GSL_VAL + t) * 2 + {0,1}], nint[k]. */
int orc_get_delta_nu_record(orc_dtt *d, double a, const double wavenum[], double out[], double mnu, double *ivl,
                            int *nint)
{
    dnu_setup S;
    int ik;
    g_qag_error = 0;
    setup_dnu(d, a, mnu, &S);
    for (ik = 0; ik < d->nk; ik++) {
        out[ik] = ic_value(d, &S, wavenum[ik], d->delta_nu_init[ik]);
        nint[ik] = 0;
    }
    if (S.integrate) {
        double *row = malloc(sizeof(double) * (size_t)S.Na);
        dnu_params P;
        memset(&P, 0, sizeof P);
        P.mnubykT = S.mnubykT;
        P.qc = S.qc;
        P.nufrac_low = d->omnu->hybnu.nufrac_low[0];
        P.light = d->light;
        P.fs = S.fs;
        for (ik = 0; ik < d->nk; ik++) {
            double res = 0, abserr;
            int j;
            for (j = 0; j < S.Na; j++) row[j] = DT(d, ik, j);
            P.k = wavenum[ik];
            orc_spline_init(&P.dt, d->scalefact, row, S.Na, S.Na <= 2);
            if (orc_qag61_record(dnu_integrand, &P, log(d->TimeTransfer), log(a), 0, S.relerr, ORC_GSL_VAL, &res,
                                 &abserr, &nint[ik], ivl + (size_t)ik * ORC_GSL_VAL * 2))
                g_qag_error = 2001;
            out[ik] += d->delta_nu_prefac * res;
            orc_spline_free(&P.dt);
        }
        free(row);
        if (P.err) g_qag_error = P.err;
    }
    free_setup(&S);
    return g_qag_error;
}

/* get_delta_nu with the recorded subdivision held fixed (the same sum of the
 * same GK61 applications, so it equals the recorded forward bit for bit). */
int orc_dnu_frozen(orc_dtt *d, double a, const double wavenum[], double out[], double mnu, const double *ivl,
                   const int *nint)
{
    dnu_setup S;
    int ik;
    g_qag_error = 0;
    setup_dnu(d, a, mnu, &S);
    for (ik = 0; ik < d->nk; ik++) out[ik] = ic_value(d, &S, wavenum[ik], d->delta_nu_init[ik]);
    if (S.integrate) {
        double *row = malloc(sizeof(double) * (size_t)S.Na);
        dnu_params P;
        memset(&P, 0, sizeof P);
        P.mnubykT = S.mnubykT;
        P.qc = S.qc;
        P.nufrac_low = d->omnu->hybnu.nufrac_low[0];
        P.light = d->light;
        P.fs = S.fs;
        for (ik = 0; ik < d->nk; ik++) {
            double res = 0;
            int j, t;
            for (j = 0; j < S.Na; j++) row[j] = DT(d, ik, j);
            P.k = wavenum[ik];
            orc_spline_init(&P.dt, d->scalefact, row, S.Na, S.Na <= 2);
            for (t = 0; t < nint[ik]; t++) {
                const double *iv = ivl + ((size_t)ik * ORC_GSL_VAL + t) * 2;
                double r, e, ra, rs;
                orc_qk61(dnu_integrand, &P, iv[0], iv[1], &r, &e, &ra, &rs);
                res += r;
            }
            out[ik] += d->delta_nu_prefac * res;
            orc_spline_free(&P.dt);
        }
        free(row);
        if (P.err) g_qag_error = P.err;
    }
    free_setup(&S);
    return g_qag_error;
}

/* Adjoint for seeds gbar[nk]: out[nk] the forward, dtot_bar[k * Na + j],
 * init_bar[nk], *mnu_bar (d/d mnu with the table's hybrid state fixed). */
int orc_get_delta_nu_adjoint(orc_dtt *d, double a, const double wavenum[], double mnu, const double gbar[],
                             double out[], double dtot_bar[], double init_bar[], double *mnu_bar)
{
    const int nk = d->nk, Na = d->ia;
    double *ivl = malloc(sizeof(double) * (size_t)nk * ORC_GSL_VAL * 2);
    int *nint = malloc(sizeof(int) * (size_t)nk);
    const double nfl = d->omnu->hybnu.nufrac_low[0];
    double mbar = 0;
    dnu_setup S;
    int ik, rc;
    rc = orc_get_delta_nu_record(d, a, wavenum, out, mnu, ivl, nint);
    setup_dnu(d, a, mnu, &S);
    memset(dtot_bar, 0, sizeof(double) * (size_t)nk * (size_t)Na);
    for (ik = 0; ik < nk; ik++) {
        const double m = S.mnubykT > 0 ? S.mnubykT : 1;
        const double lin = 1. + S.deriv_prefac * S.fsl_A0a;
        init_bar[ik] = gbar[ik] * orc_specialJ(wavenum[ik] * S.fsl_A0a / m, 0, nfl) * lin;
        if (S.mnubykT > 0)
            mbar += gbar[ik] * dJ_dmu(wavenum[ik] * S.fsl_A0a, S.mnubykT, 0, 0, nfl) * d->delta_nu_init[ik] * lin;
    }
    if (S.integrate) {
        orc_spline *basis = malloc(sizeof(orc_spline) * (size_t)Na);
        double *unit = calloc((size_t)Na * (size_t)Na, sizeof(double));
        double *row = malloc(sizeof(double) * (size_t)Na);
        double *phi = malloc(sizeof(double) * (size_t)Na);
        int j, err = 0;
        for (j = 0; j < Na; j++) {
            unit[(size_t)j * Na + j] = 1.0;
            orc_spline_init(&basis[j], d->scalefact, unit + (size_t)j * Na, Na, Na <= 2);
        }
        for (ik = 0; ik < nk; ik++) {
            orc_spline dt;
            int t;
            for (j = 0; j < Na; j++) row[j] = DT(d, ik, j);
            orc_spline_init(&dt, d->scalefact, row, Na, Na <= 2);
            for (t = 0; t < nint[ik]; t++) {
                const double *iv = ivl + ((size_t)ik * ORC_GSL_VAL + t) * 2;
                const double c = 0.5 * (iv[0] + iv[1]), hl = 0.5 * (iv[1] - iv[0]);
                int q, side;
                for (q = 0; q < 31; q++) {
                    for (side = 0; side < (q < 30 ? 2 : 1); side++) {
                        const double x = q < 30 ? (side ? c + hl * orc_xgk61[q] : c - hl * orc_xgk61[q]) : c;
                        const double w = orc_wgk61[q] * hl * d->delta_nu_prefac * gbar[ik];
                        const double fsl = orc_spline_eval(&S.fs, x, &err);
                        const double ai = exp(x);
                        const double base = fsl / (ai * orc_hubble(ai));
                        const double J = orc_specialJ(wavenum[ik] * fsl / S.mnubykT, S.qc, nfl);
                        const double Sv = orc_spline_eval(&dt, x, &err);
                        for (j = 0; j < Na; j++) phi[j] = orc_spline_eval(&basis[j], x, &err);
                        for (j = 0; j < Na; j++) dtot_bar[(size_t)ik * Na + j] += w * base * J * phi[j];
                        mbar += w * base * Sv *
                                dJ_dmu(wavenum[ik] * fsl, S.mnubykT, d->omnu->hybnu.vcrit, S.hybrid, nfl);
                    }
                }
            }
            orc_spline_free(&dt);
        }
        for (j = 0; j < Na; j++) orc_spline_free(&basis[j]);
        free(basis);
        free(unit);
        free(row);
        free(phi);
        if (err && !rc) rc = err;
    }
    *mnu_bar = mbar / d->omnu->kBtnu;
    free_setup(&S);
    free(ivl);
    free(nint);
    return rc;
}

/* ======================= adjoint of total_powerspectrum (test infrastructure) =======================
 *
 * L = sum_n gbar[n] power[n] over the nout compacted bins of
 * orc_total_powerspectrum on the same field.  With S_b = sum over the bin's
 * modes of c_m w_m |F_m|^2 (c_m the multiplicity, w_m = invwindow^2),
 * M = |F_0|^2 and N_b the bin count, power_b = S_b / (M N_b), so
 *   dL/dF_m = 2 c_m w_m gbar_b F_m / (M N_b)      (m != 0; Re and Im parts),
 *   dL/dF_0 = -2 F_0 / M * sum_n gbar_n power_n    (startslab == 0).
 * field_bar uses the field's layout (re, im interleaved).  Returns 0, or -1
 * when nout is not the forward's number of non-empty bins.  Pinned by central
 * differences of the forward (quadratic in F) in tests/test_oracle_pins.py. */
#define KVALD(n, d) ((n) <= (d) / 2 ? (n) : ((n) - (d)))
int orc_bin_power_adjoint(int dims, const double *field, int nrbins, int startslab, int nslab, const double *gbar,
                          int nout, double *field_bar)
{
    double *power = malloc(sizeof(double) * (size_t)nrbins), *keffs = malloc(sizeof(double) * (size_t)nrbins);
    double *alpha = calloc((size_t)nrbins, sizeof(double));
    long long *count = malloc(sizeof(long long) * (size_t)nrbins), *rawcount = calloc((size_t)nrbins, sizeof(long long));
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    const size_t rowlen = (size_t)(dims / 2 + 1);
    double M = 0, gp = 0;
    int n, b, i;
    if (startslab == 0) M = field[0] * field[0] + field[1] * field[1];
    n = orc_total_powerspectrum(dims, field, nrbins, startslab, nslab, power, count, keffs);
    if (n != nout) {
        free(power); free(keffs); free(alpha); free(count); free(rawcount);
        return -1;
    }
    /* raw bin counts, to map compacted outputs back to bins */
    for (i = startslab; i < startslab + nslab; i++) {
        int j, k;
        for (j = 0; j < dims; j++)
            for (k = 0; k <= dims / 2; k++) {
                const double kk = sqrt(pow(KVALD(i, dims), 2) + pow(KVALD(j, dims), 2) + pow(KVALD(k, dims), 2));
                if (kk > 0) rawcount[(int)floor(binsperunit * log(kk))] += (k == 0 || k == dims / 2) ? 1 : 2;
            }
    }
    for (b = 0, n = 0; b < nrbins; b++)
        if (rawcount[b]) {
            alpha[b] = gbar[n] / (M * (double)rawcount[b]);
            gp += gbar[n] * power[n];
            n++;
        }
    for (i = startslab; i < startslab + nslab; i++) {
        int j, k;
        for (j = 0; j < dims; j++)
            for (k = 0; k <= dims / 2; k++) {
                const size_t idx = ((size_t)(i - startslab) * (size_t)dims + (size_t)j) * rowlen + (size_t)k;
                const double kk = sqrt(pow(KVALD(i, dims), 2) + pow(KVALD(j, dims), 2) + pow(KVALD(k, dims), 2));
                double s;
                if (kk > 0) {
                    const int c = (k == 0 || k == dims / 2) ? 1 : 2;
                    const double w = invwindow(KVALD(i, dims), KVALD(j, dims), KVALD(k, dims), dims);
                    s = 2.0 * c * w * w * alpha[(int)floor(binsperunit * log(kk))];
                } else {
                    s = -2.0 * gp / M;
                }
                field_bar[2 * idx] = s * field[2 * idx];
                field_bar[2 * idx + 1] = s * field[2 * idx + 1];
            }
    }
    free(power); free(keffs); free(alpha); free(count); free(rawcount);
    return 0;
}

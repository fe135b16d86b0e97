/* quad_host.c -- host-side adaptive quadrature and interpolation.
 *
 * The reference uses GSL for these (delta_tot_table.c:404, :560-564, :597;
 * omega_nu_single.c:136, :147, :223; delta_pow.c:15).  GSL is not a dependency
 * of this library, so the same published algorithms are implemented here:
 *  - QAG (QUADPACK dqage, Piessens et al. 1983): bisect the subinterval with
 *    the largest error estimate, keep the list partially ordered by error
 *    (dqpsrt), stop on errsum <= max(epsabs, epsrel |area|), with QUADPACK's
 *    roundoff and too-small-interval detection and the 61-point Gauss-Kronrod
 *    rule with QUADPACK's error rescaling (200 |K-G| / resasc)^1.5;
 *  - natural cubic spline through a symmetric tridiagonal L D L^T solve, and
 *    its evaluation y + dx (b + dx (c + dx d)) with b, d formed from c;
 *  - linear interpolation y_lo + (x - x_lo)/h (y_hi - y_lo).
 * The device kernels in delta_nu.hip follow the same algorithms so host and
 * device agree step for step.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kn_gk61.h"
#include "kn_internal.h"

static double gk_rescale(double err, double resabs, double resasc)
{
    err = fabs(err);
    if (resasc != 0 && err != 0) {
        const double s = pow(200 * err / resasc, 1.5);
        err = s < 1 ? resasc * s : resasc;
    }
    if (resabs > DBL_MIN / (50 * DBL_EPSILON)) {
        const double floor_err = 50 * DBL_EPSILON * resabs;
        if (floor_err > err) err = floor_err;
    }
    return err;
}

/* One 61-point Gauss-Kronrod estimate on [a, b]. */
static void gk61(kn_fn f, void *p, double a, double b, double *res, double *err, double *resabs, double *resasc)
{
    double lo[31], hi[31];
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    const double fc = f(c, p);
    double sg = 0, sk = fc * kn_wgk61[30], sa = fabs(sk), sasc, mean;
    int j;
    for (j = 0; j < 15; j++) {
        const int q = 2 * j + 1;
        const double dx = h * kn_xgk61[q];
        const double f1 = f(c - dx, p), f2 = f(c + dx, p);
        lo[q] = f1;
        hi[q] = f2;
        sg += kn_wg30[j] * (f1 + f2);
        sk += kn_wgk61[q] * (f1 + f2);
        sa += kn_wgk61[q] * (fabs(f1) + fabs(f2));
    }
    for (j = 0; j < 15; j++) {
        const int q = 2 * j;
        const double dx = h * kn_xgk61[q];
        const double f1 = f(c - dx, p), f2 = f(c + dx, p);
        lo[q] = f1;
        hi[q] = f2;
        sk += kn_wgk61[q] * (f1 + f2);
        sa += kn_wgk61[q] * (fabs(f1) + fabs(f2));
    }
    mean = 0.5 * sk;
    sasc = kn_wgk61[30] * fabs(fc - mean);
    for (j = 0; j < 30; j++) sasc += kn_wgk61[j] * (fabs(lo[j] - mean) + fabs(hi[j] - mean));
    {
        const double diff = (sk - sg) * h;
        *res = sk * h;
        *resabs = sa * fabs(h);
        *resasc = sasc * fabs(h);
        *err = gk_rescale(diff, *resabs, *resasc);
    }
}

typedef struct {
    int n, limit, nrmax, imax;
    double *lo, *hi, *r, *e;
    int *ord, *lev;
} qag_list;

/* Keep ord[] descending by error for the entries QUADPACK keeps sorted. */
static void qag_sort(qag_list *w)
{
    const int last = w->n - 1;
    int nr = w->nrmax, imax = w->ord[nr], top, i, k;
    double emax, emin;
    if (last < 2) {
        w->ord[0] = 0;
        w->ord[1] = 1;
        w->imax = imax;
        return;
    }
    emax = w->e[imax];
    while (nr > 0 && emax > w->e[w->ord[nr - 1]]) {
        w->ord[nr] = w->ord[nr - 1];
        nr--;
    }
    top = last < w->limit / 2 + 2 ? last : w->limit - last + 1;
    for (i = nr + 1; i < top && emax < w->e[w->ord[i]]; i++) w->ord[i - 1] = w->ord[i];
    w->ord[i - 1] = imax;
    emin = w->e[last];
    for (k = top - 1; k > i - 2 && emin >= w->e[w->ord[k]]; k--) w->ord[k + 1] = w->ord[k];
    w->ord[k + 1] = last;
    w->imax = w->ord[nr];
    w->nrmax = nr;
}

int kn_qag61(kn_fn f, void *p, double a, double b, double epsabs, double epsrel, int limit, double *result,
             double *abserr)
{
    double buf[4 * KN_GSL_VAL];
    int ibuf[2 * KN_GSL_VAL];
    qag_list w;
    double r0, e0, ra0, rs0, tol, area, errsum;
    int it = 1, rt1 = 0, rt2 = 0, etype = 0, k;
    if (limit > KN_GSL_VAL) limit = KN_GSL_VAL;
    w.lo = buf; w.hi = buf + KN_GSL_VAL; w.r = buf + 2 * KN_GSL_VAL; w.e = buf + 3 * KN_GSL_VAL;
    w.ord = ibuf; w.lev = ibuf + KN_GSL_VAL;
    w.limit = limit; w.nrmax = 0; w.imax = 0;
    gk61(f, p, a, b, &r0, &e0, &ra0, &rs0);
    w.n = 1; w.lo[0] = a; w.hi[0] = b; w.r[0] = r0; w.e[0] = e0; w.ord[0] = 0; w.lev[0] = 0;
    *result = r0;
    *abserr = e0;
    tol = fmax(epsabs, epsrel * fabs(r0));
    if (e0 <= 50 * DBL_EPSILON * ra0 && e0 > tol) return 18;
    if ((e0 <= tol && e0 != rs0) || e0 == 0.0) return 0;
    if (limit == 1) return 11;
    area = r0;
    errsum = e0;
    do {
        const int im = w.imax;
        const double ai = w.lo[im], bi = w.hi[im], ri = w.r[im], ei = w.e[im];
        const double a1 = ai, b1 = 0.5 * (ai + bi), a2 = b1, b2 = bi;
        double r1, e1, ra1, rs1, r2, e2, ra2, rs2, r12, e12;
        gk61(f, p, a1, b1, &r1, &e1, &ra1, &rs1);
        gk61(f, p, a2, b2, &r2, &e2, &ra2, &rs2);
        r12 = r1 + r2;
        e12 = e1 + e2;
        errsum += e12 - ei;
        area += r12 - ri;
        if (rs1 != e1 && rs2 != e2) {
            if (fabs(ri - r12) <= 1.0e-5 * fabs(r12) && e12 >= 0.99 * ei) rt1++;
            if (it >= 10 && e12 > ei) rt2++;
        }
        tol = fmax(epsabs, epsrel * fabs(area));
        if (errsum > tol) {
            const double t = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
            if (rt1 >= 6 || rt2 >= 20) etype = 2;
            if (fabs(a1) <= t && fabs(b2) <= t) etype = 3;
        }
        {
            const int nw = w.n, lv = w.lev[im] + 1;
            if (e2 > e1) {
                w.lo[im] = a2; w.r[im] = r2; w.e[im] = e2;
                w.lo[nw] = a1; w.hi[nw] = b1; w.r[nw] = r1; w.e[nw] = e1;
            } else {
                w.hi[im] = b1; w.r[im] = r1; w.e[im] = e1;
                w.lo[nw] = a2; w.hi[nw] = b2; w.r[nw] = r2; w.e[nw] = e2;
            }
            w.lev[im] = lv;
            w.lev[nw] = lv;
            w.n++;
            qag_sort(&w);
        }
        it++;
    } while (it < limit && !etype && errsum > tol);
    {
        double s = 0;
        for (k = 0; k < w.n; k++) s += w.r[k];
        *result = s;
    }
    *abserr = errsum;
    if (errsum <= tol) return 0;
    if (etype == 2) return 18;
    if (etype == 3) return 21;
    if (it == limit) return 11;
    return 5;
}

size_t kn_bsearch(const double *xa, double x, size_t lo, size_t hi)
{
    while (hi > lo + 1) {
        const size_t mid = (hi + lo) / 2;
        if (xa[mid] > x) hi = mid;
        else lo = mid;
    }
    return lo;
}

void kn_cspline_init(const double *x, const double *y, int n, double *c)
{
    const int m = n - 2; /* interior unknowns c[1..n-2] */
    int i;
    double *gam, *alp, *z;
    memset(c, 0, sizeof(double) * (size_t)n);
    if (m <= 0) return;
    gam = malloc(sizeof(double) * (size_t)m * 3);
    alp = gam + m;
    z = alp + m;
    /* row i couples c[i], c[i+1], c[i+2]; diagonal 2(h_i + h_{i+1}), off-diagonal h_{i+1} */
    for (i = 0; i < m; i++) {
        const double hl = x[i + 1] - x[i], hr = x[i + 2] - x[i + 1];
        const double gl = hl != 0.0 ? 1.0 / hl : 0.0, gr = hr != 0.0 ? 1.0 / hr : 0.0;
        const double diag = 2.0 * (hr + hl);
        const double rhs = 3.0 * ((y[i + 2] - y[i + 1]) * gr - (y[i + 1] - y[i]) * gl);
        if (m == 1) {
            c[1] = rhs / diag;
            free(gam);
            return;
        }
        if (i == 0) {
            alp[0] = diag;
        } else {
            const double offp = x[i + 1] - x[i]; /* off-diagonal of row i-1 = h_i */
            alp[i] = diag - offp * gam[i - 1];
        }
        if (i < m - 1) gam[i] = hr / alp[i];
        z[i] = i == 0 ? rhs : rhs - gam[i - 1] * z[i - 1];
    }
    for (i = 0; i < m; i++) z[i] = z[i] / alp[i];
    c[m] = z[m - 1];
    for (i = m - 2; i >= 0; i--) c[i + 1] = z[i] - gam[i] * c[i + 2];
    free(gam);
}

double kn_cspline_eval(const double *x, const double *y, const double *c, int n, double xv, int *edom)
{
    size_t i;
    double h, dy, dx, b, d;
    if (xv < x[0] || xv > x[n - 1]) {
        if (edom) *edom = 1;
        return NAN;
    }
    i = kn_bsearch(x, xv, 0, (size_t)n - 1);
    h = x[i + 1] - x[i];
    dy = y[i + 1] - y[i];
    dx = xv - x[i];
    b = (dy / h) - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
    d = (c[i + 1] - c[i]) / (3.0 * h);
    return y[i] + dx * (b + dx * (c[i] + dx * d));
}

double kn_linear_eval(const double *x, const double *y, int n, double xv, int *edom)
{
    size_t i;
    double h;
    if (xv < x[0] || xv > x[n - 1]) {
        if (edom) *edom = 1;
        return NAN;
    }
    i = kn_bsearch(x, xv, 0, (size_t)n - 1);
    h = x[i + 1] - x[i];
    if (h > 0.0) return y[i] + (xv - x[i]) / h * (y[i + 1] - y[i]);
    return y[i];
}

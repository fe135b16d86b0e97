/* kn_internal.h -- declarations shared by the host C and HIP sources of
 * libkspace_neutrinos_hip.so.  Not installed; the public ABI is
 * include/kspace_neutrinos_hip.h. */
#ifndef KN_INTERNAL_H
#define KN_INTERNAL_H

#include <stddef.h>
#include "kspace_neutrinos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants of the reference (kspace_neutrino_const.h:12-19,
 * gadget_defines.h:8, omega_nu_single.c:10-14). */
#define KN_LIGHTCGS 2.99792458e10
#define KN_BOLEVK 8.61734e-5
#define KN_FLOAT_ACC 1e-6
#define KN_GSL_VAL 200
#define KN_HUBBLE 3.24077929e-18
#define KN_HBAR 6.582119e-16
#define KN_STEFAN_BOLTZMANN 5.670373e-5
#define KN_GRAVITY 6.67408e-8
#define KN_ZETA3 1.202056903159594

/* Error codes for GPU-side failures (reference codes are kept as they are). */
#define KN_ERR_HIP 3001
#define KN_ERR_ARG 3002
#define KN_ERR_NOHUBBLE 3003
#define KN_ERR_RANGE 3004
#define KN_ERR_DISABLED 3005

/* ---- host services (host_services.c) ---- */
double kn_hubble(double a);
/* The process's H in callback form (kn_set_background); NULL: the registered
 * or linked hubble_function again. */
void kn_set_hubble_ctx(double (*fn)(double a, void *ctx), void *ctx);
void kn_terminate(int code, const char *fmt, ...) __attribute__((noreturn, format(printf, 2, 3)));
void kn_message(int level, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void *kn_malloc(const char *name, size_t n);
void kn_free(void *p);
void kn_set_last_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

/* ---- background constants (omega_nu.c), shared with the device background ---- */
double kn_kbtnu(double tcmb0);           /* k_B T_nu in eV (omega_nu_single.c:19-21) */
double kn_rhocrit(double HubbleParam);   /* critical density, g cm^-3 */
double kn_rho_conversion(void);          /* (eV/c)^4 -> g cm^-3, one species */
double kn_omegag0(double tcmb0, double rhocrit); /* Omega_g at a = 1 */
/* nu_degeneracies of init_omega_nu: equal masses folded into the first */
void kn_nu_degeneracies(const double MNu[], int deg[]);
/* rho_nu_init's table range [lo, hi] in log a; 0 when no table is built */
int kn_rho_table_range(double a0, double mnu, double kBtnu, double *lo, double *hi);

/* ---- host quadrature and interpolation (quad_host.c) ---- */
typedef double (*kn_fn)(double x, void *p);
/* QUADPACK QAG with the 61-point Gauss-Kronrod rule; returns 0 on success or
 * the GSL errno of the failure (11 max iterations, 18 roundoff, 21 singular). */
int kn_qag61(kn_fn f, void *p, double a, double b, double epsabs, double epsrel, int limit, double *result,
             double *abserr);
/* Natural cubic spline (second-derivative/2 coefficients c[0..n-1]). */
void kn_cspline_init(const double *x, const double *y, int n, double *c);
/* Returns NaN and sets *edom when xv lies outside [x[0], x[n-1]]. */
double kn_cspline_eval(const double *x, const double *y, const double *c, int n, double xv, int *edom);
double kn_linear_eval(const double *x, const double *y, int n, double xv, int *edom);
size_t kn_bsearch(const double *xa, double x, size_t lo, size_t hi);

/* ---- device Hubble table (delta_nu.hip) ---- */
/* Make sure the device interpolants of 1/(a^2 H) and 1/(a H) cover [xlo, xhi] (x = ln a), with a
 * panel edge at origin (the integration's lower limit). */
void kn_htable_ensure(double origin, double xlo, double xhi);
void kn_htable_build_host(double origin, double xlo, double xhi);
/* H(e^x) from the same 1/(a H) interpolant, on the host (tests). */
double kn_htable_eval_host(double x);

/* ---- device solver for get_delta_nu (delta_nu.hip) ---- */
typedef struct {
    double mnubykT;     /* m_nu / k_B T_nu */
    double qc;          /* hybrid critical momentum, 0 when off */
    double relerr;      /* QAG relative tolerance */
    int ic_only;        /* 1: only the initial-condition piece (hybrid, all in particles) */
} kn_species;

/* The species parameters get_delta_nu derives from (d, a, mnu) (delta_tot_table.c:507-555). */
void kn_species_for(const _delta_tot_table *d, double a, double mnu, kn_species *s);

/* delta_nu_out[s*nk + k] receives get_delta_nu(..., mnu_s) for each species. */
void kn_dnu_solve(const _delta_tot_table *d, double a, const double *wavenum, int nspecies, const kn_species *sp,
                  double *delta_nu_out);

/* ---- state helpers (delta_tot_table.c) ---- */
double kn_deriv_prefac(const _delta_tot_table *d);

#ifdef __cplusplus
}
#endif
#endif

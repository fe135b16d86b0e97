/* kn_oracle.h -- CPU restatement of the kspace-neutrinos hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  It is never part of the product path.
 *
 * Every function restates the reference algorithm (file:line cited at each
 * definition in kn_oracle.c) in plain C99, compiled with strict IEEE
 * arithmetic (no fast-math, no FMA contraction).  GSL is absent from this
 * image, so QUADPACK's QAG with the 61-point Gauss-Kronrod rule, GSL's
 * natural cubic spline and linear interpolation are restated from their
 * published algorithms.  Errors that the reference routes to terminate()
 * are returned as the same integer codes instead of exiting.
 */
#ifndef KN_ORACLE_H
#define KN_ORACLE_H
#include <stddef.h>

#define ORC_NUSPECIES 3

/* ---- Omega_nu (omega_nu_single.h:51-108 semantics) ---- */
typedef struct {
    int have_table;
    double mnu;
    double loga[200];
    double rhonu[200];
    double c2[200];          /* natural-spline second-derivative/2 coefficients */
} orc_rho_nu;

typedef struct {
    int enabled;
    double nufrac_low[ORC_NUSPECIES];
    double nu_crit_time;
    double vcrit;
} orc_hybrid_nu;

typedef struct {
    orc_rho_nu tab[ORC_NUSPECIES];
    int deg[ORC_NUSPECIES];
    double rhocrit;
    double kBtnu;
    double tcmb0;
    orc_hybrid_nu hybnu;
} orc_omega_nu;

void   orc_init_omega_nu(orc_omega_nu *om, const double MNu[3], double a0, double h, double tcmb0);
double orc_get_omega_nu(const orc_omega_nu *om, double a);
double orc_get_omega_nu_nopart(const orc_omega_nu *om, double a);
double orc_get_omegag(const orc_omega_nu *om, double a);
double orc_omega_nu_single(const orc_omega_nu *om, double a, int i);
double orc_nufrac_low(double qc);
void   orc_init_hybrid_nu(orc_hybrid_nu *h, const double mnu[3], double vcrit, double light, double nu_crit_time, double kBtnu);
double orc_particle_nu_fraction(const orc_hybrid_nu *h, double a, int i);
double orc_rho_nu_exact(double a, double mnu, double kBtnu);   /* QAG 1e-9, no table */

/* ---- Hubble function of the reference tests (delta_tot_table_test.c:25-45) ---- */
void   orc_init_hubble(const orc_omega_nu *om, double Omega0, double UnitTime_in_s);
double orc_hubble(double a);

/* ---- specialJ / fslength / get_delta_tot (delta_tot_table.c:394-463, 613-617) ---- */
double orc_specialJ(double x, double qc, double nufrac_low);
double orc_fslength(double logai, double logaf, double light);
double orc_get_delta_tot(double dnu, double dcdm, double OmegaNua3, double Omeganonu, double Omeganu1, double partnu);
double orc_bessel_j0(double x);

/* ---- QAG with the 61-point Gauss-Kronrod rule (GSL/QUADPACK semantics) ---- */
typedef double (*orc_fn)(double x, void *p);
int orc_qag61(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
              int limit, double *result, double *abserr, int *nintervals);
/* One 61-point Gauss-Kronrod application (GSL qk61 semantics). */
void orc_qk61(orc_fn f, void *p, double a, double b, double *result, double *abserr,
              double *resabs, double *resasc);

/* ---- natural cubic spline / linear interpolation (GSL cspline, linear) ---- */
typedef struct {
    int n;
    int linear;
    const double *x;
    const double *y;
    double *c;               /* length n (cspline only) */
} orc_spline;
int    orc_spline_init(orc_spline *s, const double *x, const double *y, int n, int linear);
double orc_spline_eval(const orc_spline *s, double xv, int *err);
void   orc_spline_free(orc_spline *s);
size_t orc_bsearch(const double *xa, double x, size_t lo, size_t hi);

/* ---- transfer table (transfer_init.c:9-83) ---- */
typedef struct {
    int NPowerTable;
    double *logk;
    double *T_nu;
} orc_transfer;
int  orc_transfer_load(orc_transfer *t, double BoxSize, double UnitLength_in_cm, double InputSpectrum_UnitLength_in_cm, const char *file);
void orc_transfer_free(orc_transfer *t);

/* ---- integrator state (delta_tot_table.h:17-58) ---- */
typedef struct {
    int nk, nk_allocated, namax, ia;
    double delta_nu_prefac;
    int delta_tot_init_done;
    double *scalefact;       /* block: scalefact[namax], then delta_tot rows */
    double *block_delta_tot; /* delta_tot[k][a] = block_delta_tot[k*namax + a] */
    double *delta_nu_init, *delta_nu_last, *wavenum;
    const orc_omega_nu *omnu;
    double Omeganonu, light, TimeTransfer;
    int mode;                /* 0: QAG as the reference; 1: converged integral of the same
                                interpolants; 2: converged with the exact fslength */
    int last_error;          /* terminate() code of the last failure, 0 if none */
    int qag_max_intervals;   /* diagnostics: max QAG subintervals used in the last call */
} orc_dtt;

void orc_dtt_allocate(orc_dtt *d, int nk_in, double TimeTransfer, double TimeMax, double Omega0,
                      const orc_omega_nu *om, double UnitTime_in_s, double UnitLength_in_cm);
void orc_dtt_free(orc_dtt *d);
int  orc_delta_tot_init(orc_dtt *d, int nk_in, const double wavenum[], const double delta_cdm[],
                        const orc_transfer *t, double Time);
void orc_update_delta_tot(orc_dtt *d, double a, const double dcdm[], const double dnu[], int overwrite);
int  orc_get_delta_nu(orc_dtt *d, double a, const double wavenum[], double out[], double mnu);
int  orc_get_delta_nu_combined(orc_dtt *d, double a, const double wavenum[], double out[]);
int  orc_get_delta_nu_update(orc_dtt *d, double a, int nk_in, const double keff[], const double dcdm[],
                             double out[], const orc_transfer *t);

/* ---- adjoint of get_delta_nu (no reference counterpart; checker of the GPU adjoint) ---- */
int orc_qag61_record(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
                     int limit, double *result, double *abserr, int *nintervals, double *ivl);
int orc_get_delta_nu_record(orc_dtt *d, double a, const double wavenum[], double out[], double mnu, double *ivl,
                            int *nint);
int orc_dnu_frozen(orc_dtt *d, double a, const double wavenum[], double out[], double mnu, const double *ivl,
                   const int *nint);
int orc_get_delta_nu_adjoint(orc_dtt *d, double a, const double wavenum[], double mnu, const double gbar[],
                             double out[], double dtot_bar[], double init_bar[], double *mnu_bar);
double orc_specialJ_dmu(double kfsl, double mu, double vcrit, int hybrid, double nufrac_low);

/* ---- k-shell binning and rescale (powerspectrum.c:7-117, interface_gadget.c:158-194) ---- */
/* The PM Green's-function multiply that follows the neutrino rescale (Gadget-2 pmforce_periodic). */
void orc_green(int dims, double *field, int startslab, int nslab, double asmth2);
/* powerspectrum.c:56-89: raw per-bin sums of one rank's slabs (no normalisation, no compaction). */
void orc_bin_sums(int dims, const double *field, int nrbins, int startslab, int nslab, double *power,
                  long long *count, double *keffs);
int orc_total_powerspectrum(int dims, const double *field /* re,im interleaved */, int nrbins,
                            int startslab, int nslab, double *power, long long *count, double *keffs);
/* Bin index of an integer |k|^2 exactly as powerspectrum.c:40,67 evaluates it. */
int orc_psindex(long long k2, int dims, int nrbins);

/* delta_pow.c:6-37 */
typedef struct { const double *logkk; const double *ratio; int nbins; double norm; } orc_delta_pow;
double orc_get_dnudcdm(const orc_delta_pow *p, double logk);
/* In-place rescale of the slab by 1 + get_dnudcdm(log(|k| 2 pi / L)); returns 5 on NaN. */
int orc_rescale(int dims, double *field, int startslab, int nslab, double BoxSize, const orc_delta_pow *p);

/* Adjoint of orc_total_powerspectrum: field_bar = dL/dF for L = sum_n gbar[n] power[n]. */
int orc_bin_power_adjoint(int dims, const double *field, int nrbins, int startslab, int nslab, const double *gbar,
                          int nout, double *field_bar);

/* kspace_prefac of interface_common.c:139-145 */
double orc_kspace_prefac(const orc_dtt *d, double Time);

#endif

/* kspace_neutrinos_hip.h -- C ABI of the MI355X implementation of the
 * kspace-neutrinos hot path (libkspace_neutrinos_hip.so).
 *
 * The entry points below keep the reference library's C signatures so the
 * library drops into a Gadget-style PM step unchanged; each declaration cites
 * the reference declaration it replaces.  Behind them, host C code drives
 * hand-written HIP kernels for gfx950.  Additive entry points (prefix kn_)
 * cover device-resident grids, the adjoints and the batched solver; they never
 * replace the reference entry points.
 *
 * Conventions kept from the reference:
 *  - FFTW2's fftw_complex must be double precision ({double re, im}).
 *  - Errors are fatal: the library calls the host's terminate(code, fmt, ...)
 *    with the reference's error codes (2001..2024, 2, 5); GPU failures use
 *    codes 3001..3099.  kn_ entry points that return int return 0 on success.
 *  - Host services (gadget_defines.h:11-32) are imported: hubble_function,
 *    terminate, message, mymalloc_fullinfo, myfree_fullinfo.  Each is a weak
 *    reference; when the host does not define one, a built-in fallback is used
 *    (hubble_function has none: register one with kn_set_hubble_function).
 *  - One rank, one GPU: the MPI communicator arguments are accepted for ABI
 *    compatibility and not used.
 */
#ifndef KSPACE_NEUTRINOS_HIP_H
#define KSPACE_NEUTRINOS_HIP_H

#include <stddef.h>
#include <mpi.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types */

/* FFTW2 complex (powerspectrum.h:6-14 with DOUBLEPRECISION_FFTW).  Include
 * dfftw.h before this header to use FFTW's own definition. */
#ifndef FFTW_H
typedef double fftw_real;
typedef struct { fftw_real re, im; } fftw_complex;
#endif

#define NUSPECIES 3

/* omega_nu_single.h:21-28; interp/acc hold this library's spline state. */
struct _rho_nu_single {
    double *loga;
    double *rhonu;
    void *interp;
    void *acc;
    double mnu;
};
typedef struct _rho_nu_single _rho_nu_single;

/* omega_nu_single.h:51-63 */
struct _hybrid_nu {
    int enabled;
    double nufrac_low[NUSPECIES];
    double nu_crit_time;
    double vcrit;
};
typedef struct _hybrid_nu _hybrid_nu;

/* omega_nu_single.h:94-107 */
struct _omega_nu {
    _rho_nu_single *RhoNuTab[NUSPECIES];
    int nu_degeneracies[NUSPECIES];
    double rhocrit;
    double kBtnu;
    double tcmb0;
    _hybrid_nu hybnu;
};
typedef struct _omega_nu _omega_nu;

/* transfer_init.h:12-17 */
struct _transfer_init_table {
    int NPowerTable;
    double *logk;
    double *T_nu;
};
typedef struct _transfer_init_table _transfer_init_table;

/* delta_tot_table.h:17-58: field order and types unchanged; the history block
 * is scalefact[namax] followed by nk_allocated rows delta_tot[k][namax]. */
struct _delta_tot_table {
    int nk;
    int nk_allocated;
    int namax;
    int ia;
    int ThisTask;
    double delta_nu_prefac;
    int delta_tot_init_done;
    int debug;
    double **delta_tot;
    double *scalefact;
    double *delta_nu_init;
    double *delta_nu_last;
    double *wavenum;
    const _omega_nu *omnu;
    double Omeganonu;
    double light;
    double TimeTransfer;
};
typedef struct _delta_tot_table _delta_tot_table;

/* delta_pow.h:10-17; spline/acc hold this library's interpolation state. */
struct _delta_pow {
    double *logkk;
    double *delta_ratio;
    void *spline;
    void *acc;
    int nbins;
    double norm;
};
typedef struct _delta_pow _delta_pow;

/* ------------------------------------------ host services (imports) */

/* gadget_defines.h:11 and :17-32.  Weak imports, see above. */
double hubble_function(double a);
void terminate(int ierr, const char *fmt, ...);
void message(int ierr, const char *fmt, ...);
void *mymalloc_fullinfo(const char *string, size_t size, const char *func, const char *file, int line);
void myfree_fullinfo(void *ptr, const char *func, const char *file, int line);

/* ------------------------------------------------ (a) k-shell binning */

/* powerspectrum.h:30.  outfield may be device memory (the fast path, no
 * copy) or host memory (staged through HBM).  Outputs are host arrays. */
int total_powerspectrum(const int dims, fftw_complex *outfield, const int nrbins, const int startslab,
                        const int nslab, double *power, long long int *count, double *keffs,
                        MPI_Comm MYMPI_COMM_WORLD);

/* --------------------------------------- (b) free-streaming integrator */

/* delta_tot_table.h:70 */
void allocate_delta_tot_table(_delta_tot_table *d_tot, const int nk_in, const double TimeTransfer,
                              const double TimeMax, const double Omega0, const _omega_nu *const omnu,
                              const double UnitTime_in_s, const double UnitLength_in_cm, int debug);
/* delta_tot_table.h:73 */
void free_delta_tot_table(_delta_tot_table *d_tot);
/* delta_tot_table.h:85 */
void delta_tot_init(_delta_tot_table *const d_tot, const int nk_in, const double wavenum[],
                    const double delta_cdm_curr[], const _transfer_init_table *const t_init, const double Time);
/* delta_tot_table.h:90 */
void update_delta_tot(_delta_tot_table *const d_tot, const double a, const double delta_cdm_curr[],
                      const double delta_nu_curr[], const int overwrite);
/* delta_tot_table.h:104 */
void get_delta_nu_update(_delta_tot_table *const d_tot, const double a, const int nk_in, const double keff[],
                         const double P_cdm_curr[], double delta_nu_curr[], _transfer_init_table *transfer_init);
/* delta_tot_table.h:113 */
void get_delta_nu(const _delta_tot_table *const d_tot, const double a, const double wavenum[],
                  double delta_nu_curr[], const double mnu);
/* delta_tot_table.h:117 */
void get_delta_nu_combined(const _delta_tot_table *const d_tot, const double a, const double wavenum[],
                           double delta_nu_curr[]);
/* delta_tot_table.h:138 */
double specialJ(const double x, const double vcmnubylight, const double nufrac_low);
/* delta_tot_table.h:149 */
double fslength(const double logai, const double logaf, const double light);
/* delta_tot_table.h:155 */
double get_delta_tot(const double delta_nu_curr, const double delta_cdm_curr, const double OmegaNua3,
                     const double Omeganonu, const double Omeganu1, const double partnu);
/* delta_tot_table.h:121-135: state files in the reference's text formats (host only).
 * A NULL file name means "delta_tot_nu.txt" in the working directory. */
void save_delta_tot(const _delta_tot_table *const d_tot, const int iia, char *savedir);
void save_all_nu_state(const _delta_tot_table *const d_tot, char *savedir);
int save_nu_power(const _delta_tot_table *const d_tot, const double Time, const int snapnum, const char *OutputDir);
void read_all_nu_state(_delta_tot_table *const d_tot, char *savedir);

/* ---------------------------------------- Omega_nu (host, no kernel) */

/* omega_nu_single.h:116, :119, :122 */
void init_omega_nu(_omega_nu *const omnu, const double MNu[], const double a0, const double HubbleParam,
                   const double tcmb0);
double get_omega_nu(const _omega_nu *const omnu, const double a);
double get_omega_nu_nopart(const _omega_nu *const omnu, const double a);
/* omega_nu_single.h:125, :131 */
double get_omegag(const _omega_nu *const omnu, const double a);
double omega_nu_single(const _omega_nu *const rho_nu_tab, const double a, const int i);
/* omega_nu_single.h:37, :45 */
void rho_nu_init(_rho_nu_single *rho_nu_tab, double a0, const double mnu, const double HubbleParam,
                 const double kBtnu);
double rho_nu(_rho_nu_single *rho_nu_tab, const double a, const double kT);
/* omega_nu_single.h:75, :84, :87 */
void init_hybrid_nu(_hybrid_nu *const hybnu, const double mnu[], const double vcrit, const double light,
                    const double nu_crit_time, const double kBtnu);
double particle_nu_fraction(const _hybrid_nu *const hybnu, const double a, int i);
double nufrac_low(const double qc);

/* transfer_init.h:28, :31 */
void allocate_transfer_init_table(_transfer_init_table *t_init, const double BoxSize, const double UnitLength_in_cm,
                                  const double InputSpectrum_UnitLength_in_cm, const char *KspaceTransferFunction);
void free_transfer_init_table(_transfer_init_table *t_init);

/* -------------------------------------------- (c) ratio + rescale */

/* delta_pow.h:30, :39, :42 */
void init_delta_pow(_delta_pow *d_pow, double logkk[], double delta_ratio[], int nbins, double norm);
double get_dnudcdm_powerspec(_delta_pow *d_pow, double kk);
void free_d_pow(_delta_pow *d_pow);

/* interface_common.h:12-30 parameters and the module-level entry points */
extern struct __kspace_params {
    char KspaceTransferFunction[500];
    double TimeTransfer;
    double InputSpectrum_UnitLength_in_cm;
    double MNu[NUSPECIES];
    int hybrid_neutrinos_on;
    double vcrit;
    double nu_crit_time;
} kspace_params;

/* interface_common.h:37, :44, :46, :66 */
double OmegaNu(double a);
double OmegaNu_nopart(double a);
void InitOmegaNu(const double HubbleParam, const double tcmb0, MPI_Comm MYMPI_COMM_WORLD);
void allocate_kspace_memory(const int nk_in, const int ThisTask, const double BoxSize, const double UnitTime_in_s,
                            const double UnitLength_in_cm, const double Omega0, char *snapdir, const double TimeMax,
                            MPI_Comm MYMPI_COMM_WORLD);
/* interface_common.h:80 (code-agnostic entry: pre-binned P(k)) */
_delta_pow compute_neutrino_power_from_cdm(const double Time, const double keff_in[], const double P_cdm[],
                                           const long int Nmodes[], const int nk_in, MPI_Comm MYMPI_COMM_WORLD);
/* interface_common.h:86, :117: state and P_nu(k) files of the module table */
void save_nu_state(char *savedir);
int save_neutrino_power(const double Time, const int snapnum, const char *OutputDir);
/* interface_common.h:97, :108 */
void get_nu_state(double **scalefact, double **delta_tot, size_t *nk, size_t *ia);
void set_nu_state(double *scalefact, double *delta_tot, const size_t nk, const size_t ia, MPI_Comm MYMPI_COMM_WORLD);
/* interface_common.h:121 */
int particle_nu_active(double a);

/* interface_gadget.h:38: bin (a) -> integrate (b) -> rescale (c) of the
 * Fourier grid in place.  fft_of_rhogrid may be device or host memory. */
void add_nu_power_to_rhogrid(const double Time, const double BoxSize, fftw_complex *fft_of_rhogrid, const int pmgrid,
                             int slabstart_y, int nslab_y, MPI_Comm MYMPI_COMM_WORLD);

/* ================================================ additive entry points */

/* Register the Hubble function H(a) in internal time units when the host does
 * not define hubble_function() at link time (e.g. ctypes callers). */
void kn_set_hubble_function(double (*fn)(double a));

/* Synthetic Gaussian random field on the device (counter-based RNG keyed by
 * seed and mode index): layout [nslab][dims][dims/2+1], mode k gets
 * Re, Im = sqrt(P(|k| 2 pi/L)/2) N(0,1) with P = A k^ns T_BBKS(k/Gamma)^2 and
 * F[0] = (dims^3, 0).  d_field is device memory. */
int kn_synth_grid(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize, double amplitude,
                  double ns, double gamma, unsigned long long seed);

/* (a) on a device grid with device-side outputs left in HBM: power[nrbins]
 * holds the unnormalised bin sums, count/keffs the geometry sums, total_mass2
 * |F0|^2.  stream is a hipStream_t (NULL: default stream).  No host sync. */
int kn_bin_power_device(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                        double *d_power, double *d_total_mass2, void *stream);

/* (c) rescale of a device grid by 1 + norm * lininterp(log k) of the ratio
 * table (device arrays logkk/ratio of length nbins).  No host sync. */
int kn_rescale_device(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                      const double *d_logkk, const double *d_ratio, int nbins, double norm, void *stream);

/* kn_bin_power_device with no allocation: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * outputs; the sweep's bin thresholds, inverse window and per-block partial
 * sums live in a 32 MiB device array of the library, one per GPU, shared by
 * every call on that GPU (calls that overlap on one GPU race).  Limits:
 * dims <= 65536, nrbins <= 262144 (3004 beyond them).  A new (dims, nrbins)
 * copies its tables from the host synchronously on stream; otherwise the call
 * only launches on stream. */
int kn_bin_power_device_noalloc(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                                double *d_power, double *d_total_mass2, void *stream);

/* get_delta_nu keeps each table's delta_tot history in device memory at the
 * host's row stride and compares the host rows with its copy on every call:
 * only columns from a row's first changed entry on are uploaded (an appended
 * or overwritten step is one or two columns; any host write is seen).  This
 * returns how many columns the last call uploaded (diagnostics, tests). */
int kn_dnu_last_upload_columns(void);
/* Diagnostics of the last get_delta_nu's QAG items: the largest final
 * subinterval count, and the count summed over items (items into *items).
 * An item ending with n subintervals evaluated the integrand 61 (2n - 1)
 * times, the unit of work the free-streaming integral is measured in. */
int kn_dnu_last_max_intervals(void);
long long kn_dnu_last_intervals(int *items);
/* HIP-event timing of get_delta_nu's QAG kernel, on (1) or off (0, the
 * default), and the last call's kernel time in ms (measurement only). */
int kn_dnu_qag_timing(int on);
double kn_dnu_last_qag_ms(void);

/* (c) fused with the next step of a PM force pass: no reference counterpart.
 * The reference rescales the grid in add_nu_power_to_rhogrid and the caller
 * then multiplies it by the potential's Green's function (Gadget-2's
 * pmforce_periodic, right after the hook of 0002-Add-hooks-to-kspace-neutrino-code.patch:116-125):
 * two read + write sweeps.  This does both in one: for k2 > 0,
 * F *= 1 + norm lininterp(log k) (as kn_rescale_device), then
 * F *= -exp(-k2 (2 pi asmth / boxsize)^2) / k2 * (W^-1_x W^-1_y W^-1_z)^4 (W^-1
 * the per-axis CiC inverse window); the k = 0 mode of slab 0 is zeroed.
 * nbins 0 (logkk, ratio NULL) applies the Green's function alone.
 * asmth: the PM force split scale in box length units.  No host sync. */
int kn_rescale_green_device(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                            const double *d_logkk, const double *d_ratio, int nbins, double norm, double asmth,
                            void *stream);

/* Last GPU error text (empty if none). */
const char *kn_last_error(void);

/* ---- adjoint of (a): d/dF of total_powerspectrum's output (powerspectrum.c:33-117) ----
 * gbar[nout] = dL/dpower[n] for the nout compacted bins the forward returns on
 * the same field (host array).  d_field_bar (device, the field's layout)
 * receives dL/dRe F + i dL/dIm F:
 *   2 c_m W_m^-4 gbar_b F_m / (|F_0|^2 N_b) for m != 0 (c_m the multiplicity),
 *   -2 F_0 / |F_0|^2 sum_n gbar_n power_n   for the zero mode.
 * Needs the zero mode (startslab 0).  Synchronous on stream. */
int kn_bin_power_adjoint(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                         const double *gbar, int nout, fftw_complex *d_field_bar, void *stream);

/* ---- (b) one history of a group: what one get_delta_nu call integrates ----
 * (delta_tot_table.c:507-611): its stored log a values and delta_tot rows,
 * its own H(a), mass and hybrid setting.  qc, relerr and ic_only are what
 * get_delta_nu derives from a, the stored step count and the mass
 * (delta_tot_table.c:513-555); kn_batch_history_from_table fills them, and a
 * group derives them again at create, append and set_mass from a, na, mnu and
 * the hybrid fields (hybrid, nu_crit_time, nufrac_low, vcrit), so values
 * passed in are not used. */
typedef struct kn_batch_history {
    int nk;                 /* k bins */
    int na;                 /* stored steps now (>= 1) */
    int na_max;             /* capacity */
    double a;               /* current scale factor: upper limit log a */
    double a_max;           /* largest a the history will reach (H table range) */
    double time_transfer;   /* TimeTransfer: lower limit log T0 */
    double mnu, kBtnu;      /* neutrino mass (eV) and k_B T_nu (eV) */
    double light;           /* speed of light in internal units */
    double delta_nu_prefac; /* 1.5 Omega0 H0^2 / c */
    double qc, relerr;      /* derived: hybrid critical momentum (0: off) and QAG tolerance */
    double nufrac_low;      /* hybnu.nufrac_low[0] */
    double vcrit;           /* hybnu.vcrit (qc = vcrit m / kT once particles are active) */
    int hybrid;             /* hybnu.enabled */
    double nu_crit_time;    /* hybnu.nu_crit_time: particles active for a > nu_crit_time */
    int ic_only;            /* derived: 1 when only the initial-condition piece is computed */
    double (*hubble)(double a, void *ctx);
    void *hubble_ctx;
} kn_batch_history;

/* The history get_delta_nu(d, a, ..., mnu) integrates (delta_tot_table.c:507-555:
 * tolerance, hybrid qc, IC-only cases), with na = na_max = d->ia and the
 * process's H: the background set with kn_set_background
 * (kn_background_hubble_ctx), else its hubble_function (kn_hubble_ctx). */
int kn_batch_history_from_table(const _delta_tot_table *d, double a, double mnu, kn_batch_history *out);
/* The process's hubble_function (or the one registered with
 * kn_set_hubble_function) in the kn_batch_history callback form. */
double kn_hubble_ctx(double a, void *ctx);

/* ---- (f)4: the background, Omega_nu(a) and H(a), tabulated on the device ----
 * The reference computes Omega_nu(a) on the host (omega_nu_single.c:16-205)
 * and takes H from the N-body code's hubble_function (gadget-2 patch
 * 0003:38-55), which the integrals sample on the host.  A kn_background holds
 * both on the device:
 *   - the density tables of rho_nu_init (omega_nu_single.c:117-153) for each
 *     distinct mass: 200 log-a points from a0/1.2 to 120 kT/m, each a
 *     QAG-GK61 Fermi-Dirac energy integral at relative error 1e-9 (one
 *     wavefront each), and their natural cubic spline;
 *   - Omega_nu(a) as get_omega_nu (:55-67), with the series and
 *     relativistic limits of rho_nu (:156-205);
 *   - H(a) = Hubble sqrt(Omega0/a^3 + OmegaLambda + (1 - OmegaLambda - Omega0)/a^2
 *            + Omega_nu(a) - Omega_nu(1)/a^3 + Omega_g(a) if radiation);
 * and builds the Chebyshev panels of 1/(a^2 H) and 1/(a H) that the device
 * integrands read (split where H jumps) on the device.  A group history or
 * the per-call path whose H is a background never calls the host's
 * hubble_function.  Return codes: 0, 2001 (a density integral failed), 3002
 * bad arguments. */
typedef struct kn_background_params {
    double Omega0;          /* All.Omega0: matter, neutrinos included */
    double OmegaLambda;     /* All.OmegaLambda */
    double Hubble;          /* All.Hubble: H0 in internal time units */
    double HubbleParam;     /* h */
    double tcmb0;           /* CMB temperature today, K */
    double a0;              /* lowest a of the density tables (init_omega_nu's a0) */
    double MNu[NUSPECIES];  /* neutrino masses, eV */
    int radiation;          /* 1: photons, Omega_g(a) = get_omegag (omega_nu_single.c:77-81) */
} kn_background_params;
typedef struct kn_background kn_background;
int kn_background_create(kn_background **out, const kn_background_params *p);
void kn_background_destroy(kn_background *bg);
/* Omega_nu(a) and H(a) at n scale factors: device arrays, on stream. */
int kn_background_omega_nu(const kn_background *bg, int n, const double *d_a, double *d_out, void *stream);
int kn_background_hubble(const kn_background *bg, int n, const double *d_a, double *d_out, void *stream);
/* H(a) on the host from the same tables, in the kn_batch_history callback
 * form (ctx: the kn_background).  A group history whose hubble is this
 * function gets its H panels built on the device. */
double kn_background_hubble_ctx(double a, void *ctx);
/* Species i's density table: host copies of its log a and rho_nu values
 * (200 each).  Returns the table length, 0 when species i has no table (a
 * mass folded into an equal one, or one the limits cover), -3002 for bad
 * arguments. */
int kn_background_rho_table(const kn_background *bg, int i, double *loga, double *rhonu);
/* The process's H becomes bg's (NULL: the host's hubble_function again):
 * get_delta_nu's H panels are then built on the device, and the library's
 * host uses of H (the initial-condition piece) take bg's host form.  bg must
 * outlive its use. */
int kn_set_background(const kn_background *bg);

/* ---- (b) grouped: many get_delta_nu integrals on the per-call kernels ----
 * No reference counterpart (the reference evaluates one table per
 * get_delta_nu call, delta_tot_table.c:507-611); additive.  Histories
 * (kn_batch_history) live in device memory and are solved with the kernels of
 * get_delta_nu, either in merged launches (the default: one grid per phase
 * covering every history, so the launch count does not grow with the number
 * of histories) or with each history's launch sequence on one of up to 8 HIP
 * streams forked from the caller's stream (longest histories first, each to
 * the least-loaded stream), so the small per-history grids overlap.  Every
 * solve records its final QAG subdivisions; kn_group_adjoint differentiates
 * the last solve with them held fixed.  Return codes: 0, 2001 (QAG or
 * interpolation failure, the history is named in kn_last_error), 3002 bad
 * arguments, 3004 capacity. */
typedef struct kn_group kn_group;
int kn_group_create(kn_group **out, int nhist, const kn_batch_history *hist);
void kn_group_destroy(kn_group *g);
int kn_group_views(kn_group *g, double **scalefact, double **delta_tot, double **wavenum, double **delta_nu_init,
                   double **delta_nu, double **gbar, double **dtot_bar, double **init_bar, double **mnu_bar);
int kn_group_layout(const kn_group *g, long long *off_k, long long *off_y, long long *off_s);
int kn_group_solve(kn_group *g, void *stream);   /* synchronous */
int kn_group_adjoint(kn_group *g, void *stream); /* synchronous */
/* One more stored step per history (delta_tot_table.c:177-191, the step
 * get_delta_nu_update keeps): d_loga[nhist] (the new log a) and
 * d_rows[items] (the new delta_tot column, at each history's off_k) in device
 * memory.  The device checks every history (capacity, the H table range,
 * built up to max(a, a_max), and a log a after the last stored one) before
 * anything is written: 3004 or 3002 then, with the history named.  It also
 * computes the new fslength(T0, a) and re-derives tolerance, qc and the
 * IC-only state as get_delta_nu does (delta_tot_table.c:513-555). */
int kn_group_append(kn_group *g, const double *d_loga, const double *d_rows, void *stream);
/* Launch sizes for na_max stored steps (on = 1) or for the stored steps (0,
 * the default).  On the capacity plan the grids cover na_max steps and both
 * the integral's and the IC-only phases, and the kernels read from device
 * memory what applies, so device appends leave the launches (and a captured
 * graph) valid; merged launches only (kn_group_set_streams(g, n > 0) is 3002). */
int kn_group_set_capacity_plan(kn_group *g, int on);
/* kn_group_append with no host synchronisation, on the capacity plan (3002
 * otherwise): one launch checks and appends every history; a history whose
 * check fails is left as it was and its code is reported by the next
 * kn_group_status. */
int kn_group_append_async(kn_group *g, const double *d_loga, const double *d_rows, void *stream);
/* Capture [append +] solve [+ adjoint] on a non-null stream.  Without append
 * (d_loga = d_rows = NULL) the graph is replayable until the next append or
 * autotune (new masses and new values in the device arrays are read at
 * replay).  With d_loga and d_rows the group goes on its capacity plan and
 * every replay first appends the step held in those arrays at replay time,
 * so one graph serves every PM step until a history is full. */
int kn_group_graph_capture(kn_group *g, int with_adjoint, const double *d_loga, const double *d_rows, void *stream);
int kn_group_graph_launch(kn_group *g, void *stream);
int kn_group_status(kn_group *g);
/* History h's final QAG subdivisions of the last solve (held fixed by the
 * adjoint), in QUADPACK's interval-list order: ivl[(k * 200 + t) * 2 + {0,1}]
 * for t < nint[k]; nint is 0 for initial-condition-only items.  Host arrays of
 * nk * 400 doubles and nk ints. */
int kn_group_subdivision(kn_group *g, int h, double *ivl, int *nint);
int kn_group_set_mass(kn_group *g, int h, double mnu);
/* (f)2: history h's stored steps to a host table (delta_tot rows, scalefact,
 * delta_nu_init, wavenum, nk and ia set; delta_nu_last gets the last solve's
 * delta_nu when allocated) and from one (same nk, at most na_max steps,
 * integrated to a >= the last stored step from then on, re-derived as after
 * an append), so that the reference's state files (save_all_nu_state /
 * read_all_nu_state, delta_tot_table.c:254-352) and get_nu_state /
 * set_nu_state (interface_common.c:150-180) checkpoint a device-resident
 * group.  3002 for a table that does not fit or a non-increasing history,
 * 3004 for a outside [last stored step, the H table's end]. */
int kn_group_export(kn_group *g, int h, _delta_tot_table *d);
int kn_group_import(kn_group *g, int h, const _delta_tot_table *d, double a);
/* Pick the launch strategy by timing the forward: merged launches (0) or
 * per-history sequences on 1, 2, 4 or 8 streams. */
int kn_group_autotune(kn_group *g, void *stream);
/* The strategy in use: 0 merged launches, n > 0 per-history sequences on n
 * streams.  Setting it (0..8; 3002 otherwise) makes a captured graph stale. */
int kn_group_streams(const kn_group *g);
int kn_group_set_streams(kn_group *g, int nstream);

/* ---- adjoint of get_delta_nu (delta_tot_table.c:507-611) ----
 * gbar[nk] = dL/d delta_nu.  Outputs (host arrays, any may be NULL): the
 * forward delta_nu[nk], dL/d delta_tot[k][0..ia) (row stride ia), dL/d
 * delta_nu_init[nk] and dL/d mnu (the function's mnu argument; the hybrid
 * state of the table held fixed).  The QAG subdivision is that of the
 * forward evaluation, so the result is the exact derivative of the computed
 * quadrature.  Returns 0, or 2001 when the forward's QAG fails. */
int kn_get_delta_nu_adjoint(const _delta_tot_table *d, double a, const double *wavenum, double mnu,
                            const double *gbar, double *delta_nu, double *dtot_bar, double *init_bar,
                            double *mnu_bar);

#ifdef __cplusplus
}
#endif
#endif

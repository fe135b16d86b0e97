/* interface.c -- module-level state and the PM-step entry points.
 *
 * Same API and flow as interface_common.c:16-197 and interface_gadget.c:75-194
 * of the reference, for one rank:
 *   add_nu_power_to_rhogrid: (a) bin the grid (GPU) -> delta_cdm = sqrt(P L^3),
 *   k *= 2 pi / L -> get_delta_nu_update (GPU integrator) -> ratio
 *   delta_nu/delta_cdm in log k with norm kspace_prefac -> (c) rescale (GPU).
 * The grid stays in HBM between the two sweeps; only the nrbins-long
 * spectra cross PCIe.  With a host grid the data is staged once in and out.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "kn_internal.h"

struct __kspace_params kspace_params;

static _transfer_init_table transfer_init;
_delta_tot_table delta_tot_table;
static _omega_nu omeganu_table;
double *delta_cdm_curr;

/* device helpers (powerspectrum.hip, rescale.hip) */
void kn_bin_power_host_sums(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                            double *power, double *total_mass2);
void kn_bin_geometry_host(int dims, int nrbins, int startslab, int nslab, long long *count, double *keff_sum);
int kn_finish_powerspectrum(int nrbins, double total_mass2, double *power, long long *count, double *keffs);
fftw_complex *kn_grid_to_device(fftw_complex *grid, size_t nmodes, int *staged);
void kn_grid_release(fftw_complex *host, fftw_complex *dev, size_t nmodes, int staged, int copy_back);
int kn_rescale_from_host_table(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                               const double *logkk, const double *ratio, int nbins, double norm);

double OmegaNu(double a) { return get_omega_nu(&omeganu_table, a); }

double OmegaNu_nopart(double a) { return get_omega_nu_nopart(&omeganu_table, a); }

void InitOmegaNu(const double HubbleParam, const double tcmb0, MPI_Comm comm)
{
    (void)comm;
    init_omega_nu(&omeganu_table, kspace_params.MNu, kspace_params.TimeTransfer, HubbleParam, tcmb0);
}

void allocate_kspace_memory(const int nk_in, const int ThisTask, const double BoxSize, const double UnitTime_in_s,
                            const double UnitLength_in_cm, const double Omega0, char *snapdir, const double TimeMax,
                            MPI_Comm comm)
{
    (void)comm;
    if (kspace_params.hybrid_neutrinos_on)
        init_hybrid_nu(&omeganu_table.hybnu, kspace_params.MNu, kspace_params.vcrit, KN_LIGHTCGS / 1e5,
                       kspace_params.nu_crit_time, omeganu_table.kBtnu);
    allocate_transfer_init_table(&transfer_init, BoxSize, UnitLength_in_cm, kspace_params.InputSpectrum_UnitLength_in_cm,
                                 kspace_params.KspaceTransferFunction);
    delta_tot_table.ThisTask = ThisTask;
    allocate_delta_tot_table(&delta_tot_table, nk_in, kspace_params.TimeTransfer, TimeMax, Omega0, &omeganu_table,
                             UnitTime_in_s, UnitLength_in_cm, 0);
    /* resume from a saved history when the caller names one (interface_common.c:93-95) */
    if (ThisTask == 0 && snapdir != NULL) read_all_nu_state(&delta_tot_table, snapdir);
    delta_cdm_curr = kn_malloc("temp_power_spectrum", 3 * (size_t)nk_in * sizeof(double));
}

/* interface_common.c:41-52: the module table's state and P_nu(k) files (rank 0 writes) */
void save_nu_state(char *savefile)
{
    if (delta_tot_table.ThisTask == 0) save_all_nu_state(&delta_tot_table, savefile);
}

int save_neutrino_power(const double Time, const int snapnum, const char *OutputDir)
{
    if (delta_tot_table.ThisTask != 0) return 0;
    return save_nu_power(&delta_tot_table, Time, snapnum, OutputDir);
}

/* interface_common.c:125-148 */
static _delta_pow compute_neutrino_power_internal(const double Time, double *keff, double *dcdm, double *dnu,
                                                  const int nk_nonzero)
{
    _delta_pow d_pow;
    int i;
    get_delta_nu_update(&delta_tot_table, Time, nk_nonzero, keff, dcdm, dnu, &transfer_init);
    kn_message(0, "neutrino power ready: %d k bins; k=%g: delta_nu=%g, delta_cdm=%g\n", nk_nonzero,
               keff[1], dnu[1], dcdm[1]);
    for (i = 0; i < nk_nonzero; i++) {
        keff[i] = log(keff[i]);
        dcdm[i] = dnu[i] / dcdm[i];
    }
    {
        const double OmegaNu_nop = get_omega_nu_nopart(&omeganu_table, Time);
        const double omega_hybrid = get_omega_nu(&omeganu_table, Time) - OmegaNu_nop;
        const double kspace_prefac = OmegaNu_nop / (delta_tot_table.Omeganonu / pow(Time, 3) + omega_hybrid);
        init_delta_pow(&d_pow, keff, dcdm, nk_nonzero, kspace_prefac);
    }
    return d_pow;
}

/* interface_common.c:106-123 */
_delta_pow compute_neutrino_power_from_cdm(const double Time, const double keff_in[], const double P_cdm[],
                                           const long int Nmodes[], const int nk_in, MPI_Comm comm)
{
    double *dnu = delta_cdm_curr + nk_in;
    double *keff = delta_cdm_curr + 2 * nk_in;
    int i, n = 0;
    (void)comm;
    for (i = 0; i < nk_in; i++) {
        if (Nmodes[i] == 0) continue;
        delta_cdm_curr[n] = sqrt(P_cdm[i]);
        keff[n] = keff_in[i];
        n++;
    }
    return compute_neutrino_power_internal(Time, keff, delta_cdm_curr, dnu, n);
}

/* interface_gadget.c:75-102 with the binning of a device-resident grid */
static _delta_pow compute_neutrino_power_spectrum(const double Time, const double BoxSize, const fftw_complex *d_grid,
                                                  const int pmgrid, int slabstart, int nslab)
{
    const int nk_alloc = delta_tot_table.nk_allocated;
    double *dnu = delta_cdm_curr + nk_alloc;
    double *keff = delta_cdm_curr + 2 * nk_alloc;
    long long *count = kn_malloc("temp_modecount", (size_t)nk_alloc * sizeof(long long));
    const double scale = pow(BoxSize, -3);
    double total_mass2;
    int i, nk_in;
    kn_bin_power_host_sums(d_grid, pmgrid, nk_alloc, slabstart, nslab, delta_cdm_curr, &total_mass2);
    kn_bin_geometry_host(pmgrid, nk_alloc, slabstart, nslab, count, keff);
    nk_in = kn_finish_powerspectrum(nk_alloc, total_mass2, delta_cdm_curr, count, keff);
    kn_free(count);
    for (i = 0; i < nk_in; i++) {
        delta_cdm_curr[i] = sqrt(delta_cdm_curr[i] / scale);
        keff[i] *= (2 * M_PI / BoxSize);
    }
    return compute_neutrino_power_internal(Time, keff, delta_cdm_curr, dnu, nk_in);
}

/* interface_gadget.c:158-194 */
void add_nu_power_to_rhogrid(const double Time, const double BoxSize, fftw_complex *fft_of_rhogrid, const int pmgrid,
                             int slabstart_y, int nslab_y, MPI_Comm comm)
{
    const size_t nmodes = (size_t)nslab_y * (size_t)pmgrid * (size_t)(pmgrid / 2 + 1);
    int staged = 0;
    fftw_complex *d_grid = kn_grid_to_device(fft_of_rhogrid, nmodes, &staged);
    _delta_pow d_pow;
    (void)comm;
    d_pow = compute_neutrino_power_spectrum(Time, BoxSize, d_grid, pmgrid, slabstart_y, nslab_y);
    if (kn_rescale_from_host_table(d_grid, pmgrid, slabstart_y, nslab_y, BoxSize, d_pow.logkk, d_pow.delta_ratio,
                                   d_pow.nbins, d_pow.norm))
        kn_terminate(5, "NaN in the neutrino/CDM power ratio\n");
    kn_grid_release(fft_of_rhogrid, d_grid, nmodes, staged, 1);
    kn_message(0, "neutrino rescale of the grid done\n");
    free_d_pow(&d_pow);
}

/* interface_common.c:150-180 */
void get_nu_state(double **scalefact, double **delta_tot, size_t *nk, size_t *ia)
{
    size_t i, k;
    *nk = (size_t)delta_tot_table.nk;
    *ia = (size_t)delta_tot_table.ia;
    *scalefact = kn_malloc("tmp_scales", (*ia) * sizeof(double));
    *delta_tot = kn_malloc("tmp_delta", (*nk) * (*ia) * sizeof(double));
    for (i = 0; i < *ia; i++) (*scalefact)[i] = delta_tot_table.scalefact[i];
    for (k = 0; k < *nk; k++)
        for (i = 0; i < *ia; i++) (*delta_tot)[k * (*ia) + i] = delta_tot_table.delta_tot[k][i];
}

void set_nu_state(double *scalefact, double *delta_tot, const size_t nk, const size_t ia, MPI_Comm comm)
{
    size_t i, k;
    (void)comm;
    delta_tot_table.nk = (int)nk;
    delta_tot_table.ia = (int)ia;
    for (i = 0; i < ia; i++) delta_tot_table.scalefact[i] = scalefact[i];
    for (k = 0; k < nk; k++)
        for (i = 0; i < ia; i++) delta_tot_table.delta_tot[k][i] = delta_tot[k * ia + i];
}

int particle_nu_active(double a) { return particle_nu_fraction(&omeganu_table.hybnu, a, 0) == 0. ? 0 : 1; }

// dnu_device.h -- device building blocks of the free-streaming integral (b).
//
// Included by delta_nu.hip.  The integrand, interpolants and adaptive rule
// follow delta_tot_table.c:375-611 of the reference:
//   delta_nu(k) = IC(k) + delta_nu_prefac * int_{log T0}^{log a} I_k(x) dx,
//   I_k(x) = fsl(x) / (e^x H(e^x)) * specialJ(k fsl(x) / (m/kT), qc, nufrac_low) * S_k(x),
// with fsl a natural cubic spline through 16 Na fslength values on a uniform
// log-a grid and S_k a natural cubic spline (linear for Na = 2) through the
// stored delta_tot[k][.], integrated by QUADPACK QAG with the 61-point
// Gauss-Kronrod rule (relative tolerance 1e-6, at most 200 subintervals).
// One wavefront runs one QAG: its lanes evaluate the 61 nodes of both halves
// of the bisected interval in one pass; the interval list lives in LDS.
#ifndef KN_DNU_DEVICE_H
#define KN_DNU_DEVICE_H

#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include <vector>

#include "kn_gk61.h"

namespace kn_dnu {

constexpr int kLimit = 200;   // QAG subinterval limit (GSL_VAL)
constexpr int kHnc = 10;      // Chebyshev coefficients per function per H panel
constexpr int kHStride = 2 * kHnc;  // one panel: 1/(a^2 H) coefficients, then 1/(a H)
// A panel whose fit misses H by more than this (relative, at test points
// between the interpolation nodes) holds a jump or a kink of H: it is split
// there into two fits, kept out of line (at most kHSplitMax per table).
constexpr double kHSplitTol = 1e-12;
constexpr int kHSplitMax = 4;
// One piece of a split panel: centre and scale of its own t = (x - centre) *
// scale, then both functions' coefficients.
constexpr int kHPiece = 2 + kHStride;
// One split record: the split point, the piece left of it, the piece right of it.
constexpr int kHSplitStride = 1 + 2 * kHPiece;

__constant__ double c_xgk[31];
__constant__ double c_wgk[31];
__constant__ double c_wg[15];

// ---- H(a) factors as functions of x = ln a: Chebyshev panels of width 1/inv_w from origin x0 ----
// Each panel holds the coefficients of 1/(a^2 H) (the fslength integrand,
// delta_tot_table.c:378-383) and of 1/(a H) (the delta_nu integrand's factor,
// :492-500), so neither integrand forms an exponential.  H may be only
// piecewise smooth (the host's neutrino density switches from its spline
// table to the non-relativistic series at a = 100 kT/m, omega_nu_single.c:
// 171-183, where H jumps by ~1e-8 relative): a panel that holds such a point
// is split there into two fits, stored after the npanel regular panels; the
// table lists the split panels' indices, so the regular evaluation loads
// nothing more than before.
struct HTab {
    const double *coef;  // npanel * kHStride, then nsplit * kHSplitStride
    double x0, inv_w;
    int npanel, nsplit;
    int split_p[kHSplitMax];
};

// Both functions' interpolants at the kHnc Chebyshev points of [lo, hi];
// hub(a) = H(a).  Host and device (the device background's panels).
template <typename Hub>
__host__ __device__ inline void h_fit(double lo, double hi, Hub hub, double *coef)
{
    const double half = 0.5 * (hi - lo), mid = lo + half;
    double g2[kHnc], g1[kHnc];
    for (int j = 0; j < kHnc; j++) {
        const double x = mid + half * cos(M_PI * (j + 0.5) / kHnc);
        const double h = hub(exp(x));
        g2[j] = exp(-2.0 * x) / h;
        g1[j] = exp(-x) / h;
    }
    for (int k = 0; k < kHnc; k++) {
        double s2 = 0.0, s1 = 0.0;
        for (int j = 0; j < kHnc; j++) {
            const double c = cos(M_PI * k * (j + 0.5) / kHnc);
            s2 += g2[j] * c;
            s1 += g1[j] * c;
        }
        coef[k] = (k == 0 ? 1.0 : 2.0) * s2 / kHnc;
        coef[kHnc + k] = (k == 0 ? 1.0 : 2.0) * s1 / kHnc;
    }
}

// Clenshaw sum of one function's coefficients at t in [-1, 1].
template <typename P>
__host__ __device__ __forceinline__ double h_clenshaw(P c, double t)
{
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = kHnc - 1; k >= 1; k--) {
        const double b0 = fma(2.0 * t, b1, c[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(t, b1, c[0] - b2);
}

// A split panel's value: the piece left or right of the split point.
template <typename P>
__host__ __device__ __forceinline__ double h_split_eval(P rec, double x, int which)
{
    P piece = rec + 1 + (x >= rec[0] ? kHPiece : 0);
    return h_clenshaw(piece + 2 + which * kHnc, (x - piece[0]) * piece[1]);
}

// The largest relative miss of the fits `coef` of [lo, hi], at the ends and
// half-way between adjacent interpolation nodes.
template <typename Hub>
__host__ __device__ inline double h_fit_err(const double *coef, double lo, double hi, Hub hub)
{
    const double half = 0.5 * (hi - lo), mid = lo + half;
    double worst = 0.0;
    for (int j = -1; j < kHnc; j++) {
        const double t = j < 0 ? -1.0 : (j == kHnc - 1 ? 1.0 : cos(M_PI * (j + 1) / kHnc));
        const double x = mid + half * t;
        const double h = hub(exp(x));
        const double e2 = fabs(h_clenshaw(coef, t) / (exp(-2.0 * x) / h) - 1.0);
        const double e1 = fabs(h_clenshaw(coef + kHnc, t) / (exp(-x) / h) - 1.0);
        worst = fmax(worst, fmax(e2, e1));
    }
    return worst;
}

// One piece [lo, hi] of a split panel into piece[kHPiece].
template <typename Hub>
__host__ __device__ inline void h_piece(double lo, double hi, Hub hub, double *piece)
{
    const double half = 0.5 * (hi - lo);
    piece[0] = lo + half;
    piece[1] = half > 0.0 ? 1.0 / half : 0.0;
    h_fit(lo, hi, hub, piece + 2);
}

// The panel [lo, hi] into coef[kHStride].  Where the fit misses H by
// more than kHSplitTol the panel is also split, into rec[kHSplitStride], at
// the point that bisection finds (the half whose own fit misses more is
// kept, until both halves fit or the bracket is a few units in the last
// place wide).  Returns the fit's miss (above kHSplitTol: rec was written).
template <typename Hub>
__host__ __device__ inline double h_panel(double lo, double hi, Hub hub, double *coef, double *rec)
{
    h_fit(lo, hi, hub, coef);
    const double err = h_fit_err(coef, lo, hi, hub);
    if (err <= kHSplitTol) return err;
    double l = lo, r = hi, tmp[kHStride];
    for (;;) {
        const double m = l + 0.5 * (r - l);
        if (!(m > l && m < r)) break;
        h_fit(l, m, hub, tmp);
        const double el = h_fit_err(tmp, l, m, hub);
        h_fit(m, r, hub, tmp);
        const double er = h_fit_err(tmp, m, r, hub);
        if (el <= kHSplitTol && er <= kHSplitTol) {
            l = r = m;
            break;
        }
        if (el >= er) r = m;
        else l = m;
    }
    const double xs = l + 0.5 * (r - l);
    rec[0] = xs;
    h_piece(lo, xs, hub, rec + 1);
    h_piece(xs, hi, hub, rec + 1 + kHPiece);
    return err;
}

// Host: split records of one table, at most kHSplitMax, the panels whose
// fit misses most kept (the others keep their one-piece fit).
struct HSplits {
    int n = 0;
    int p[kHSplitMax];
    double err[kHSplitMax];
    double rec[kHSplitMax][kHSplitStride];
    // offers panel q's record; returns whether it was kept
    bool offer(int q, double e, const double *r)
    {
        int slot = n;
        if (n == kHSplitMax) {
            slot = 0;
            for (int i = 1; i < n; i++)
                if (err[i] < err[slot]) slot = i;
            if (err[slot] >= e) return false;
        } else {
            n++;
        }
        p[slot] = q;
        err[slot] = e;
        for (int i = 0; i < kHSplitStride; i++) rec[slot][i] = r[i];
        return true;
    }
};

// Host: the whole table of npanel panels of width w from x0, then the split
// records; H.split_p and H.nsplit from them.
template <typename Hub>
inline void h_table(double x0, double w, int npanel, Hub hub, std::vector<double> &coef, HTab &H)
{
    coef.assign((size_t)npanel * kHStride, 0.0);
    HSplits sp;
    double rec[kHSplitStride];
    for (int p = 0; p < npanel; p++) {
        const double e = h_panel(x0 + p * w, x0 + (p + 1) * w, hub, &coef[(size_t)p * kHStride], rec);
        if (e > kHSplitTol) sp.offer(p, e, rec);
    }
    for (int i = 0; i < sp.n; i++) coef.insert(coef.end(), sp.rec[i], sp.rec[i] + kHSplitStride);
    H.x0 = x0;
    H.inv_w = 1.0 / w;
    H.npanel = npanel;
    H.nsplit = sp.n;
    for (int i = 0; i < kHSplitMax; i++) H.split_p[i] = i < sp.n ? sp.p[i] : -1;
}

// Evaluation of one function (which = 0: 1/(a^2 H), 1: 1/(a H)) at x.
__host__ __device__ __forceinline__ double h_cheb(const double *coef, double x0, double inv_w, int npanel, int nsplit,
                                                  const int *split_p, double x, int which)
{
    const double u = (x - x0) * inv_w;
    int p = (int)floor(u);
    p = p < 0 ? 0 : (p >= npanel ? npanel - 1 : p);
    for (int s = 0; s < nsplit; s++)
        if (p == split_p[s]) return h_split_eval(coef + (size_t)npanel * kHStride + (size_t)s * kHSplitStride, x, which);
    return h_clenshaw(coef + (size_t)p * kHStride + which * kHnc, 2.0 * (u - p) - 1.0);
}

// Device form: the table lives in global memory, so its loads are global
// loads even where the pointer came from a descriptor in memory.  A lane in
// a split panel (index compared with the table's list, no load) takes the
// out-of-line pieces.
typedef const __attribute__((address_space(1))) double gconst_double;

// x's panel p (clamped to the table), its split record s (-1: none) and the
// local coordinate t in [-1, 1]; one expression for every evaluation form.
struct HLoc {
    int p, s;
    double t;
};
__device__ __forceinline__ HLoc h_locate(const HTab &H, double x)
{
    const double u = (x - H.x0) * H.inv_w;
    int p = (int)floor(u);
    p = p < 0 ? 0 : (p >= H.npanel ? H.npanel - 1 : p);
    int s = -1;
    if (H.nsplit > 0)  // the table's (uniform): most tables have no split panel
#pragma unroll
        for (int i = 0; i < kHSplitMax; i++)
            if (i < H.nsplit && p == H.split_p[i]) s = i;
    HLoc L;
    L.p = p;
    L.s = s;
    L.t = 2.0 * (u - p) - 1.0;
    return L;
}

__device__ __forceinline__ gconst_double *h_split_rec(const HTab &H, int s)
{
    return (gconst_double *)(H.coef + (size_t)H.npanel * kHStride + (size_t)s * kHSplitStride);
}

__device__ __forceinline__ double h_cheb_dev(const HTab &H, double x, int which)
{
    const HLoc L = h_locate(H, x);
    if (H.nsplit > 0 && L.s >= 0) return h_split_eval(h_split_rec(H, L.s), x, which);
    return h_clenshaw((gconst_double *)(H.coef + (size_t)L.p * kHStride + which * kHnc), L.t);
}

__device__ __forceinline__ double inv_a2h(const HTab &H, double x) { return h_cheb_dev(H, x, 0); }
__device__ __forceinline__ double inv_ah(const HTab &H, double x) { return h_cheb_dev(H, x, 1); }

// ---- the background: Omega_nu(a) and H(a), host and device ----
// The neutrino density of omega_nu_single.c:55-205 (a table of kRhoTab log-a
// points per distinct mass, each a Fermi-Dirac energy integral, interpolated
// by a natural cubic spline; the relativistic and non-relativistic limits
// outside it) and the Hubble function of the N-body code's hook (gadget-2
// patch 0003:38-55, with the kspace neutrinos and, optionally, photons).
constexpr int kRhoTab = 200;     // omega_nu_single.c:14
constexpr double kNuSw = 100.0;  // omega_nu_single.h: the series holds for kT/(a m) < 1/100
constexpr int kNuSpecies = 3;

struct BgDev {
    double Omega0, OmegaLambda, Hubble;  // Hubble: H0 in internal time units
    double rhocrit, kBtnu;               // g cm^-3 (h included); eV
    double conv;                         // (eV/c)^4 -> g cm^-3 for one species (omega_nu_single.c:100-114)
    double og;                           // Omega_g at a = 1 (omega_nu_single.c:77-81)
    double onu1;                         // Omega_nu(1)
    int radiation;                       // add Omega_g(a) (INCLUDE_RADIATION)
    int deg[kNuSpecies];                 // equal masses folded into the first (init_omega_nu)
    int tab[kNuSpecies];                 // 1: species i has a density table
    double mnu[kNuSpecies];
    const double *loga[kNuSpecies], *rho[kNuSpecies], *c2[kNuSpecies];  // kRhoTab each
};

// omega_nu_single.c:89-96: the integrand of rho_nu_init, q^2 eps(q) f0(q).
struct FdEnergy {
    double am, kT;
    __host__ __device__ double operator()(double q) const
    {
        const double eps = sqrt(q * q + am * am);
        const double f0 = 1. / (exp(q / kT) + 1);
        return q * q * eps * f0;
    }
};

// omega_nu_single.c:156-160: heavily non-relativistic, zeta(3), zeta(5) and
// zeta(7) terms.
__host__ __device__ inline double bg_nonrel(double a, double kT, double am, double r2, double conv)
{
    return am * (kT * kT * kT) / (a * a * a * a) *
           (1.5 * 1.202056903159594 + r2 * 45. / 4. * 1.0369277551433704 + 2835. / 32. * r2 * r2 * 1.0083492773819229 +
            80325 / 32. * r2 * r2 * r2 * 1.0020083928260826) *
           conv;
}

// omega_nu_single.c:164-167: heavily relativistic.
__host__ __device__ inline double bg_rel(double a, double kT, double conv)
{
    return 7 * pow(M_PI * kT / a, 4) / 120. * conv;
}

// Natural cubic spline at v (knots x, values y, second-derivative halves c;
// the interval by bisection, as gsl_interp_eval's cspline).
__host__ __device__ inline double bg_spline(const double *x, const double *y, const double *c, int n, double v)
{
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        const int mid = (hi + lo) / 2;
        if (x[mid] > v) hi = mid;
        else lo = mid;
    }
    const double h = x[lo + 1] - x[lo], dy = y[lo + 1] - y[lo], dx = v - x[lo];
    const double b = (dy / h) - h * (c[lo + 1] + 2.0 * c[lo]) / 3.0;
    const double d = (c[lo + 1] - c[lo]) / (3.0 * h);
    return y[lo] + dx * (b + dx * (c[lo] + dx * d));
}

// Natural cubic spline coefficients c[n] of (x, y): the tridiagonal system of
// the interior second derivatives by forward elimination and back
// substitution; work holds 3 (n - 2) doubles.
__host__ __device__ inline void bg_spline_init(const double *x, const double *y, int n, double *c, double *work)
{
    const int m = n - 2;
    for (int i = 0; i < n; i++) c[i] = 0.0;
    if (m <= 0) return;
    double *gam = work, *alp = work + m, *z = work + 2 * m;
    for (int i = 0; i < m; i++) {
        const double hl = x[i + 1] - x[i], hr = x[i + 2] - x[i + 1];
        const double gl = hl != 0.0 ? 1.0 / hl : 0.0, gr = hr != 0.0 ? 1.0 / hr : 0.0;
        const double diag = 2.0 * (hr + hl);
        const double rhs = 3.0 * ((y[i + 2] - y[i + 1]) * gr - (y[i + 1] - y[i]) * gl);
        if (m == 1) {
            c[1] = rhs / diag;
            return;
        }
        alp[i] = i == 0 ? diag : diag - (x[i + 1] - x[i]) * gam[i - 1];
        if (i < m - 1) gam[i] = hr / alp[i];
        z[i] = i == 0 ? rhs : rhs - gam[i - 1] * z[i - 1];
    }
    for (int i = 0; i < m; i++) z[i] = z[i] / alp[i];
    c[m] = z[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = z[i] - gam[i] * c[i + 2];
}

// omega_nu_single.c:171-205: the physical density of species s at a.
__host__ __device__ inline double bg_rho_nu(const BgDev &b, int s, double a)
{
    const double kT = b.kBtnu, am = a * b.mnu[s];
    const double r2 = kT * kT / am / am;
    if (kNuSw * kNuSw * r2 < 1) return bg_nonrel(a, kT, am, r2, b.conv);
    if (am < 1e-6 * kT) return bg_rel(a, kT, b.conv);
    const double la = log(a);
    if (!b.tab[s] || la < b.loga[s][0])
        return am < 1e-4 * kT ? bg_rel(a, kT, b.conv) : bg_nonrel(a, kT, am, r2, b.conv);
    return bg_spline(b.loga[s], b.rho[s], b.c2[s], kRhoTab, la);
}

// omega_nu_single.c:55-67 (get_omega_nu).
__host__ __device__ inline double bg_omega_nu(const BgDev &b, double a)
{
    double rho = 0;
    for (int i = 0; i < kNuSpecies; i++)
        if (b.deg[i] > 0) rho += b.deg[i] * bg_rho_nu(b, i, a);
    return rho / b.rhocrit;
}

// gadget-2 patch 0003:38-55: matter + Lambda + curvature, the neutrinos as
// Omega_nu(a) in place of their share of Omega0/a^3, photons when enabled.
__host__ __device__ inline double bg_hubble(const BgDev &b, double a)
{
    double omega_tot = b.Omega0 / pow(a, 3) + b.OmegaLambda + (1. - b.OmegaLambda - b.Omega0) / (a * a);
    omega_tot += bg_omega_nu(b, a) - b.onu1 / pow(a, 3);
    if (b.radiation) omega_tot += b.og / pow(a, 4);
    return b.Hubble * sqrt(omega_tot);
}

struct BgHub {
    BgDev b;
    __host__ __device__ double operator()(double a) const { return bg_hubble(b, a); }
};

// ---- the split search of h_panel by one wavefront (device-built tables) ----
// Two intervals' fits at once: lanes 0-19 sample H at the two intervals'
// kHnc interpolation nodes while lanes 20-41 sample it at their 11 check
// points (those of h_fit_err); lanes 0-39 then form one coefficient each
// (interval, function, k) and lanes 20-41 the fits' misses.  Every lane gets
// both intervals' largest misses (el, er; with check only).
struct HFit2Lds {
    double g[2][2][kHnc];  // samples: interval, function (1/(a^2 H), 1/(a H)), node
    double c[2][2][kHnc];  // coefficients
    double e[2][kHnc + 1];  // misses at the check points
};

template <typename Hub>
__device__ void h_fit2_wave(double lo0, double hi0, double lo1, double hi1, const Hub &hub, HFit2Lds &L, bool check,
                            double &el, double &er)
{
    const int lane = threadIdx.x & 63;
    double f2 = 0.0, f1 = 0.0, t = 0.0;
    if (lane < 2 * kHnc || (check && lane < 2 * kHnc + 2 * (kHnc + 1))) {
        const int iv = lane < 2 * kHnc ? lane / kHnc : (lane - 2 * kHnc) / (kHnc + 1);
        const double lo = iv ? lo1 : lo0, hi = iv ? hi1 : hi0;
        const double half = 0.5 * (hi - lo), mid = lo + half;
        if (lane < 2 * kHnc) {
            t = cos(M_PI * (lane % kHnc + 0.5) / kHnc);
        } else {
            const int j = (lane - 2 * kHnc) % (kHnc + 1) - 1;
            t = j < 0 ? -1.0 : (j == kHnc - 1 ? 1.0 : cos(M_PI * (j + 1) / kHnc));
        }
        const double x = mid + half * t;
        const double h = hub(exp(x));
        f2 = exp(-2.0 * x) / h;
        f1 = exp(-x) / h;
        if (lane < 2 * kHnc) {
            L.g[iv][0][lane % kHnc] = f2;
            L.g[iv][1][lane % kHnc] = f1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < 4 * kHnc) {
        const int iv = lane / (2 * kHnc), fn = (lane / kHnc) & 1, k = lane % kHnc;
        double s = 0.0;
        for (int j = 0; j < kHnc; j++) s += L.g[iv][fn][j] * cos(M_PI * k * (j + 0.5) / kHnc);
        L.c[iv][fn][k] = (k == 0 ? 1.0 : 2.0) * s / kHnc;
    }
    __builtin_amdgcn_wave_barrier();
    if (!check) return;
    if (lane >= 2 * kHnc && lane < 2 * kHnc + 2 * (kHnc + 1)) {
        const int iv = (lane - 2 * kHnc) / (kHnc + 1), jj = (lane - 2 * kHnc) % (kHnc + 1);
        const double e2 = fabs(h_clenshaw(L.c[iv][0], t) / f2 - 1.0);
        const double e1 = fabs(h_clenshaw(L.c[iv][1], t) / f1 - 1.0);
        L.e[iv][jj] = fmax(e2, e1);
    }
    __builtin_amdgcn_wave_barrier();
    el = 0.0;
    er = 0.0;
    for (int j = 0; j <= kHnc; j++) {
        el = fmax(el, L.e[0][j]);
        er = fmax(er, L.e[1][j]);
    }
    __builtin_amdgcn_wave_barrier();
}

// h_panel's split of [lo, hi] (bisection until both halves fit, then both
// pieces' fits) by one wavefront into rec[kHSplitStride].  Every lane keeps
// the same bracket, so the loop ends together.
template <typename Hub>
__device__ void h_split_wave(double lo, double hi, const Hub &hub, double *rec, HFit2Lds &L)
{
    double l = lo, r = hi, el, er;
    for (;;) {
        const double m = l + 0.5 * (r - l);
        if (!(m > l && m < r)) break;
        h_fit2_wave(l, m, m, r, hub, L, true, el, er);
        if (el <= kHSplitTol && er <= kHSplitTol) {
            l = r = m;
            break;
        }
        if (el >= er) r = m;
        else l = m;
    }
    const double xs = l + 0.5 * (r - l);
    h_fit2_wave(lo, xs, xs, hi, hub, L, false, el, er);
    const int lane = threadIdx.x & 63;
    if (lane < 2 * kHStride) {
        const int iv = lane / kHStride, i = lane % kHStride;
        rec[1 + iv * kHPiece + 2 + i] = L.c[iv][i / kHnc][i % kHnc];
    } else if (lane < 2 * kHStride + 2) {
        const int iv = lane - 2 * kHStride;
        const double plo = iv ? xs : lo, phi = iv ? hi : xs, half = 0.5 * (phi - plo);
        rec[1 + iv * kHPiece] = plo + half;
        rec[1 + iv * kHPiece + 1] = half > 0.0 ? 1.0 / half : 0.0;
    } else if (lane == 2 * kHStride + 2) {
        rec[0] = xs;
    }
    __builtin_amdgcn_wave_barrier();
}

// 1/d from the hardware estimate refined by one Newton step (relative error
// a few units in the last place, against ten instructions for a division);
// d finite and normal.
__device__ __forceinline__ double rcp_fast(double d)
{
    const double r = __builtin_amdgcn_rcp(d);
    return fma(r, fma(-d, r, 1.0), r);
}

// ---- special functions (delta_tot_table.c:417-463) ----
struct JPar {
    double qc, nufrac_low;
    double en[20];  // exp(-n qc) for n = 1..19, from the host
};

__device__ __forceinline__ double sph_j0(double x)
{
    if (fabs(x) < 0.5) {
        const double y = x * x;
        return 1.0 + y * (-1.0 / 6.0 + y * (1.0 / 120.0 + y * (-1.0 / 5040.0 + y * (1.0 / 362880.0 +
                         y * (-1.0 / 39916800.0 + y * (1.0 / 6227020800.0))))));
    }
    return sin(x) / x;
}

// The hybrid series of specialJ (delta_tot_table.c:435-463),
//   sum_{n=1}^{19} (-1)^{n+1} e^{-n qc} [A_n qc j0 + B_n cos(qc x)] / (n^2 + x^2)^2,
//   A_n = n^2 + n^3 qc + n qc x^2 - x^2,  B_n = 2n + n^2 qc + qc x^2,
// with each bracket regrouped by powers of n,
//   A_n qc j0 + B_n c = n^2 al + n^3 be + n ga + de,
//   al = qc j0 + qc c, be = qc^2 j0, ga = qc^2 x^2 j0 + 2c, de = qc x^2 c - qc x^2 j0,
// so that a term is three fused multiply-adds, the squared denominator's
// reciprocal and one more fused multiply-add.  SJTerms holds one point's
// per-x factors; specialJ and specialJ2 run the same operations per point.
struct SJTerms {
    double y, al, be, ga, de, integ;
};
__device__ __forceinline__ SJTerms sj_terms(double x, double qc)
{
    double s, c;
    sincos(qc * x, &s, &c);
    const double j0 = fabs(qc * x) < 0.5 ? sph_j0(qc * x) : s / (qc * x);
    const double y = x * x, qj = qc * j0;
    return {y, qj + qc * c, qc * qj, qc * (y * qj) + 2.0 * c, qc * y * c - y * qj, 0.0};
}
__device__ __forceinline__ void sj_term(SJTerms &t, int n, double w)
{
    const double fn = (double)n, n2 = fn * fn, n3 = n2 * fn;
    const double ii = fma(n2, t.al, fma(n3, t.be, fma(fn, t.ga, t.de)));
    const double d = n2 + t.y;
    t.integ = fma(w * rcp_fast(d * d), ii, t.integ);
}
__device__ __forceinline__ double sj_fit(double x)
{
    if (x <= 0.) return 1.;
    const double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    return (1. + 0.0168 * x2 + 0.0407 * x4) / (1. + 2.1734 * x2 + 1.6787 * exp(4.1811 * log(x)) + 0.1467 * x8);
}

__device__ __forceinline__ double specialJ(double x, const JPar &P)
{
    if (P.qc > 0) {
        SJTerms t = sj_terms(x, P.qc);
#pragma unroll 2
        for (int n = 1; n < 20; n++) sj_term(t, n, ((n & 1) ? 1.0 : -1.0) * P.en[n]);
        return t.integ / (1.5 * 1.202056903159594 * (1 - P.nufrac_low));
    }
    return sj_fit(x);
}

// specialJ at two points, the two series in one loop (two independent
// chains per term); each point's operations are specialJ's.
__device__ __forceinline__ void specialJ2(double x1, double x2, const JPar &P, double &j1, double &j2)
{
    if (P.qc > 0) {
        SJTerms t1 = sj_terms(x1, P.qc), t2 = sj_terms(x2, P.qc);
#pragma unroll 2
        for (int n = 1; n < 20; n++) {
            const double w = ((n & 1) ? 1.0 : -1.0) * P.en[n];
            sj_term(t1, n, w);
            sj_term(t2, n, w);
        }
        j1 = t1.integ / (1.5 * 1.202056903159594 * (1 - P.nufrac_low));
        j2 = t2.integ / (1.5 * 1.202056903159594 * (1 - P.nufrac_low));
        return;
    }
    j1 = sj_fit(x1);
    j2 = sj_fit(x2);
}

// ---- natural cubic spline evaluation, GSL cspline/linear semantics ----
__device__ __forceinline__ double cspline_val(double xl, double xr, double yl, double yr, double ci, double ci1,
                                              double x)
{
    const double h = xr - xl, ih = rcp_fast(h);
    const double dy = yr - yl;
    const double dx = x - xl;
    const double b = dy * ih - h * (ci1 + 2.0 * ci) * (1.0 / 3.0);
    const double d = (ci1 - ci) * ih * (1.0 / 3.0);
    return yl + dx * (b + dx * (ci + dx * d));
}

__device__ __forceinline__ double cspline_at(const double *xa, const double *ya, const double *c, int i, double x)
{
    return cspline_val(xa[i], xa[i + 1], ya[i], ya[i + 1], c[i], c[i + 1], x);
}

__device__ __forceinline__ double linear_at(const double *xa, const double *ya, int i, double x)
{
    const double xl = xa[i], h = xa[i + 1] - xl;
    if (h > 0.0) return ya[i] + (x - xl) / h * (ya[i + 1] - ya[i]);
    return ya[i];
}

// The fslength table's knots are x0 + i (xa - x0) / (n - 1.) (delta_tot_table.c:582),
// the expression the fslength kernels store them with: knot i recomputed here
// equals the stored one, so the spline is evaluated without loading knots.
struct UniformKnots {
    double x0, span, den;  // x0, xa - x0, n - 1.
    __device__ __forceinline__ double at(int i) const { return x0 + i * span / den; }
};

// settle() on those knots: i and its interval [xl, xr]
__device__ __forceinline__ int settle_uniform(const UniformKnots &u, int n, int i, double x, double &xl, double &xr)
{
    i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
    xl = u.at(i);
    while (i > 0 && xl > x) xl = u.at(--i);
    xr = u.at(i + 1);
    while (i < n - 2 && xr <= x) {
        xl = xr;
        xr = u.at(++i + 1);
    }
    return i;
}

// Largest i <= n-2 with xa[i] <= x, starting from a guess (GSL bsearch result).
__device__ __forceinline__ int settle(const double *xa, int n, int i, double x)
{
    i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
    while (i > 0 && xa[i] > x) i--;
    while (i < n - 2 && xa[i + 1] <= x) i++;
    return i;
}

// ---- integrands ----
struct FslIntegrand {  // delta_tot_table.c:378-383: 1 / a / (a H(a))
    HTab H;
    __device__ __forceinline__ double operator()(double x) const { return inv_a2h(H, x); }
};

struct DnuIntegrand {  // delta_tot_table.c:492-500
    HTab H;
    const double *fy, *fc;  // fslength spline on the uniform knots fk
    UniformKnots fk;
    int nfs;
    double fx0, fxhi, finv_h;    // first and last knot, 1 / spacing
    const double *sx, *sy, *sc;  // delta_tot spline of this k
    const int *sbucket;
    int ns, nsb, linear;
    double sx0, sxhi, sinv_bw;   // first and last stored step
    double k, mnubykT, k_by_m;  // k_by_m = k / mnubykT
    const JPar *J;  // in device memory: the series weights are read where used
    int *edom;
    __device__ __forceinline__ double operator()(double x) const
    {
        if (x < fx0 || x > fxhi || x < sx0 || x > sxhi) {
            atomicOr(edom, 1);  // GSL_EDOM in gsl_interp_eval -> terminate(2001)
            return 0.0;
        }
        double xl, xr;
        const int fi = settle_uniform(fk, nfs, (int)((x - fx0) * finv_h), x, xl, xr);
        const double fsl = cspline_val(xl, xr, fy[fi], fy[fi + 1], fc[fi], fc[fi + 1], x);
        int q = (int)((x - sx0) * sinv_bw);
        q = q < 0 ? 0 : (q > nsb ? nsb : q);
        const int si = settle(sx, ns, sbucket[q], x);
        const double dtot = linear ? linear_at(sx, sy, si, x) : cspline_at(sx, sy, sc, si, x);
        const double specJ = specialJ(k_by_m * fsl, *J);
        return fsl * inv_ah(H, x) * specJ * dtot;  // fsl / (a H) specJ dtot
    }
    // Two points with operator()'s arithmetic, in straight-line code so that
    // their dependent loads (bucket, knot walk, coefficients, H panel) overlap;
    // use1 / use2 say which results are used.  A used point outside the
    // tables flags GSL_EDOM and gives 0; the indices of every point are
    // clamped into the tables, so an unused or outside point reads in bounds.
    __device__ __forceinline__ void eval2(double x1, double x2, bool use1, bool use2, double &r1, double &r2) const
    {
        const bool in1 = !(x1 < fx0 || x1 > fxhi || x1 < sx0 || x1 > sxhi);
        const bool in2 = !(x2 < fx0 || x2 > fxhi || x2 < sx0 || x2 > sxhi);
        if ((use1 && !in1) || (use2 && !in2)) atomicOr(edom, 1);
        double xl1, xr1, xl2, xr2;
        const int fi1 = settle_uniform(fk, nfs, (int)((x1 - fx0) * finv_h), x1, xl1, xr1);
        const int fi2 = settle_uniform(fk, nfs, (int)((x2 - fx0) * finv_h), x2, xl2, xr2);
        int q1 = (int)((x1 - sx0) * sinv_bw), q2 = (int)((x2 - sx0) * sinv_bw);
        q1 = q1 < 0 ? 0 : (q1 > nsb ? nsb : q1);
        q2 = q2 < 0 ? 0 : (q2 > nsb ? nsb : q2);
        const int b1 = sbucket[q1], b2 = sbucket[q2];
        // the delta_tot spline's values at the bucket's knot and the next two,
        // loaded before the knot walk (which moves at most one knot up but
        // rarely); the walk's result picks the pair
        const int i1 = b1 < 0 ? 0 : (b1 > ns - 2 ? ns - 2 : b1), i2 = b2 < 0 ? 0 : (b2 > ns - 2 ? ns - 2 : b2);
        const int u1 = i1 + 2 < ns ? i1 + 2 : ns - 1, u2 = i2 + 2 < ns ? i2 + 2 : ns - 1;
        const double y10 = sy[i1], y11 = sy[i1 + 1], y12 = sy[u1], y20 = sy[i2], y21 = sy[i2 + 1], y22 = sy[u2];
        double c10 = 0.0, c11 = 0.0, c12 = 0.0, c20 = 0.0, c21 = 0.0, c22 = 0.0;
        if (!linear) {
            c10 = sc[i1], c11 = sc[i1 + 1], c12 = sc[u1];
            c20 = sc[i2], c21 = sc[i2 + 1], c22 = sc[u2];
        }
        const double fsl1 = cspline_val(xl1, xr1, fy[fi1], fy[fi1 + 1], fc[fi1], fc[fi1 + 1], x1);
        const double fsl2 = cspline_val(xl2, xr2, fy[fi2], fy[fi2 + 1], fc[fi2], fc[fi2 + 1], x2);
        const double ih1 = inv_ah(H, x1), ih2 = inv_ah(H, x2);
        const int si1 = settle(sx, ns, b1, x1), si2 = settle(sx, ns, b2, x2);
        double dt1, dt2;
        if (linear) {
            dt1 = linear_at(sx, sy, si1, x1);
            dt2 = linear_at(sx, sy, si2, x2);
        } else {
            dt1 = si1 == i1       ? cspline_val(sx[si1], sx[si1 + 1], y10, y11, c10, c11, x1)
                  : si1 == i1 + 1 ? cspline_val(sx[si1], sx[si1 + 1], y11, y12, c11, c12, x1)
                                  : cspline_at(sx, sy, sc, si1, x1);
            dt2 = si2 == i2       ? cspline_val(sx[si2], sx[si2 + 1], y20, y21, c20, c21, x2)
                  : si2 == i2 + 1 ? cspline_val(sx[si2], sx[si2 + 1], y21, y22, c21, c22, x2)
                                  : cspline_at(sx, sy, sc, si2, x2);
        }
        double j1, j2;
        specialJ2(k_by_m * fsl1, k_by_m * fsl2, *J, j1, j2);
        r1 = in1 ? fsl1 * ih1 * j1 * dt1 : 0.0;
        r2 = in2 ? fsl2 * ih2 * j2 * dt2 : 0.0;
    }
};

// The integrand at two points of one lane: f.eval2 where the integrand has
// it, else two calls.
template <class F>
__device__ __forceinline__ void eval_pair(const F &f, double x1, double x2, bool use1, bool use2, double &r1,
                                          double &r2)
{
    r1 = use1 ? f(x1) : 0.0;
    r2 = use2 ? f(x2) : 0.0;
}
template <>
__device__ __forceinline__ void eval_pair<DnuIntegrand>(const DnuIntegrand &f, double x1, double x2, bool use1,
                                                        bool use2, double &r1, double &r2)
{
    f.eval2(x1, x2, use1, use2, r1, r2);
    r1 = use1 ? r1 : 0.0;
    r2 = use2 ? r2 : 0.0;
}

// ---- wave reductions over 32-lane halves ----
__device__ __forceinline__ double half_sum(double v)
{
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// t^1.5 rounded once: t sqrt(t) carried in two parts (the products' and the
// square root's rounding errors recovered exactly by fma), then summed.  The
// plain t * sqrt(t) rounds twice and differs from the correctly rounded
// value in about a quarter of all t, which could reorder QUADPACK's error
// list against a libm pow(t, 1.5).
__device__ __forceinline__ double pow_three_halves(double t)
{
#pragma clang fp contract(off)  // p must be the rounded product for e to be its error
    const double s = sqrt(t);
    const double r = fma(-s, s, t);  // t - s^2, exact
    const double p = t * s;
    const double e = fma(t, s, -p);  // t s - p, exact
    return p + (e + t * r / (2 * s));
}

__device__ __forceinline__ double gk_rescale(double err, double resabs, double resasc)
{
    err = fabs(err);
    if (resasc != 0 && err != 0) {
        // QUADPACK's rescale: resasc min(1, (200 err / resasc)^1.5)
        const double t = 200 * err / resasc;
        const double s = pow_three_halves(t);
        err = s < 1 ? resasc * s : resasc;
    }
    if (resabs > DBL_MIN / (50 * DBL_EPSILON)) {
        const double floor_err = 50 * DBL_EPSILON * resabs;
        if (floor_err > err) err = floor_err;
    }
    return err;
}

struct GK {
    double res, err, resabs, resasc;
};

// 61-point Gauss-Kronrod on [a1,b1] (lanes 0-31) and [a2,b2] (lanes 32-63).
template <class F>
__device__ __forceinline__ void gk61_pair(const F &f, double a1, double b1, double a2, double b2, GK &g1, GK &g2)
{
#pragma clang fp contract(off)  // every product and sum rounded as written (gk61_q4's sums match bit for bit)
    const int lane = threadIdx.x & 63, hf = lane >> 5, j = lane & 31;
    const double a = hf ? a2 : a1, b = hf ? b2 : b1;
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    // lane 30 takes the centre (c_xgk[30] = 0: c - 0 is c) in the same
    // evaluation as the other lanes' left nodes, so a wavefront runs the
    // integrand twice, not three times
    double f1 = 0.0, f2 = 0.0;
    {
        const double dx = j <= 30 ? h * c_xgk[j] : 0.0;
        eval_pair(f, c - dx, c + dx, j <= 30, j < 30, f1, f2);
    }
    const double wk = j <= 30 ? c_wgk[j] : 0.0;
    const double wgv = (j < 30 && (j & 1)) ? c_wg[j >> 1] : 0.0;
    const double sk = half_sum(j == 30 ? wk * f1 : wk * (f1 + f2));
    const double sg = half_sum(wgv * (f1 + f2));
    const double sa = half_sum(j == 30 ? wk * fabs(f1) : wk * (fabs(f1) + fabs(f2)));
    const double mean = sk * 0.5;
    const double sasc = half_sum(j == 30 ? wk * fabs(f1 - mean) : (j < 30 ? wk * (fabs(f1 - mean) + fabs(f2 - mean)) : 0.0));
    GK r;
    r.res = sk * h;
    r.resabs = sa * fabs(h);
    r.resasc = sasc * fabs(h);
    r.err = gk_rescale((sk - sg) * h, r.resabs, r.resasc);
    // broadcast each half's result to every lane
    g1.res = __shfl(r.res, 0, 64);
    g1.err = __shfl(r.err, 0, 64);
    g1.resabs = __shfl(r.resabs, 0, 64);
    g1.resasc = __shfl(r.resasc, 0, 64);
    g2.res = __shfl(r.res, 32, 64);
    g2.err = __shfl(r.err, 32, 64);
    g2.resabs = __shfl(r.resabs, 32, 64);
    g2.resasc = __shfl(r.resasc, 32, 64);
}

// The first 61-point Gauss-Kronrod estimate on [a, b] by four lanes of a
// wavefront, lanes l, l ^ 16, l ^ 32, l ^ 48 (quarter q = (l >> 4) & 3; the
// other 15 lanes of each quarter take 15 other intervals, so that lanes of
// one load instruction evaluate nearby points of nearby intervals).  Quarter
// q evaluates gk61_pair's lanes j and j + 16 for j = q, q + 4, q + 8, q + 12
// and adds their terms as gk61_pair's butterfly does (xor 16, 8 and 4 within
// the quarter), then xor 2 and xor 1 across quarters (lane xor 32, 16): the
// sums, and so the estimate, are gk61_pair's bit for bit, in all four lanes.
template <class F>
__device__ __forceinline__ GK gk61_q4(const F &f, double a, double b)
{
#pragma clang fp contract(off)  // as gk61_pair
    const int q = (threadIdx.x >> 4) & 3;
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    double f01[4], f02[4], f11[4], f12[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int j0 = q + 4 * r, j1 = j0 + 16;
        const double dx0 = h * c_xgk[j0];
        const double dx1 = j1 <= 30 ? h * c_xgk[j1] : 0.0;
        f01[r] = f(c - dx0);
        f02[r] = f(c + dx0);
        f11[r] = j1 <= 30 ? f(c - dx1) : 0.0;
        f12[r] = j1 < 30 ? f(c + dx1) : 0.0;
    }
    double tk[4], tg[4], ta[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int j0 = q + 4 * r, j1 = j0 + 16;
        const double wk0 = c_wgk[j0], wk1 = j1 <= 30 ? c_wgk[j1] : 0.0;
        const double wg0 = (j0 & 1) ? c_wg[j0 >> 1] : 0.0;
        const double wg1 = (j1 < 30 && (j1 & 1)) ? c_wg[j1 >> 1] : 0.0;
        tk[r] = wk0 * (f01[r] + f02[r]) + (j1 == 30 ? wk1 * f11[r] : wk1 * (f11[r] + f12[r]));
        tg[r] = wg0 * (f01[r] + f02[r]) + wg1 * (f11[r] + f12[r]);
        ta[r] = wk0 * (fabs(f01[r]) + fabs(f02[r])) +
                (j1 == 30 ? wk1 * fabs(f11[r]) : wk1 * (fabs(f11[r]) + fabs(f12[r])));
    }
    // xor 8 (j, j + 8: r, r + 2), xor 4 (j, j + 4: r 0, 1), then across quarters
    auto sum4 = [](const double(&t)[4]) {
        double v = (t[0] + t[2]) + (t[1] + t[3]);
        v += __shfl_xor(v, 32, 64);
        v += __shfl_xor(v, 16, 64);
        return v;
    };
    const double sk = sum4(tk), sg = sum4(tg), sa = sum4(ta);
    const double mean = sk * 0.5;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int j0 = q + 4 * r, j1 = j0 + 16;
        const double wk0 = c_wgk[j0], wk1 = j1 <= 30 ? c_wgk[j1] : 0.0;
        tk[r] = wk0 * (fabs(f01[r] - mean) + fabs(f02[r] - mean)) +
                (j1 == 30 ? wk1 * fabs(f11[r] - mean)
                          : (j1 < 30 ? wk1 * (fabs(f11[r] - mean) + fabs(f12[r] - mean)) : 0.0));
    }
    const double sasc = sum4(tk);
    GK r;
    r.res = sk * h;
    r.resabs = sa * fabs(h);
    r.resasc = sasc * fabs(h);
    r.err = gk_rescale((sk - sg) * h, r.resabs, r.resasc);
    return r;
}

// Interval list of one wavefront (QUADPACK workspace), in LDS.  QUADPACK's
// level counts are left out: only QAGS's extrapolation reads them.
struct QagWs {
    double lo[kLimit], hi[kLimit], r[kLimit], e[kLimit];
    int ord[kLimit];
};

// Keep ord[] partially sorted by error, as QUADPACK's dqpsrt.  All lanes run
// it on identical values.
__device__ __forceinline__ void qag_sort(QagWs &w, int n, int limit, int &nrmax, int &imax)
{
    const int last = n - 1;
    int nr = nrmax, im = w.ord[nr], top, i, k;
    if (last < 2) {
        w.ord[0] = 0;
        w.ord[1] = 1;
        imax = im;
        return;
    }
    const double emax = w.e[im];
    while (nr > 0 && emax > w.e[w.ord[nr - 1]]) {
        w.ord[nr] = w.ord[nr - 1];
        nr--;
    }
    top = last < limit / 2 + 2 ? last : limit - last + 1;
    for (i = nr + 1; i < top && emax < w.e[w.ord[i]]; i++) w.ord[i - 1] = w.ord[i];
    w.ord[i - 1] = im;
    const double emin = w.e[last];
    for (k = top - 1; k > i - 2 && emin >= w.e[w.ord[k]]; k--) w.ord[k + 1] = w.ord[k];
    w.ord[k + 1] = last;
    imax = w.ord[nr];
    nrmax = nr;
}

// qag_sort with nrmax = 0 (QAG never moves it; only QAGS's extrapolation
// does), by all 64 lanes of the wavefront: the two linear scans become
// wavefront searches for the first position where the scan would stop, and
// the shifts become parallel moves (each lane reads its entry before any
// lane writes; a wavefront's LDS accesses complete in program order).  The
// list comes out exactly as qag_sort leaves it.  Returns ord[0].
__device__ __forceinline__ int qag_sort_wave(QagWs &w, int n, int limit)
{
    const int lane = threadIdx.x & 63;
    const int last = n - 1;
    const int im = w.ord[0];
    if (last < 2) {
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            w.ord[0] = 0;
            w.ord[1] = 1;
        }
        return im;
    }
    const double emax = w.e[im];
    __builtin_amdgcn_wave_barrier();  // every lane has read ord[0] before it moves
    const int top = last < limit / 2 + 2 ? last : limit - last + 1;
    // i = first p in [1, top) where !(emax < e[ord[p]]), else top
    int i = top;
    for (int p0 = 1; p0 < top; p0 += 64) {
        const int p = p0 + lane;
        const unsigned long long m = __ballot(p < top && !(emax < w.e[w.ord[p]]));
        if (m) {
            i = p0 + __ffsll((long long)m) - 1;
            break;
        }
    }
    // ord[p - 1] = ord[p] for p in [1, i), then ord[i - 1] = im
    for (int p0 = 1; p0 < i; p0 += 64) {
        const int p = p0 + lane;
        const int v = p < i ? w.ord[p] : 0;
        __builtin_amdgcn_wave_barrier();
        if (p < i) w.ord[p - 1] = v;
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) w.ord[i - 1] = im;
    __builtin_amdgcn_wave_barrier();
    // k = largest p in [i - 1, top - 1] where !(emin >= e[ord[p]]), else i - 2
    const double emin = w.e[last];
    int k = i - 2;
    for (int p0 = top - 1; p0 > i - 2; p0 -= 64) {
        const int p = p0 - lane;
        const unsigned long long m = __ballot(p > i - 2 && !(emin >= w.e[w.ord[p]]));
        if (m) {
            k = p0 - (__ffsll((long long)m) - 1);
            break;
        }
    }
    // ord[p + 1] = ord[p] for p in [k + 1, top - 1], then ord[k + 1] = last
    for (int p0 = top - 1; p0 > k; p0 -= 64) {
        const int p = p0 - lane;
        const int v = p > k ? w.ord[p] : 0;
        __builtin_amdgcn_wave_barrier();
        if (p > k) w.ord[p + 1] = v;
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) w.ord[k + 1] = last;
    __builtin_amdgcn_wave_barrier();
    return w.ord[0];
}

// The same error order held by the wavefront in registers while the list
// is short: lane p keeps ord[p] (os) and that interval's error (es).  While
// fewer than limit / 2 + 2 intervals are listed, dqpsrt's top is the last
// index, so the whole order is kept sorted; the two scans are wavefront
// searches on es and the two shifts are moves between neighbouring lanes.
// ord[0] (im) was just bisected: emax is the error of the half it kept,
// emin that of the half appended at index last.  Returns the new ord[0].
#ifndef KN_WAVE_ORD
#define KN_WAVE_ORD 64
#endif
constexpr int kWaveOrd = KN_WAVE_ORD;  // most intervals held in registers
__device__ __forceinline__ int wave_ord_insert(double &es, int &os, int im, double emax, double emin, int last)
{
    const int p = threadIdx.x & 63;
    if (last < 2) {
        es = p == 0 ? emax : emin;
        os = p;
        return 0;
    }
    const int top = last;
    // i = first p in [1, top) where !(emax < e[ord[p]]), else top
    const unsigned long long m1 = __ballot(p >= 1 && p < top && !(emax < es));
    const int i = m1 ? __ffsll((long long)m1) - 1 : top;
    // ord[p - 1] = ord[p] for p in [1, i), then ord[i - 1] = im
    const int dn = p < 63 ? p + 1 : 63;
    const double es_dn = __shfl(es, dn, 64);
    const int os_dn = __shfl(os, dn, 64);
    if (p < i - 1) {
        es = es_dn;
        os = os_dn;
    } else if (p == i - 1) {
        es = emax;
        os = im;
    }
    // k = largest p in [i - 1, top - 1] where !(emin >= e[ord[p]]), else i - 2
    const unsigned long long m2 = __ballot(p >= i - 1 && p <= top - 1 && !(emin >= es));
    const int k = m2 ? 63 - __builtin_clzll(m2) : i - 2;
    // ord[p + 1] = ord[p] for p in [k + 1, top - 1], then ord[k + 1] = last
    const int up = p > 0 ? p - 1 : 0;
    const double es_up = __shfl(es, up, 64);
    const int os_up = __shfl(os, up, 64);
    if (p > k + 1 && p <= top) {
        es = es_up;
        os = os_up;
    } else if (p == k + 1) {
        es = emin;
        os = last;
    }
    return __shfl(os, 0, 64);
}

// The order after appending interval n - 1: in registers while it fits,
// else (once, when it outgrows them) copied to w.ord and kept there by
// qag_sort_wave.  w holds the bisected interval's new values already.
__device__ __forceinline__ int ord_insert(QagWs &w, double &es, int &os, bool &in_lds, int im, double emax,
                                          double emin, int n, int limit)
{
    if (!in_lds && n <= kWaveOrd && n - 1 < limit / 2 + 2) return wave_ord_insert(es, os, im, emax, emin, n - 1);
    if (!in_lds) {
        const int p = threadIdx.x & 63;
        if (p < n - 1) w.ord[p] = os;
        __builtin_amdgcn_wave_barrier();
        in_lds = true;
    }
    return qag_sort_wave(w, n, limit);
}

// ord[r] (r < the interval count), from wherever the order is kept; every
// lane of the wavefront calls it.
__device__ __forceinline__ int ord_rank(const QagWs &w, int os, bool in_lds, int r)
{
    const int v = __shfl(os, r & 63, 64);
    return in_lds ? w.ord[r] : v;
}

// QAG (QUADPACK dqage semantics, GSL's qag.c control flow) by one wavefront.
// Returns 0 on success or the GSL errno (11 iteration limit, 18 roundoff,
// 21 bad integrand); *result is the sum over the final subintervals.  Every
// lane keeps the same scalar state; lane 0 alone writes the intervals, and
// all lanes keep the error order together (qag_sort_wave), whose wavefront
// operations order lane 0's writes before the other lanes' reads.
template <class F>
__device__ int wave_qag61(const F &f, double a, double b, double epsrel, int limit, QagWs &w, double *result,
                          int *nint)
{
    const bool lead = (threadIdx.x & 63) == 0;
    GK g0, gx;
    gk61_pair(f, a, b, a, b, g0, gx);
    *result = g0.res;
    *nint = 1;
    if (lead) {
        w.lo[0] = a;
        w.hi[0] = b;
        w.r[0] = g0.res;
        w.e[0] = g0.err;
        w.ord[0] = 0;
    }
    int imax = __shfl(0, 0, 64);
    double es = g0.err;  // the error order in registers (ord_insert)
    int os = 0;
    bool ord_lds = false;
    double tol = epsrel * fabs(g0.res);
    if (g0.err <= 50 * DBL_EPSILON * g0.resabs && g0.err > tol) return 18;
    if ((g0.err <= tol && g0.err != g0.resasc) || g0.err == 0.0) return 0;
    if (limit == 1) return 11;
    double area = g0.res, errsum = g0.err;
    int n = 1, it = 1, rt1 = 0, rt2 = 0, etype = 0;
    do {
        const int im = imax;
        const double ai = w.lo[im], bi = w.hi[im], ri = w.r[im], ei = w.e[im];
        const double a1 = ai, b1 = 0.5 * (ai + bi), a2 = b1, b2 = bi;
        GK g1, g2;
        gk61_pair(f, a1, b1, a2, b2, g1, g2);
        const double r12 = g1.res + g2.res, e12 = g1.err + g2.err;
        errsum += e12 - ei;
        area += r12 - ri;
        if (g1.resasc != g1.err && g2.resasc != g2.err) {
            if (fabs(ri - r12) <= 1.0e-5 * fabs(r12) && e12 >= 0.99 * ei) rt1++;
            if (it >= 10 && e12 > ei) rt2++;
        }
        tol = epsrel * fabs(area);
        if (errsum > tol) {
            const double t = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
            if (rt1 >= 6 || rt2 >= 20) etype = 2;
            if (fabs(a1) <= t && fabs(b2) <= t) etype = 3;
        }
        if (lead) {
            if (g2.err > g1.err) {
                w.lo[im] = a2;
                w.r[im] = g2.res;
                w.e[im] = g2.err;
                w.lo[n] = a1;
                w.hi[n] = b1;
                w.r[n] = g1.res;
                w.e[n] = g1.err;
            } else {
                w.hi[im] = b1;
                w.r[im] = g1.res;
                w.e[im] = g1.err;
                w.lo[n] = a2;
                w.hi[n] = b2;
                w.r[n] = g2.res;
                w.e[n] = g2.err;
            }
        }
        const bool hi2 = g2.err > g1.err;
        imax = ord_insert(w, es, os, ord_lds, im, hi2 ? g2.err : g1.err, hi2 ? g1.err : g2.err, n + 1, limit);
        n++;
        it++;
    } while (it < limit && !etype && errsum > tol);
    double s = 0.0;
    for (int k = 0; k < n; k++) s += w.r[k];
    *result = s;
    *nint = n;
    if (errsum <= tol) return 0;
    if (etype == 2) return 18;
    if (etype == 3) return 21;
    if (it == limit) return 11;
    return 5;
}

// ---- QAG by W wavefronts with speculative bisections ----
// wave_qag61's algorithm on one block of W wavefronts.  The chain of
// bisections is serial (each choice waits for the previous estimates), so
// the other wavefronts evaluate ahead: every pass each wavefront evaluates
// the two halves of one interval -- the interval to bisect now, then the two
// halves it will be cut into (one of them is next when the cut leaves it the
// largest error, as in the long chains of oscillatory integrands), then the
// next interval(s) in QUADPACK's error order (ord[]) -- and wavefront 0 runs
// wave_qag61's update for as long as the interval to bisect next has its
// halves evaluated (an interval is known by its bounds: the list holds
// disjoint intervals).  The evaluations are gk61_pair's and the update is
// wave_qag61's, so the decisions and results are wave_qag61's bit for bit;
// an evaluation that the update does not reach is only work lost.
// Estimated errors of the intervals the bisection now in hand will create,
// as fractions of its error, when they compete for a slot with the next
// intervals in error order.  The chains of the oscillatory hybrid integrand
// alternate between the next interval in order and a half of the one just
// bisected, mostly the right half, sometimes twice running.
constexpr double kRightErr = 0.1, kLeftErr = 0.05, kRightRightErr = 0.05;
struct QagSlot {
    int used, eval;    // holds an interval; evaluate it this pass
    double lo, hi;     // the interval
    GK g1, g2;         // its halves
};

template <int W>
struct QagBlock {
    QagWs ws;
    QagSlot slot[W];
    int rc, n;         // return code (-1: running) and interval count, from wavefront 0
};

template <int W, class F>
__device__ int block_qag61(const F &f, double a, double b, double epsrel, int limit, QagBlock<W> &S, double *result,
                           int *nint)
{
    const int wave = threadIdx.x >> 6;
    const bool lead = threadIdx.x == 0;  // lane 0 of wavefront 0 writes the list and the slots
    QagWs &w = S.ws;
    // pass 0: wavefront 0 the first estimate; wavefront 1 the halves of [a, b] (the first bisection)
    GK g0, gx;
    if (wave == 0) {
        gk61_pair(f, a, b, a, b, g0, gx);
    } else if constexpr (W > 1) {
        if (wave == 1) {
            const double m = 0.5 * (a + b);
            GK g1, g2;
            gk61_pair(f, a, m, m, b, g1, g2);
            if ((threadIdx.x & 63) == 0) {
                S.slot[1].used = 1;
                S.slot[1].lo = a;
                S.slot[1].hi = b;
                S.slot[1].g1 = g1;
                S.slot[1].g2 = g2;
            }
        }
    }
    if (lead)
        for (int q = 0; q < W; q++)
            if (q != 1) S.slot[q].used = 0;
    __syncthreads();
    // wavefront 0's scalar state (identical in its 64 lanes)
    double area = 0.0, errsum = 0.0, tol = 0.0;
    int n = 1, it = 1, rt1 = 0, rt2 = 0, etype = 0, imax = 0, rc = -1;
    double es = g0.err;  // the error order in registers (ord_insert)
    int os = 0;
    bool ord_lds = false;
    if (wave == 0) {
        if (lead) {
            w.lo[0] = a;
            w.hi[0] = b;
            w.r[0] = g0.res;
            w.e[0] = g0.err;
            w.ord[0] = 0;
            }
        tol = epsrel * fabs(g0.res);
        if (g0.err <= 50 * DBL_EPSILON * g0.resabs && g0.err > tol) rc = 18;
        else if ((g0.err <= tol && g0.err != g0.resasc) || g0.err == 0.0) rc = 0;
        else if (limit == 1) rc = 11;
        area = g0.res;
        errsum = g0.err;
        imax = __shfl(0, 0, 64);
    }
    for (;;) {
        if (wave == 0) {
            // the bisections whose halves are at hand, exactly as wave_qag61's loop
            double pa = 0.0, pb = 0.0, pe = 0.0;  // the last interval looked at: the next to bisect
            while (rc < 0) {
                const int im = imax;
                const double ai = w.lo[im], bi = w.hi[im], ri = w.r[im], ei = w.e[im];
                pa = ai, pb = bi, pe = ei;
                int s = -1;
#pragma unroll
                for (int q = 0; q < W; q++)
                    if (S.slot[q].used && S.slot[q].lo == ai && S.slot[q].hi == bi) s = q;
                // every lane has read the interval before lane 0 rewrites it below
                __builtin_amdgcn_wave_barrier();
                if (s < 0) break;
                const double a1 = ai, b1 = 0.5 * (ai + bi), a2 = b1, b2 = bi;
                const GK g1 = S.slot[s].g1, g2 = S.slot[s].g2;
                const double r12 = g1.res + g2.res, e12 = g1.err + g2.err;
                errsum += e12 - ei;
                area += r12 - ri;
                if (g1.resasc != g1.err && g2.resasc != g2.err) {
                    if (fabs(ri - r12) <= 1.0e-5 * fabs(r12) && e12 >= 0.99 * ei) rt1++;
                    if (it >= 10 && e12 > ei) rt2++;
                }
                tol = epsrel * fabs(area);
                if (errsum > tol) {
                    const double t = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
                    if (rt1 >= 6 || rt2 >= 20) etype = 2;
                    if (fabs(a1) <= t && fabs(b2) <= t) etype = 3;
                }
                if (lead) {
                            if (g2.err > g1.err) {
                        w.lo[im] = a2;
                        w.r[im] = g2.res;
                        w.e[im] = g2.err;
                        w.lo[n] = a1;
                        w.hi[n] = b1;
                        w.r[n] = g1.res;
                        w.e[n] = g1.err;
                    } else {
                        w.hi[im] = b1;
                        w.r[im] = g1.res;
                        w.e[im] = g1.err;
                        w.lo[n] = a2;
                        w.hi[n] = b2;
                        w.r[n] = g2.res;
                        w.e[n] = g2.err;
                    }
                }
                const bool hi2 = g2.err > g1.err;
                imax = ord_insert(w, es, os, ord_lds, im, hi2 ? g2.err : g1.err, hi2 ? g1.err : g2.err, n + 1, limit);
                n++;
                it++;
                if (!(it < limit && !etype && errsum > tol)) {
                    rc = errsum <= tol ? 0 : etype == 2 ? 18 : etype == 3 ? 21 : it == limit ? 11 : 5;
                }
            }
            // the next pass's intervals: slot 0 the one to bisect now (never
            // evaluated yet: it would have been used above); slots 1.. the
            // W - 1 largest by error of the next W - 1 intervals in error
            // order, the two halves of the one bisected now and its right
            // half's right half, the new intervals' errors estimated as
            // fractions of its error.  An interval a slot already holds stays
            // there unevaluated; a slot not chosen keeps what it holds (the
            // list only ever splits its intervals, so a held interval is
            // either still listed with those halves or gone).
            // Lane c of wavefront 0 weighs candidate c.
            if (lead) {
                S.rc = rc;
                S.n = n;
            }
            if (rc < 0) {  // wavefront 0's state: uniform
                const int lane = threadIdx.x & 63;
                const double lo = pa, hi = pb, m = 0.5 * (lo + hi), e0 = pe;  // w's values at imax
                // the extra candidates, in lane order W - 1.. and by estimated
                // error: the right half, the left half, the right half's right half
                const double eR = kRightErr * e0, eL = kLeftErr * e0, eRR = kRightRightErr * e0;
                double cl = 0.0, ch = 0.0, ce = 0.0;
                bool has = false;
                const int ri = ord_rank(w, os, ord_lds, lane + 1 < n ? lane + 1 : 0);
                if (lane < W - 1) {
                    if (lane + 1 < n) {
                        const int i = ri;
                        cl = w.lo[i];
                        ch = w.hi[i];
                        ce = w.e[i];
                        has = true;
                    }
                } else if (lane == W - 1) {
                    cl = m, ch = hi, ce = eR, has = true;
                } else if (lane == W) {
                    cl = lo, ch = m, ce = eL, has = true;
                } else if (lane == W + 1) {
                    cl = 0.5 * (m + hi), ch = hi, ce = eRR, has = true;
                }
                // rank by error among the candidates, the lower lane first on
                // ties: the intervals in error order come sorted and so do the
                // extras, so an extra's rank counts the intervals at or above
                // its error, and an interval's the extras above its error
                const bool ordc = lane < W - 1 && has;
                const int nR = __popcll(__ballot(ordc && ce >= eR)), nL = __popcll(__ballot(ordc && ce >= eL));
                const int nRR = __popcll(__ballot(ordc && ce >= eRR));
                const int rank = lane < W - 1 ? lane + (eR > ce) + (eL > ce) + (eRR > ce)
                                 : lane == W - 1 ? nR
                                 : lane == W     ? nL + 1
                                                 : nRR + 2;
                const bool take = has && rank < W - 1;
                int held = 0;  // the slot (1..) that already holds this candidate
#pragma unroll
                for (int q = 1; q < W; q++)
                    if (S.slot[q].used && S.slot[q].lo == cl && S.slot[q].hi == ch) held = q;
                const bool keep = take && held > 0, place = take && held == 0;
                unsigned kept = 0;  // bit q: slot q keeps a chosen interval
#pragma unroll
                for (int q = 1; q < W; q++)
                    if (__ballot(keep && held == q)) kept |= 1u << q;
                const unsigned long long pl = __ballot(place);
                // the j-th candidate to place (in lane order) goes to the j-th slot not kept
                const int j = __popcll(pl & ((1ull << lane) - 1));
                int slot = 0, free = 0;
#pragma unroll
                for (int q = 1; q < W; q++)
                    if (!(kept >> q & 1)) {
                        if (free == j) slot = q;
                        free++;
                    }
                unsigned fresh = 0;
#pragma unroll
                for (int q = 1; q < W; q++)
                    if (__ballot(place && slot == q)) fresh |= 1u << q;
                __builtin_amdgcn_wave_barrier();  // every lane has read the slots before any lane writes
                if (place && slot > 0) {
                    QagSlot &sl = S.slot[slot];
                    sl.used = 1;
                    sl.lo = cl;
                    sl.hi = ch;
                }
                if (lane == 0) {
                    S.slot[0].used = 1;
                    S.slot[0].lo = lo;
                    S.slot[0].hi = hi;
                    S.slot[0].eval = 1;
                } else if (lane < W) {
                    S.slot[lane].eval = (int)(fresh >> lane & 1);
                }
            }
        }
        __syncthreads();
        if (S.rc >= 0) break;
        {
            QagSlot &sl = S.slot[wave];
            if (sl.eval) {
                const double lo = sl.lo, hi = sl.hi, m = 0.5 * (lo + hi);
                GK g1, g2;
                gk61_pair(f, lo, m, m, hi, g1, g2);
                if ((threadIdx.x & 63) == 0) {
                    sl.g1 = g1;
                    sl.g2 = g2;
                }
            }
        }
        __syncthreads();
    }
    const int nn = S.n;
    double s = 0.0;
    for (int k = 0; k < nn; k++) s += w.r[k];
    *result = s;
    *nint = nn;
    return S.rc;
}

// ---- natural cubic spline coefficients, windowed parallel solve ----
// The interior system h_{i-1} c_{i-1} + 2(h_{i-1}+h_i) c_i + h_i c_{i+1} = g_i is
// 2:1 diagonally dominant, so the influence of a row decays at least as
// 0.268^distance.  A thread solves rows [lo-halo, hi+halo) with zero
// boundary values at the window edges and keeps rows [lo, hi): the edge error
// reaching the kept rows is below 0.268^48 ~ 1e-27 relative.
constexpr int kHalo = 48;

// Rows per batch of loads: a thread issues the loads of kSB rows at once
// and then runs their (serial) elimination steps, so the window's memory
// round trips are m / kSB instead of m.  Same arithmetic, step by step, as
// the one-row-at-a-time loops.
constexpr int kSB = 16;

__device__ inline void spline_window(const double *x, const double *y, int n, int lo, int hi, double *c)
{
    // unknowns c[1..n-2]; c[0] = c[n-1] = 0
    if (n < 3) {
        for (int i = lo; i < hi; i++) c[i] = 0.0;
        return;
    }
    int a = lo - kHalo, b = hi + kHalo;  // window of interior unknowns [a, b)
    if (a < 1) a = 1;
    if (b > n - 1) b = n - 1;
    const int m = b - a;
    // forward elimination (L D L^T of the window), then back substitution into c;
    // the window is at most hi-lo+2*kHalo rows
    double gam_prev = 0.0, z_prev = 0.0;
    double gam[64 + 2 * kHalo + 2], zz[64 + 2 * kHalo + 2];
    for (int r0 = 0; r0 < m; r0 += kSB) {
        double xs[kSB + 2], ys[kSB + 2];  // x, y at rows a+r0-1 .. a+r0+kSB (only up to b is read)
#pragma unroll
        for (int j = 0; j < kSB + 2; j++) {
            const int i = a + r0 - 1 + j;
            xs[j] = i <= b ? x[i] : 0.0;
            ys[j] = i <= b ? y[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kSB; j++) {
            const int r = r0 + j;  // unknown c[a+r], row couples c[i-1], c[i], c[i+1]
            if (r < m) {
                const double hl = xs[j + 1] - xs[j], hr = xs[j + 2] - xs[j + 1];
                const double gl = hl != 0.0 ? 1.0 / hl : 0.0, gr = hr != 0.0 ? 1.0 / hr : 0.0;
                const double diag = 2.0 * (hr + hl);
                const double rhs = 3.0 * ((ys[j + 2] - ys[j + 1]) * gr - (ys[j + 1] - ys[j]) * gl);
                const double alp = r == 0 ? diag : diag - hl * gam_prev;
                const double g = hr / alp;
                const double z = r == 0 ? rhs : rhs - gam_prev * z_prev;
                gam[r] = g;
                zz[r] = z / alp;
                gam_prev = g;
                z_prev = z;
            }
        }
    }
    double cnext = 0.0;
    for (int r1 = m; r1 > 0; r1 -= kSB) {
        double gs[kSB], zs[kSB];
#pragma unroll
        for (int j = 0; j < kSB; j++) {
            const int r = r1 - 1 - j;
            gs[j] = r >= 0 ? gam[r] : 0.0;
            zs[j] = r >= 0 ? zz[r] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kSB; j++) {
            const int r = r1 - 1 - j;
            if (r >= 0) {
                const double ci = zs[j] - (r == m - 1 ? 0.0 : gs[j] * cnext);
                const int i = a + r;
                if (i >= lo && i < hi) c[i] = ci;
                cnext = ci;
            }
        }
    }
    if (lo == 0) c[0] = 0.0;
    if (hi >= n) c[n - 1] = 0.0;
    if (lo <= n - 1 && n - 1 < hi) c[n - 1] = 0.0;
}

// ---- natural-spline rows by wavefront scans -------------------------------
// One wavefront solves a window of up to 64 kRun interior rows of
// M c = rhs (rows i = 1..n-2, c_0 = c_{n-1} = 0), lane l taking kRun
// consecutive rows in registers, and keeps the rows [lo, hi) that lie kHalo
// rows or more inside the window (zero values are assumed just outside it,
// as in spline_window).  M's L D L^T factors
//   alp_i = 2(hl_i + hr_i) - hl_i gam_{i-1},  gam_i = hr_i / alp_i
// are those of the elimination from row 1; the gam recurrence contracts by
// at most (hl/alp)(hr/alp) <= 1/4 per row, so a lane that starts it kWarm
// rows before its run (alp = diagonal there) reaches them to below
// 0.25^32 ~ 5e-20 relative.  The forward substitution
// z_i = rhs_i - gam_{i-1} z_{i-1} and the back substitution
// c_i = z_i / alp_i - gam_i c_{i+1} are first-order linear recurrences: each
// lane composes its run into one affine map, a 64-lane scan of the maps
// gives every lane its incoming value, and the lane replays its run.
constexpr int kWarm = 32;
constexpr int kRun = 10;                        // rows per lane
constexpr int kKeep = 64 * kRun - 2 * kHalo;    // rows kept per window (544)

// Right-hand side of the natural-spline system (GSL cspline): 3 (dy/h right - dy/h left).
struct SplineRhs {
    const double *x, *y;
    __device__ __forceinline__ double operator()(int i) const
    {
        const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
        const double gl = hl != 0.0 ? 1.0 / hl : 0.0, gr = hr != 0.0 ? 1.0 / hr : 0.0;
        return 3.0 * ((y[i + 1] - y[i]) * gr - (y[i] - y[i - 1]) * gl);
    }
};

struct ArrayRhs {
    const double *r;
    __device__ __forceinline__ double operator()(int i) const { return r[i]; }
};

// A lane's share of a window: rows [s, s + len) of the window's interior
// rows [wa, wb).
struct TriLane {
    int wa, wb, s, len;
};

__device__ __forceinline__ TriLane tri_lane(int n, int lo, int hi)
{
    const int lane = threadIdx.x & 63;
    TriLane t;
    t.wa = lo - kHalo;
    t.wb = hi + kHalo;  // window of interior rows [wa, wb)
    if (t.wa < 1) t.wa = 1;
    if (t.wb > n - 1) t.wb = n - 1;
    t.s = t.wa + lane * kRun;
    t.len = t.s < t.wb ? (t.wb - t.s < kRun ? t.wb - t.s : kRun) : 0;
    return t;
}

// The lane's factors gam (g) and alp (al) over its run, and gam_{s-1}; past
// the run al = 1, g = 0.
__device__ __forceinline__ void tri_factor(const double *x, const TriLane &t, double (&g)[kRun], double (&al)[kRun],
                                           double &g_before)
{
    const int s = t.s, wb = t.wb, len = t.len;
    int a = s - kWarm;
    if (a < 1) a = 1;
    double gp = 0.0;
    for (int r = a; r < s && r < wb; r++) {
        const double hl = x[r] - x[r - 1], hr = x[r + 1] - x[r];
        const double diag = 2.0 * (hr + hl);
        const double alr = r == a ? diag : diag - hl * gp;
        gp = hr / alr;
    }
    g_before = gp;
#pragma unroll
    for (int j = 0; j < kRun; j++) {
        if (j < len) {
            const int i = s + j;
            const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
            const double diag = 2.0 * (hr + hl);
            al[j] = i == a ? diag : diag - hl * gp;
            g[j] = hr / al[j];
            gp = g[j];
        } else {
            al[j] = 1.0;
            g[j] = 0.0;
        }
    }
}

// Called by all 64 lanes of a wavefront together, with hi - lo <= kKeep and
// n >= 3, given the lane's factors gam and 1 / alp.
// ADJ_RHS: instead of storing the solution z, add its image under the
// transpose of the natural-spline right-hand side's y-derivative onto out
// (the second half of the spline adjoint) for the kept rows, with the
// reciprocal knot spacings ih (as SplineRhsIh's).
// OUT: where the kept rows go (set for the solution, add for ADJ_RHS); IH:
// the reciprocal spacings.  Each kept row receives exactly one set or add.
struct OutDirect {
    double *o;
    __device__ __forceinline__ void set(int i, double v) const { o[i] = v; }
    __device__ __forceinline__ void add(int i, double v) const { o[i] += v; }
};
struct IhDirect {
    const double *ih;
    __device__ __forceinline__ double operator()(int i) const { return ih[i]; }
};

template <class RHS, bool ADJ_RHS, class OUT = OutDirect, class IH = IhDirect>
__device__ __forceinline__ void tri_solve(const double *x, int n, const RHS &rhs, int lo, int hi, const OUT &out,
                                          const TriLane &t, const double (&g)[kRun], const double (&ial)[kRun],
                                          double g_before, const IH &ih)
{
    const int lane = threadIdx.x & 63;
    const int wa = t.wa, wb = t.wb, s = t.s, len = t.len;
    double zz[kRun];
    double rv[kRun];
#pragma unroll
    for (int j = 0; j < kRun; j++) rv[j] = j < len ? rhs(s + j) : 0.0;
    // forward: z_i = rhs_i - gam_{i-1} z_{i-1}, z_{wa-1} = 0
    double A = 0.0, B = 1.0;
#pragma unroll
    for (int j = 0; j < kRun; j++) {
        if (j < len) {
            const double gm = s + j == wa ? 0.0 : (j == 0 ? g_before : g[j - 1]);
            A = rv[j] - gm * A;
            B = -gm * B;
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Ap = __shfl_up(A, d, 64), Bp = __shfl_up(B, d, 64);
        if (lane >= d) {
            A = A + B * Ap;
            B = B * Bp;
        }
    }
    double z = __shfl_up(A, 1, 64);
    if (lane == 0) z = 0.0;
#pragma unroll
    for (int j = 0; j < kRun; j++) {
        if (j < len) {
            const double gm = s + j == wa ? 0.0 : (j == 0 ? g_before : g[j - 1]);
            z = rv[j] - gm * z;
            zz[j] = z * ial[j];
        } else {
            zz[j] = 0.0;
        }
    }
    // back: c_i = zz_i - gam_i c_{i+1}, c_{wb} = 0
    A = 0.0;
    B = 1.0;
#pragma unroll
    for (int j = kRun - 1; j >= 0; j--) {
        if (j < len) {
            A = zz[j] - g[j] * A;
            B = -g[j] * B;
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double An = __shfl_down(A, d, 64), Bn = __shfl_down(B, d, 64);
        if (lane + d < 64) {
            A = A + B * An;
            B = B * Bn;
        }
    }
    double c = __shfl_down(A, 1, 64);
    if (lane == 63) c = 0.0;
    if (!ADJ_RHS) {
#pragma unroll
        for (int j = kRun - 1; j >= 0; j--) {
            if (j < len) {
                const int i = s + j;
                c = zz[j] - g[j] * c;
                if (i >= lo && i < hi) out.set(i, c);
            }
        }
        if (lane == 0) {
            if (lo == 0) out.set(0, 0.0);
            if (lo <= n - 1 && n - 1 < hi) out.set(n - 1, 0.0);
        }
        return;
    }
    // ADJ_RHS: out (+)= d(rhs)/dy^T z for the kept rows, with
    // rhs_i = 3 ((y_{i+1} - y_i) / h_i - (y_i - y_{i-1}) / h_{i-1}) and z_0 = z_{n-1} = 0
    double zv[kRun];
#pragma unroll
    for (int j = kRun - 1; j >= 0; j--) {
        if (j < len) {
            c = zz[j] - g[j] * c;
            zv[j] = c;
        } else {
            zv[j] = 0.0;
        }
    }
    // z just before and after this lane's run (0 at the natural ends; a
    // window's inner ends are never next to a kept row)
    double zp = __shfl_up(zv[kRun - 1], 1, 64), zn = __shfl_down(zv[0], 1, 64);
    if (lane == 0 || s - 1 < 1) zp = 0.0;
    if (lane == 63 || s + len >= wb) zn = 0.0;
#pragma unroll
    for (int j = 0; j < kRun; j++) {
        const int i = s + j;
        if (j < len && i >= lo && i < hi) {
            const double zl = j > 0 ? zv[j - 1] : zp, zr = j + 1 < len ? zv[j + 1] : zn;
            const double wl = 3.0 * ih(i - 1), wr = 3.0 * ih(i);
            double acc = 0.0;
            if (i - 1 >= 1) acc += zl * wl;
            acc -= zv[j] * (wr + wl);
            if (i + 1 <= n - 2) acc += zr * wr;
            out.add(i, acc);
        }
    }
    if (lane == 0 && lo == 0 && len > 0) out.add(0, zv[0] * (3.0 * ih(0)));  // row 0: the z_1 term
    if (lo <= n - 1 && n - 1 < hi) {                                         // row n-1: the z_{n-2} term
#pragma unroll
        for (int j = 0; j < kRun; j++)
            if (j < len && s + j == n - 2) out.add(n - 1, zv[j] * (3.0 * ih(n - 2)));
    }
}

// One window [lo, hi) of natural-spline rows by one wavefront (factors
// formed in place).
template <class RHS>
__device__ inline void tridiag_window_wave(const double *x, int n, const RHS &rhs, int lo, int hi, double *out)
{
    const int lane = threadIdx.x & 63;
    if (n < 3) {
        for (int i = lo + lane; i < hi; i += 64) out[i] = 0.0;
        return;
    }
    const TriLane t = tri_lane(n, lo, hi);
    double g[kRun], al[kRun], gb;
    tri_factor(x, t, g, al, gb);
#pragma unroll
    for (int j = 0; j < kRun; j++) al[j] = 1.0 / al[j];
    tri_solve<RHS, false>(x, n, rhs, lo, hi, OutDirect{out}, t, g, al, gb, IhDirect{nullptr});
}

// Factors of the knots x[0..n) stored once for every row that shares them:
// window w at fac + w kFacStride, [j][lane] for gam, then 1 / alp, then
// gam_{s-1} per lane (from the values tri_factor forms in each wavefront).
constexpr int kFacStride = (2 * kRun + 1) * 64;

__host__ __device__ inline int fac_windows(int n) { return (n + kKeep - 1) / kKeep; }

__device__ inline void tri_factor_store(const double *x, int n, int w, double *fac)
{
    const int lane = threadIdx.x & 63;
    const int lo = w * kKeep, hi = lo + kKeep < n ? lo + kKeep : n;
    const TriLane t = tri_lane(n, lo, hi);
    double g[kRun], al[kRun], gb;
    tri_factor(x, t, g, al, gb);
    double *f = fac + (size_t)w * kFacStride;
#pragma unroll
    for (int j = 0; j < kRun; j++) {
        f[j * 64 + lane] = g[j];
        f[(kRun + j) * 64 + lane] = 1.0 / al[j];
    }
    f[2 * kRun * 64 + lane] = gb;
}

// A window's rows staged in LDS: the span [wa - 1, wb] of a window and its
// two neighbours, twice (kTriLds doubles per wavefront).
constexpr int kTriSpan = 64 * kRun + 2;
constexpr int kTriLds = 2 * kTriSpan;
constexpr int kTriPass = (kTriSpan + 63) / 64;  // 64-wide passes over a span

// A wavefront's LDS writes ordered before its lanes' reads of them (LDS
// operations of one wavefront complete in order; the fences keep the
// compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// dst[k] = src[k] for k < cnt (cnt <= kTriSpan) by the 64 lanes: every load
// issued before the first LDS store, so that the passes' loads are in flight
// together.
__device__ __forceinline__ void stage_span(double *dst, const double *__restrict__ src, int cnt)
{
    const int lane = threadIdx.x & 63;
    double v[kTriPass];
#pragma unroll
    for (int p = 0; p < kTriPass; p++) {
        const int k = p * 64 + lane;
        v[p] = k < cnt ? src[k] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < kTriPass; p++) {
        const int k = p * 64 + lane;
        if (k < cnt) dst[k] = v[p];
    }
}

// SplineRhsIh's values, bit for bit, from staged y and ih (ys[k] = y[base + k]).
struct SplineRhsLds {
    const double *ys, *ihs;
    int base;
    __device__ __forceinline__ double operator()(int i) const
    {
        const int k = i - base;
        return 3.0 * ((ys[k + 1] - ys[k]) * ihs[k] - (ys[k] - ys[k - 1]) * ihs[k - 1]);
    }
};
struct ArrayRhsLds {
    const double *rs;
    int base;
    __device__ __forceinline__ double operator()(int i) const { return rs[i - base]; }
};
struct OutLds {  // kept row i at o[i - lo]; add stores the row's one contribution
    double *o;
    int lo;
    __device__ __forceinline__ void set(int i, double v) const { o[i - lo] = v; }
    __device__ __forceinline__ void add(int i, double v) const { o[i - lo] = v; }
};
struct IhLds {
    const double *ihs;
    int base;
    __device__ __forceinline__ double operator()(int i) const { return ihs[i - base]; }
};

// tridiag_window_wave for window w = lo / kKeep with the stored factors
// (fac_windows(n) windows, then the reciprocal spacings), the window staged
// through the wavefront's LDS (lds: kTriLds doubles): lane l solves rows
// s .. s + kRun - 1, kRun rows from its neighbours', so read from global
// memory directly each load's 64 lanes would touch ~40 cache lines (and a
// CU's wavefronts together more lines than its L1 holds); staged, the rows
// come in and go out by loads and stores of consecutive doubles across the
// lanes.  src: the y row (natural-spline coefficients set on out) or, with
// ADJ_RHS, the right-hand side (its adjoint added onto out).
template <bool ADJ_RHS>
__device__ inline void tri_window_staged(const double *x, int n, const double *__restrict__ src, int lo, int hi,
                                         double *__restrict__ out, const double *__restrict__ fac, double *lds)
{
    const int lane = threadIdx.x & 63;
    if (n < 3) {
        if (!ADJ_RHS)
            for (int i = lo + lane; i < hi; i += 64) out[i] = 0.0;
        return;
    }
    const TriLane t = tri_lane(n, lo, hi);
    const double *f = fac + (size_t)(lo / kKeep) * kFacStride;
    const double *ih = fac + (size_t)fac_windows(n) * kFacStride;
    double *sa = lds, *sb = lds + kTriSpan;
    const int base = t.wa - 1;
    // ih[wa - 1 .. wb - 1]; y[wa - 1 .. wb] or rhs[wa .. wb - 1]
    stage_span(sb, ih + base, t.wb - base);
    if (ADJ_RHS) stage_span(sa, src + t.wa, t.wb - t.wa);
    else stage_span(sa, src + base, t.wb - base + 1);
    wave_lds_sync();
    double g[kRun], ial[kRun];
#pragma unroll
    for (int j = 0; j < kRun; j++) {
        g[j] = f[j * 64 + lane];
        ial[j] = f[(kRun + j) * 64 + lane];
    }
    const double gb = f[2 * kRun * 64 + lane];
    const OutLds o{sa, lo};  // the staged rows are all read before the first kept row is written
    if (ADJ_RHS)
        tri_solve<ArrayRhsLds, true>(x, n, ArrayRhsLds{sa, t.wa}, lo, hi, o, t, g, ial, gb, IhLds{sb, base});
    else
        tri_solve<SplineRhsLds, false>(x, n, SplineRhsLds{sa, sb, base}, lo, hi, o, t, g, ial, gb,
                                       IhLds{sb, base});
    wave_lds_sync();
    const int cnt = hi - lo;
    if (ADJ_RHS) {  // out += the staged contributions, every load issued first
        double v[kTriPass];
#pragma unroll
        for (int p = 0; p < kTriPass; p++) {
            const int k = p * 64 + lane;
            v[p] = k < cnt ? out[lo + k] : 0.0;
        }
#pragma unroll
        for (int p = 0; p < kTriPass; p++) {
            const int k = p * 64 + lane;
            if (k < cnt) out[lo + k] = v[p] + sa[k];
        }
    } else {
#pragma unroll
        for (int p = 0; p < kTriPass; p++) {
            const int k = p * 64 + lane;
            if (k < cnt) out[lo + k] = sa[k];
        }
    }
}

// Right-hand side of the natural-spline system with the reciprocal knot
// spacings ih[i] = 1 / (x[i+1] - x[i]) (0 for a zero spacing) stored: the
// values of SplineRhs, bit for bit.
struct SplineRhsIh {
    const double *ih, *y;
    __device__ __forceinline__ double operator()(int i) const
    {
        return 3.0 * ((y[i + 1] - y[i]) * ih[i] - (y[i] - y[i - 1]) * ih[i - 1]);
    }
};

// ---- derivatives for the adjoint ----------------------------------------

// j0'(z) = (z cos z - sin z) / z^2, Taylor series below |z| = 0.5
__device__ __forceinline__ double sph_j0_prime(double z)
{
    if (fabs(z) < 0.5) {
        const double y = z * z;
        // d/dz of 1 - z^2/6 + z^4/120 - z^6/5040 + z^8/362880 - z^10/39916800 + z^12/6227020800
        return z * (-2.0 / 6.0 + y * (4.0 / 120.0 + y * (-6.0 / 5040.0 + y * (8.0 / 362880.0 +
                    y * (-10.0 / 39916800.0 + y * (12.0 / 6227020800.0))))));
    }
    double s, c;
    sincos(z, &s, &c);
    return (z * c - s) / (z * z);
}

// specialJ(x, qc) with its partial derivatives d/dx and d/dqc.
__device__ inline double specialJ_d(double x, const JPar &P, double &dJdx, double &dJdqc)
{
    if (P.qc > 0) {
        const double qc = P.qc;
        const double z = qc * x;
        double s, c;
        sincos(z, &s, &c);
        const double j0 = fabs(z) < 0.5 ? sph_j0(z) : s / z;
        const double j0p = sph_j0_prime(z);
        double J = 0.0, Jx = 0.0, Jq = 0.0;
        for (int n = 1; n < 20; n++) {
            const double nn = (double)(n * n), d = nn + x * x, rd = rcp_fast(d);
            const double A = (n * n + n * n * n * qc + n * qc * x * x - x * x);
            const double B = (2 * n + n * n * qc + qc * x * x);
            const double ii = A * qc * j0 + B * c;
            const double Ax = 2.0 * n * qc * x - 2.0 * x, Bx = 2.0 * qc * x;
            const double Aq = (double)(n * n * n) + n * x * x, Bq = nn + x * x;
            const double iix = Ax * qc * j0 + A * qc * qc * j0p + Bx * c - B * qc * s;
            const double iiq = Aq * qc * j0 + A * (j0 + z * j0p) + Bq * c - B * x * s;
            const double sgn = (n & 1) ? 1.0 : -1.0;
            const double w = sgn * P.en[n] * (rd * rd);
            J += w * ii;
            Jx += w * (iix - 4.0 * x * rd * ii);
            Jq += w * (iiq - n * ii);
        }
        const double Z = 1.5 * 1.202056903159594 * (1 - P.nufrac_low);
        dJdx = Jx / Z;
        dJdqc = Jq / Z;
        return J / Z;
    }
    dJdqc = 0.0;
    if (x <= 0.) {
        dJdx = 0.0;
        return 1.;
    }
    const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x7 = x4 * x3, x8 = x4 * x4;
    const double p = exp(4.1811 * log(x));
    const double N = 1. + 0.0168 * x2 + 0.0407 * x4;
    const double D = 1. + 2.1734 * x2 + 1.6787 * p + 0.1467 * x8;
    const double Nx = 0.0336 * x + 0.1628 * x3;
    const double Dx = 4.3468 * x + 1.6787 * 4.1811 * p / x + 1.1736 * x7;
    dJdx = (Nx * D - N * Dx) / (D * D);
    return N / D;
}

// Weights of S(x) = y_i + dx (b + dx (c_i + dx d)) on y_i, y_{i+1}, c_i, c_{i+1}.
__device__ __forceinline__ void cspline_grad(const double *xa, int i, double x, double &wy0, double &wy1, double &wc0,
                                             double &wc1)
{
    const double h = xa[i + 1] - xa[i], dx = x - xa[i], ih = rcp_fast(h);
    const double t = dx * ih, dx3 = dx * dx * dx * ih * (1.0 / 3.0);
    wy0 = 1.0 - t;
    wy1 = t;
    wc0 = -dx * h * (2.0 / 3.0) + dx * dx - dx3;
    wc1 = -dx * h * (1.0 / 3.0) + dx3;
}

// Windowed solve of the natural-spline system M z = rhs (interior rows
// 1..n-2, z[0] = z[n-1] = 0); same window scheme as spline_window.
__device__ inline void tridiag_window(const double *x, const double *rhs, int n, int lo, int hi, double *z_out)
{
    if (n < 3) {
        for (int i = lo; i < hi; i++) z_out[i] = 0.0;
        return;
    }
    int a = lo - kHalo, b = hi + kHalo;
    if (a < 1) a = 1;
    if (b > n - 1) b = n - 1;
    const int m = b - a;
    double gam[64 + 2 * kHalo + 2], zz[64 + 2 * kHalo + 2];
    double gam_prev = 0.0, z_prev = 0.0;
    for (int r0 = 0; r0 < m; r0 += kSB) {
        double xs[kSB + 2], rs[kSB];  // x at rows a+r0-1 .. a+r0+kSB (up to b), rhs at a+r0 .. a+r0+kSB-1
#pragma unroll
        for (int j = 0; j < kSB + 2; j++) {
            const int i = a + r0 - 1 + j;
            xs[j] = i <= b ? x[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kSB; j++) rs[j] = r0 + j < m ? rhs[a + r0 + j] : 0.0;
#pragma unroll
        for (int j = 0; j < kSB; j++) {
            const int r = r0 + j;
            if (r < m) {
                const double hl = xs[j + 1] - xs[j], hr = xs[j + 2] - xs[j + 1];
                const double diag = 2.0 * (hr + hl);
                const double alp = r == 0 ? diag : diag - hl * gam_prev;
                const double zr = r == 0 ? rs[j] : rs[j] - gam_prev * z_prev;
                gam[r] = hr / alp;
                zz[r] = zr / alp;
                gam_prev = gam[r];
                z_prev = zr;
            }
        }
    }
    double next = 0.0;
    for (int r1 = m; r1 > 0; r1 -= kSB) {
        double gs[kSB], zs[kSB];
#pragma unroll
        for (int j = 0; j < kSB; j++) {
            const int r = r1 - 1 - j;
            gs[j] = r >= 0 ? gam[r] : 0.0;
            zs[j] = r >= 0 ? zz[r] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kSB; j++) {
            const int r = r1 - 1 - j;
            if (r >= 0) {
                const double v = zs[j] - (r == m - 1 ? 0.0 : gs[j] * next);
                const int i = a + r;
                if (i >= lo && i < hi) z_out[i] = v;
                next = v;
            }
        }
    }
    if (lo == 0) z_out[0] = 0.0;
    if (lo <= n - 1 && n - 1 < hi) z_out[n - 1] = 0.0;
}

}  // namespace kn_dnu

#endif

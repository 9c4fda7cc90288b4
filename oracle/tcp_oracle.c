/*
 * tcp_oracle.c -- FP64 CPU restatement of the TCP_allpair force-and-integrate
 * path (TEST INFRASTRUCTURE: the checker, never the thing measured or shipped).
 *
 * Each function cites the reference lines it restates.  Expressions keep the
 * reference's evaluation order (Fortran left-to-right for * and /) so that,
 * built without FMA contraction and against the same libm, the direct-sum and
 * integrator results can agree with the compiled reference bit for bit.
 *
 * Parity status per function:
 *   tcpo_ewald_self, tcpo_psi, tcpo_force_direct_self, tcpo_force_direct_ext,
 *   tcpo_remove_rigidmtn, tcpo_vverlet1, tcpo_vverlet2 (NOTEMP/BEREND/ISOKIN),
 *   tcpo_step (reciprocal part off): pinned against the compiled reference.
 *   tcpo_vverlet2 (STOCH): pinned by the closed form with supplied noise (the
 *   reference's RANDOM_NUMBER stream is unseeded).
 *   tcpo_spme, tcpo_bspline, tcpo_bspline_modulus: pinned by mathematics
 *   (partition of unity, spline modulus closed form, convergence to
 *   tcpo_ewald_kspace); FFTE is absent so Force_SPME4 cannot run here.
 *   tcpo_ewald_kspace: pinned by the Madelung constant of rock salt and by
 *   finite differences of its own energy.
 *   tcpo_force_allpair, tcpo_force_rows: pinned against tcpo_force_direct_self
 *   (same rows, same pair maths, a different summation order).
 */
#include "tcp_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NBIN TCPO_NBIN
static const double PI = 3.1415926535897931;      /* data.f90:23 */
static const double SQRTPI = 1.7724538509055159;  /* data.f90:24 */

int tcpo_version(void) { return 1; }

static double alpha_of(const tcpo_params* p)
{
    return p->alpha > 0.0 ? p->alpha : 3.1 / p->rc; /* force.f90:23 */
}

/* force.f90:18-24, 44-45 */
int tcpo_ewald_self(const tcpo_params* p, double* alpha, double* rc2, double* Uo)
{
    double rc_cut = fmin(fmin(p->box[0], p->box[1]), p->box[2]) / 2.0;
    if (p->rc > rc_cut) return -1; /* " *** error at cut-off radius *** " */
    double a = alpha_of(p);
    *alpha = a;
    *rc2 = p->rc * p->rc;
    double u = -((double)p->ne * (p->q[0] * p->q[0])) - (double)p->ni * (p->q[1] * p->q[1]);
    *Uo = u * a / SQRTPI;
    return 0;
}

/* force.f90:27-42: PSI(i,j,k) = exp(-pi^2 m^2/alpha^2)/m^2 for m != 0 */
int tcpo_psi(const tcpo_params* p, double alpha, double* psi)
{
    const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
    const double pi2 = PI * PI, a2 = alpha * alpha;
    for (int k = 0; k < K3; ++k) {
        int iz = (k + K3 / 2) % K3 - K3 / 2;
        for (int j = 0; j < K2; ++j) {
            int iy = (j + K2 / 2) % K2 - K2 / 2;
            for (int i = 0; i < K1; ++i) {
                int ix = (i + K1 / 2) % K1 - K1 / 2;
                size_t idx = ((size_t)k * K2 + j) * K1 + i;
                int m = ix * ix + iy * iy + iz * iz;
                if (m == 0) { psi[idx] = 0.0; continue; }
                double m2 = (double)ix * (double)ix / (p->box[0] * p->box[0])
                          + (double)iy * (double)iy / (p->box[1] * p->box[1])
                          + (double)iz * (double)iz / (p->box[2] * p->box[2]);
                psi[idx] = exp(-(m2 * pi2 / a2)) / m2;
            }
        }
    }
    return 0;
}

/* Minimum-image separation and r^2 as force.f90:89-94 (DNINT = round half away). */
static inline double min_image(const double* box, const double* xi, const double* xj, double* xr)
{
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double x = xi[k] - xj[k];
        xr[k] = x - box[k] * round(x / box[k]);
        r2 = r2 + xr[k] * xr[k];
    }
    return r2;
}

/* One pair inside the cut-off, force.f90:99-110 (bare) and 130-144 (Deutsch e-i).
 * Returns df (eps included) and adds the energy term q1 q2 (erfc + R)/r to *U.
 * reg selects the regularised form: R = -exp(-r/lambda) (Deutsch, the reference)
 * or, with TCPO_KELBG, R = -exp(-s^2) + sqrt(pi) s erfc(s), s = r/lambda, for
 * which -d/dr[(erfc + R)/r] = (g + erfc - exp(-s^2))/r^2 since dR/dr = sqrt(pi) erfc(s)/lambda. */
static inline double pair_term(const tcpo_params* p, double a, double q1, double q2, int reg,
                               double r2, double* U, double* rdf_bin, double dx)
{
    double r = sqrt(r2);
    int idx = (int)(r / dx) + 1;
    if (idx <= NBIN) rdf_bin[idx - 1] += 2.0;
    double x = 1.0 - erf(a * r);
    double g = 2.0 * exp(-(a * a * r2)) * a * r / SQRTPI;
    if (!reg) {
        *U = *U + q1 * q2 * x / r;
        return p->eps * q1 * q2 * (g + x) / r2 / r;
    }
    if (p->flags & TCPO_KELBG) {
        double s = r / p->rs;
        double y2 = exp(-(s * s));
        *U = *U + q1 * q2 * (x - y2 + SQRTPI * s * erfc(s)) / r;
        return p->eps * q1 * q2 * (g + x - y2) / r2 / r;
    }
    double y = exp(-r / p->rs);
    *U = *U + q1 * q2 * (x - y) / r;
    return p->eps * q1 * q2 * ((g + x - y) / r2 / r - y / r2 / p->rs);
}

/* Force_direct_self, force.f90:52-180 */
int tcpo_force_direct_self(const tcpo_params* p, double alpha, double rc2,
                           const double* x, double* f, double* Ur, double* rdf)
{
    const int ne = p->ne, ni = p->ni, n = ne + ni;
    const double* xe = x;
    const double* xi = x + 3 * (size_t)ne;
    double* fe = f;
    double* fi = f + 3 * (size_t)ne;
    double dummy[3 * NBIN];
    double* h = rdf ? rdf : dummy;
    memset(f, 0, sizeof(double) * 3 * (size_t)n);
    double U = 0.0, xr[3];
    /* electron-electron, force.f90:85-113 */
    for (int i = 0; i < ne - 1; ++i)
        for (int j = i + 1; j < ne; ++j) {
            double r2 = min_image(p->box, xe + 3 * i, xe + 3 * j, xr);
            if (r2 < rc2) {
                double df = pair_term(p, alpha, p->q[0], p->q[0], (p->flags & TCPO_REG_EE) != 0, r2,
                                      &U, h, p->dx_ee);
                for (int k = 0; k < 3; ++k) {
                    fe[3 * i + k] = fe[3 * i + k] + df * xr[k];
                    fe[3 * j + k] = fe[3 * j + k] - df * xr[k];
                }
            }
        }
    /* electron-ion, force.f90:116-147 */
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ni; ++j) {
            double r2 = min_image(p->box, xe + 3 * i, xi + 3 * j, xr);
            if (r2 < rc2) {
                double df = pair_term(p, alpha, p->q[0], p->q[1], 1, r2, &U, h + NBIN, p->dx_ei);
                for (int k = 0; k < 3; ++k) {
                    fe[3 * i + k] = fe[3 * i + k] + df * xr[k];
                    fi[3 * j + k] = fi[3 * j + k] - df * xr[k];
                }
            }
        }
    /* ion-ion, force.f90:149-178 */
    for (int i = 0; i < ni - 1; ++i)
        for (int j = i + 1; j < ni; ++j) {
            double r2 = min_image(p->box, xi + 3 * i, xi + 3 * j, xr);
            if (r2 < rc2) {
                double df = pair_term(p, alpha, p->q[1], p->q[1], 0, r2, &U, h + 2 * NBIN, p->dx_ii);
                for (int k = 0; k < 3; ++k) {
                    fi[3 * i + k] = fi[3 * i + k] + df * xr[k];
                    fi[3 * j + k] = fi[3 * j + k] - df * xr[k];
                }
            }
        }
    *Ur = U;
    return 0;
}

/* Force_direct_ext, force.f90:183-343 */
int tcpo_force_direct_ext(const tcpo_params* p, double alpha, double rc2,
                          int ne_o, int ni_o, const double* x_own, double* f_own,
                          int ne_g, int ni_g, const double* xg, double* gemit,
                          double* Ur, double* rdf)
{
    const double* xe = x_own;
    const double* xi = x_own + 3 * (size_t)ne_o;
    double* fe = f_own;
    double* fi = f_own + 3 * (size_t)ne_o;
    const double* ge = xg;
    const double* gi = xg + 3 * (size_t)ne_g;
    double* me = gemit;
    double* mi = gemit + 3 * (size_t)ne_g;
    double dummy[3 * NBIN];
    double* h = rdf ? rdf : dummy;
    memset(gemit, 0, sizeof(double) * 3 * (size_t)(ne_g + ni_g));
    double U = 0.0, xr[3];
    struct { const double *a, *b; double *fa, *fb; int na, nb, sa, sb, ei, hist; double dx; } blk[4] = {
        {xe, ge, fe, me, ne_o, ne_g, 0, 0, (p->flags & TCPO_REG_EE) != 0, 0, p->dx_ee}, /* e-e_g :213-241 */
        {xe, gi, fe, mi, ne_o, ni_g, 0, 1, 1, 1, p->dx_ei}, /* e-i_g :244-275 */
        {xi, ge, fi, me, ni_o, ne_g, 1, 0, 1, 1, p->dx_ei}, /* i-e_g :278-309 */
        {xi, gi, fi, mi, ni_o, ni_g, 1, 1, 0, 2, p->dx_ii}, /* i-i_g :312-341 */
    };
    for (int b = 0; b < 4; ++b) {
        double q1 = p->q[blk[b].sa], q2 = p->q[blk[b].sb];
        for (int i = 0; i < blk[b].na; ++i)
            for (int j = 0; j < blk[b].nb; ++j) {
                double r2 = min_image(p->box, blk[b].a + 3 * i, blk[b].b + 3 * j, xr);
                if (r2 < rc2) {
                    double df = pair_term(p, alpha, q1, q2, blk[b].ei, r2, &U,
                                          h + blk[b].hist * NBIN, blk[b].dx);
                    for (int k = 0; k < 3; ++k) {
                        blk[b].fa[3 * i + k] = blk[b].fa[3 * i + k] + df * xr[k];
                        blk[b].fb[3 * j + k] = blk[b].fb[3 * j + k] - df * xr[k];
                    }
                }
            }
    }
    *Ur = U;
    return 0;
}

/* Neumaier two-sum accumulation */
static inline void nsum(double* s, double* c, double v)
{
    double t = *s + v;
    if (fabs(*s) >= fabs(v)) *c += (*s - t) + v;
    else *c += (v - t) + *s;
    *s = t;
}

/* Row i of the all-pair sum: f3 = sum_j F_ij, *u = sum_j U_ij over every j != i
 * inside the cut-off (compensated sums in j order), +1 per ordered pair into h.
 * The pair maths of pair_term, with the energy terms kept apart per pair. */
static void pair_row(const tcpo_params* p, double alpha, double rc2, const double* x, int i,
                     double* f3, double* u, double* h)
{
    const int ne = p->ne, n = p->ne + p->ni;
    int si = i >= ne;
    double fs[3] = {0, 0, 0}, fc[3] = {0, 0, 0}, us = 0.0, uc = 0.0, xr[3];
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        double r2 = min_image(p->box, x + 3 * (size_t)i, x + 3 * (size_t)j, xr);
        if (!(r2 < rc2)) continue;
        int sj = j >= ne;
        int hsel = si + sj; /* 0 ee, 1 ei, 2 ii */
        double dx = hsel == 0 ? p->dx_ee : (hsel == 1 ? p->dx_ei : p->dx_ii);
        double U = 0.0;
        /* +1 per ordered pair == +2 per unordered pair */
        double r = sqrt(r2);
        int idx = (int)(r / dx) + 1;
        if (h && idx <= NBIN) h[hsel * NBIN + idx - 1] += 1.0;
        double xx = 1.0 - erf(alpha * r);
        double g = 2.0 * exp(-(alpha * alpha * r2)) * alpha * r / SQRTPI;
        double q1 = p->q[si], q2 = p->q[sj], df;
        const int reg = si != sj || (si == 0 && sj == 0 && (p->flags & TCPO_REG_EE));
        if (!reg) {
            U = q1 * q2 * xx / r;
            df = p->eps * q1 * q2 * (g + xx) / r2 / r;
        } else if (p->flags & TCPO_KELBG) {
            double s = r / p->rs;
            double y2 = exp(-(s * s));
            U = q1 * q2 * (xx - y2 + SQRTPI * s * erfc(s)) / r;
            df = p->eps * q1 * q2 * (g + xx - y2) / r2 / r;
        } else {
            double y = exp(-r / p->rs);
            U = q1 * q2 * (xx - y) / r;
            df = p->eps * q1 * q2 * ((g + xx - y) / r2 / r - y / r2 / p->rs);
        }
        nsum(&us, &uc, U);
        for (int k = 0; k < 3; ++k) nsum(&fs[k], &fc[k], df * xr[k]);
    }
    for (int k = 0; k < 3; ++k) f3[k] = fs[k] + fc[k];
    *u = us + uc;
}

int tcpo_force_allpair(const tcpo_params* p, double alpha, double rc2,
                       const double* x, double* f, double* Ur, double* rdf, int nthreads)
{
    const int n = p->ne + p->ni;
    double* usum = (double*)calloc((size_t)n, sizeof(double));
    if (!usum) return -2;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    double* hist = (double*)calloc((size_t)nthreads * 3 * NBIN, sizeof(double));
    if (!hist) { free(usum); return -2; }
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double* h = hist + (size_t)tid * 3 * NBIN;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; ++i) pair_row(p, alpha, rc2, x, i, f + 3 * (size_t)i, usum + i, h);
    }
    double s = 0.0, c = 0.0;
    for (int i = 0; i < n; ++i) nsum(&s, &c, usum[i]);
    *Ur = 0.5 * (s + c);
    if (rdf)
        for (int t = 0; t < nthreads; ++t)
            for (int b = 0; b < 3 * NBIN; ++b) rdf[b] += hist[(size_t)t * 3 * NBIN + b];
    free(hist);
    free(usum);
    return 0;
}

int tcpo_force_rows(const tcpo_params* p, double alpha, double rc2, const double* x, int nrows,
                    const int* rows, double* f, double* u, int nthreads)
{
    const int n = p->ne + p->ni;
    for (int k = 0; k < nrows; ++k)
        if (rows[k] < 0 || rows[k] >= n) return -1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int k = 0; k < nrows; ++k) pair_row(p, alpha, rc2, x, rows[k], f + 3 * (size_t)k, u + k, NULL);
    return 0;
}

/* ---------------------------------------------------------------- SPME --- */

/* Cardinal B-spline M_n(u) on [0, n] by the recursion of Essmann et al. (1995), eq. 4.1 */
static double bspline_M(int n, double u)
{
    if (u <= 0.0 || u >= (double)n) return 0.0;
    if (n == 2) return 1.0 - fabs(u - 1.0);
    return (u * bspline_M(n - 1, u) + ((double)n - u) * bspline_M(n - 1, u - 1.0)) / (double)(n - 1);
}

/* Weights of grid points floor(u) - p/2 + 1 + i, i = 0..p-1:  M_p(f + p - 1 - i), f = frac(u).
 * For p = 4 the explicit cubic forms of force.f90:703-738 are used (z = 1 - f). */
int tcpo_bspline(int order, double u, double* w, double* dw)
{
    if (order < 2 || order > 16 || (order & 1)) return -1;
    double f = u - floor(u);
    if (order == 4) {
        double z = floor(u) - u + 1.0;
        w[0] = z * z * z / 6.0;
        dw[0] = -(3.0 * (z * z)) / 6.0;
        z = z + 1.0;
        w[1] = (-3.0 * z * z * z + 12.0 * z * z - 12.0 * z + 4.0) / 6.0;
        dw[1] = (9.0 * z * z - 24.0 * z + 12.0) / 6.0;
        z = z + 1.0;
        w[2] = (3.0 * z * z * z - 24.0 * z * z + 60.0 * z - 44.0) / 6.0;
        dw[2] = (-9.0 * z * z + 48.0 * z - 60.0) / 6.0;
        z = z + 1.0;
        w[3] = (-z * z * z + 12.0 * z * z - 48.0 * z + 64.0) / 6.0;
        dw[3] = (3.0 * z * z - 24.0 * z + 48.0) / 6.0;
        return 0;
    }
    for (int i = 0; i < order; ++i) {
        double t = f + (double)(order - 1 - i);
        w[i] = bspline_M(order, t);
        dw[i] = bspline_M(order - 1, t) - bspline_M(order - 1, t - 1.0);
    }
    return 0;
}

/* |b(m)|^2 = 1/|sum_{k=0}^{p-2} M_p(k+1) exp(2 pi i m k/K)|^2; for p = 4 this is
 * 36/(4 + 2 cos(2 pi m/K))^2, the factor of force.f90:834-839. */
int tcpo_bspline_modulus(int order, int K, double* bsq)
{
    if (order < 2 || (order & 1)) return -1;
    for (int m = 0; m < K; ++m) {
        if (order == 4) {
            double z = 4.0 + 2.0 * cos(2.0 * PI * (double)m / (double)K);
            bsq[m] = 36.0 / (z * z);
            continue;
        }
        double re = 0.0, im = 0.0;
        for (int k = 0; k <= order - 2; ++k) {
            double mk = bspline_M(order, (double)(k + 1));
            double th = 2.0 * PI * (double)m * (double)k / (double)K;
            re += mk * cos(th);
            im += mk * sin(th);
        }
        bsq[m] = 1.0 / (re * re + im * im);
    }
    return 0;
}

/* In-place DFT of n points at the given stride: sum_k a_k exp(sign 2 pi i j k/n), unnormalised. */
static void dft_line(double complex* a, int n, size_t stride, int sign, double complex* tmp)
{
    for (int k = 0; k < n; ++k) tmp[k] = a[(size_t)k * stride];
    if ((n & (n - 1)) == 0) {
        /* iterative radix-2, bit reversal then butterflies */
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { double complex t = tmp[i]; tmp[i] = tmp[j]; tmp[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; ++k) {
                    double th = sign * 2.0 * PI * (double)k / (double)len;
                    double complex w = cos(th) + I * sin(th);
                    double complex u = tmp[i + k], v = tmp[i + k + len / 2] * w;
                    tmp[i + k] = u + v;
                    tmp[i + k + len / 2] = u - v;
                }
        }
        for (int k = 0; k < n; ++k) a[(size_t)k * stride] = tmp[k];
    } else {
        for (int j = 0; j < n; ++j) {
            double complex s = 0.0;
            for (int k = 0; k < n; ++k) {
                double th = sign * 2.0 * PI * (double)(((long)j * k) % n) / (double)n;
                s += tmp[k] * (cos(th) + I * sin(th));
            }
            a[(size_t)j * stride] = s;
        }
    }
}

static void dft3d(double complex* a, const int* K, int sign)
{
    const int K1 = K[0], K2 = K[1], K3 = K[2];
    int kmax = K1 > K2 ? K1 : K2;
    kmax = kmax > K3 ? kmax : K3;
    double complex* tmp = (double complex*)malloc(sizeof(double complex) * 2 * (size_t)kmax);
    for (int k = 0; k < K3; ++k)
        for (int j = 0; j < K2; ++j) dft_line(a + ((size_t)k * K2 + j) * K1, K1, 1, sign, tmp);
    for (int k = 0; k < K3; ++k)
        for (int i = 0; i < K1; ++i) dft_line(a + (size_t)k * K2 * K1 + i, K2, (size_t)K1, sign, tmp);
    for (int j = 0; j < K2; ++j)
        for (int i = 0; i < K1; ++i) dft_line(a + (size_t)j * K1 + i, K3, (size_t)K1 * K2, sign, tmp);
    free(tmp);
}

/* Force_SPME4 (force.f90:652-929) + SPME_SUM (:932-969), order p:
 *   Q = sum q w w w;  Qinv = IDFT_norm(Q);  Qinv *= PSI * B;  Qfin = DFT(Qinv);
 *   Epme = K^3 sum Re(Q Qfin);  F = -eps K^3 sum dQdx Re(Qfin)/(V pi);  Um = eps Epme/(2 pi V). */
static int spme_core(const tcpo_params* p, double alpha, const double* x, const double* qp,
                     double* frec, double* pot, double* Um);

int tcpo_spme(const tcpo_params* p, double alpha, const double* x, double* frec, double* Um)
{
    return spme_core(p, alpha, x, NULL, frec, NULL, Um);
}

/* qp: per-particle charges (NULL: the species charges); pot[m] = dUm/dq_m */
static int spme_core(const tcpo_params* p, double alpha, const double* x, const double* qp,
                     double* frec, double* pot, double* Um)
{
    const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2], P = p->order, n = p->ne + p->ni;
    if (P < 2 || P > 16 || (P & 1) || K1 < P || K2 < P || K3 < P) return -1;
    const size_t G = (size_t)K1 * K2 * K3;
    const double V = p->box[0] * p->box[1] * p->box[2];
    double* Q = (double*)calloc(G, sizeof(double));
    double* psi = (double*)malloc(G * sizeof(double));
    double complex* A = (double complex*)malloc(G * sizeof(double complex));
    double* b1 = (double*)malloc(sizeof(double) * (K1 + K2 + K3));
    if (!Q || !psi || !A || !b1) { free(Q); free(psi); free(A); free(b1); return -2; }
    double* b2 = b1 + K1;
    double* b3 = b2 + K2;
    tcpo_psi(p, alpha, psi);
    tcpo_bspline_modulus(P, K1, b1);
    tcpo_bspline_modulus(P, K2, b2);
    tcpo_bspline_modulus(P, K3, b3);
    double w[3][16], dw[3][16];
    int base[3];
    /* spread, force.f90:696-820 */
    for (int m = 0; m < n; ++m) {
        double q1 = qp ? qp[m] : p->q[m >= p->ne];
        for (int d = 0; d < 3; ++d) {
            double u = (double)p->K[d] * (x[3 * (size_t)m + d] + p->box[d] * 0.5) / p->box[d];
            tcpo_bspline(P, u, w[d], dw[d]);
            base[d] = (int)floor(u) - P / 2 + 1;
        }
        for (int i = 0; i < P; ++i) {
            int ix = ((base[0] + i) % K1 + K1) % K1;
            for (int j = 0; j < P; ++j) {
                int iy = ((base[1] + j) % K2 + K2) % K2;
                for (int k = 0; k < P; ++k) {
                    int iz = ((base[2] + k) % K3 + K3) % K3;
                    Q[((size_t)iz * K2 + iy) * K1 + ix] += q1 * w[0][i] * w[1][j] * w[2][k];
                }
            }
        }
    }
    /* force.f90:823-851: inverse (normalised), convolution, forward */
    for (size_t g = 0; g < G; ++g) A[g] = Q[g];
    dft3d(A, p->K, +1);
    const double invG = 1.0 / (double)G;
    for (int k = 0; k < K3; ++k)
        for (int j = 0; j < K2; ++j)
            for (int i = 0; i < K1; ++i) {
                size_t g = ((size_t)k * K2 + j) * K1 + i;
                A[g] = A[g] * invG * psi[g] * b1[i] * b2[j] * b3[k];
            }
    dft3d(A, p->K, -1);
    /* force.f90:854-865 */
    double e = 0.0, ec = 0.0;
    for (size_t g = 0; g < G; ++g) nsum(&e, &ec, Q[g] * creal(A[g]));
    double Epme = (e + ec) * (double)G;
    double q1g = p->eps * (double)G;
    /* gather, force.f90:868-919 */
    for (int m = 0; m < n; ++m) {
        double q1 = qp ? qp[m] : p->q[m >= p->ne];
        double s[3];
        for (int d = 0; d < 3; ++d) {
            double u = (double)p->K[d] * (x[3 * (size_t)m + d] + p->box[d] * 0.5) / p->box[d];
            tcpo_bspline(P, u, w[d], dw[d]);
            base[d] = (int)floor(u) - P / 2 + 1;
            s[d] = (double)p->K[d] / p->box[d];
        }
        double fx = 0.0, fy = 0.0, fz = 0.0, ph = 0.0;
        for (int i = 0; i < P; ++i) {
            int ix = ((base[0] + i) % K1 + K1) % K1;
            for (int j = 0; j < P; ++j) {
                int iy = ((base[1] + j) % K2 + K2) % K2;
                for (int k = 0; k < P; ++k) {
                    int iz = ((base[2] + k) % K3 + K3) % K3;
                    double c = creal(A[((size_t)iz * K2 + iy) * K1 + ix]);
                    ph += w[0][i] * w[1][j] * w[2][k] * c;
                    fx -= q1g * (q1 * dw[0][i] * w[1][j] * w[2][k] * s[0]) * c / V / PI;
                    fy -= q1g * (q1 * w[0][i] * dw[1][j] * w[2][k] * s[1]) * c / V / PI;
                    fz -= q1g * (q1 * w[0][i] * w[1][j] * dw[2][k] * s[2]) * c / V / PI;
                }
            }
        }
        frec[3 * (size_t)m + 0] = fx;
        frec[3 * (size_t)m + 1] = fy;
        frec[3 * (size_t)m + 2] = fz;
        if (pot) pot[m] = q1g * ph / V / PI;  /* dUm/dq_m = eps K^3 sum W Qfin/(pi V) */
    }
    *Um = p->eps * Epme / V / PI / 2.0; /* force.f90:921 */
    free(Q); free(psi); free(A); free(b1);
    return 0;
}

/* Direct reciprocal Ewald sum: E = 1/(2 pi V) sum_{m != 0} exp(-pi^2 m^2/alpha^2)/m^2 |S(m)|^2,
 * S(m) = sum_j q_j exp(2 pi i m.r_j), F_j = (2 eps q_j/V) sum_m G(m) m Im(conj(S) e^{2 pi i m.r_j}). */
static int kspace_core(const tcpo_params* p, double alpha, const double* x, const double* qp,
                       double* frec, double* pot, double* Um);

int tcpo_ewald_kspace(const tcpo_params* p, double alpha, const double* x, double* frec, double* Um)
{
    return kspace_core(p, alpha, x, NULL, frec, NULL, Um);
}

/* pot[j] = dUm/dq_j = eps/(pi V) sum_m G Re(conj(S) e^{2 pi i m.r_j}) */
static int kspace_core(const tcpo_params* p, double alpha, const double* x, const double* qp,
                       double* frec, double* pot, double* Um)
{
    const int n = p->ne + p->ni, km = p->kmax, nk = 2 * km + 1;
    if (km < 1) return -1;
    const double V = p->box[0] * p->box[1] * p->box[2];
    /* e^{2 pi i m x_d / L_d} for m = -km..km, per particle and axis */
    double complex* ph = (double complex*)malloc(sizeof(double complex) * (size_t)n * 3 * nk);
    if (!ph) return -2;
    for (int j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d)
            for (int m = -km; m <= km; ++m) {
                double th = 2.0 * PI * (double)m * x[3 * (size_t)j + d] / p->box[d];
                ph[((size_t)j * 3 + d) * nk + (m + km)] = cos(th) + I * sin(th);
            }
    double* fs = (double*)calloc((size_t)n * 4, sizeof(double));
    double* ps = fs + 3 * (size_t)n;
    double e = 0.0, ec = 0.0;
    for (int mx = -km; mx <= km; ++mx)
        for (int my = -km; my <= km; ++my)
            for (int mz = -km; mz <= km; ++mz) {
                if (mx == 0 && my == 0 && mz == 0) continue;
                double kx = mx / p->box[0], ky = my / p->box[1], kz = mz / p->box[2];
                double m2 = kx * kx + ky * ky + kz * kz;
                double G = exp(-PI * PI * m2 / (alpha * alpha)) / m2;
                double complex S = 0.0;
                for (int j = 0; j < n; ++j) {
                    const double complex* pj = ph + (size_t)j * 3 * nk;
                    S += (qp ? qp[j] : p->q[j >= p->ne]) * pj[mx + km] * pj[nk + my + km] * pj[2 * nk + mz + km];
                }
                nsum(&e, &ec, G * (creal(S) * creal(S) + cimag(S) * cimag(S)));
                for (int j = 0; j < n; ++j) {
                    const double complex* pj = ph + (size_t)j * 3 * nk;
                    double complex ej = pj[mx + km] * pj[nk + my + km] * pj[2 * nk + mz + km];
                    double s = cimag(conj(S) * ej) * G * (qp ? qp[j] : p->q[j >= p->ne]);
                    fs[3 * (size_t)j + 0] += s * kx;
                    fs[3 * (size_t)j + 1] += s * ky;
                    fs[3 * (size_t)j + 2] += s * kz;
                    ps[j] += creal(conj(S) * ej) * G;
                }
            }
    for (size_t t = 0; t < (size_t)n * 3; ++t) frec[t] = 2.0 * p->eps / V * fs[t];
    if (pot)
        for (int j = 0; j < n; ++j) pot[j] = p->eps / (PI * V) * ps[j];
    *Um = p->eps * (e + ec) / (2.0 * PI * V);
    free(fs);
    free(ph);
    return 0;
}

/* ------------------------------------------------------------- adjoint --- */

/* U = eps (Ur + Uo) + Um is homogeneous of degree 2 in the charges, so
 * dU/dq_i = eps (phi_i - 2 alpha q_i/sqrt(pi)) + dUm/dq_i with
 * phi_i = sum_j q_j (erfc(alpha r) + R)/r.  lambda enters only R:
 *   Deutsch R = -e^{-s}:            d(R/r)/dlambda = -e^{-s}/lambda^2,
 *   Kelbg   R = -e^{-s^2} + sqrt(pi) s erfc(s): d(R/r)/dlambda = -sqrt(pi) erfc(s)/lambda^2,
 * and dF_i/dlambda = -eps sum_j q_i q_j h'(r) r_ij/r with h = d(R/r)/dlambda:
 *   h' = e^{-s}/lambda^3 (Deutsch), 2 e^{-s^2}/lambda^3 (Kelbg). */
int tcpo_adjoint(const tcpo_params* p, const double* x, const double* qp, double* U,
                 double* dUdx, double* dUdq, double* dUdl, double* dFdl, int nthreads)
{
    const int ne = p->ne, n = p->ne + p->ni;
    double alpha, rc2, Uo;
    int rc = tcpo_ewald_self(p, &alpha, &rc2, &Uo);
    if (rc) return rc;
    double q2sum = 0.0;
    for (int i = 0; i < n; ++i) {
        const double qi = qp ? qp[i] : p->q[i >= ne];
        q2sum += qi * qi;
    }
    Uo = -q2sum * alpha / SQRTPI;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    double* ui = (double*)calloc((size_t)n * 2, sizeof(double));
    double* li = ui + n;
    if (!ui) return -2;
    const double lam = p->rs;
    const int kelbg = (p->flags & TCPO_KELBG) != 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int i = 0; i < n; ++i) {
        const int si = i >= ne;
        const double qi = qp ? qp[i] : p->q[si];
        double phs = 0.0, phc = 0.0, us = 0.0, uc = 0.0, ls = 0.0, lc = 0.0;
        double fs[3] = {0, 0, 0}, fc[3] = {0, 0, 0}, gs[3] = {0, 0, 0}, xr[3];
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double r2 = min_image(p->box, x + 3 * (size_t)i, x + 3 * (size_t)j, xr);
            if (!(r2 < rc2)) continue;
            const int sj = j >= ne;
            const double qj = qp ? qp[j] : p->q[sj];
            const int reg = si != sj || (si == 0 && sj == 0 && (p->flags & TCPO_REG_EE));
            const double r = sqrt(r2);
            const double ec = erfc(alpha * r);
            const double g = 2.0 * alpha * r * exp(-(alpha * alpha * r2)) / SQRTPI;
            double R = 0.0, num = g + ec, h = 0.0, hp = 0.0;
            if (reg) {
                const double s = r / lam;
                if (kelbg) {
                    const double y2 = exp(-(s * s));
                    R = -y2 + SQRTPI * s * erfc(s);
                    num -= y2;
                    h = -SQRTPI * erfc(s) / (lam * lam);
                    hp = 2.0 * y2 / (lam * lam * lam);
                } else {
                    const double y = exp(-s);
                    R = -y;
                    num -= y * (1.0 + s);
                    h = -y / (lam * lam);
                    hp = y / (lam * lam * lam);
                }
            }
            const double phi = (ec + R) / r;
            nsum(&phs, &phc, qj * phi);
            nsum(&us, &uc, qi * qj * phi);
            nsum(&ls, &lc, qi * qj * h);
            const double df = p->eps * qi * qj * num / (r2 * r);
            for (int k = 0; k < 3; ++k) {
                nsum(&fs[k], &fc[k], df * xr[k]);
                gs[k] -= p->eps * qi * qj * hp * xr[k] / r;
            }
        }
        dUdq[i] = p->eps * ((phs + phc) - 2.0 * alpha * qi / SQRTPI);
        for (int k = 0; k < 3; ++k) {
            dUdx[3 * (size_t)i + k] = -(fs[k] + fc[k]);
            if (dFdl) dFdl[3 * (size_t)i + k] = gs[k];
        }
        ui[i] = us + uc;
        li[i] = ls + lc;
    }
    double s = 0.0, c = 0.0, sl = 0.0, cl = 0.0;
    for (int i = 0; i < n; ++i) {
        nsum(&s, &c, ui[i]);
        nsum(&sl, &cl, li[i]);
    }
    double Um = 0.0;
    double* frec = (double*)calloc((size_t)n * 4, sizeof(double));
    double* pot = frec + 3 * (size_t)n;
    if (p->recip == TCPO_RECIP_SPME) rc = spme_core(p, alpha, x, qp, frec, pot, &Um);
    else if (p->recip == TCPO_RECIP_EWALD) rc = kspace_core(p, alpha, x, qp, frec, pot, &Um);
    if (!rc)
        for (int i = 0; i < n; ++i) {
            dUdq[i] += pot[i];
            for (int k = 0; k < 3; ++k) dUdx[3 * (size_t)i + k] -= frec[3 * (size_t)i + k];
        }
    *U = p->eps * (0.5 * (s + c) + Uo) + Um;
    *dUdl = p->eps * 0.5 * (sl + cl);
    free(frec);
    free(ui);
    return rc;
}

/* ------------------------------------------------------------ integrate --- */

/* REMOVE_RIGIDMTN, force.f90:986-1004 */
int tcpo_remove_rigidmtn(const tcpo_params* p, double* f)
{
    const int n = p->ne + p->ni;
    double F[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) F[k] = F[k] + f[3 * (size_t)i + k];
    for (int k = 0; k < 3; ++k) F[k] = F[k] / (double)n;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) f[3 * (size_t)i + k] = f[3 * (size_t)i + k] - F[k];
    return 0;
}

/* VVerletNotemp1 (vverlet.f90:148-167), VVerletStoch1 (:34-55) */
int tcpo_vverlet1(const tcpo_params* p, int thermo, double dt, double fric,
                  double* x, double* v, const double* f)
{
    const int n = p->ne + p->ni;
    for (int i = 0; i < n; ++i) {
        double xm = p->xm[i >= p->ne];
        for (int k = 0; k < 3; ++k) {
            size_t t = 3 * (size_t)i + k;
            if (thermo == TCPO_STOCH)
                v[t] = v[t] * (1.0 - 0.5 * fric * dt / xm) + 0.5 * dt * f[t] / xm;
            else
                v[t] = v[t] + 0.5 * dt * f[t] / xm;
            double xx = x[t] + dt * v[t];
            x[t] = xx - p->box[k] * round(xx / p->box[k]);
        }
    }
    return 0;
}

/* id = INT(mv2/3/dv) + 1, counted when id <= Ndist+1 (vverlet.f90:201-204); the
 * quotient is range-checked before the conversion, so a vanishing bin width
 * (T -> 0) counts nothing instead of converting an out-of-range value */
static void vdf_add(double* vdf, double mv2, double dv)
{
    double t = mv2 / 3.0 / dv;
    if (t >= 0.0 && t < (double)NBIN) vdf[(int)t] += 1.0;
}

/* VVerletNotemp2 (:190-219), VVerletBerend2 (:245-294), VVerletIsokin2 (:319-368),
 * VVerletStoch2 (:85-123) */
int tcpo_vverlet2(const tcpo_params* p, int thermo, double dt, double Te, double Ti,
                  double tau, double fric, double* v, double* f, const double* eta,
                  double* mv2el, double* mv2ion, double* vdf)
{
    double dummy[2 * NBIN];
    double* h = vdf ? vdf : dummy;
    double mv2s[2];
    for (int s = 0; s < 2; ++s) {
        const int lo = s ? p->ne : 0, hi = s ? p->ne + p->ni : p->ne;
        const double xm = p->xm[s], T = s ? Ti : Te, dv = s ? p->dv_i : p->dv_e;
        const int ntot = s ? p->ni : p->ne;
        double acc = 0.0;
        if (thermo == TCPO_BEREND || thermo == TCPO_ISOKIN) {
            for (int i = lo; i < hi; ++i)
                for (int k = 0; k < 3; ++k) {
                    size_t t = 3 * (size_t)i + k;
                    v[t] = v[t] + 0.5 * dt * f[t] / xm;
                    acc = acc + xm * (v[t] * v[t]);
                }
            double T0 = acc / (double)(3 * ntot);
            double lambda = thermo == TCPO_BEREND ? sqrt(1.0 + dt * (T / T0 - 1.0) / tau)
                                                  : sqrt(T / T0);
            acc = 0.0;
            for (int i = lo; i < hi; ++i) {
                size_t t = 3 * (size_t)i;
                for (int k = 0; k < 3; ++k) v[t + k] = v[t + k] * lambda;
                double mv2 = (v[t] * v[t] + v[t + 1] * v[t + 1] + v[t + 2] * v[t + 2]) * xm;
                acc = acc + mv2;
                vdf_add(h + s * NBIN, mv2, dv);
            }
        } else {
            double beta = sqrt(2.0 * fric * T / dt);
            for (int i = lo; i < hi; ++i) {
                size_t t = 3 * (size_t)i;
                for (int k = 0; k < 3; ++k) {
                    if (thermo == TCPO_STOCH) {
                        f[t + k] = f[t + k] + eta[t + k] * beta;
                        v[t + k] = (v[t + k] + 0.5 * dt * f[t + k] / xm) / (1.0 + 0.5 * fric * dt / xm);
                    } else {
                        v[t + k] = v[t + k] + 0.5 * dt * f[t + k] / xm;
                    }
                }
                double mv2 = (v[t] * v[t] + v[t + 1] * v[t + 1] + v[t + 2] * v[t + 2]) * xm;
                acc = acc + mv2;
                vdf_add(h + s * NBIN, mv2, dv);
            }
        }
        mv2s[s] = acc;
    }
    *mv2el = mv2s[0];
    *mv2ion = mv2s[1];
    return 0;
}

/* main.f90:87-153 */
int tcpo_step(const tcpo_params* p, int thermo, double dt, double Te, double Ti,
              double tau, double fric, double* x, double* v, double* f,
              double* rdf, double* vdf, double* out, int fast, int nthreads)
{
    const int n = p->ne + p->ni;
    double alpha, rc2, Uo, Ur, Um = 0.0;
    int rc = tcpo_ewald_self(p, &alpha, &rc2, &Uo);
    if (rc) return rc;
    if (thermo == TCPO_STOCH) return -3; /* noise needs an eta source: use the pieces */
    tcpo_vverlet1(p, thermo, dt, fric, x, v, f);
    double* frec = (double*)calloc((size_t)n * 3, sizeof(double));
    if (!frec) return -2;
    if (p->recip == TCPO_RECIP_SPME) rc = tcpo_spme(p, alpha, x, frec, &Um);
    else if (p->recip == TCPO_RECIP_EWALD) rc = tcpo_ewald_kspace(p, alpha, x, frec, &Um);
    if (rc) { free(frec); return rc; }
    if (fast) tcpo_force_allpair(p, alpha, rc2, x, f, &Ur, rdf, nthreads);
    else tcpo_force_direct_self(p, alpha, rc2, x, f, &Ur, rdf);
    if (p->recip != TCPO_RECIP_NONE)
        for (size_t t = 0; t < (size_t)n * 3; ++t) f[t] = f[t] + frec[t]; /* force.f90:953-964 */
    free(frec);
    double Epot = (Ur + Uo) * p->eps; /* main.f90:138 */
    tcpo_remove_rigidmtn(p, f);
    double mv2el, mv2ion;
    tcpo_vverlet2(p, thermo, dt, Te, Ti, tau, fric, v, f, NULL, &mv2el, &mv2ion, vdf);
    out[0] = Ur; out[1] = Uo; out[2] = Um; out[3] = Epot; out[4] = mv2el; out[5] = mv2ion;
    return 0;
}

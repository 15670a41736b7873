/*
 * shtk_oracle.c -- CPU restatement of SHTnsKit.jl's spherical-harmonic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the FP64 checker the GPU library is
 * compared against, and the timed CPU baseline.  Nothing in the product path
 * (shtnskit.jl_amd/) links, loads or calls it; only tests/, the smoke check and
 * the cpu_baseline leg of bench.py do.
 *
 * It restates the reference algorithm step for step, including the costs the
 * GPU design removes (per-step sqrt/divide in the recurrence, no north/south
 * folding), so that it is both an oracle and a fair "reference CPU path":
 *
 *   gauss_legendre            src/legendre.jl:956-1003
 *   plm_norm_row              src/legendre.jl:233-322   (Plm_norm_row!, scaled seed
 *                                                         and scaled recurrence)
 *   plm_dtheta_over_sinth_row src/legendre.jl:326-338, 470-555
 *   m order / threading       src/buffer_utils.jl:16-52 (balanced_m_order, @threads :static)
 *   scale conventions         src/SHTnsKit.jl:141-177, src/normalization.jl:114-334
 *   scalar synthesis          src/core_transforms.jl:294-344, 749-869, src/kernels.jl:57-65
 *   scalar analysis           src/core_transforms.jl:169-207, 655-746, src/kernels.jl:33-40
 *   vector (sphtor)           src/sphtor_transforms.jl:162-289, 455-612, src/kernels.jl:155-215
 *   batch                     src/batch_transforms.jl:372-553
 *   adjoints                  src/core_transforms.jl:477-652, src/sphtor_transforms.jl:385-452
 *
 * The phi FFT is FFTW in the reference (unavailable here); this file uses its own
 * mixed-radix complex FFT with the same conventions (forward unnormalized
 * exp(-2 pi i jk/N), inverse with 1/N), src/fftutils.jl:283-305, 323-432.
 *
 * Array layout follows Julia's column-major convention:
 *   grid  f[i + nlat*j (+ nlat*nlon*k)]                 (latitude fastest, north first)
 *   coeff a[l + (lmax+1)*m (+ (lmax+1)*(mmax+1)*k)]     complex as (re, im) doubles
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

typedef struct {
    int lmax, mmax, mres, nlat, nlon;
    int norm;        /* 0 orthonormal, 1 fourpi, 2 schmidt */
    int cs_phase;    /* 1 = Condon-Shortley phase */
    int real_norm;   /* 1 = SHTns REAL_NORM (m>0 stored sqrt(2) larger) */
    int robert_form; /* 1 = Robert form for vector fields */
    int phi_quad;    /* 0 = :dft, 1 = :quad */
    const double *x; /* cos(theta), north first, length nlat */
    const double *w; /* Gauss weights, length nlat */
} ora_cfg;

static int g_threads = 0; /* 0: OpenMP default */

void ora_set_threads(int n) { g_threads = n; }

int ora_max_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre nodes and weights: src/legendre.jl:956-1003.  Ascending x.  */
/* ------------------------------------------------------------------------ */
int ora_gauss_legendre(int n, double *x, double *w) {
    if (n <= 0) return -1;
    int half = (n + 1) / 2;
    for (int k = 1; k <= half; ++k) {
        double z = cos(M_PI * (k - 0.25) / (n + 0.5));
        double z1, pnm1, pn, pd;
        for (int it = 0; it < 50; ++it) {
            pnm1 = 1.0;
            pn = z;
            for (int l = 2; l <= n; ++l) {
                double pnp1 = ((2.0 * l - 1.0) * z * pn - (l - 1.0) * pnm1) / l;
                pnm1 = pn;
                pn = pnp1;
            }
            pd = n * (z * pn - pnm1) / (z * z - 1.0);
            z1 = z;
            z -= pn / pd;
            if (fabs(z - z1) < 1e-15) break;
        }
        pnm1 = 1.0;
        pn = z;
        for (int l = 2; l <= n; ++l) {
            double pnp1 = ((2.0 * l - 1.0) * z * pn - (l - 1.0) * pnm1) / l;
            pnm1 = pn;
            pn = pnp1;
        }
        pd = n * (z * pn - pnm1) / (z * z - 1.0);
        x[k - 1] = -z;
        x[n - k] = z;
        double wk = 2.0 / ((1.0 - z * z) * pd * pd);
        w[k - 1] = wk;
        w[n - k] = wk;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Orthonormal + CS associated Legendre rows: src/legendre.jl:197-322.        */
/* ------------------------------------------------------------------------ */
#define INV_SQRT_4PI 0.28209479177387814
#define PLM_BITS 480

static inline double sin_from_cos(double x) {
    double v = (1.0 - x) * (1.0 + x);
    return sqrt(v > 0.0 ? v : 0.0);
}

/* _plm_sectoral_seed (legendre.jl:252-271): pmm * 2^e = Pbar_m^m. */
static void plm_sectoral_seed(double s, int m, double *pmm_out, int *e_out) {
    double pmm = INV_SQRT_4PI;
    int e = 0;
    const double tiny = ldexp(1.0, -PLM_BITS);
    const int shift = 2 * PLM_BITS;
    const double up = ldexp(1.0, shift);
    int k = 1;
    while (k <= m) {
        int kend = k + 7 < m ? k + 7 : m;
        for (int j = k; j <= kend; ++j) pmm = -sqrt((2.0 * j + 1.0) / (2.0 * j)) * s * pmm;
        while (fabs(pmm) < tiny && pmm != 0.0) {
            pmm *= up;
            e -= shift;
        }
        k = kend + 1;
    }
    *pmm_out = pmm;
    *e_out = e;
}

static inline void plm_ab(int l, int m, double *a, double *b) {
    /* legendre.jl:277-278, recomputed every step as the reference does */
    *a = sqrt(((2.0 * l - 1.0) * (2.0 * l + 1.0)) / ((double)(l - m) * (double)(l + m)));
    *b = sqrt(((2.0 * l + 1.0) * (double)(l - 1 - m) * (double)(l - 1 + m)) /
              ((2.0 * l - 3.0) * (double)(l - m) * (double)(l + m)));
}

static void plm_recurrence(double *P, double x, int lfirst, int lmax, int m) {
    for (int l = lfirst; l <= lmax; ++l) {
        double a, b;
        plm_ab(l, m, &a, &b);
        P[l] = a * x * P[l - 1] - b * P[l - 2];
    }
}

/* Fills P[l] = Pbar_l^m(x) for l = m..lmax, zero for l < m.  P has lmax+1 entries. */
void ora_plm_row(double *P, double x, int lmax, int m) {
    for (int l = 0; l <= lmax; ++l) P[l] = 0.0;
    if (m < 0 || lmax < m) return;
    double s = sin_from_cos(x);
    double pmm;
    int e;
    plm_sectoral_seed(s, m, &pmm, &e);
    if (e == 0) {
        P[m] = pmm;
        if (lmax == m) return;
        P[m + 1] = sqrt(2.0 * m + 3.0) * x * pmm;
        plm_recurrence(P, x, m + 2, lmax, m);
        return;
    }
    /* _plm_norm_row_scaled! (legendre.jl:287-322) */
    double factor = ldexp(1.0, e);
    P[m] = pmm * factor;
    if (lmax == m) return;
    double p2 = pmm;
    double p1 = sqrt(2.0 * m + 3.0) * x * pmm;
    P[m + 1] = p1 * factor;
    const double tiny = ldexp(1.0, -PLM_BITS);
    const double huge = ldexp(1.0, PLM_BITS);
    const int shift = 2 * PLM_BITS;
    const double up = ldexp(1.0, shift), down = ldexp(1.0, -shift);
    int l = m + 2;
    while (l <= lmax) {
        double a, b;
        plm_ab(l, m, &a, &b);
        double p = a * x * p1 - b * p2;
        P[l] = p * factor;
        p2 = p1;
        p1 = p;
        l += 1;
        double mag = fabs(p2) > fabs(p1) ? fabs(p2) : fabs(p1);
        if (mag > huge) {
            p2 *= down;
            p1 *= down;
            e += shift;
            if (e == 0) break;
            factor = ldexp(1.0, e);
        } else if (mag < tiny && mag != 0.0) {
            p2 *= up;
            p1 *= up;
            e -= shift;
            factor = ldexp(1.0, e);
        }
    }
    plm_recurrence(P, x, l, lmax, m);
}

/* P, dP/dtheta and P/sin(theta): legendre.jl:326-338 (zonal) and 470-555.
 * Pbuf must hold lmax+2 doubles. */
void ora_plm_dtheta_over_sinth_row(double *P, double *dP, double *Ps, double *Pbuf,
                                   double x, int lmax, int m) {
    for (int l = 0; l <= lmax; ++l) {
        dP[l] = 0.0;
        Ps[l] = 0.0;
    }
    double sinth = sin_from_cos(x);
    int interior = (lmax >= m) && (sinth >= 100.0 * 2.220446049250313e-16);
    if (m == 0) {
        ora_plm_row(P, x, lmax, m);
        if (interior) {
            double d = 0.0;
            for (int l = 1; l <= lmax; ++l) {
                d = sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0)) * (x * d - (double)l * sinth * P[l - 1]);
                dP[l] = d;
            }
            double inv = 1.0 / sinth;
            for (int l = 0; l <= lmax; ++l) Ps[l] = P[l] * inv;
        }
        return;
    }
    if (interior) {
        ora_plm_row(Pbuf, x, lmax + 1, m);
        for (int l = 0; l <= lmax; ++l) P[l] = Pbuf[l];
    } else {
        ora_plm_row(P, x, lmax, m);
    }
    if (lmax < m) return;
    if (!interior) {
        if (m == 1) {
            for (int l = 1; l <= lmax; ++l) {
                double mag = 0.5 * INV_SQRT_4PI * sqrt((2.0 * l + 1.0) * l * (l + 1.0));
                if (x > 0) {
                    dP[l] = -mag;
                    Ps[l] = -mag;
                } else {
                    dP[l] = ((l + 1) & 1 ? -1.0 : 1.0) * mag;
                    Ps[l] = (l & 1 ? -1.0 : 1.0) * mag;
                }
            }
        }
        return;
    }
    double inv = 1.0 / sinth;
    for (int l = m; l <= lmax; ++l) {
        double v1 = ((double)(l + 1) * (l + 1) - (double)m * m) / (4.0 * (l + 1) * (l + 1) - 1.0);
        double dl1m = sqrt(v1 > 0 ? v1 : 0.0);
        double dlm = 0.0;
        if (!(l == 0 || l == m)) {
            double v = ((double)l * l - (double)m * m) / (4.0 * l * l - 1.0);
            dlm = sqrt(v > 0 ? v : 0.0);
        }
        double Pm1 = l > 0 ? Pbuf[l - 1] : 0.0;
        double Pp1 = Pbuf[l + 1];
        dP[l] = ((double)l * dl1m * Pp1 - (double)(l + 1) * dlm * Pm1) * inv;
        Ps[l] = Pbuf[l] * inv;
    }
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (any length).  Conventions of src/fftutils.jl.    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    cplx *wf; /* exp(-2 pi i t / n) */
    cplx *wb; /* exp(+2 pi i t / n) */
} fft_tab;

static void fft_tab_init(fft_tab *t, int n) {
    t->n = n;
    t->wf = (cplx *)malloc(sizeof(cplx) * n);
    t->wb = (cplx *)malloc(sizeof(cplx) * n);
    for (int k = 0; k < n; ++k) {
        /* angle reduced to the first octant pattern via long double for accuracy */
        long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        double c = (double)cosl(ang), s = (double)sinl(ang);
        t->wf[k] = c - I * s;
        t->wb[k] = c + I * s;
    }
}

static void fft_tab_free(fft_tab *t) {
    free(t->wf);
    free(t->wb);
}

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* X[k] = sum_j x[j*stride] W^{jk}, W = table root of order n (n divides N). */
static void fft_rec(const cplx *x, int stride, cplx *X, int n, const cplx *W, int N, cplx *tmp) {
    if (n == 1) {
        X[0] = x[0];
        return;
    }
    int p = smallest_factor(n), q = n / p;
    for (int r = 0; r < p; ++r) fft_rec(x + (size_t)r * stride, stride * p, X + (size_t)r * q, q, W, N, tmp);
    int step_n = N / n, step_p = N / p;
    cplx *t = tmp, *o = tmp + p;
    for (int k = 0; k < q; ++k) {
        for (int r = 0; r < p; ++r) t[r] = X[r * q + k] * W[((size_t)r * k * step_n) % N];
        for (int s = 0; s < p; ++s) {
            cplx acc = 0;
            for (int r = 0; r < p; ++r) acc += t[r] * W[((size_t)r * s * step_p) % N];
            o[s] = acc;
        }
        for (int s = 0; s < p; ++s) X[k + q * s] = o[s];
    }
}

/* In-place transform of one row held contiguously.  inverse: includes 1/n. */
static void fft_row(cplx *row, const fft_tab *t, int inverse, cplx *work, cplx *tmp) {
    int n = t->n;
    fft_rec(row, 1, work, n, inverse ? t->wb : t->wf, n, tmp);
    double sc = inverse ? 1.0 / n : 1.0;
    for (int k = 0; k < n; ++k) row[k] = work[k] * sc;
}

void ora_fft(double *data, int n, int inverse) {
    fft_tab t;
    fft_tab_init(&t, n);
    cplx *work = (cplx *)malloc(sizeof(cplx) * n);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * 2 * n);
    fft_row((cplx *)data, &t, inverse, work, tmp);
    free(work);
    free(tmp);
    fft_tab_free(&t);
}

/* Phi transform of nlat rows of a column-major (nlat, nlon) complex matrix. */
static void fft_phi(cplx *F, int nlat, int nlon, int inverse) {
    fft_tab t;
    fft_tab_init(&t, nlon);
#pragma omp parallel num_threads(ora_max_threads())
    {
        cplx *row = (cplx *)malloc(sizeof(cplx) * nlon);
        cplx *work = (cplx *)malloc(sizeof(cplx) * nlon);
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * 2 * nlon);
#pragma omp for schedule(static)
        for (int i = 0; i < nlat; ++i) {
            for (int j = 0; j < nlon; ++j) row[j] = F[i + (size_t)nlat * j];
            fft_row(row, &t, inverse, work, tmp);
            for (int j = 0; j < nlon; ++j) F[i + (size_t)nlat * j] = row[j];
        }
        free(row);
        free(work);
        free(tmp);
    }
    fft_tab_free(&t);
}

/* ------------------------------------------------------------------------ */
/* Conventions: src/SHTnsKit.jl:141-177, src/normalization.jl:114-153.       */
/* ------------------------------------------------------------------------ */
static double phi_inv_scale(const ora_cfg *c) {
    return c->phi_quad ? c->nlon / (2.0 * M_PI) : (double)c->nlon;
}
static double analysis_phi_scale(const ora_cfg *c) {
    return (2.0 * M_PI / c->nlon) * c->nlon / phi_inv_scale(c);
}
/* a_canonical = s * a_configured */
double ora_coef_scale(const ora_cfg *c, int l, int m) {
    double nrm = 1.0;
    if (c->norm == 1) nrm = sqrt(4.0 * M_PI);
    else if (c->norm == 2) nrm = sqrt(4.0 * M_PI / (2.0 * l + 1.0));
    double rs = (c->real_norm && m > 0) ? 1.0 / sqrt(2.0) : 1.0;
    double ph = (c->cs_phase) ? 1.0 : ((m & 1) ? -1.0 : 1.0);
    return nrm * rs * ph;
}
static int canonical(const ora_cfg *c) { return c->norm == 0 && !c->real_norm && c->cs_phase; }

/* balanced_m_order: src/buffer_utils.jl:16-25 */
static int *m_order(const ora_cfg *c, int *n_out) {
    int n = c->mmax / c->mres + 1;
    int *o = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        int im = (i % 2 == 0) ? i / 2 : (n - 1) - i / 2;
        o[i] = im * c->mres;
    }
    *n_out = n;
    return o;
}

static inline cplx A(const double *a, int lmax, int l, int m) {
    size_t k = (size_t)l + (size_t)(lmax + 1) * m;
    return a[2 * k] + I * a[2 * k + 1];
}
static inline void setA(double *a, int lmax, int l, int m, cplx v) {
    size_t k = (size_t)l + (size_t)(lmax + 1) * m;
    a[2 * k] = creal(v);
    a[2 * k + 1] = cimag(v);
}

/* ------------------------------------------------------------------------ */
/* Scalar synthesis (batch of nf fields sharing each Legendre row, as         */
/* src/batch_transforms.jl:489-553 does; nf=1 is core_transforms.jl:294-344). */
/* out: real (nlat, nlon, nf) if real_output, else complex.                  */
/* ------------------------------------------------------------------------ */
int ora_synthesis(const ora_cfg *c, const double *alm, double *out, int nf, int real_output, int ltr) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    if (ltr < 0 || ltr > lmax) ltr = lmax;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    double inv_scale = phi_inv_scale(c);
    int no;
    int *order = m_order(c, &no);
    cplx *F = (cplx *)calloc(nspat * nf, sizeof(cplx));
#pragma omp parallel num_threads(ora_max_threads())
    {
        double *P = (double *)malloc(sizeof(double) * (lmax + 1));
#pragma omp for schedule(static)
        for (int idx = 0; idx < no; ++idx) {
            int m = order[idx];
            for (int i = 0; i < nlat; ++i) {
                ora_plm_row(P, c->x[i], lmax, m);
                for (int k = 0; k < nf; ++k) {
                    const double *a = alm + 2 * nalm * k;
                    cplx acc = 0;
                    for (int l = m; l <= ltr; ++l) {
                        double s = canonical(c) ? 1.0 : ora_coef_scale(c, l, m);
                        acc += P[l] * (s * A(a, lmax, l, m));
                    }
                    F[nspat * k + i + (size_t)nlat * m] = inv_scale * acc;
                }
            }
        }
        free(P);
    }
    free(order);
    for (int k = 0; k < nf; ++k) {
        cplx *Fk = F + nspat * k;
        if (real_output)
            for (int m = 1; m <= mmax; ++m)
                for (int i = 0; i < nlat; ++i) Fk[i + (size_t)nlat * (nlon - m)] = conj(Fk[i + (size_t)nlat * m]);
        fft_phi(Fk, nlat, nlon, 1);
        for (size_t t = 0; t < nspat; ++t) {
            if (real_output) out[nspat * k + t] = creal(Fk[t]);
            else {
                out[2 * (nspat * k + t)] = creal(Fk[t]);
                out[2 * (nspat * k + t) + 1] = cimag(Fk[t]);
            }
        }
    }
    free(F);
    return 0;
}

/* Scalar analysis of real fields: core_transforms.jl:169-207 (nf=1) and the
 * batch form batch_transforms.jl:372-429 (row shared, scale applied last).
 * ora_analysis_cplx takes complex fields (interleaved re/im): the same m-loop
 * over the bins m >= 0 of the complex FFT (core_transforms.jl:196-206 with a
 * complex eltype). */
static int analysis_fourier(const ora_cfg *c, cplx *F, double *alm, int nf, int ltr);

int ora_analysis(const ora_cfg *c, const double *f, double *alm, int nf, int ltr) {
    size_t nspat = (size_t)c->nlat * c->nlon;
    cplx *F = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    for (size_t t = 0; t < nspat * nf; ++t) F[t] = f[t];
    return analysis_fourier(c, F, alm, nf, ltr);
}

int ora_analysis_cplx(const ora_cfg *c, const double *f, double *alm, int nf, int ltr) {
    size_t nspat = (size_t)c->nlat * c->nlon;
    cplx *F = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    for (size_t t = 0; t < nspat * nf; ++t) F[t] = f[2 * t] + I * f[2 * t + 1];
    return analysis_fourier(c, F, alm, nf, ltr);
}

/* Takes ownership of F (grid values, transformed and freed here). */
static int analysis_fourier(const ora_cfg *c, cplx *F, double *alm, int nf, int ltr) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    if (ltr < 0 || ltr > lmax) ltr = lmax;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    for (int k = 0; k < nf; ++k) fft_phi(F + nspat * k, nlat, nlon, 0);
    memset(alm, 0, sizeof(double) * 2 * nalm * nf);
    double scale_phi = analysis_phi_scale(c);
    int no;
    int *order = m_order(c, &no);
#pragma omp parallel num_threads(ora_max_threads())
    {
        double *P = (double *)malloc(sizeof(double) * (lmax + 1));
        cplx *acc = (cplx *)malloc(sizeof(cplx) * (lmax + 1) * nf);
#pragma omp for schedule(static)
        for (int idx = 0; idx < no; ++idx) {
            int m = order[idx];
            for (size_t t = 0; t < (size_t)(lmax + 1) * nf; ++t) acc[t] = 0;
            for (int i = 0; i < nlat; ++i) {
                ora_plm_row(P, c->x[i], lmax, m);
                double wi = c->w[i];
                for (int k = 0; k < nf; ++k) {
                    cplx Fi = F[nspat * k + i + (size_t)nlat * m];
                    cplx *ak = acc + (size_t)(lmax + 1) * k;
                    if (nf == 1)
                        for (int l = m; l <= ltr; ++l) ak[l] += (wi * P[l] * scale_phi) * Fi;
                    else
                        for (int l = m; l <= ltr; ++l) ak[l] += (wi * P[l]) * Fi;
                }
            }
            for (int k = 0; k < nf; ++k)
                for (int l = m; l <= ltr; ++l) {
                    cplx v = acc[(size_t)(lmax + 1) * k + l];
                    if (nf != 1) v *= scale_phi;
                    if (!canonical(c)) v /= ora_coef_scale(c, l, m);
                    setA(alm + 2 * nalm * k, lmax, l, m, v);
                }
        }
        free(P);
        free(acc);
    }
    free(order);
    free(F);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Vector (spheroidal/toroidal) transforms: sphtor_transforms.jl:162-289,   */
/* 455-612; kernels.jl:155-215.  Batch loops fields (batch_transforms.jl:    */
/* 689-815 routes each field through one plan).                              */
/* ------------------------------------------------------------------------ */
int ora_synthesis_sphtor(const ora_cfg *c, const double *S, const double *T, double *Vt, double *Vp, int nf,
                         int real_output, int ltr) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    if (ltr < 0 || ltr > lmax) ltr = lmax;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    double inv_scale = phi_inv_scale(c);
    int no;
    int *order = m_order(c, &no);
    cplx *Ft = (cplx *)calloc(nspat * nf, sizeof(cplx));
    cplx *Fp = (cplx *)calloc(nspat * nf, sizeof(cplx));
#pragma omp parallel num_threads(ora_max_threads())
    {
        double *P = (double *)malloc(sizeof(double) * (lmax + 1));
        double *dP = (double *)malloc(sizeof(double) * (lmax + 1));
        double *Ps = (double *)malloc(sizeof(double) * (lmax + 1));
        double *Pb = (double *)malloc(sizeof(double) * (lmax + 2));
#pragma omp for schedule(static)
        for (int idx = 0; idx < no; ++idx) {
            int m = order[idx];
            for (int i = 0; i < nlat; ++i) {
                ora_plm_dtheta_over_sinth_row(P, dP, Ps, Pb, c->x[i], lmax, m);
                for (int k = 0; k < nf; ++k) {
                    cplx gt = 0, gp = 0;
                    for (int l = (m > 1 ? m : 1); l <= ltr; ++l) {
                        double s = canonical(c) ? 1.0 : ora_coef_scale(c, l, m);
                        cplx Sl = s * A(S + 2 * nalm * k, lmax, l, m);
                        cplx Tl = s * A(T + 2 * nalm * k, lmax, l, m);
                        double dY = dP[l], Yo = Ps[l];
                        gt += dY * Sl - I * (double)m * Yo * Tl;
                        gp += I * (double)m * Yo * Sl + dY * Tl;
                    }
                    Ft[nspat * k + i + (size_t)nlat * m] = inv_scale * gt;
                    Fp[nspat * k + i + (size_t)nlat * m] = inv_scale * gp;
                }
            }
        }
        free(P);
        free(dP);
        free(Ps);
        free(Pb);
    }
    free(order);
    for (int k = 0; k < nf; ++k) {
        cplx *Fs[2] = {Ft + nspat * k, Fp + nspat * k};
        double *outs[2] = {Vt, Vp};
        for (int q = 0; q < 2; ++q) {
            cplx *Fk = Fs[q];
            if (real_output)
                for (int m = 1; m <= mmax; ++m)
                    for (int i = 0; i < nlat; ++i)
                        Fk[i + (size_t)nlat * (nlon - m)] = conj(Fk[i + (size_t)nlat * m]);
            fft_phi(Fk, nlat, nlon, 1);
            for (int j = 0; j < nlon; ++j)
                for (int i = 0; i < nlat; ++i) {
                    size_t t = i + (size_t)nlat * j;
                    cplx v = Fk[t];
                    if (c->robert_form) v *= sin_from_cos(c->x[i]);
                    if (real_output) outs[q][nspat * k + t] = creal(v);
                    else {
                        outs[q][2 * (nspat * k + t)] = creal(v);
                        outs[q][2 * (nspat * k + t) + 1] = cimag(v);
                    }
                }
        }
    }
    free(Ft);
    free(Fp);
    return 0;
}

int ora_analysis_sphtor(const ora_cfg *c, const double *Vt, const double *Vp, double *S, double *T, int nf,
                        int ltr) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    if (ltr < 0 || ltr > lmax) ltr = lmax;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    cplx *Ft = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    cplx *Fp = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    for (size_t t = 0; t < nspat * nf; ++t) {
        Ft[t] = Vt[t];
        Fp[t] = Vp[t];
    }
    for (int k = 0; k < nf; ++k) {
        fft_phi(Ft + nspat * k, nlat, nlon, 0);
        fft_phi(Fp + nspat * k, nlat, nlon, 0);
    }
    memset(S, 0, sizeof(double) * 2 * nalm * nf);
    memset(T, 0, sizeof(double) * 2 * nalm * nf);
    double scale_phi = analysis_phi_scale(c);
    int no;
    int *order = m_order(c, &no);
#pragma omp parallel num_threads(ora_max_threads())
    {
        double *P = (double *)malloc(sizeof(double) * (lmax + 1));
        double *dP = (double *)malloc(sizeof(double) * (lmax + 1));
        double *Ps = (double *)malloc(sizeof(double) * (lmax + 1));
        double *Pb = (double *)malloc(sizeof(double) * (lmax + 2));
        cplx *accS = (cplx *)malloc(sizeof(cplx) * (lmax + 1));
        cplx *accT = (cplx *)malloc(sizeof(cplx) * (lmax + 1));
#pragma omp for schedule(static)
        for (int idx = 0; idx < no; ++idx) {
            int m = order[idx];
            for (int k = 0; k < nf; ++k) {
                for (int l = 0; l <= lmax; ++l) accS[l] = accT[l] = 0;
                for (int i = 0; i < nlat; ++i) {
                    ora_plm_dtheta_over_sinth_row(P, dP, Ps, Pb, c->x[i], lmax, m);
                    double wi = c->w[i];
                    cplx Fti = Ft[nspat * k + i + (size_t)nlat * m], Fpi = Fp[nspat * k + i + (size_t)nlat * m];
                    if (c->robert_form) {
                        double st = sin_from_cos(c->x[i]);
                        if (st > 0) {
                            Fti /= st;
                            Fpi /= st;
                        }
                    }
                    for (int l = (m > 1 ? m : 1); l <= ltr; ++l) {
                        double coeff = wi * scale_phi / ((double)l * (l + 1));
                        cplx term = I * (double)m * Ps[l];
                        double dY = dP[l];
                        accS[l] += coeff * (Fti * dY + conj(term) * Fpi);
                        accT[l] += coeff * (-conj(term) * Fti + dY * Fpi);
                    }
                }
                for (int l = m; l <= ltr; ++l) {
                    double s = canonical(c) ? 1.0 : ora_coef_scale(c, l, m);
                    setA(S + 2 * nalm * k, lmax, l, m, accS[l] / s);
                    setA(T + 2 * nalm * k, lmax, l, m, accT[l] / s);
                }
            }
        }
        free(P);
        free(dP);
        free(Ps);
        free(Pb);
        free(accS);
        free(accT);
    }
    free(order);
    free(Ft);
    free(Fp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoints: core_transforms.jl:477-652, sphtor_transforms.jl:385-452.       */
/* ------------------------------------------------------------------------ */

/* _adjoint_synthesis: real cotangent fbar (nlat,nlon,nf) -> almbar. */
int ora_adjoint_synthesis(const ora_cfg *c, const double *fbar, double *almbar, int nf, int real_output) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    cplx *F = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    for (size_t t = 0; t < nspat * nf; ++t) F[t] = fbar[t];
    for (int k = 0; k < nf; ++k) fft_phi(F + nspat * k, nlat, nlon, 0);
    memset(almbar, 0, sizeof(double) * 2 * nalm * nf);
    double phiadj = phi_inv_scale(c) / nlon;
    double *P = (double *)malloc(sizeof(double) * (lmax + 1));
    for (int k = 0; k < nf; ++k)
        for (int m = 0; m <= mmax; m += c->mres) {
            double wm = ((m == 0 || !real_output) ? 1.0 : 2.0) * phiadj;
            for (int i = 0; i < nlat; ++i) {
                ora_plm_row(P, c->x[i], lmax, m);
                cplx Fi = wm * F[nspat * k + i + (size_t)nlat * m];
                for (int l = m; l <= lmax; ++l) {
                    size_t q = 2 * (nalm * k + l + (size_t)(lmax + 1) * m);
                    cplx v = P[l] * Fi;
                    almbar[q] += creal(v);
                    almbar[q + 1] += cimag(v);
                }
            }
            if (!canonical(c))
                for (int l = m; l <= lmax; ++l) {
                    size_t q = 2 * (nalm * k + l + (size_t)(lmax + 1) * m);
                    double s = ora_coef_scale(c, l, m);
                    almbar[q] *= s;
                    almbar[q + 1] *= s;
                }
        }
    free(P);
    free(F);
    return 0;
}

/* _adjoint_analysis: almbar -> complex (nlat, nlon, nf) cotangent. */
int ora_adjoint_analysis(const ora_cfg *c, const double *almbar, double *out_cplx, int nf) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    cplx *F = (cplx *)calloc(nspat * nf, sizeof(cplx));
    double phiadj = nlon * analysis_phi_scale(c);
    double *P = (double *)malloc(sizeof(double) * (lmax + 1));
    for (int k = 0; k < nf; ++k) {
        const double *a = almbar + 2 * nalm * k;
        for (int m = 0; m <= mmax; m += c->mres)
            for (int i = 0; i < nlat; ++i) {
                ora_plm_row(P, c->x[i], lmax, m);
                cplx acc = 0;
                for (int l = m; l <= lmax; ++l) {
                    cplx v = A(a, lmax, l, m);
                    if (!canonical(c)) v /= ora_coef_scale(c, l, m);
                    acc += P[l] * v;
                }
                F[nspat * k + i + (size_t)nlat * m] = (phiadj * c->w[i]) * acc;
            }
        fft_phi(F + nspat * k, nlat, nlon, 1);
    }
    for (size_t t = 0; t < nspat * nf; ++t) {
        out_cplx[2 * t] = creal(F[t]);
        out_cplx[2 * t + 1] = cimag(F[t]);
    }
    free(P);
    free(F);
    return 0;
}

int ora_adjoint_synthesis_sphtor(const ora_cfg *c, const double *Vtb, const double *Vpb, double *Sb, double *Tb,
                                 int nf, int real_output) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    cplx *Ft = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    cplx *Fp = (cplx *)malloc(sizeof(cplx) * nspat * nf);
    for (int k = 0; k < nf; ++k)
        for (int j = 0; j < nlon; ++j)
            for (int i = 0; i < nlat; ++i) {
                size_t t = nspat * k + i + (size_t)nlat * j;
                double sc = c->robert_form ? sin_from_cos(c->x[i]) : 1.0;
                Ft[t] = sc * Vtb[t];
                Fp[t] = sc * Vpb[t];
            }
    for (int k = 0; k < nf; ++k) {
        fft_phi(Ft + nspat * k, nlat, nlon, 0);
        fft_phi(Fp + nspat * k, nlat, nlon, 0);
    }
    memset(Sb, 0, sizeof(double) * 2 * nalm * nf);
    memset(Tb, 0, sizeof(double) * 2 * nalm * nf);
    double phiadj = phi_inv_scale(c) / nlon;
    double *P = (double *)malloc(sizeof(double) * (lmax + 1));
    double *dP = (double *)malloc(sizeof(double) * (lmax + 1));
    double *Ps = (double *)malloc(sizeof(double) * (lmax + 1));
    double *Pb = (double *)malloc(sizeof(double) * (lmax + 2));
    for (int k = 0; k < nf; ++k)
        for (int m = 0; m <= mmax; m += c->mres) {
            double wm = ((m == 0 || !real_output) ? 1.0 : 2.0) * phiadj;
            for (int i = 0; i < nlat; ++i) {
                ora_plm_dtheta_over_sinth_row(P, dP, Ps, Pb, c->x[i], lmax, m);
                cplx Fti = Ft[nspat * k + i + (size_t)nlat * m], Fpi = Fp[nspat * k + i + (size_t)nlat * m];
                for (int l = (m > 1 ? m : 1); l <= lmax; ++l) {
                    double dY = dP[l];
                    cplx term = I * (double)m * Ps[l];
                    cplx vs = wm * (dY * Fti + conj(term) * Fpi);
                    cplx vt = wm * (-conj(term) * Fti + dY * Fpi);
                    size_t q = 2 * (nalm * k + l + (size_t)(lmax + 1) * m);
                    Sb[q] += creal(vs);
                    Sb[q + 1] += cimag(vs);
                    Tb[q] += creal(vt);
                    Tb[q + 1] += cimag(vt);
                }
            }
            if (!canonical(c))
                for (int l = m; l <= lmax; ++l) {
                    size_t q = 2 * (nalm * k + l + (size_t)(lmax + 1) * m);
                    double s = ora_coef_scale(c, l, m);
                    Sb[q] *= s;
                    Sb[q + 1] *= s;
                    Tb[q] *= s;
                    Tb[q + 1] *= s;
                }
        }
    free(P);
    free(dP);
    free(Ps);
    free(Pb);
    free(Ft);
    free(Fp);
    return 0;
}

int ora_adjoint_analysis_sphtor(const ora_cfg *c, const double *Sb, const double *Tb, double *Vt_cplx,
                                double *Vp_cplx, int nf) {
    int lmax = c->lmax, mmax = c->mmax, nlat = c->nlat, nlon = c->nlon;
    size_t nalm = (size_t)(lmax + 1) * (mmax + 1), nspat = (size_t)nlat * nlon;
    cplx *Ft = (cplx *)calloc(nspat * nf, sizeof(cplx));
    cplx *Fp = (cplx *)calloc(nspat * nf, sizeof(cplx));
    double phiadj = nlon * analysis_phi_scale(c);
    double *P = (double *)malloc(sizeof(double) * (lmax + 1));
    double *dP = (double *)malloc(sizeof(double) * (lmax + 1));
    double *Ps = (double *)malloc(sizeof(double) * (lmax + 1));
    double *Pb = (double *)malloc(sizeof(double) * (lmax + 2));
    for (int k = 0; k < nf; ++k) {
        for (int m = 0; m <= mmax; m += c->mres)
            for (int i = 0; i < nlat; ++i) {
                ora_plm_dtheta_over_sinth_row(P, dP, Ps, Pb, c->x[i], lmax, m);
                cplx st = 0, sp = 0;
                for (int l = (m > 1 ? m : 1); l <= lmax; ++l) {
                    double s = canonical(c) ? 1.0 : ora_coef_scale(c, l, m);
                    cplx Sv = A(Sb + 2 * nalm * k, lmax, l, m) / s;
                    cplx Tv = A(Tb + 2 * nalm * k, lmax, l, m) / s;
                    double dY = dP[l];
                    double ll1 = (double)l * (l + 1);
                    cplx term = I * (double)m * Ps[l];
                    st += (dY * Sv + conj(term) * Tv) / ll1;
                    sp += (-conj(term) * Sv + dY * Tv) / ll1;
                }
                Ft[nspat * k + i + (size_t)nlat * m] = phiadj * c->w[i] * st;
                Fp[nspat * k + i + (size_t)nlat * m] = phiadj * c->w[i] * sp;
            }
        fft_phi(Ft + nspat * k, nlat, nlon, 1);
        fft_phi(Fp + nspat * k, nlat, nlon, 1);
    }
    for (int k = 0; k < nf; ++k)
        for (int j = 0; j < nlon; ++j)
            for (int i = 0; i < nlat; ++i) {
                size_t t = nspat * k + i + (size_t)nlat * j;
                double sc = 1.0;
                if (c->robert_form) {
                    double st = sin_from_cos(c->x[i]);
                    sc = st > 0 ? 1.0 / st : 1.0;
                }
                Vt_cplx[2 * t] = sc * creal(Ft[t]);
                Vt_cplx[2 * t + 1] = sc * cimag(Ft[t]);
                Vp_cplx[2 * t] = sc * creal(Fp[t]);
                Vp_cplx[2 * t + 1] = sc * cimag(Fp[t]);
            }
    free(P);
    free(dP);
    free(Ps);
    free(Pb);
    free(Ft);
    free(Fp);
    return 0;
}

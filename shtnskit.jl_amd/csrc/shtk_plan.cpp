// Host side of libshtk_mi355x.so: plan construction and the C ABI of include/shtk.h.
//
// A plan is the device-resident counterpart of SHTnsKit's SHTConfig
// (src/config.jl:170-206, create_gauss_config :633-655) for one Gauss grid:
// nodes and weights, the FP64 recurrence coefficients of every order, the
// sectoral seeds with their scale counts, the coefficient convention map
// (src/normalization.jl:142-153) and the FFT twiddles.  Transforms enqueue the
// kernels of shtk_legendre.hip and shtk_fft.hip on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "shtk_internal.h"

namespace {

thread_local std::string g_err;
thread_local int g_errcode = SHTK_OK;

int fail(int code, const std::string &msg) {
    g_err = msg;
    g_errcode = code;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? SHTK_ENOMEM : SHTK_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                          \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

template <typename T>
int upload(T **dst, const std::vector<T> &src, const char *what) {
    if (src.empty()) {
        *dst = nullptr;
        return SHTK_OK;
    }
    HIP_TRY(hipMalloc((void **)dst, sizeof(T) * src.size()), what);
    HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice), what);
    return SHTK_OK;
}

// Gauss-Legendre nodes (ascending) and weights: Newton iteration on P_n from the
// Abramowitz-Stegun initial guess, tolerance 1e-15, at most 50 steps, weights
// 2/((1-z^2) P_n'(z)^2), symmetric fill -- the scheme of src/legendre.jl:956-1003.
void gauss_nodes(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const int half = (n + 1) / 2;
    for (int k = 1; k <= half; ++k) {
        double z = std::cos(M_PI * (k - 0.25) / (n + 0.5));
        auto legendre = [n](double zz, double &pn, double &pnm1) {
            pnm1 = 1.0;
            pn = zz;
            for (int l = 2; l <= n; ++l) {
                const double pnp1 = ((2.0 * l - 1.0) * zz * pn - (l - 1.0) * pnm1) / l;
                pnm1 = pn;
                pn = pnp1;
            }
        };
        double pn, pnm1, pd;
        for (int it = 0; it < 50; ++it) {
            legendre(z, pn, pnm1);
            pd = n * (z * pn - pnm1) / (z * z - 1.0);
            const double z1 = z;
            z -= pn / pd;
            if (std::fabs(z - z1) < 1e-15) break;
        }
        legendre(z, pn, pnm1);
        pd = n * (z * pn - pnm1) / (z * z - 1.0);
        x[k - 1] = -z;
        x[n - k] = z;
        const double wk = 2.0 / ((1.0 - z * z) * pd * pd);
        w[k - 1] = wk;
        w[n - k] = wk;
    }
}

double coef_scale(const shtk_desc &d, int l, int m) {
    double nrm = 1.0;
    if (d.norm == SHTK_NORM_FOURPI) nrm = std::sqrt(4.0 * M_PI);
    else if (d.norm == SHTK_NORM_SCHMIDT) nrm = std::sqrt(4.0 * M_PI / (2.0 * l + 1.0));
    const double rs = (d.real_norm && m > 0) ? 1.0 / std::sqrt(2.0) : 1.0;
    const double ph = d.cs_phase ? 1.0 : ((m & 1) ? -1.0 : 1.0);
    return nrm * rs * ph;
}

std::vector<int> factor_radix(int n) {
    std::vector<int> r;
    const int cand[] = {8, 4, 2, 3, 5};
    for (int c : cand)
        while (n % c == 0) {
            r.push_back(c);
            n /= c;
        }
    if (n != 1) r.clear();
    return r;
}

void free_plan(shtk_plan *p) {
    if (!p) return;
    void *ptrs[] = {p->x, p->s, p->w, p->wscaled, p->wadj, p->ab, p->aboff, p->dd, p->seedv, p->seedk,
                    p->convscale, p->ll1inv, p->tw, p->ws};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
}

int validate(const shtk_desc *d) {
    if (!d) return fail(SHTK_EINVAL, "desc is NULL");
    if (d->lmax < 0) return fail(SHTK_EINVAL, "lmax must be >= 0");
    if (d->mmax < 0 || d->mmax > d->lmax) return fail(SHTK_EINVAL, "mmax must satisfy 0 <= mmax <= lmax");
    if (d->mres < 1) return fail(SHTK_EINVAL, "mres must be >= 1");
    if (d->nlat < d->lmax + 1)
        return fail(SHTK_EINVAL, "Gauss-Legendre quadrature requires nlat >= lmax+1");
    if (d->nlon < 2 * d->mmax + 1) return fail(SHTK_EINVAL, "nlon must be >= 2*mmax+1");
    if (d->norm < 0 || d->norm > 2) return fail(SHTK_EINVAL, "unknown normalization");
    if (d->phi_scale != SHTK_PHI_DFT && d->phi_scale != SHTK_PHI_QUAD)
        return fail(SHTK_EINVAL, "unknown phi scaling mode");
    if (d->nlat > 8192) return fail(SHTK_EUNSUPPORTED, "nlat > 8192 is not supported by this backend");
    return SHTK_OK;
}

struct DeviceGuard {
    int prev = -1;
    bool switched = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) switched = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (switched) (void)hipSetDevice(prev);
    }
};

// Workspace layout: [F0][F1][spec0][spec1], each region 256-byte aligned.
struct WsLayout {
    size_t f_bytes, spec_bytes, total;
};

WsLayout ws_layout(const shtk_plan *p, int nf, int vector) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    WsLayout L;
    L.f_bytes = al(sizeof(double2) * (size_t)nf * p->nm * p->d.nlat);
    L.spec_bytes = p->canonical ? 0 : al(sizeof(double2) * (size_t)nf * (p->d.lmax + 1) * (p->d.mmax + 1));
    L.total = (vector ? 2 : 1) * (L.f_bytes + L.spec_bytes);
    return L;
}

int ensure_ws(shtk_plan *p, size_t bytes) {
    if (p->ws_bytes >= bytes) return SHTK_OK;
    if (p->ws) {
        HIP_TRY(hipDeviceSynchronize(), "workspace resize sync");
        HIP_TRY(hipFree(p->ws), "workspace free");
        p->ws = nullptr;
        p->ws_bytes = 0;
    }
    HIP_TRY(hipMalloc(&p->ws, bytes), "workspace allocation");
    p->ws_bytes = bytes;
    return SHTK_OK;
}

shtk::LegArgs base_args(const shtk_plan *p, int ltr) {
    shtk::LegArgs a{};
    a.lmax = p->d.lmax;
    a.mmax = p->d.mmax;
    a.mres = p->d.mres;
    a.nlat = p->d.nlat;
    a.np = p->np;
    a.nm = p->nm;
    a.ltr = ltr;
    a.x = p->x;
    a.s = p->s;
    a.ab = p->ab;
    a.aboff = p->aboff;
    a.seedv = p->seedv;
    a.seedk = p->seedk;
    a.dd = p->dd;
    a.ll1inv = p->ll1inv;
    a.spec_fstride = (long long)(p->d.lmax + 1) * (p->d.mmax + 1);
    return a;
}

shtk::FftArgs fft_args(const shtk_plan *p) {
    shtk::FftArgs f{};
    f.nlat = p->d.nlat;
    f.nlon = p->d.nlon;
    f.np = p->np;
    f.nm = p->nm;
    f.mres = p->d.mres;
    f.mmax = p->d.mmax;
    f.nrad = (int)p->radix.size();
    for (size_t q = 0; q < p->radix.size() && q < 24; ++q) f.radix[q] = p->radix[q];
    f.tw = p->tw;
    f.s = p->s;
    return f;
}

int check_common(shtk_plan *p, int nfields, int dtype, std::initializer_list<const void *> ptrs) {
    if (!p) return fail(SHTK_EINVAL, "plan is NULL");
    if (nfields < 1) return fail(SHTK_EINVAL, "nfields must be >= 1");
    if (dtype != SHTK_F64 && dtype != SHTK_F32) return fail(SHTK_EINVAL, "dtype must be SHTK_F64 or SHTK_F32");
    for (const void *q : ptrs)
        if (!q) return fail(SHTK_EINVAL, "array pointer is NULL");
    if (p->radix.empty())
        return fail(SHTK_EUNSUPPORTED, "nlon=" + std::to_string(p->d.nlon) +
                                           " has a prime factor above 5; the LDS FFT supports 2^a 3^b 5^c");
    return SHTK_OK;
}

int norm_ltr(const shtk_plan *p, int ltr, int *out) {
    if (ltr < 0) ltr = p->d.lmax;
    if (ltr > p->d.lmax) return fail(SHTK_EINVAL, "ltr must be <= lmax");
    *out = ltr;
    return SHTK_OK;
}

}  // namespace

extern "C" {

const char *shtk_version(void) { return "0.1.0"; }

int shtk_last_error(char *buf, size_t len) {
    if (buf && len) {
        std::snprintf(buf, len, "%s", g_err.c_str());
    }
    return g_errcode;
}

int shtk_stream_sync(void *stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "stream synchronize");
    return SHTK_OK;
}

int shtk_plan_create(const shtk_desc *desc, shtk_plan **out) {
    if (!out) return fail(SHTK_EINVAL, "plan output pointer is NULL");
    *out = nullptr;
    int rc = validate(desc);
    if (rc) return rc;
    shtk_plan *p = new (std::nothrow) shtk_plan();
    if (!p) return fail(SHTK_ENOMEM, "host allocation failed");
    p->d = *desc;
    const shtk_desc &d = p->d;
    HIP_TRY(hipGetDevice(&p->device), "hipGetDevice");
    p->np = (d.nlat + 1) / 2;
    p->nm = d.mmax / d.mres + 1;
    p->canonical = d.norm == SHTK_NORM_ORTHONORMAL && !d.real_norm && d.cs_phase;
    p->inv_scale = d.phi_scale == SHTK_PHI_QUAD ? d.nlon / (2.0 * M_PI) : (double)d.nlon;
    p->scale_phi = (2.0 * M_PI / d.nlon) * d.nlon / p->inv_scale;

    std::vector<double> xa, wa;
    gauss_nodes(d.nlat, xa, wa);
    p->hx.assign(xa.rbegin(), xa.rend());  // north first (src/config.jl:639)
    p->hw.assign(wa.rbegin(), wa.rend());

    const int np = p->np, nm = p->nm, lmax = d.lmax;
    std::vector<double> x(np), s(np), w(np), ws(np), wadj(np);
    for (int q = 0; q < np; ++q) {
        x[q] = p->hx[q];
        const double v = (1.0 - x[q]) * (1.0 + x[q]);
        s[q] = std::sqrt(v > 0 ? v : 0.0);
        w[q] = p->hw[q];
        ws[q] = w[q] * p->scale_phi;
        wadj[q] = w[q] * p->scale_phi;
    }

    // Recurrence coefficients per order, l = m .. lmax+1 (src/legendre.jl:242, 277-278).
    std::vector<double2> ab;
    std::vector<double2> dd;
    std::vector<long long> aboff(nm);
    for (int im = 0; im < nm; ++im) {
        const int m = im * d.mres;
        aboff[im] = (long long)ab.size();
        for (int l = m; l <= lmax + 1; ++l) {
            double a = 0.0, b = 0.0;
            if (l == m + 1) {
                a = std::sqrt(2.0 * m + 3.0);
            } else if (l >= m + 2) {
                a = std::sqrt(((2.0 * l - 1.0) * (2.0 * l + 1.0)) / ((double)(l - m) * (double)(l + m)));
                b = std::sqrt(((2.0 * l + 1.0) * (double)(l - 1 - m) * (double)(l - 1 + m)) /
                              ((2.0 * l - 3.0) * (double)(l - m) * (double)(l + m)));
            }
            ab.push_back(make_double2(a, b));
            // sin(theta) dP_l/dtheta = l d_{l+1,m} P_{l+1} - (l+1) d_{l,m} P_{l-1}  (src/legendre.jl:542-548)
            const double v1 = ((double)(l + 1) * (l + 1) - (double)m * m) / (4.0 * (l + 1) * (l + 1) - 1.0);
            const double dl1 = std::sqrt(v1 > 0 ? v1 : 0.0);
            double dl = 0.0;
            if (!(l == 0 || l == m)) {
                const double v = ((double)l * l - (double)m * m) / (4.0 * l * l - 1.0);
                dl = std::sqrt(v > 0 ? v : 0.0);
            }
            dd.push_back(make_double2((double)l * dl1, (double)(l + 1) * dl));
        }
    }

    // Sectoral seeds P_m^m(x_p) = (-1)^m sqrt(1/4pi) prod_k sqrt((2k+1)/(2k)) sin(theta)
    // (src/legendre.jl:252-271), carried as mantissa * 2^e and stored in the
    // scaled representation of the kernels.
    std::vector<double> seedv((size_t)nm * np);
    std::vector<int> seedk((size_t)nm * np);
    for (int q = 0; q < np; ++q) {
        double mant = 0.28209479177387814;
        int e = 0;
        for (int m = 0; m <= d.mmax; ++m) {
            if (m > 0) {
                mant = (-std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s[q]) * mant;
                int de = 0;
                mant = std::frexp(mant, &de);
                e += de;
            }
            if (m % d.mres) continue;
            const size_t idx = (size_t)(m / d.mres) * np + q;
            if (mant == 0.0) {
                seedv[idx] = 0.0;
                seedk[idx] = 0;
            } else if (e > shtk::SCALE_LIVE_MIN) {
                seedv[idx] = std::ldexp(mant, e);
                seedk[idx] = 0;
            } else {
                const int k = (shtk::SCALE_LIVE_MIN - e) / shtk::SCALE_SHIFT + 1;
                seedv[idx] = std::ldexp(mant, e + k * shtk::SCALE_SHIFT);
                seedk[idx] = k;
            }
        }
    }

    std::vector<double> conv;
    if (!p->canonical) {
        conv.assign((size_t)(lmax + 1) * (d.mmax + 1), 1.0);
        for (int m = 0; m <= d.mmax; ++m)
            for (int l = m; l <= lmax; ++l) conv[(size_t)l + (size_t)(lmax + 1) * m] = coef_scale(d, l, m);
    }
    std::vector<double> ll1(lmax + 1, 0.0);
    for (int l = 1; l <= lmax; ++l) ll1[l] = 1.0 / ((double)l * (l + 1));

    std::vector<double2> tw(d.nlon);
    for (int t = 0; t < d.nlon; ++t) {
        const long double ang = 2.0L * 3.141592653589793238462643383279502884L * t / d.nlon;
        tw[t] = make_double2((double)cosl(ang), -(double)sinl(ang));
    }
    p->radix = factor_radix(d.nlon);

    if ((rc = upload(&p->x, x, "plan tables")) || (rc = upload(&p->s, s, "plan tables")) ||
        (rc = upload(&p->w, w, "plan tables")) || (rc = upload(&p->wscaled, ws, "plan tables")) ||
        (rc = upload(&p->wadj, wadj, "plan tables")) || (rc = upload(&p->ab, ab, "plan tables")) ||
        (rc = upload(&p->aboff, aboff, "plan tables")) || (rc = upload(&p->dd, dd, "plan tables")) ||
        (rc = upload(&p->seedv, seedv, "plan tables")) || (rc = upload(&p->seedk, seedk, "plan tables")) ||
        (rc = upload(&p->convscale, conv, "plan tables")) || (rc = upload(&p->ll1inv, ll1, "plan tables")) ||
        (rc = upload(&p->tw, tw, "plan tables"))) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return SHTK_OK;
}

int shtk_plan_destroy(shtk_plan *plan) {
    if (!plan) return SHTK_OK;
    DeviceGuard g(plan->device);
    if (plan->ws) (void)hipDeviceSynchronize();
    free_plan(plan);
    return SHTK_OK;
}

int shtk_plan_nodes(const shtk_plan *plan, double *x, double *w) {
    if (!plan) return fail(SHTK_EINVAL, "plan is NULL");
    for (int i = 0; i < plan->d.nlat; ++i) {
        if (x) x[i] = plan->hx[i];
        if (w) w[i] = plan->hw[i];
    }
    return SHTK_OK;
}

size_t shtk_workspace_size(const shtk_plan *plan, int nfields, int vector) {
    if (!plan || nfields < 1) return 0;
    return ws_layout(plan, nfields, vector).total;
}

int shtk_synthesis(shtk_plan *p, const void *alm, void *f, int nfields, int dtype, int real_output, int ltr,
                   void *stream) {
    int rc = check_common(p, nfields, dtype, {alm, f});
    if (rc) return rc;
    if (real_output != 1)
        return fail(SHTK_EUNSUPPORTED, "complex-output synthesis is not implemented by this backend");
    if ((rc = norm_ltr(p, ltr, &ltr))) return rc;
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard g(p->device);
    hipStream_t st = (hipStream_t)stream;
    const WsLayout L = ws_layout(p, nfields, 0);
    if ((rc = ensure_ws(p, L.total))) return rc;
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws;
    shtk::LegArgs a = base_args(p, ltr);
    a.in0 = alm;
    a.io_f32 = dtype == SHTK_F32;
    if (!p->canonical) {
        void *spec = ws + L.f_bytes;
        HIP_TRY(shtk::launch_spec_scale(alm, spec, a.io_f32, 0, p->convscale, 0, p->d.lmax, p->d.mmax, p->d.mres,
                                        a.spec_fstride, nfields, st),
                "coefficient conversion");
        a.in0 = spec;
        a.io_f32 = 0;
    }
    a.F0 = F0;
    a.g = p->inv_scale / p->d.nlon;
    a.mfac_mode = 0;
    HIP_TRY(shtk::launch_leg_synth(a, nfields, 0, st), "Legendre synthesis launch");
    shtk::FftArgs fa = fft_args(p);
    fa.F = F0;
    fa.grid = f;
    fa.io_f32 = dtype == SHTK_F32;
    fa.mode = 0;
    HIP_TRY(shtk::launch_fft(fa, nfields, 1, st), "phi synthesis launch");
    return SHTK_OK;
}

int shtk_analysis(shtk_plan *p, const void *f, void *alm, int nfields, int dtype, int ltr, void *stream) {
    int rc = check_common(p, nfields, dtype, {alm, f});
    if (rc) return rc;
    if ((rc = norm_ltr(p, ltr, &ltr))) return rc;
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard g(p->device);
    hipStream_t st = (hipStream_t)stream;
    const WsLayout L = ws_layout(p, nfields, 0);
    if ((rc = ensure_ws(p, L.total))) return rc;
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws;
    shtk::FftArgs fa = fft_args(p);
    fa.F = F0;
    fa.grid = (void *)f;
    fa.io_f32 = dtype == SHTK_F32;
    fa.mode = 0;
    HIP_TRY(shtk::launch_fft(fa, nfields, 0, st), "phi analysis launch");
    shtk::LegArgs a = base_args(p, ltr);
    a.F0 = F0;
    a.wlat = p->wscaled;
    a.g = 1.0;
    a.mfac_mode = 0;
    a.out0 = alm;
    a.io_f32 = dtype == SHTK_F32;
    void *spec = ws + L.f_bytes;
    if (!p->canonical) {
        a.out0 = spec;
        a.io_f32 = 0;
    }
    HIP_TRY(shtk::launch_leg_anal(a, nfields, 0, st), "Legendre analysis launch");
    if (!p->canonical)
        HIP_TRY(shtk::launch_spec_scale(spec, alm, 0, dtype == SHTK_F32, p->convscale, 1, p->d.lmax, p->d.mmax,
                                        p->d.mres, a.spec_fstride, nfields, st),
                "coefficient conversion");
    return SHTK_OK;
}

int shtk_adjoint_synthesis(shtk_plan *p, const void *fbar, void *almbar, int nfields, int dtype, int real_output,
                           void *stream) {
    int rc = check_common(p, nfields, dtype, {fbar, almbar});
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard g(p->device);
    hipStream_t st = (hipStream_t)stream;
    const WsLayout L = ws_layout(p, nfields, 0);
    if ((rc = ensure_ws(p, L.total))) return rc;
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws;
    shtk::FftArgs fa = fft_args(p);
    fa.F = F0;
    fa.grid = (void *)fbar;
    fa.io_f32 = dtype == SHTK_F32;
    HIP_TRY(shtk::launch_fft(fa, nfields, 0, st), "phi adjoint launch");
    shtk::LegArgs a = base_args(p, p->d.lmax);
    a.F0 = F0;
    a.wlat = nullptr;
    a.g = p->inv_scale / p->d.nlon;
    a.mfac_mode = real_output ? 1 : 0;
    a.out0 = almbar;
    a.io_f32 = dtype == SHTK_F32;
    void *spec = ws + L.f_bytes;
    if (!p->canonical) {
        a.out0 = spec;
        a.io_f32 = 0;
    }
    HIP_TRY(shtk::launch_leg_anal(a, nfields, 0, st), "Legendre adjoint launch");
    if (!p->canonical)
        HIP_TRY(shtk::launch_spec_scale(spec, almbar, 0, dtype == SHTK_F32, p->convscale, 0, p->d.lmax,
                                        p->d.mmax, p->d.mres, a.spec_fstride, nfields, st),
                "coefficient conversion");
    return SHTK_OK;
}

int shtk_adjoint_analysis(shtk_plan *p, const void *almbar, void *fbar, int nfields, int dtype, void *stream) {
    int rc = check_common(p, nfields, dtype, {fbar, almbar});
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard g(p->device);
    hipStream_t st = (hipStream_t)stream;
    const WsLayout L = ws_layout(p, nfields, 0);
    if ((rc = ensure_ws(p, L.total))) return rc;
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws;
    shtk::LegArgs a = base_args(p, p->d.lmax);
    a.in0 = almbar;
    a.io_f32 = dtype == SHTK_F32;
    if (!p->canonical) {
        void *spec = ws + L.f_bytes;
        HIP_TRY(shtk::launch_spec_scale(almbar, spec, a.io_f32, 0, p->convscale, 1, p->d.lmax, p->d.mmax,
                                        p->d.mres, a.spec_fstride, nfields, st),
                "coefficient conversion");
        a.in0 = spec;
        a.io_f32 = 0;
    }
    a.F0 = F0;
    a.wlat = p->wadj;
    a.g = 1.0;
    a.mfac_mode = 2;
    HIP_TRY(shtk::launch_leg_synth(a, nfields, 0, st), "Legendre adjoint launch");
    shtk::FftArgs fa = fft_args(p);
    fa.F = F0;
    fa.grid = fbar;
    fa.io_f32 = dtype == SHTK_F32;
    HIP_TRY(shtk::launch_fft(fa, nfields, 1, st), "phi adjoint launch");
    return SHTK_OK;
}

int shtk_synthesis_sphtor(shtk_plan *p, const void *S, const void *T, void *Vt, void *Vp, int nfields, int dtype,
                          int real_output, int ltr, void *stream) {
    int rc = check_common(p, nfields, dtype, {S, T, Vt, Vp});
    if (rc) return rc;
    (void)real_output;
    (void)ltr;
    (void)stream;
    return fail(SHTK_EUNSUPPORTED, "vector synthesis is not implemented yet");
}

int shtk_analysis_sphtor(shtk_plan *p, const void *Vt, const void *Vp, void *S, void *T, int nfields, int dtype,
                         int ltr, void *stream) {
    int rc = check_common(p, nfields, dtype, {S, T, Vt, Vp});
    if (rc) return rc;
    (void)ltr;
    (void)stream;
    return fail(SHTK_EUNSUPPORTED, "vector analysis is not implemented yet");
}

int shtk_adjoint_synthesis_sphtor(shtk_plan *p, const void *Vtb, const void *Vpb, void *Sb, void *Tb, int nfields,
                                  int dtype, int real_output, void *stream) {
    int rc = check_common(p, nfields, dtype, {Vtb, Vpb, Sb, Tb});
    if (rc) return rc;
    (void)real_output;
    (void)stream;
    return fail(SHTK_EUNSUPPORTED, "vector adjoint is not implemented yet");
}

int shtk_adjoint_analysis_sphtor(shtk_plan *p, const void *Sb, const void *Tb, void *Vtb, void *Vpb, int nfields,
                                 int dtype, void *stream) {
    int rc = check_common(p, nfields, dtype, {Vtb, Vpb, Sb, Tb});
    if (rc) return rc;
    (void)stream;
    return fail(SHTK_EUNSUPPORTED, "vector adjoint is not implemented yet");
}

}  // extern "C"

// Legendre stage of the spherical harmonic transforms, FP64, for gfx950.
//
// Reference behaviour (SHTnsKit.jl):
//   rows       src/legendre.jl:233-322 (orthonormal + Condon-Shortley recurrence,
//              scaled exponents for seeds that underflow), 470-555 (d/dtheta and
//              1/sin(theta) rows), 326-338 (zonal d/dtheta)
//   synthesis  src/kernels.jl:57-65 and 155-169, orchestrated per m by
//              src/core_transforms.jl:749-869 and src/sphtor_transforms.jl:455-525
//   analysis   src/kernels.jl:33-40 and 205-215, orchestrated per m by
//              src/core_transforms.jl:655-746 and src/sphtor_transforms.jl:528-612
//
// Design (MI355X-first, not a translation of the per-(m, latitude) loops):
//   * Gauss nodes come in +/- x pairs; P_l^m(-x) = (-1)^(l-m) P_l^m(x), so each
//     recurrence chain serves a north/south latitude pair and the even/odd
//     degree sums give both hemispheres (half the recurrence work).
//   * The recurrence coefficients (a_l, b_l) are precomputed once per plan in
//     FP64 and read as wave-uniform scalar loads, so a step is mul + mul + fma.
//   * Seeds that underflow FP64 carry a 2^(512k) scale; scaled chains run the
//     recurrence without contributing (their physical value is < 2^-359) until
//     they come back into range.  Once every lane of a wave is live the wave
//     leaves the checked loop for a branch-free loop.
//   * Synthesis: one lane per latitude pair, 256-lane workgroups per
//     (field, m, latitude block); no cross-lane traffic.
//   * Analysis: one workgroup per (field, m); each lane holds K latitude-pair
//     chains, sums them in registers, and the per-degree sums over lanes are
//     formed through an LDS transpose every 8 degrees.
#include "shtk_internal.h"

namespace shtk {

namespace {

constexpr double kHi = 0x1p100;
constexpr double kDown = 0x1p-512;

__device__ __forceinline__ double mfac_of(int mode, int m) {
    if (m == 0 || mode == 0) return 1.0;
    return mode == 1 ? 2.0 : 0.5;
}

template <typename T>
__device__ __forceinline__ double2 ld_c(const T *p, size_t idx) {
    return make_double2((double)p[2 * idx], (double)p[2 * idx + 1]);
}

template <typename T>
__device__ __forceinline__ void st_c(T *p, size_t idx, double re, double im) {
    p[2 * idx] = (T)re;
    p[2 * idx + 1] = (T)im;
}

// ---------------------------------------------------------------------------
// Scalar synthesis: F[f][im][i] = g * mfac(m) * wlat * sum_l P_l^m(x_i) a_lm
// ---------------------------------------------------------------------------
template <typename IT>
__global__ __launch_bounds__(256) void k_leg_synth(const LegArgs a) {
    const int im = blockIdx.y;
    const int f = blockIdx.z;
    const int m = im * a.mres;
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool act = p < a.np;
    const int pc = act ? p : 0;
    const double x = a.x[pc];
    const size_t sidx = (size_t)im * a.np + pc;
    double cur = act ? a.seedv[sidx] : 0.0;
    int k = act ? a.seedk[sidx] : 0;
    double prev = 0.0;
    const double2 *__restrict__ ab = a.ab + a.aboff[im] - m;
    const IT *__restrict__ al = (const IT *)a.in0 + 2 * ((size_t)f * a.spec_fstride + (size_t)(a.lmax + 1) * m);
    double er = 0.0, ei = 0.0, orr = 0.0, oi = 0.0;
    const int ltr = a.ltr;
    int l = m;

    if (!__all(k == 0)) {
        double c = (k == 0) ? 1.0 : 0.0;
        while (l <= ltr) {
            const int lend = min(l + SCALE_CHECK, ltr + 1);
            for (; l < lend; ++l) {
                const double v = cur * c;
                const double ar = (double)al[2 * l], ai = (double)al[2 * l + 1];
                if ((l - m) & 1) {
                    orr = fma(v, ar, orr);
                    oi = fma(v, ai, oi);
                } else {
                    er = fma(v, ar, er);
                    ei = fma(v, ai, ei);
                }
                const double2 co = ab[l + 1];
                const double nx = fma(co.x * x, cur, -(co.y * prev));
                prev = cur;
                cur = nx;
            }
            if (k > 0 && fmax(fabs(cur), fabs(prev)) > kHi) {
                cur *= kDown;
                prev *= kDown;
                --k;
                c = (k == 0) ? 1.0 : 0.0;
            }
            if (__all(k == 0)) break;
        }
    }
    if (l <= ltr && ((l - m) & 1)) {
        orr = fma(cur, (double)al[2 * l], orr);
        oi = fma(cur, (double)al[2 * l + 1], oi);
        const double2 co = ab[l + 1];
        const double nx = fma(co.x * x, cur, -(co.y * prev));
        prev = cur;
        cur = nx;
        ++l;
    }
    for (; l + 1 <= ltr; l += 2) {
        const double ar0 = (double)al[2 * l], ai0 = (double)al[2 * l + 1];
        const double ar1 = (double)al[2 * l + 2], ai1 = (double)al[2 * l + 3];
        const double2 c0 = ab[l + 1], c1 = ab[l + 2];
        er = fma(cur, ar0, er);
        ei = fma(cur, ai0, ei);
        const double n0 = fma(c0.x * x, cur, -(c0.y * prev));
        orr = fma(n0, ar1, orr);
        oi = fma(n0, ai1, oi);
        const double n1 = fma(c1.x * x, n0, -(c1.y * cur));
        prev = n0;
        cur = n1;
    }
    if (l <= ltr) {
        er = fma(cur, (double)al[2 * l], er);
        ei = fma(cur, (double)al[2 * l + 1], ei);
    }
    if (act) {
        const double sc = a.g * mfac_of(a.mfac_mode, m) * (a.wlat ? a.wlat[p] : 1.0);
        double2 *Fo = a.F0 + ((size_t)f * a.nm + im) * a.nlat;
        Fo[p] = make_double2((er + orr) * sc, (ei + oi) * sc);
        const int ps = a.nlat - 1 - p;
        if (ps != p) Fo[ps] = make_double2((er - orr) * sc, (ei - oi) * sc);
    }
}

// ---------------------------------------------------------------------------
// Scalar analysis: a_lm = g * mfac(m) * sum_i wlat_i P_l^m(x_i) F[f][im][i]
// ---------------------------------------------------------------------------
template <int K, int W, typename OT>
__global__ __launch_bounds__(64 * W) void k_leg_anal(const LegArgs a) {
    constexpr int NT = 64 * W, CH = 8, TPL = NT / CH;
    __shared__ double2 buf[CH * NT];
    const int im = blockIdx.x;
    const int f = blockIdx.y;
    const int m = im * a.mres;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int lmax = a.lmax, ltr = a.ltr, np = a.np, nlat = a.nlat;

    double x[K], cur[K], prev[K], gpr[K], gpi[K], gmr[K], gmi[K], c[K];
    int kk[K];
    const double2 *Fi = a.F0 + ((size_t)f * a.nm + im) * nlat;
    bool live = true;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = wv * 64 * K + lane + 64 * j;
        const bool act = p < np;
        const int pc = act ? p : 0;
        x[j] = a.x[pc];
        const size_t sidx = (size_t)im * np + pc;
        cur[j] = act ? a.seedv[sidx] : 0.0;
        kk[j] = act ? a.seedk[sidx] : 0;
        prev[j] = 0.0;
        c[j] = kk[j] == 0 ? 1.0 : 0.0;
        live = live && kk[j] == 0;
        double2 FN = make_double2(0.0, 0.0), FS = make_double2(0.0, 0.0);
        double wl = 0.0;
        if (act) {
            FN = Fi[p];
            const int ps = nlat - 1 - p;
            if (ps != p) FS = Fi[ps];
            wl = a.wlat ? a.wlat[p] : 1.0;
        }
        gpr[j] = wl * (FN.x + FS.x);
        gpi[j] = wl * (FN.y + FS.y);
        gmr[j] = wl * (FN.x - FS.x);
        gmi[j] = wl * (FN.y - FS.y);
    }
    bool wave_live = __all(live);

    const double2 *__restrict__ ab = a.ab + a.aboff[im] - m;
    OT *out = (OT *)a.out0 + 2 * ((size_t)f * a.spec_fstride + (size_t)(lmax + 1) * m);
    const double gm = a.g * mfac_of(a.mfac_mode, m);

    // zero the entries of this column (and of skipped columns m+1..m+mres-1) that no chunk writes
    for (int q = t; q < (lmax + 1) * a.mres; q += NT) {
        const int col = q / (lmax + 1), l = q % (lmax + 1);
        if (m + col > a.mmax) continue;
        if (col == 0 && l >= m && l <= ltr) continue;
        st_c(out, (size_t)col * (lmax + 1) + l, 0.0, 0.0);
    }

    for (int l0 = m; l0 <= ltr; l0 += CH) {
        if (wave_live) {
#pragma unroll
            for (int s = 0; s < CH; ++s) {
                const int l = l0 + s;
                if (l <= ltr) {
                    double pr = 0.0, pi = 0.0;
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        pr = fma(cur[j], (s & 1) ? gmr[j] : gpr[j], pr);
                        pi = fma(cur[j], (s & 1) ? gmi[j] : gpi[j], pi);
                    }
                    buf[s * NT + t] = make_double2(pr, pi);
                    const double2 co = ab[l + 1];
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const double nx = fma(co.x * x[j], cur[j], -(co.y * prev[j]));
                        prev[j] = cur[j];
                        cur[j] = nx;
                    }
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < CH; ++s) {
                const int l = l0 + s;
                if (l <= ltr) {
                    double pr = 0.0, pi = 0.0;
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const double v = cur[j] * c[j];
                        pr = fma(v, (s & 1) ? gmr[j] : gpr[j], pr);
                        pi = fma(v, (s & 1) ? gmi[j] : gpi[j], pi);
                    }
                    buf[s * NT + t] = make_double2(pr, pi);
                    const double2 co = ab[l + 1];
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const double nx = fma(co.x * x[j], cur[j], -(co.y * prev[j]));
                        prev[j] = cur[j];
                        cur[j] = nx;
                    }
                }
            }
            bool all_live = true;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (kk[j] > 0 && fmax(fabs(cur[j]), fabs(prev[j])) > kHi) {
                    cur[j] *= kDown;
                    prev[j] *= kDown;
                    kk[j] -= 1;
                    c[j] = kk[j] == 0 ? 1.0 : 0.0;
                }
                all_live = all_live && kk[j] == 0;
            }
            wave_live = __all(all_live);
        }
        __syncthreads();
        {
            const int li = t / TPL, q = t % TPL;
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int jj = 0; jj < CH; ++jj) {
                const double2 v = buf[li * NT + q + TPL * jj];
                sr += v.x;
                si += v.y;
            }
#pragma unroll
            for (int msk = TPL / 2; msk >= 1; msk >>= 1) {
                sr += __shfl_xor(sr, msk, TPL);
                si += __shfl_xor(si, msk, TPL);
            }
            const int l = l0 + li;
            if (q == 0 && l <= ltr) st_c(out, (size_t)l, gm * sr, gm * si);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Spectral conversion: out = in * scale (or in / scale), zero for l < m and
// for orders m that are not multiples of mres; dtype conversion on the way.
// ---------------------------------------------------------------------------
template <typename IT, typename OT>
__global__ void k_spec_scale(const IT *in, OT *out, const double *scale, int divide, int lmax, int mmax,
                             int mres, long long per_field, long long total) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const long long r = q % per_field;
    const int l = (int)(r % (lmax + 1)), m = (int)(r / (lmax + 1));
    double re = 0.0, im = 0.0;
    if (l >= m && m % mres == 0 && m <= mmax) {
        const double s = scale ? scale[r] : 1.0;
        re = (double)in[2 * q];
        im = (double)in[2 * q + 1];
        if (divide) {
            re /= s;
            im /= s;
        } else {
            re *= s;
            im *= s;
        }
    }
    out[2 * q] = (OT)re;
    out[2 * q + 1] = (OT)im;
}

template <int K, int W>
hipError_t launch_anal_kw(const LegArgs &a, int nfields, hipStream_t st) {
    dim3 grid(a.nm, nfields);
    if (a.io_f32)
        hipLaunchKernelGGL((k_leg_anal<K, W, float>), grid, dim3(64 * W), 0, st, a);
    else
        hipLaunchKernelGGL((k_leg_anal<K, W, double>), grid, dim3(64 * W), 0, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_leg_synth(const LegArgs &a, int nfields, int vector, hipStream_t st) {
    if (vector) return hipErrorNotSupported;
    dim3 grid((a.np + 255) / 256, a.nm, nfields);
    if (a.io_f32)
        hipLaunchKernelGGL(k_leg_synth<float>, grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_leg_synth<double>, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_leg_anal(const LegArgs &a, int nfields, int vector, hipStream_t st) {
    if (vector) return hipErrorNotSupported;
    // chains per lane K and waves per workgroup W cover all latitude pairs
    const int np = a.np;
    if (np <= 64) return launch_anal_kw<1, 1>(a, nfields, st);
    if (np <= 128) return launch_anal_kw<2, 1>(a, nfields, st);
    if (np <= 256) return launch_anal_kw<2, 2>(a, nfields, st);
    if (np <= 512) return launch_anal_kw<4, 2>(a, nfields, st);
    if (np <= 1024) return launch_anal_kw<4, 4>(a, nfields, st);
    if (np <= 2048) return launch_anal_kw<8, 4>(a, nfields, st);
    return hipErrorNotSupported;
}

hipError_t launch_spec_scale(const void *in, void *out, int in_f32, int out_f32, const double *scale, int divide,
                             int lmax, int mmax, int mres, long long per_field, int nfields, hipStream_t st) {
    const long long total = per_field * nfields;
    dim3 grid((unsigned)((total + 255) / 256));
    if (in_f32 && out_f32)
        hipLaunchKernelGGL((k_spec_scale<float, float>), grid, dim3(256), 0, st, (const float *)in, (float *)out,
                           scale, divide, lmax, mmax, mres, per_field, total);
    else if (in_f32)
        hipLaunchKernelGGL((k_spec_scale<float, double>), grid, dim3(256), 0, st, (const float *)in,
                           (double *)out, scale, divide, lmax, mmax, mres, per_field, total);
    else if (out_f32)
        hipLaunchKernelGGL((k_spec_scale<double, float>), grid, dim3(256), 0, st, (const double *)in,
                           (float *)out, scale, divide, lmax, mmax, mres, per_field, total);
    else
        hipLaunchKernelGGL((k_spec_scale<double, double>), grid, dim3(256), 0, st, (const double *)in,
                           (double *)out, scale, divide, lmax, mmax, mres, per_field, total);
    return hipGetLastError();
}

}  // namespace shtk

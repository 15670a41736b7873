// Longitude (phi) transforms between the grid and the Fourier intermediate, FP64.
//
// Reference conventions (src/fftutils.jl:283-432, src/core_transforms.jl:294-344):
// forward DFT unnormalised with exp(-2 pi i jk/N); synthesis fills the
// Hermitian mirror F[N-m] = conj(F[m]) and keeps the real part of the inverse
// DFT.  The inverse here applies no 1/N: the Legendre stage folds
// phi_inv_scale/nlon into its output factor.
//
// Design: one workgroup per (field, G consecutive latitude pairs), G = 1, 2
// or 4.  The two real rows of a north/south pair travel as one complex row
// z = f_N + i f_S, so a single complex FFT of length N serves both rows (no
// separate real-FFT pass).  The G rows stay resident in LDS (16 G N bytes)
// through a Stockham autosort FFT with radix 8/4/2/3/5 passes; twiddles come
// from a per-plan table.  The grid is latitude-fastest, so G adjacent pairs
// turn the longitude-major row accesses into runs of G elements.
#include "shtk_internal.h"

namespace shtk {

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by sign*i
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 a) {
    return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_small(double2 *v) {
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const double2 t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if constexpr (R == 8) {
        double2 e[4] = {v[0], v[2], v[4], v[6]};
        double2 o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<4, SIGN>(e);
        dft_small<4, SIGN>(o);
        constexpr double h = 0.70710678118654752440;
        // W8^k = exp(SIGN * 2 pi i k / 8)
        const double2 w1 = make_double2(h, SIGN * h);
        const double2 w3 = make_double2(-h, SIGN * h);
        const double2 o1 = cmul(o[1], w1), o2 = mul_si<SIGN>(o[2]), o3 = cmul(o[3], w3);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else if constexpr (R == 3) {
        constexpr double c1 = -0.5, s1 = 0.86602540378443864676;
        const double2 a = v[0], b = v[1], c = v[2];
        const double2 sum = cadd(b, c), dif = csub(b, c);
        const double2 t = make_double2(fma(c1, sum.x, a.x), fma(c1, sum.y, a.y));
        // SIGN*i*s1*(b-c)
        const double2 u = mul_si<SIGN>(make_double2(s1 * dif.x, s1 * dif.y));
        v[0] = cadd(a, sum);
        v[1] = cadd(t, u);
        v[2] = csub(t, u);
    } else if constexpr (R == 5) {
        constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        constexpr double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
        const double2 a = v[0];
        const double2 b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        const double2 b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        const double2 t1 = make_double2(a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y);
        const double2 t2 = make_double2(a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y);
        const double2 u1 = mul_si<SIGN>(make_double2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
        const double2 u2 = mul_si<SIGN>(make_double2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
        v[0] = make_double2(a.x + b1.x + b2.x, a.y + b1.y + b2.y);
        v[1] = cadd(t1, u1);
        v[4] = csub(t1, u1);
        v[2] = cadd(t2, u2);
        v[3] = csub(t2, u2);
    }
}

// One Stockham pass of radix R over the G LDS rows z[g N + .] (length N
// each).  Every thread stages all its inputs in registers before the barrier,
// then writes.
template <int R, int SIGN, int G>
__device__ __forceinline__ void stockham_pass(double2 *z, int N, int Ns, const double2 *tw, int tid, int nt) {
    constexpr int MAXB = (8 * G) / R + 1;  // butterflies per thread: G N / (R nt) with nt >= N / 8
    double2 v[MAXB][R];
    const int nb = N / R, tot = G * nb;
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int jj = tid + q * nt;
        if (jj < tot) {
            const int g = jj / nb, j = jj - g * nb;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = z[g * N + j + r * nb];
        }
    }
    __syncthreads();
    const int stride = N / (Ns * R);
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int jj = tid + q * nt;
        if (jj < tot) {
            const int g = jj / nb, j = jj - g * nb;
            const int k = j % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 w = tw[r * k * stride];
                if (SIGN > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            dft_small<R, SIGN>(v[q]);
            const int base = g * N + (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) z[base + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

template <int SIGN, int G>
__device__ void fft_lds(double2 *z, const FftArgs &a, int tid, int nt) {
    const int N = a.nlon;
    int Ns = 1;
    for (int q = 0; q < a.nrad; ++q) {
        const int R = (int)((a.radix >> (4 * q)) & 15ull);
        switch (R) {
            case 8: stockham_pass<8, SIGN, G>(z, N, Ns, a.tw, tid, nt); break;
            case 4: stockham_pass<4, SIGN, G>(z, N, Ns, a.tw, tid, nt); break;
            case 2: stockham_pass<2, SIGN, G>(z, N, Ns, a.tw, tid, nt); break;
            case 3: stockham_pass<3, SIGN, G>(z, N, Ns, a.tw, tid, nt); break;
            default: stockham_pass<5, SIGN, G>(z, N, Ns, a.tw, tid, nt); break;
        }
        Ns *= R;
    }
}

__device__ __forceinline__ double row_factor(const FftArgs &a, int p) {
    if (a.mode == 1) {
        const double st = a.s[p];
        return st > 0.0 ? 1.0 / st : 1.0;
    }
    return a.mode == 2 ? a.s[p] : 1.0;
}

// Synthesis: one workgroup per (G consecutive latitude pairs, field).  North
// rows p0..p0+G-1 and their south partners are adjacent in memory, so the
// Fourier columns are read and the grid rows written in G-element runs.
// Blocks run with at most NT threads: 256 (rows up to 2048 points, any G) or
// 1024 (longer rows, G = 1).
template <typename GT, int G, int NT>
__global__ __launch_bounds__(NT) void k_fft_synth(const FftArgs a) {
    extern __shared__ double2 z[];
    const int p0 = blockIdx.x * G;
    const int f = blockIdx.y;
    const int N = a.nlon, nlat = a.nlat, np = a.np;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int q = tid; q < G * N; q += nt) z[q] = make_double2(0.0, 0.0);
    __syncthreads();
    const double2 *Ff = a.F + (size_t)f * a.nm * nlat;
    // z = f_N + i f_S has the spectrum Z_m = F_N + i F_S, Z_{N-m} = conj(F_N) + i conj(F_S):
    // north terms first (unit (im, g), G consecutive latitudes per order) ...
    for (int u = tid; u < a.nm * G; u += nt) {
        const int g = u % G, im = u / G;
        const int p = p0 + g;
        if (p >= np) continue;
        const double2 FN = Ff[(size_t)im * nlat + p];
        const int m = im * a.mres;
        double2 *zr = z + g * N;
        if (m == 0) {
            zr[0].x = FN.x;
        } else {
            zr[m] = FN;
            zr[N - m] = make_double2(FN.x, -FN.y);
        }
    }
    __syncthreads();
    // ... then south terms (the G partner latitudes, also consecutive)
    for (int u = tid; u < a.nm * G; u += nt) {
        const int g = u % G, im = u / G;
        const int p = p0 + (G - 1 - g);
        if (p >= np) continue;
        const int lat = nlat - 1 - p;
        const bool hasS = lat != p;
        const double2 FS = hasS ? Ff[(size_t)im * nlat + lat] : make_double2(0.0, 0.0);
        const int m = im * a.mres;
        double2 *zr = z + (p - p0) * N;
        if (m == 0) {
            zr[0].y = FS.x;
        } else {
            zr[m].x -= FS.y;
            zr[m].y += FS.x;
            zr[N - m].x += FS.y;
            zr[N - m].y += FS.x;
        }
    }
    __syncthreads();
    fft_lds<+1, G>(z, a, tid, nt);
    GT *gr = (GT *)a.grid + (size_t)f * nlat * N;
    // unit (j, h, g): longitude j, north (h = 0) or south row of pair p0 + g
    for (int u = tid; u < N * 2 * G; u += nt) {
        const int g = u % G, h = (u / G) & 1, j = u / (2 * G);
        const int p = h == 0 ? p0 + g : p0 + (G - 1 - g);
        if (p >= np) continue;
        const int lat = h == 0 ? p : nlat - 1 - p;
        if (h == 1 && lat == p) continue;
        const double2 v = z[(p - p0) * N + j];
        gr[(size_t)j * nlat + lat] = (GT)((h == 0 ? v.x : v.y) * row_factor(a, p));
    }
}

// Analysis: the mirror of the synthesis; z = f_N + i f_S per pair, forward
// FFT, then F_N = (Z_m + conj Z_{N-m}) / 2, F_S = (Z_m - conj Z_{N-m}) / (2i).
template <typename GT, int G, int NT>
__global__ __launch_bounds__(NT) void k_fft_anal(const FftArgs a) {
    extern __shared__ double2 z[];
    const int p0 = blockIdx.x * G;
    const int f = blockIdx.y;
    const int N = a.nlon, nlat = a.nlat, np = a.np;
    const int tid = threadIdx.x, nt = blockDim.x;
    const GT *gr = (const GT *)a.grid + (size_t)f * nlat * N;
    for (int u = tid; u < N * 2 * G; u += nt) {
        const int g = u % G, h = (u / G) & 1, j = u / (2 * G);
        const int p = h == 0 ? p0 + g : p0 + (G - 1 - g);
        if (p >= np) continue;
        const int lat = h == 0 ? p : nlat - 1 - p;
        double v = 0.0;
        if (h == 0 || lat != p) v = (double)gr[(size_t)j * nlat + lat] * row_factor(a, p);
        double2 *zr = z + (p - p0) * N + j;
        if (h == 0) zr->x = v;
        else zr->y = v;
    }
    __syncthreads();
    fft_lds<-1, G>(z, a, tid, nt);
    double2 *Ff = a.F + (size_t)f * a.nm * nlat;
    for (int u = tid; u < a.nm * 2 * G; u += nt) {
        const int g = u % G, h = (u / G) & 1, im = u / (2 * G);
        const int p = h == 0 ? p0 + g : p0 + (G - 1 - g);
        if (p >= np) continue;
        const int lat = h == 0 ? p : nlat - 1 - p;
        if (h == 1 && lat == p) continue;
        const int m = im * a.mres;
        const double2 *zr = z + (p - p0) * N;
        const double2 Zm = zr[m];
        const double2 Zc = zr[(N - m) % N];
        Ff[(size_t)im * nlat + lat] = h == 0 ? make_double2(0.5 * (Zm.x + Zc.x), 0.5 * (Zm.y - Zc.y))
                                             : make_double2(0.5 * (Zm.y + Zc.y), -0.5 * (Zm.x - Zc.x));
    }
}

template <int G, int NT>
hipError_t launch_fft_g(const FftArgs &a, int nfields, int inverse, int nt, hipStream_t st) {
    const size_t lds = sizeof(double2) * (size_t)G * a.nlon;
    dim3 grid((a.np + G - 1) / G, nfields);
    const void *fn = inverse ? (a.io_f32 ? (const void *)k_fft_synth<float, G, NT>
                                         : (const void *)k_fft_synth<double, G, NT>)
                             : (a.io_f32 ? (const void *)k_fft_anal<float, G, NT>
                                         : (const void *)k_fft_anal<double, G, NT>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (inverse) {
        if (a.io_f32)
            hipLaunchKernelGGL((k_fft_synth<float, G, NT>), grid, dim3(nt), lds, st, a);
        else
            hipLaunchKernelGGL((k_fft_synth<double, G, NT>), grid, dim3(nt), lds, st, a);
    } else {
        if (a.io_f32)
            hipLaunchKernelGGL((k_fft_anal<float, G, NT>), grid, dim3(nt), lds, st, a);
        else
            hipLaunchKernelGGL((k_fft_anal<double, G, NT>), grid, dim3(nt), lds, st, a);
    }
    return hipGetLastError();
}

}  // namespace

// Pairs per block: 4 or 2 when the rows are short enough (N <= 2048, G N
// complex values within 128 KB of LDS) and the grid keeps at least two blocks
// per CU, else 1.
hipError_t launch_fft(const FftArgs &a, int nfields, int inverse, hipStream_t st) {
    if (a.nrad <= 0) return hipErrorNotSupported;
    const int N = a.nlon;
    int nt = N / 8;
    if (nt < 64) nt = 64;
    if (nt > 1024) nt = 1024;
    nt = (nt + 63) / 64 * 64;
    if (N > 8 * nt) return hipErrorNotSupported;
    if (sizeof(double2) * (size_t)N > 160 * 1024) return hipErrorNotSupported;
    const long long pairs = (long long)a.np * nfields;
    int G = 1;
    if (a.fft_pairs > 0) G = a.fft_pairs;
    else if (pairs / 4 >= 512) G = 4;
    else if (pairs / 2 >= 512) G = 2;
    if (nt > 256 || sizeof(double2) * (size_t)G * N > 128 * 1024) G = 1;
    if (nt > 256) return launch_fft_g<1, 1024>(a, nfields, inverse, nt, st);
    switch (G) {
        case 4: return launch_fft_g<4, 256>(a, nfields, inverse, nt, st);
        case 2: return launch_fft_g<2, 256>(a, nfields, inverse, nt, st);
        default: return launch_fft_g<1, 256>(a, nfields, inverse, nt, st);
    }
}

}  // namespace shtk

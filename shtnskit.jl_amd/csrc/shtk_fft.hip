// Longitude (phi) transforms between the grid and the Fourier intermediate, FP64.
//
// Reference conventions (src/fftutils.jl:283-432, src/core_transforms.jl:294-344):
// forward DFT unnormalised with exp(-2 pi i jk/N); synthesis fills the
// Hermitian mirror F[N-m] = conj(F[m]) and keeps the real part of the inverse
// DFT.  The inverse here applies no 1/N: the Legendre stage folds
// phi_inv_scale/nlon into its output factor.
//
// Design: one workgroup per (field, latitude pair).  The two real rows of a
// north/south pair travel as one complex row z = f_N + i f_S, so a single
// complex FFT of length N serves both rows (no separate real-FFT pass).  The
// row stays resident in LDS (16 N bytes) through a Stockham autosort FFT with
// radix 8/4/2/3/5 passes; twiddles come from a per-plan table.  Consecutive
// latitude pairs are mapped to the same XCD so the L2 merges the strided
// longitude-major stores of neighbouring rows.
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

// One Stockham pass of radix R over the LDS row z (length N).  Every thread
// stages all its inputs in registers before the barrier, then writes.
template <int R, int SIGN>
__device__ __forceinline__ void stockham_pass(double2 *z, int N, int Ns, const double2 *tw, int tid, int nt) {
    constexpr int MAXB = 8 / R + 1;  // at most N/8 butterflies' worth of values per thread
    double2 v[MAXB][R];
    const int nb = N / R;
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int j = tid + q * nt;
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = z[j + r * nb];
        }
    }
    __syncthreads();
    const int stride = N / (Ns * R);
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int j = tid + q * nt;
        if (j < nb) {
            const int k = j % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 w = tw[r * k * stride];
                if (SIGN > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            dft_small<R, SIGN>(v[q]);
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) z[base + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

template <int SIGN>
__device__ void fft_lds(double2 *z, const FftArgs &a, int tid, int nt) {
    const int N = a.nlon;
    int Ns = 1;
    for (int q = 0; q < a.nrad; ++q) {
        const int R = (int)((a.radix >> (4 * q)) & 15ull);
        switch (R) {
            case 8: stockham_pass<8, SIGN>(z, N, Ns, a.tw, tid, nt); break;
            case 4: stockham_pass<4, SIGN>(z, N, Ns, a.tw, tid, nt); break;
            case 2: stockham_pass<2, SIGN>(z, N, Ns, a.tw, tid, nt); break;
            case 3: stockham_pass<3, SIGN>(z, N, Ns, a.tw, tid, nt); break;
            default: stockham_pass<5, SIGN>(z, N, Ns, a.tw, tid, nt); break;
        }
        Ns *= R;
    }
}

// Block index -> latitude pair, keeping runs of consecutive pairs on one XCD
// (workgroups are dispatched round-robin over the 8 XCDs).
__device__ __forceinline__ int pair_of_block(int b, int np) {
    const int per = (np + 7) / 8;
    const int xcd = b & 7, slot = b >> 3;
    const int p = xcd * per + slot;
    return p;
}

template <typename GT>
__global__ __launch_bounds__(1024) void k_fft_synth(const FftArgs a) {
    extern __shared__ double2 z[];
    const int p = pair_of_block(blockIdx.x, a.np);
    if (p >= a.np) return;
    const int f = blockIdx.y;
    const int N = a.nlon, nlat = a.nlat;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int iN = p, iS = nlat - 1 - p;
    const bool hasS = iS != iN;
    for (int q = tid; q < N; q += nt) z[q] = make_double2(0.0, 0.0);
    __syncthreads();
    const double2 *Ff = a.F + (size_t)f * a.nm * nlat;
    for (int im = tid; im < a.nm; im += nt) {
        const int m = im * a.mres;
        const double2 FN = Ff[(size_t)im * nlat + iN];
        const double2 FS = hasS ? Ff[(size_t)im * nlat + iS] : make_double2(0.0, 0.0);
        if (m == 0) {
            z[0] = make_double2(FN.x, FS.x);
        } else {
            z[m] = make_double2(FN.x - FS.y, FN.y + FS.x);
            z[N - m] = make_double2(FN.x + FS.y, FS.x - FN.y);
        }
    }
    __syncthreads();
    fft_lds<+1>(z, a, tid, nt);
    GT *g = (GT *)a.grid + (size_t)f * nlat * N;
    double sc = 1.0;
    if (a.mode == 1) {
        const double st = a.s[p];
        sc = st > 0.0 ? 1.0 / st : 1.0;
    } else if (a.mode == 2) {
        sc = a.s[p];
    }
    for (int j = tid; j < N; j += nt) {
        const double2 v = z[j];
        g[(size_t)j * nlat + iN] = (GT)(v.x * sc);
        if (hasS) g[(size_t)j * nlat + iS] = (GT)(v.y * sc);
    }
}

template <typename GT>
__global__ __launch_bounds__(1024) void k_fft_anal(const FftArgs a) {
    extern __shared__ double2 z[];
    const int p = pair_of_block(blockIdx.x, a.np);
    if (p >= a.np) return;
    const int f = blockIdx.y;
    const int N = a.nlon, nlat = a.nlat;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int iN = p, iS = nlat - 1 - p;
    const bool hasS = iS != iN;
    const GT *g = (const GT *)a.grid + (size_t)f * nlat * N;
    double sc = 1.0;
    if (a.mode == 1) {
        const double st = a.s[p];
        sc = st > 0.0 ? 1.0 / st : 1.0;
    } else if (a.mode == 2) {
        sc = a.s[p];
    }
    for (int j = tid; j < N; j += nt) {
        const double vN = (double)g[(size_t)j * nlat + iN];
        const double vS = hasS ? (double)g[(size_t)j * nlat + iS] : 0.0;
        z[j] = make_double2(vN * sc, vS * sc);
    }
    __syncthreads();
    fft_lds<-1>(z, a, tid, nt);
    double2 *Ff = a.F + (size_t)f * a.nm * nlat;
    for (int im = tid; im < a.nm; im += nt) {
        const int m = im * a.mres;
        const double2 Zm = z[m];
        const double2 Zc = z[(N - m) % N];  // conj taken below
        // FN = (Zm + conj(Zc)) / 2, FS = (Zm - conj(Zc)) / (2i)
        const double2 FN = make_double2(0.5 * (Zm.x + Zc.x), 0.5 * (Zm.y - Zc.y));
        const double2 FS = make_double2(0.5 * (Zm.y + Zc.y), -0.5 * (Zm.x - Zc.x));
        Ff[(size_t)im * nlat + iN] = FN;
        if (hasS) Ff[(size_t)im * nlat + iS] = FS;
    }
}

}  // namespace

hipError_t launch_fft(const FftArgs &a, int nfields, int inverse, hipStream_t st) {
    if (a.nrad <= 0) return hipErrorNotSupported;
    const int N = a.nlon;
    int nt = N / 8;
    if (nt < 64) nt = 64;
    if (nt > 1024) nt = 1024;
    nt = (nt + 63) / 64 * 64;
    if (N > 8 * nt) return hipErrorNotSupported;
    const size_t lds = sizeof(double2) * (size_t)N;
    if (lds > 160 * 1024) return hipErrorNotSupported;
    const int per = (a.np + 7) / 8;
    dim3 grid(per * 8, nfields);
    if (lds > 64 * 1024) {
        const void *fn = inverse ? (a.io_f32 ? (const void *)k_fft_synth<float> : (const void *)k_fft_synth<double>)
                                 : (a.io_f32 ? (const void *)k_fft_anal<float> : (const void *)k_fft_anal<double>);
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (inverse) {
        if (a.io_f32)
            hipLaunchKernelGGL(k_fft_synth<float>, grid, dim3(nt), lds, st, a);
        else
            hipLaunchKernelGGL(k_fft_synth<double>, grid, dim3(nt), lds, st, a);
    } else {
        if (a.io_f32)
            hipLaunchKernelGGL(k_fft_anal<float>, grid, dim3(nt), lds, st, a);
        else
            hipLaunchKernelGGL(k_fft_anal<double>, grid, dim3(nt), lds, st, a);
    }
    return hipGetLastError();
}

}  // namespace shtk

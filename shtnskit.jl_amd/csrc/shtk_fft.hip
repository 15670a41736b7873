// Longitude (phi) transforms between the grid and the Fourier intermediate, FP64.
//
// Reference conventions (src/fftutils.jl:283-432, src/core_transforms.jl:294-344):
// forward DFT unnormalised with exp(-2 pi i jk/N); synthesis fills the
// Hermitian mirror F[N-m] = conj(F[m]) and keeps the real part of the inverse
// DFT.  The inverse here applies no 1/N: the Legendre stage folds
// phi_inv_scale/nlon into its output factor.
//
// Design: persistent workgroups over tiles of 2G adjacent latitudes (the grid
// is latitude-fastest, so a tile reads and writes 2G-latitude runs at every
// longitude, and the order-major Fourier columns give 2G-latitude runs at every
// order).  Two real rows travel as one complex row z = f_a + i f_b (latitudes
// k and 2G-1-k of the tile), so a single complex FFT of length N serves both
// rows (no separate real-FFT pass).  The rows stay resident in LDS (16 N bytes
// each) through a Stockham autosort FFT with radix 8/4/2/3/5 passes and generic
// passes for other small primes; lengths with larger prime factors run
// Bluestein's chirp convolution inside the same LDS row (of the convolution
// length).  Twiddles come from per-plan tables.  Consecutive tiles are mapped to the same XCD so
// the L2 merges the strided longitude-major stores of neighbouring rows.
#include <algorithm>

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
// values per thread in one Stockham pass: the row length over the threads of a row
constexpr int FFT_VPT = 8;

// LDS position of row element i: one padding slot per 16 elements, which
// spreads the strided 16-byte accesses of the Stockham passes (stride R or
// R Ns elements) over all LDS banks.
__device__ __forceinline__ int zp(int i) { return i + (i >> 4); }

// P2: Ns is a power of two, so the butterfly's position j splits into j / Ns and
// j mod Ns by a shift and a mask (a division by a run-time value is a sequence
// of some twenty instructions).
template <int R, int SIGN, bool P2>
__device__ __forceinline__ void stockham_pass(double2 *z, int N, int Ns, const double2 *tw, int tid, int nt) {
    // nt >= N / FFT_VPT (launch_fft), so a thread has at most ceil(FFT_VPT / R) butterflies
    constexpr int MAXB = (FFT_VPT + R - 1) / R;
    const int nb = N / R;
    const int stride = N / (Ns * R);
    const int sh = __builtin_ctz((unsigned)Ns);
    auto kmod = [&](int j) { return P2 ? (j & (Ns - 1)) : j % Ns; };
    auto kdiv = [&](int j) { return P2 ? (j >> sh) : j / Ns; };
    // base twiddle w = W^(k stride) of each butterfly (the others are its
    // powers): independent of the LDS data, so its load overlaps the barrier
    double2 w1[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int j = tid + q * nt;
        w1[q] = (Ns > 1 && j < nb) ? tw[kmod(j) * stride] : make_double2(1.0, 0.0);
    }
    double2 v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int j = tid + q * nt;
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = z[zp(j + r * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int j = tid + q * nt;
        if (j < nb) {
            const int k = kmod(j);
            if (Ns > 1) {
                double2 w = w1[q];
                if (SIGN > 0) w.y = -w.y;
                double2 wr = w;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[q][r] = cmul(v[q][r], wr);
                    if (r + 1 < R) wr = cmul(wr, w);
                }
            }
            dft_small<R, SIGN>(v[q]);
            const int base = kdiv(j) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) z[zp(base + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// Generic radix-R pass (R a prime above 5): every output is the R-term sum
// out(j, r) = sum_s z[j + s nb] W^(s (k + r Ns) stride), k = j mod Ns, with the
// twiddles read from the table (exact angles, no accumulated rounding).  A
// thread computes up to FFT_VPT outputs (nt >= N / FFT_VPT).
template <int SIGN>
__device__ void stockham_gen(double2 *z, int N, int Ns, int R, const double2 *tw, int tid, int nt) {
    const int nb = N / R, stride = N / (Ns * R);
    double2 out[FFT_VPT];
#pragma unroll
    for (int q = 0; q < FFT_VPT; ++q) {
        const int e = tid + q * nt;
        out[q] = make_double2(0.0, 0.0);
        if (e < N) {
            const int j = e % nb, r = e / nb, k = j % Ns;
            const int step = (k + r * Ns) * stride;  // < N
            double2 acc = z[zp(j)];
            int t = 0;
            for (int s = 1; s < R; ++s) {
                t += step;
                if (t >= N) t -= N;
                double2 w = tw[t];
                if (SIGN > 0) w.y = -w.y;
                const double2 v = z[zp(j + s * nb)];
                acc = make_double2(fma(v.x, w.x, fma(-v.y, w.y, acc.x)), fma(v.x, w.y, fma(v.y, w.x, acc.y)));
            }
            out[q] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < FFT_VPT; ++q) {
        const int e = tid + q * nt;
        if (e < N) {
            const int j = e % nb, r = e / nb, k = j % Ns;
            z[zp((j / Ns) * Ns * R + k + r * Ns)] = out[q];
        }
    }
    __syncthreads();
}

// All passes of a length-N transform of the LDS row z (unnormalised).  GEN =
// false compiles the radix 2/3/4/5/8 butterflies alone (the common lengths keep
// their registers free of the generic pass's state).
template <int SIGN, bool GEN>
__device__ void fft_passes(double2 *z, int N, int nrad, unsigned long long radix, const double2 *tw, int tid,
                           int nt) {
    int Ns = 1;
    for (int q = 0; q < nrad; ++q) {
        const int R = (int)((radix >> (FFT_RADIX_BITS * q)) & ((1ull << FFT_RADIX_BITS) - 1));
        const bool p2 = (Ns & (Ns - 1)) == 0;
        switch (R) {
            case 8:
                if (p2) stockham_pass<8, SIGN, true>(z, N, Ns, tw, tid, nt);
                else stockham_pass<8, SIGN, false>(z, N, Ns, tw, tid, nt);
                break;
            case 4:
                if (p2) stockham_pass<4, SIGN, true>(z, N, Ns, tw, tid, nt);
                else stockham_pass<4, SIGN, false>(z, N, Ns, tw, tid, nt);
                break;
            case 2:
                if (p2) stockham_pass<2, SIGN, true>(z, N, Ns, tw, tid, nt);
                else stockham_pass<2, SIGN, false>(z, N, Ns, tw, tid, nt);
                break;
            case 3: stockham_pass<3, SIGN, false>(z, N, Ns, tw, tid, nt); break;
            default:
                if constexpr (GEN) {
                    if (R != 5) {
                        stockham_gen<SIGN>(z, N, Ns, R, tw, tid, nt);
                        break;
                    }
                }
                stockham_pass<5, SIGN, false>(z, N, Ns, tw, tid, nt);
                break;
        }
        Ns *= R;
    }
}

// Length-nlon DFT of the LDS row z (entries at zp(j), j < nlon; the row holds
// nfft + nfft/16 + 1 slots).  Direct passes, or (GEN only) Bluestein: chirp,
// zero-pad to nfft, cyclic convolution with the kernel spectrum (forward FFT,
// pointwise product, inverse FFT), chirp.  The caller's barrier precedes; one
// follows.
template <int SIGN, bool GEN>
__device__ void fft_lds(double2 *z, const FftArgs &a, int tid, int nt) {
    if (!GEN || !a.blue) {
        fft_passes<SIGN, GEN>(z, a.nlon, a.nrad, a.radix, a.tw, tid, nt);
        return;
    }
    const int N = a.nlon, M = a.nfft;
    for (int j = tid; j < M; j += nt) {
        if (j < N) {
            double2 c = a.chirp[j];
            if (SIGN > 0) c.y = -c.y;
            z[zp(j)] = cmul(z[zp(j)], c);
        } else {
            z[zp(j)] = make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    fft_passes<-1, GEN>(z, M, a.nrad, a.radix, a.tw, tid, nt);
    const double2 *bh = a.bhat + (SIGN > 0 ? M : 0);
    for (int j = tid; j < M; j += nt) z[zp(j)] = cmul(z[zp(j)], bh[j]);
    __syncthreads();
    fft_passes<+1, GEN>(z, M, a.nrad, a.radix, a.tw, tid, nt);
    for (int j = tid; j < N; j += nt) {
        double2 c = a.chirp[j];
        if (SIGN > 0) c.y = -c.y;
        z[zp(j)] = cmul(z[zp(j)], c);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Tiles of 2G adjacent latitudes, held as G packed complex rows (G = 4 for
// nlon <= 2560, less for longer rows: the G complex rows must fit in LDS).  The
// grid and the Fourier workspace both keep the latitude index fastest, so a
// tile is one contiguous run of 2G latitudes at every longitude / order: lanes
// are grouped by 2G over one longitude / order (lane k: latitude L0 + k), which
// gives 16G-byte grid runs and 32G-byte Fourier runs (a tile of G north/south
// pairs would give two runs of half that length).  Lane k and its partner
// k ^ (2G-1) share packed row min(k, 2G-1-k); the split into the N/S halves of a
// latitude pair happens in the Legendre kernels, which address both latitudes
// of a pair directly.
//
// Work distribution: the (field, tile) units are dealt to the 8 XCDs in
// contiguous ranges and every workgroup of an XCD sweeps that range with a
// stride of the XCD's workgroup count, so neighbouring tiles (which share
// cache lines) run on one XCD at the same time.
// ---------------------------------------------------------------------------
// Value of lane k ^ (L - 1) within each group of L = 2, 4 or 8 lanes (the other
// latitude of a packed row), through DPP: quad_perm [1,0,3,2], quad_perm
// [3,2,1,0] or row_half_mirror.
template <int L>
__device__ __forceinline__ double tile_partner(double v) {
    static_assert(L == 2 || L == 4 || L == 8, "tile of 2, 4 or 8 latitudes");
    constexpr int ctrl = L == 2 ? 0xB1 : L == 4 ? 0x1B : 0x141;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), ctrl, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

struct FftTile {
    int total, per, nslot, ntile;
};

__device__ __forceinline__ FftTile fft_tiles(const FftArgs &a, int G, int nfields) {
    FftTile t;
    t.ntile = (a.nlat + 2 * G - 1) / (2 * G);
    t.total = t.ntile * nfields;
    t.per = (t.total + 7) / 8;
    t.nslot = gridDim.x >> 3;
    return t;
}

// Row factor of latitude lat (pair index p = min(lat, nlat - 1 - lat)).
__device__ __forceinline__ double fft_row_scale(const FftArgs &a, int lat) {
    const int p = min(lat, a.nlat - 1 - lat);
    if (a.mode == 1) {
        const double st = a.s[p];
        return st > 0.0 ? 1.0 / st : 1.0;
    }
    return a.mode == 2 ? a.s[p] : 1.0;
}

// Synthesis: Fourier data -> grid rows.  Row g of the tile holds the packed
// rows z = f_a + i f_b of latitudes L0 + g and L0 + 2G - 1 - g at positions zp(j)
// (row stride RS = N + N/16 + 1 complex: the rows of one longitude fall in
// different LDS banks).
template <int G, typename GT, bool GEN>
__global__ __launch_bounds__(1024) void k_fft_synth_t(const FftArgs a, int nfields) {
    extern __shared__ double2 zs[];
    constexpr int L = 2 * G;
    const int N = a.nlon, nlat = a.nlat, RS = a.nfft + (a.nfft >> 4) + 1;
    // orders 0 .. N/2 - 1 in steps of 1 fill every bin of the row except N / 2
    const bool full = !a.blue && a.mres == 1 && 2 * a.nm == N;
    const int tid = threadIdx.x, nt = blockDim.x, tr = nt / G;
    const int row = tid / tr, lt = tid % tr;
    const int k = tid % L;                // lane role within its group of 2G
    const int gk = k < G ? k : L - 1 - k;  // packed row of this lane within the tile
    const FftTile T = fft_tiles(a, G, nfields);
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    for (int u = slot; u < T.per; u += T.nslot) {
        const int ug = xcd * T.per + u;
        if (ug >= T.total) break;
        const int f = ug / T.ntile, lat = (ug % T.ntile) * L + k;
        const bool live = lat < nlat;
        if (full) {
            // every bin but N / 2 is written below (the previous tile's barrier
            // ordered its reads before these writes)
            if (tid < G) zs[tid * RS + zp(N / 2)] = make_double2(0.0, 0.0);
        } else {
            for (int q = tid; q < G * RS; q += nt) zs[q] = make_double2(0.0, 0.0);
            __syncthreads();
        }
        // packed rows: lane k < G writes z[m], its partner z[N - m]
        const double2 *Ff = a.F + (size_t)f * a.nm * nlat + lat;
        double2 *z = zs + gk * RS;
        constexpr int PF = 8;  // loads in flight per thread
        for (int e0 = 0; e0 < a.nm * L; e0 += PF * nt) {
            double2 ld[PF];
#pragma unroll
            for (int q = 0; q < PF; ++q) {
                const int im = (e0 + tid + q * nt) / L;
                ld[q] = live && im < a.nm ? Ff[(size_t)im * nlat] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int q = 0; q < PF; ++q) {
                const int im = (e0 + tid + q * nt) / L;
                const double2 mine = ld[q];
                const double2 other = make_double2(tile_partner<L>(mine.x), tile_partner<L>(mine.y));
                const double2 Fa = k < G ? mine : other, Fb = k < G ? other : mine;
                const int m = im * a.mres;
                if (im < a.nm) {
                    if (m == 0) {
                        if (k < G) z[0] = make_double2(Fa.x, Fb.x);
                    } else if (k < G) {
                        z[zp(m)] = make_double2(Fa.x - Fb.y, Fa.y + Fb.x);
                    } else {
                        z[zp(N - m)] = make_double2(Fa.x + Fb.y, Fb.x - Fa.y);
                    }
                }
            }
        }
        __syncthreads();
        fft_lds<+1, GEN>(zs + row * RS, a, lt, tr);
        const int nrow = a.nrow;
        GT *gf = (GT *)a.grid + (size_t)f * nrow * N + lat;
        if (live && lat < nrow) {
            const double sc = fft_row_scale(a, lat);
            // this lane's latitude is the real (k < G) or imaginary part of the packed row
            const double *zr = (const double *)z + (k < G ? 0 : 1);
            for (int e = tid; e < N * L; e += nt) {
                const int j = e / L;
                gf[(size_t)j * nrow] = (GT)(zr[2 * zp(j)] * sc);
            }
        }
        __syncthreads();
    }
}

// Complex-valued synthesis (real_output = false, src/core_transforms.jl:294-344):
// each latitude is its own complex row with the bins m = 0 .. mmax of the
// Fourier data (no Hermitian mirror) and the grid element is the complex value.
// Tiles of G adjacent latitudes, one LDS row each.
template <int G, typename GT, bool GEN>
__global__ __launch_bounds__(1024) void k_fft_synth_c(const FftArgs a, int nfields) {
    extern __shared__ double2 zs[];
    const int N = a.nlon, nlat = a.nlat, RS = a.nfft + (a.nfft >> 4) + 1;
    const int tid = threadIdx.x, nt = blockDim.x, tr = nt / G;
    const int row = tid / tr, lt = tid % tr;
    const int k = tid % G;  // latitude of this lane within the tile
    const int ntile = (nlat + G - 1) / G, total = ntile * nfields;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int f = u / ntile, lat = (u % ntile) * G + k;
        const bool live = lat < nlat;
        for (int q = tid; q < G * RS; q += nt) zs[q] = make_double2(0.0, 0.0);
        __syncthreads();
        const double2 *Ff = a.F + (size_t)f * a.nm * nlat + lat;
        double2 *z = zs + k * RS;
        for (int e = tid; e < a.nm * G; e += nt) {
            const int im = e / G;
            if (live) z[zp(im * a.mres)] = Ff[(size_t)im * nlat];
        }
        __syncthreads();
        fft_lds<+1, GEN>(zs + row * RS, a, lt, tr);
        const int nrow = a.nrow;
        GT *gf = (GT *)a.grid + 2 * ((size_t)f * nrow * N + lat);
        if (live && lat < nrow) {
            const double sc = fft_row_scale(a, lat);
            for (int e = tid; e < N * G; e += nt) {
                const int j = e / G;
                const double2 v = z[zp(j)];
                gf[2 * (size_t)j * nrow] = (GT)(v.x * sc);
                gf[2 * (size_t)j * nrow + 1] = (GT)(v.y * sc);
            }
        }
        __syncthreads();
    }
}

// Complex-valued analysis (analysis of a complex field, src/core_transforms.jl:
// 169-207 with a complex eltype): each latitude's complex row is transformed
// and the bins m = 0 .. mmax are kept, the columns the m-loop reads
// (Fph[:, m+1]).  Tiles of G adjacent latitudes, one LDS row each.
template <int G, typename GT, bool GEN>
__global__ __launch_bounds__(1024) void k_fft_anal_c(const FftArgs a, int nfields) {
    extern __shared__ double2 zs[];
    const int N = a.nlon, nlat = a.nlat, RS = a.nfft + (a.nfft >> 4) + 1;
    const int tid = threadIdx.x, nt = blockDim.x, tr = nt / G;
    const int row = tid / tr, lt = tid % tr;
    const int k = tid % G;  // latitude of this lane within the tile
    const int ntile = (nlat + G - 1) / G, total = ntile * nfields;
    for (int u = blockIdx.x; u < total; u += gridDim.x) {
        const int f = u / ntile, lat = (u % ntile) * G + k;
        const bool live = lat < nlat;
        const int nrow = a.nrow;
        const bool has_row = live && lat < nrow;  // rows past nrow read as zero
        const double sc = live ? fft_row_scale(a, lat) : 0.0;
        const GT *gf = (const GT *)a.grid + 2 * ((size_t)f * nrow * N + lat);
        double2 *z = zs + k * RS;
        for (int e = tid; e < N * G; e += nt) {
            const int j = e / G;
            z[zp(j)] = has_row ? make_double2((double)gf[2 * (size_t)j * nrow] * sc,
                                              (double)gf[2 * (size_t)j * nrow + 1] * sc)
                               : make_double2(0.0, 0.0);
        }
        __syncthreads();
        fft_lds<-1, GEN>(zs + row * RS, a, lt, tr);
        if (live) {
            double2 *Ff = a.F + (size_t)f * a.nm * nlat + lat;
            for (int e = tid; e < a.nm * G; e += nt) {
                const int im = e / G;
                Ff[(size_t)im * nlat] = z[zp(im * a.mres)];
            }
        }
        __syncthreads();
    }
}

// Analysis: grid rows -> Fourier data, F_a = (Z_m + conj Z_{N-m}) / 2,
// F_b = (Z_m - conj Z_{N-m}) / (2i) of the packed row Z = f_a + i f_b.
template <int G, typename GT, bool GEN>
__global__ __launch_bounds__(1024) void k_fft_anal_t(const FftArgs a, int nfields) {
    extern __shared__ double2 zs[];
    constexpr int L = 2 * G;
    const int N = a.nlon, nlat = a.nlat, RS = a.nfft + (a.nfft >> 4) + 1;
    const int tid = threadIdx.x, nt = blockDim.x, tr = nt / G;
    const int row = tid / tr, lt = tid % tr;
    const int k = tid % L;
    const int gk = k < G ? k : L - 1 - k;
    const FftTile T = fft_tiles(a, G, nfields);
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    for (int u = slot; u < T.per; u += T.nslot) {
        const int ug = xcd * T.per + u;
        if (ug >= T.total) break;
        const int f = ug / T.ntile, lat = (ug % T.ntile) * L + k;
        const bool live = lat < nlat;
        const double sc = live ? fft_row_scale(a, lat) : 0.0;
        const int nrow = a.nrow;
        const bool has_row = live && lat < nrow;  // rows past nrow read as zero
        const GT *gf = (const GT *)a.grid + (size_t)f * nrow * N + lat;
        double2 *z = zs + gk * RS;
        // this lane's latitude is the real (k < G) or imaginary part of the packed row
        double *zr = (double *)z + (k < G ? 0 : 1);
        constexpr int PF = 8;  // loads in flight per thread
        for (int e0 = 0; e0 < N * L; e0 += PF * nt) {
            double ld[PF];
#pragma unroll
            for (int q = 0; q < PF; ++q) {
                const int j = (e0 + tid + q * nt) / L;
                ld[q] = has_row && j < N ? (double)gf[(size_t)j * nrow] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < PF; ++q) {
                const int j = (e0 + tid + q * nt) / L;
                if (j < N) zr[2 * zp(j)] = ld[q] * sc;
            }
        }
        __syncthreads();
        fft_lds<-1, GEN>(zs + row * RS, a, lt, tr);
        if (live) {
            double2 *Ff = a.F + (size_t)f * a.nm * nlat + lat;
            for (int e = tid; e < a.nm * L; e += nt) {
                const int im = e / L;
                const int m = im * a.mres;
                const double2 Zm = z[zp(m)];
                const double2 Zc = z[zp((N - m) % N)];
                Ff[(size_t)im * nlat] = k < G ? make_double2(0.5 * (Zm.x + Zc.x), 0.5 * (Zm.y - Zc.y))
                                              : make_double2(0.5 * (Zm.y + Zc.y), -0.5 * (Zm.x - Zc.x));
            }
        }
        __syncthreads();
    }
}

}  // namespace

namespace {
// Threads per LDS row: nfft / FFT_VPT rounded up to whole wavefronts (every
// pass has at most FFT_VPT values per thread); 0 if the row does not fit.
int fft_row_threads(const FftArgs &a, size_t *row_bytes) {
    const int NF = a.nfft;
    *row_bytes = sizeof(double2) * (size_t)(NF + (NF >> 4) + 1);
    const int tr = (std::max(64, (NF + FFT_VPT - 1) / FFT_VPT) + 63) / 64 * 64;
    if (NF < 1 || *row_bytes > 160 * 1024 || tr > 1024) return 0;
    return tr;
}

// The general kernels (generic prime passes, Bluestein) only where the length needs them.
bool fft_general(const FftArgs &a) {
    if (a.blue) return true;
    for (int q = 0; q < a.nrad; ++q) {
        const int R = (int)((a.radix >> (FFT_RADIX_BITS * q)) & ((1ull << FFT_RADIX_BITS) - 1));
        if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8) return true;
    }
    return false;
}

template <int G, typename GT, bool GEN>
hipError_t launch_tile_g(const FftArgs &a, int nfields, int inverse, int tr, size_t row, hipStream_t st) {
    const int nt = G * tr;
    const size_t lds = row * G;
    const void *fn = inverse ? (const void *)k_fft_synth_t<G, GT, GEN> : (const void *)k_fft_anal_t<G, GT, GEN>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int ntile = (a.nlat + 2 * G - 1) / (2 * G);
    const int total = ntile * nfields;
    // persistent grid: as many workgroups as fit the chip (LDS-limited), a multiple of 8
    const int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
    int grid = std::min(256 * per_cu, (total + 7) / 8 * 8);
    grid = std::max(8, grid / 8 * 8);
    if (inverse) hipLaunchKernelGGL((k_fft_synth_t<G, GT, GEN>), dim3(grid), dim3(nt), lds, st, a, nfields);
    else hipLaunchKernelGGL((k_fft_anal_t<G, GT, GEN>), dim3(grid), dim3(nt), lds, st, a, nfields);
    return hipGetLastError();
}

template <int G>
hipError_t launch_tile(const FftArgs &a, int nfields, int inverse, int tr, size_t row, hipStream_t st) {
    const bool gen = fft_general(a);
    if (a.io_f32)
        return gen ? launch_tile_g<G, float, true>(a, nfields, inverse, tr, row, st)
                   : launch_tile_g<G, float, false>(a, nfields, inverse, tr, row, st);
    return gen ? launch_tile_g<G, double, true>(a, nfields, inverse, tr, row, st)
               : launch_tile_g<G, double, false>(a, nfields, inverse, tr, row, st);
}

template <int G, typename GT, bool GEN>
hipError_t launch_cplx_g(const FftArgs &a, int nfields, int inverse, int tr, size_t row, hipStream_t st) {
    const size_t lds = row * G;
    const void *fn = inverse ? (const void *)k_fft_synth_c<G, GT, GEN> : (const void *)k_fft_anal_c<G, GT, GEN>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int total = (a.nlat + G - 1) / G * nfields;
    const int grid = std::max(1, std::min(256 * (int)std::max<size_t>(1, (160 * 1024) / lds), total));
    if (inverse) hipLaunchKernelGGL((k_fft_synth_c<G, GT, GEN>), dim3(grid), dim3(G * tr), lds, st, a, nfields);
    else hipLaunchKernelGGL((k_fft_anal_c<G, GT, GEN>), dim3(grid), dim3(G * tr), lds, st, a, nfields);
    return hipGetLastError();
}

template <typename GT, bool GEN>
hipError_t launch_cplx_t(const FftArgs &a, int nfields, int inverse, hipStream_t st) {
    size_t row = 0;
    const int tr = fft_row_threads(a, &row);
    if (!tr) return hipErrorNotSupported;
    if (4 * row <= 160 * 1024 && 4 * tr <= 1024) return launch_cplx_g<4, GT, GEN>(a, nfields, inverse, tr, row, st);
    if (2 * row <= 160 * 1024 && 2 * tr <= 1024) return launch_cplx_g<2, GT, GEN>(a, nfields, inverse, tr, row, st);
    return launch_cplx_g<1, GT, GEN>(a, nfields, inverse, tr, row, st);
}
}  // namespace

hipError_t launch_fft_cplx(const FftArgs &a, int nfields, int inverse, hipStream_t st) {
    const bool gen = fft_general(a);
    if (a.io_f32)
        return gen ? launch_cplx_t<float, true>(a, nfields, inverse, st) : launch_cplx_t<float, false>(a, nfields, inverse, st);
    return gen ? launch_cplx_t<double, true>(a, nfields, inverse, st) : launch_cplx_t<double, false>(a, nfields, inverse, st);
}

hipError_t launch_fft(const FftArgs &a, int nfields, int inverse, hipStream_t st) {
    size_t row = 0;
    const int tr = fft_row_threads(a, &row);
    if (!tr) return hipErrorNotSupported;
    // widest tile that fits LDS and still gives every CU work (a.tile_pairs > 0 forces it)
    const long long units4 = (long long)((a.nlat + 7) / 8) * nfields, units2 = (long long)((a.nlat + 3) / 4) * nfields;
    int G = a.tile_pairs;
    if (G <= 0) G = units4 >= 512 ? 4 : units2 >= 256 ? 2 : 1;
    if (G >= 4 && 4 * row <= 160 * 1024 && 4 * tr <= 1024) return launch_tile<4>(a, nfields, inverse, tr, row, st);
    if (G >= 2 && 2 * row <= 160 * 1024 && 2 * tr <= 1024) return launch_tile<2>(a, nfields, inverse, tr, row, st);
    return launch_tile<1>(a, nfields, inverse, tr, row, st);
}

}  // namespace shtk

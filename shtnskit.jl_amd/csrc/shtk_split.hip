// Batched split-precision (Ozaki) Legendre synthesis on the int8 matrix cores
// of gfx950.
//
// For one order m and a group of 16 fields the synthesis contraction is two
// GEMMs (even and odd l - m, N/S folding as in the FP64 kernels):
//     E/O[p, c] = sum_l P_l(x_p) a_l[c],   p = latitude pair, c = real column.
// The reference computes the same sums one latitude at a time in FP64
// (src/kernels.jl:57-65 inside src/core_transforms.jl:823-869 and
// src/batch_transforms.jl:489-553).  Here:
//   * P_l(x_p) is generated in FP64 by the P-form recurrence (scaled exponents
//     as src/legendre.jl:287-322), scaled by a per-pair power of two 2^(46-ep)
//     (|P| < 2^ep, one bit of headroom) and rounded to an integer X = S + D,
//     S = sum_k s_k 256^(5-k) with signed digits s_k = byte_k(X + 2^47) - 128
//     and D = 0x8080808080;
//   * the field column is scaled by 2^(46-eb) and written as balanced signed
//     base-256 digits b_t (Y = sum_t b_t 256^(5-t));
//   * sum_l X Y = sum_{k+t<=5} 256^(10-k-t) sum_l s_k b_t + D sum_l Y + (dropped
//     low-order products); each sum_l s_k b_t is one v_mfma_i32_16x16x64_i8 chain
//     with exact int32 accumulation, and the 6 order levels combine in FP64.
// Worst per-order relative L2 error of this scheme at lmax = 1023 (numpy
// emulation, l^-1.5 spectra): 7.8e-13.
#include "shtk_internal.h"

namespace shtk {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr double kHiS = 0x1p100;
constexpr double kDownS = 0x1p-512;
constexpr double kBias = 551911719040.0;  // 0x8080808080

__device__ __forceinline__ int oz_nch(int lmax, int m) { return (lmax - m + 128) / 128; }

// ---------------------------------------------------------------------------
// Plan-time seeds at the quarter starts and per-pair exponents: one thread per
// (order, pair) runs the scaled P-form recurrence over the whole order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_oz_seed(const OzSeedArgs a) {
    const int im = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.np) return;
    const int m = im * a.mres, lmax = a.lmax;
    const int Kq = 32 * oz_nch(lmax, m);
    const double x = a.x[p];
    const double2 *ab = a.ab + a.coff[im] - m;
    const size_t sidx = (size_t)im * a.np + p;
    double cur = a.sectv[sidx], prev = 0.0;
    int k = a.sectk[sidx];
    int emax = -100000;
    int g = 0;
    for (int l = m; l <= lmax; ++l) {
        if (l == m + g * Kq) {
            const size_t q = ((size_t)im * 4 + g) * a.np + p;
            a.q0[q] = prev;
            a.q1[q] = cur;
            a.qk[q] = k;
            ++g;
        }
        if (k == 0 && cur != 0.0) emax = max(emax, ilogb(cur));
        const double2 co = ab[l + 1];
        const double nx = fma(co.x * x, cur, -(co.y * prev));
        prev = cur;
        cur = nx;
        if (k > 0 && ((l - m) & (SCALE_CHECK - 1)) == SCALE_CHECK - 1 && fmax(fabs(cur), fabs(prev)) > kHiS) {
            cur *= kDownS;
            prev *= kDownS;
            --k;
        }
    }
    for (; g < 4; ++g) {
        const size_t q = ((size_t)im * 4 + g) * a.np + p;
        a.q0[q] = 0.0;
        a.q1[q] = 0.0;
        a.qk[q] = 0;
    }
    // |P| < 2^ep; rows below 2^-200 everywhere contribute nothing at FP64 accuracy
    a.ep[sidx] = emax < -201 ? -200 : min(emax + 1, 64);
}

// ---------------------------------------------------------------------------
// Field operand: one block per (order, field group).  Column maxima give the
// exponents eb; every thread then writes one 4-byte word of each of the 6
// digit planes, in the B-operand layout
//   [chunk][parity][slice t][column tile ct][lane][16 bytes],
// lane = c + 16 g (column c of the tile, quarter g), byte j <-> degree
// l = m + g Kq + 32 chunk + 2 j + parity.
// ---------------------------------------------------------------------------
template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bdig(const OzArgs a) {
    __shared__ double cmax[256];
    __shared__ int ebs[OZ_COLS];
    __shared__ unsigned long long csum[2][OZ_COLS];
    const int im = blockIdx.x % a.nm, fg = blockIdx.x / a.nm;
    const int m = im * a.mres, lmax = a.lmax;
    const int nch = oz_nch(lmax, m), Kq = 32 * nch;
    const int lend = min(lmax, a.ltr);
    const int tid = threadIdx.x;
    const IT *base = (const IT *)a.spec_in;
    auto colptr = [&](int c, bool *ok) {
        const int f = fg * OZ_FIELDS + (c >> 1);
        *ok = f < a.nf;
        return base + 2 * ((size_t)(*ok ? f : 0) * a.spec_fstride + (size_t)(lmax + 1) * m) + (c & 1);
    };
    {
        const int c = tid & (OZ_COLS - 1), sub = tid >> 5;
        bool ok;
        const IT *src = colptr(c, &ok);
        double mx = 0.0;
        if (ok)
            for (int l = m + sub; l <= lend; l += 8) mx = fmax(mx, fabs((double)src[2 * l]));
        cmax[tid] = mx;
    }
    if (tid < 2 * OZ_COLS) csum[tid >> 5][tid & 31] = 0ull;
    __syncthreads();
    if (tid < OZ_COLS) {
        double v = 0.0;
        for (int s = 0; s < 8; ++s) v = fmax(v, cmax[s * OZ_COLS + tid]);
        ebs[tid] = v > 0.0 ? ilogb(v) + 1 : 0;
    }
    __syncthreads();
    unsigned *dst = (unsigned *)(a.bdig + ((size_t)fg * a.ch_total + a.chbase[im]) * OZ_CHUNK_V4);
    const int units = nch * 2 * 2 * 64 * 4;
    for (int u = tid; u < units; u += 256) {
        const int d = u & 3, lane = (u >> 2) & 63, ct = (u >> 8) & 1, par = (u >> 9) & 1, rho = u >> 10;
        const int g = lane >> 4, c = (lane & 15) + 16 * ct;
        bool ok;
        const IT *src = colptr(c, &ok);
        const double sc = ldexp(1.0, 46 - ebs[c]);
        unsigned dig[OZ_SLICES];
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t) dig[t] = 0u;
        long long ysum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = 4 * d + q;
            const int l = m + g * Kq + 32 * rho + 2 * j + par;
            long long Y = 0;
            if (ok && l <= lend) Y = llrint((double)src[2 * l] * sc);
            ysum += Y;
#pragma unroll
            for (int t = OZ_SLICES - 1; t >= 0; --t) {
                const long long dd = ((Y + 128) & 255) - 128;
                Y = (Y - dd) >> 8;
                dig[t] |= ((unsigned)dd & 255u) << (8 * q);
            }
        }
        atomicAdd(&csum[par][c], (unsigned long long)ysum);
        const size_t chunk0 = (size_t)rho * OZ_CHUNK_V4 * 4;
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t)
            dst[chunk0 + ((size_t)((par * OZ_SLICES + t) * 2 + ct) * 64 + lane) * 4 + d] = dig[t];
    }
    __syncthreads();
    if (tid < 2 * OZ_COLS) {
        const int par = tid >> 5, c = tid & 31;
        a.colsum[(((size_t)fg * a.nm + im) * 2 + par) * OZ_COLS + c] = (double)(long long)csum[par][c];
    }
    if (tid < OZ_COLS) a.eb[((size_t)fg * a.nm + im) * OZ_COLS + tid] = ebs[tid];
}

// byte b of v0..v3 packed into one word (v0 in the low byte)
__device__ __forceinline__ unsigned pack4(unsigned v0, unsigned v1, unsigned v2, unsigned v3, int b) {
    const unsigned sel = (unsigned)b | ((unsigned)(b + 4) << 8) | 0x0c0c0000u;
    const unsigned t01 = __builtin_amdgcn_perm(v1, v0, sel);
    const unsigned t23 = __builtin_amdgcn_perm(v3, v2, sel);
    return __builtin_amdgcn_perm(t23, t01, 0x05040100u);
}

// ---------------------------------------------------------------------------
// Synthesis: one block of 4 waves per (64 pairs, order, field group); wave w
// owns pairs 16w..16w+15 of the block (MFMA rows), lane group g its quarter of
// the degrees.  Per chunk: coefficients and B operands are staged in LDS, each
// lane runs 32 recurrence steps and packs its digits into the A operands,
// then 84 MFMAs accumulate the 2 parities x 2 column tiles x 6 levels.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_synth_oz(const OzArgs a) {
    __shared__ double2 abch[4 * 32];
    __shared__ v4i bch[OZ_CHUNK_V4];
    const int gx = (a.np + 63) / 64;
    int pbk, unit;
    if (!unit_of_block(a.xcd, gx, a.nm * a.nfg, &unit, &pbk)) return;
    const int im = unit % a.nm, fg = unit / a.nm;
    const int m = im * a.mres, lmax = a.lmax;
    const int nch = oz_nch(lmax, m), Kq = 32 * nch;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int p = pbk * 64 + w * 16 + r;
    const bool act = p < a.np;
    const double x = act ? a.x[p] : 0.0;
    const size_t sq = ((size_t)im * 4 + g) * a.np + (act ? p : 0);
    double prev = act ? a.q0[sq] : 0.0, cur = act ? a.q1[sq] : 0.0;
    int kk = act ? a.qk[sq] : 0;
    const double scl = ldexp(1.0, 46 - (act ? a.ep[(size_t)im * a.np + p] : 0));
    const int lq = m + g * Kq;
    const double2 *ab = a.ab + a.coff[im] - m;
    const v4i *bsrc = (const v4i *)a.bdig + ((size_t)fg * a.ch_total + a.chbase[im]) * OZ_CHUNK_V4;

    v4i acc[2][2][OZ_SLICES];
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int v = 0; v < OZ_SLICES; ++v) acc[par][ct][v] = v4i{0, 0, 0, 0};

    for (int rho = 0; rho < nch; ++rho) {
        __syncthreads();
        if (tid < 128) {
            const int l = m + (tid >> 5) * Kq + 32 * rho + (tid & 31);
            abch[tid] = ab[min(l, lmax) + 1];
        }
        for (int q = tid; q < OZ_CHUNK_V4; q += 256) bch[q] = bsrc[(size_t)rho * OZ_CHUNK_V4 + q];
        __syncthreads();

        v4i A[2][OZ_SLICES];
        const int lc = lq + 32 * rho;
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            unsigned h[8], lo[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = blk * 8 + q;
                const double v = (kk == 0 && lc + i <= lmax) ? cur : 0.0;
                const double y = fma(v, scl, 0x1p47);
                const double hi = floor(y * 0x1p-24);
                const double lw = fma(hi, -0x1p24, y);
                h[q] = (unsigned)hi ^ 0x808080u;
                lo[q] = (unsigned)lw ^ 0x808080u;
                const double2 co = abch[g * 32 + i];
                const double nx = fma(co.x * x, cur, -(co.y * prev));
                prev = cur;
                cur = nx;
            }
            if (kk > 0 && fmax(fabs(cur), fabs(prev)) > kHiS) {
                cur *= kDownS;
                prev *= kDownS;
                --kk;
            }
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                // slices 0..2 are bytes 2..0 of the high word, 3..5 of the low word
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    A[par][s][blk] = (int)pack4(h[par], h[par + 2], h[par + 4], h[par + 6], 2 - s);
                    A[par][s + 3][blk] = (int)pack4(lo[par], lo[par + 2], lo[par + 4], lo[par + 6], 2 - s);
                }
            }
        }
#pragma unroll
        for (int par = 0; par < 2; ++par)
#pragma unroll
            for (int t = 0; t < OZ_SLICES; ++t)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const v4i B = bch[((par * OZ_SLICES + t) * 2 + ct) * 64 + lane];
#pragma unroll
                    for (int s = 0; s + t < OZ_SLICES; ++s)
                        acc[par][ct][s + t] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(A[par][s], B, acc[par][ct][s + t], 0, 0, 0);
                }
    }

    // epilogue: lane holds rows 4 (lane >> 4) + i, column (lane & 15) + 16 ct
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int c = (lane & 15) + 16 * ct;
        const int f = fg * OZ_FIELDS + (c >> 1);
        const size_t cidx = ((size_t)fg * a.nm + im) * OZ_COLS + c;
        const double cs0 = a.colsum[(((size_t)fg * a.nm + im) * 2 + 0) * OZ_COLS + c];
        const double cs1 = a.colsum[(((size_t)fg * a.nm + im) * 2 + 1) * OZ_COLS + c];
        const int ebc = a.eb[cidx];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pp = pbk * 64 + w * 16 + 4 * (lane >> 4) + i;
            const bool live = pp < a.np;
            const int epp = live ? a.ep[(size_t)im * a.np + pp] : 0;
            double E = kBias * cs0, O = kBias * cs1;
#pragma unroll
            for (int v = 0; v < OZ_SLICES; ++v) {
                const double wv = ldexp(1.0, 80 - 8 * v);
                E = fma((double)acc[0][ct][v][i], wv, E);
                O = fma((double)acc[1][ct][v][i], wv, O);
            }
            const double sc = a.g * ldexp(1.0, epp - 46 + ebc - 46);
            const double N = (E + O) * sc, S = (E - O) * sc;
            const double Np = __shfl_xor(N, 1), Sp = __shfl_xor(S, 1);
            if ((lane & 1) == 0 && live && f < a.nf) {
                double2 *Fo = a.F0 + ((size_t)f * a.nm + im) * a.nlat;
                Fo[pp] = make_double2(N, Np);
                const int ps = a.nlat - 1 - pp;
                if (ps != pp) Fo[ps] = make_double2(S, Sp);
            }
        }
    }
}

}  // namespace

hipError_t launch_oz_seed(const OzSeedArgs &a, hipStream_t st) {
    dim3 grid((a.np + 255) / 256, a.nm);
    hipLaunchKernelGGL(k_oz_seed, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_oz_synth(const OzArgs &a, hipStream_t st) {
    if (a.in_f32)
        hipLaunchKernelGGL(k_oz_bdig<float>, dim3((unsigned)a.nm * a.nfg), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_oz_bdig<double>, dim3((unsigned)a.nm * a.nfg), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int gx = (a.np + 63) / 64;
    hipLaunchKernelGGL(k_synth_oz, dim3(grid_for_units(a.xcd, gx, a.nm * a.nfg)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace shtk

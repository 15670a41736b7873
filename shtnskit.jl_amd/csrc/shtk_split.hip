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
// 2^52 + 2^47: adding it to |v| < 2^46 leaves the integer 2^47 + round(v) in the
// low 48 mantissa bits (exact: the sum stays in [2^52, 2^53))
constexpr double kMagic = 0x1p52 + 0x1p47;
constexpr int OZ_ROWB = 272;               // analysis LDS row stride (256 pairs + 16 bytes)

__device__ __forceinline__ int oz_nch(int lmax, int m) { return (lmax - m + 128) / 128; }

// ---------------------------------------------------------------------------
// Plan-time seeds at the quarter starts and per-pair exponents: one thread per
// (order, pair) runs the scaled P-form recurrence over the whole order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_oz_seed(const OzSeedArgs a) {
    const int im = blockIdx.y;
    const int p0 = blockIdx.x * 256 + threadIdx.x;
    const bool act = p0 < a.np;
    const int p = act ? p0 : 0;
    const int m = im * a.mres, lmax = a.lmax;
    const int Kq = 32 * oz_nch(lmax, m);
    const double x = a.x[p];
    const double2 *ab = a.ab + a.coff[im] - m;
    int *emaxl = a.emaxl + a.coff[im] - m;
    const size_t sidx = (size_t)im * a.np + p;
    double cur = act ? a.sectv[sidx] : 0.0, prev = 0.0;
    int k = act ? a.sectk[sidx] : 0;
    int emax = -100000;
    int g = 0;
    for (int l = m; l <= lmax; ++l) {
        if (act && l == m + g * Kq) {
            const size_t q = ((size_t)im * 4 + g) * a.np + p;
            a.q0[q] = prev;
            a.q1[q] = cur;
            a.qk[q] = k;
            ++g;
        }
        const int el = (act && k == 0 && cur != 0.0) ? ilogb(cur) : -100000;
        emax = max(emax, el);
        // wavefront maximum per degree (all lanes run the same degrees)
        int wm = el;
#pragma unroll
        for (int msk = 32; msk >= 1; msk >>= 1) wm = max(wm, __shfl_xor(wm, msk));
        if ((threadIdx.x & 63) == 0 && wm > -100000) atomicMax(&emaxl[l], wm);
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
    if (!act) return;
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
// Field operand of the synthesis.  k_oz_bexp: one block per (order, field
// group) finds the column exponents eb (|a| < 2^eb over the order's degrees)
// and clears the column sums.  k_oz_bdig: one block per (field group, chunk)
// writes one 4-byte word of each of the 6 digit planes per thread-unit, in the
// B-operand layout
//   [chunk][parity][slice t][column tile ct][lane][16 bytes],
// lane = c + 16 g (column c of the tile, quarter g), byte j <-> degree
// l = m + g Kq + 32 chunk + 2 j + parity, and adds its integer column sums.
// ---------------------------------------------------------------------------
template <typename IT>
__device__ __forceinline__ const IT *oz_col(const OzArgs &a, int fg, int m, int c, bool *ok) {
    const int f = fg * OZ_FIELDS + (c >> 1);
    *ok = f < a.nf;
    return (const IT *)a.spec_in + 2 * ((size_t)(*ok ? f : 0) * a.spec_fstride + (size_t)(a.lmax + 1) * m) + (c & 1);
}

template <typename IT>
__device__ __forceinline__ void oz_bexp_body(const OzArgs &a, int im, int fg, double *cmax) {
    const int m = im * a.mres;
    const int lend = min(a.lmax, a.ltr);
    const int tid = threadIdx.x;
    const int c = tid & (OZ_COLS - 1), sub = tid >> 5;
    bool ok;
    const IT *src = oz_col<IT>(a, fg, m, c, &ok);
    double mx = 0.0;
    if (ok)
        for (int l = m + sub; l <= lend; l += 8) mx = fmax(mx, fabs((double)src[2 * l]));
    cmax[tid] = mx;
    __syncthreads();
    if (tid < OZ_COLS) {
        double v = 0.0;
        for (int s = 0; s < 8; ++s) v = fmax(v, cmax[s * OZ_COLS + tid]);
        a.eb[((size_t)fg * a.nm + im) * OZ_COLS + tid] = v > 0.0 ? ilogb(v) + 1 : 0;
    }
    if (tid < 2 * OZ_COLS) a.colsum[((size_t)fg * a.nm + im) * 2 * OZ_COLS + tid] = 0ull;
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bexp(const OzArgs a) {
    __shared__ double cmax[256];
    oz_bexp_body<IT>(a, blockIdx.x % a.nm, blockIdx.x / a.nm, cmax);
}

template <typename IT>
__device__ __forceinline__ void oz_bdig_body(const OzArgs &a, int gc, int fg, unsigned long long (*csum)[OZ_COLS]) {
    const int im = a.chm[gc], rho = gc - a.chbase[im];
    const int m = im * a.mres, lmax = a.lmax;
    const int Kq = 32 * oz_nch(lmax, m);
    const int lend = min(lmax, a.ltr);
    const int tid = threadIdx.x;
    if (tid < 2 * OZ_COLS) csum[tid >> 5][tid & 31] = 0ull;
    __syncthreads();
    unsigned *dst = (unsigned *)(a.bdig + ((size_t)fg * a.ch_total + gc) * OZ_CHUNK_V4);
    const int *ebm = a.eb + ((size_t)fg * a.nm + im) * OZ_COLS;
    for (int u = tid; u < 1024; u += 256) {
        const int d = u & 3, lane = (u >> 2) & 63, ct = (u >> 8) & 1, par = u >> 9;
        const int g = lane >> 4, c = (lane & 15) + 16 * ct;
        bool ok;
        const IT *src = oz_col<IT>(a, fg, m, c, &ok);
        const double sc = ldexp(1.0, 46 - ebm[c]);
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
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t)
            dst[((size_t)((par * OZ_SLICES + t) * 2 + ct) * 64 + lane) * 4 + d] = dig[t];
    }
    __syncthreads();
    if (tid < 2 * OZ_COLS)
        atomicAdd(a.colsum + ((size_t)fg * a.nm + im) * 2 * OZ_COLS + tid, csum[tid >> 5][tid & 31]);
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bdig(const OzArgs a) {
    __shared__ unsigned long long csum[2][OZ_COLS];
    oz_bdig_body<IT>(a, blockIdx.x % a.ch_total, blockIdx.x / a.ch_total, csum);
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
__device__ __forceinline__ void synth_oz_body(const OzArgs &a, int pbk, int im, int fg, double2 *abch, v4i *bch) {
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

    // the next chunk's coefficients and B operands travel in registers while the
    // current chunk computes (OZ_CHUNK_V4 / 256 = 6 operands per thread)
    constexpr int NB = OZ_CHUNK_V4 / 256;
    v4i bnext[NB];
    double2 abnext = make_double2(0.0, 0.0);
    auto fetch = [&](int rho) {
#pragma unroll
        for (int k = 0; k < NB; ++k) bnext[k] = bsrc[(size_t)rho * OZ_CHUNK_V4 + tid + 256 * k];
        if (tid < 128) {
            const int l = m + (tid >> 5) * Kq + 32 * rho + (tid & 31);
            abnext = ab[min(l, lmax) + 1];
        }
    };
    fetch(0);
    for (int rho = 0; rho < nch; ++rho) {
        __syncthreads();
        if (tid < 128) abch[tid] = abnext;
#pragma unroll
        for (int k = 0; k < NB; ++k) bch[tid + 256 * k] = bnext[k];
        __syncthreads();
        if (rho + 1 < nch) fetch(rho + 1);

        v4i A[2][OZ_SLICES];
        // a lane still in the scaled range contributes zeros (scale 0 below);
        // degrees beyond lmax meet zero field digits, so they need no mask
        double sl = kk == 0 ? scl : 0.0;
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            unsigned h[8], lo[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = blk * 8 + q;
                // U = 2^47 + round(P 2^(46-ep)) lands in the low 48 bits of the mantissa
                const unsigned long long bits = (unsigned long long)__double_as_longlong(fma(cur, sl, kMagic));
                h[q] = (unsigned)(bits >> 32) ^ 0x8080u;
                lo[q] = (unsigned)bits ^ 0x80808080u;
                const double2 co = abch[g * 32 + i];
                const double nx = fma(co.x * x, cur, -(co.y * prev));
                prev = cur;
                cur = nx;
            }
            if (kk > 0 && fmax(fabs(cur), fabs(prev)) > kHiS) {
                cur *= kDownS;
                prev *= kDownS;
                --kk;
                sl = kk == 0 ? scl : 0.0;
            }
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                // slices 0, 1 are bytes 1, 0 of the high word; 2..5 bytes 3..0 of the low word
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    A[par][s][blk] = (int)pack4(h[par], h[par + 2], h[par + 4], h[par + 6], 1 - s);
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    A[par][s + 2][blk] = (int)pack4(lo[par], lo[par + 2], lo[par + 4], lo[par + 6], 3 - s);
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
        const double cs0 = (double)(long long)a.colsum[(((size_t)fg * a.nm + im) * 2 + 0) * OZ_COLS + c];
        const double cs1 = (double)(long long)a.colsum[(((size_t)fg * a.nm + im) * 2 + 1) * OZ_COLS + c];
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

__global__ __launch_bounds__(256) void k_synth_oz(const OzArgs a) {
    __shared__ double2 abch[4 * 32];
    __shared__ v4i bch[OZ_CHUNK_V4];
    const int gx = (a.np + 63) / 64;
    int pbk, unit;
    if (!unit_of_block(a.xcd, gx, a.nm * a.nfg, &unit, &pbk)) return;
    synth_oz_body(a, pbk, unit % a.nm, unit / a.nm, abch, bch);
}

// ---------------------------------------------------------------------------
// Per-degree digit scales for the analysis: |P_l| < 2^epl over all pairs, and
// asc = c_l 2^(46 - epl) so that one FMA turns Q_l into the digit value.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_oz_asc(const OzAscArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int e = a.emaxl[i];
    const int ep = e < -201 ? -200 : min(e + 1, 64);
    a.epl[i] = ep;
    a.asc[i] = a.qc[i].y * ldexp(1.0, 46 - ep);
}

// ---------------------------------------------------------------------------
// Analysis field operand.  G+- = w (F_N +- F_S) per pair (even degrees use G+,
// odd G-).  k_oz_gexp: one block per (order, field group) finds the column
// exponents eg per parity.  k_oz_gdig: one block per (kstep of 64 pairs,
// order, field group) writes balanced digits in the B-operand layout
//   [kstep][parity][slice t][column tile ct][lane][16 bytes]
// and the kstep's integer column sums.  Within a half of 256 pairs, LDS
// position q (kstep 64 (ks & 3) + 16 g + byte j) holds local pair
// (q & 3) * 64 + (q >> 2): the order in which the analysis kernel's lanes write
// their digits without bank conflicts.
// ---------------------------------------------------------------------------
constexpr int OZ_MAX_KS = 32;  // np <= 2048

__device__ __forceinline__ double oz_gval(const OzArgs &a, int fg, int im, int c, int p, int par) {
    const int f = fg * OZ_FIELDS + (c >> 1);
    if (f >= a.nf || p >= a.np) return 0.0;
    const double2 *Fi = a.F0 + ((size_t)f * a.nm + im) * a.nlat;
    const double2 FN = Fi[p];
    const int ps = a.nlat - 1 - p;
    const double2 FS = ps != p ? Fi[ps] : make_double2(0.0, 0.0);
    const double n = (c & 1) ? FN.y : FN.x, s = (c & 1) ? FS.y : FS.x;
    const double wl = a.wlat[p];
    return par ? wl * (n - s) : wl * (n + s);
}

__device__ __forceinline__ int oz_pair_of_slot(int ks, int slot) {
    const int q = 64 * (ks & 3) + slot;
    return (ks >> 2) * 256 + (q & 3) * 64 + (q >> 2);
}

__device__ __forceinline__ void oz_gexp_body(const OzArgs &a, int im, int fg, double (*cmax)[256]) {
    const int tid = threadIdx.x;
    const int c = tid & (OZ_COLS - 1), sub = tid >> 5;
    double m0 = 0.0, m1 = 0.0;
    for (int p = sub; p < a.np; p += 8) {
        m0 = fmax(m0, fabs(oz_gval(a, fg, im, c, p, 0)));
        m1 = fmax(m1, fabs(oz_gval(a, fg, im, c, p, 1)));
    }
    cmax[0][tid] = m0;
    cmax[1][tid] = m1;
    __syncthreads();
    if (tid < 2 * OZ_COLS) {
        const int par = tid >> 5, cc = tid & 31;
        double v = 0.0;
        for (int s = 0; s < 8; ++s) v = fmax(v, cmax[par][s * OZ_COLS + cc]);
        a.eg[((size_t)fg * a.nm + im) * 2 * OZ_COLS + tid] = v > 0.0 ? ilogb(v) + 1 : 0;
    }
}

__global__ __launch_bounds__(256) void k_oz_gexp(const OzArgs a) {
    __shared__ double cmax[2][256];
    oz_gexp_body(a, blockIdx.x % a.nm, blockIdx.x / a.nm, cmax);
}

__device__ __forceinline__ void oz_gdig_body(const OzArgs &a, int ks, int im, int fg,
                                             unsigned long long (*csum)[OZ_COLS]) {
    const int tid = threadIdx.x;
    if (tid < 2 * OZ_COLS) csum[tid >> 5][tid & 31] = 0ull;
    __syncthreads();
    const size_t ubase = ((size_t)fg * a.nm + im) * a.nks;
    const int *egm = a.eg + ((size_t)fg * a.nm + im) * 2 * OZ_COLS;
    unsigned *dst = (unsigned *)(a.gdig + (ubase + ks) * OZ_CHUNK_V4);
    for (int u = tid; u < 1024; u += 256) {
        const int d = u & 3, lane = (u >> 2) & 63, ct = (u >> 8) & 1, par = u >> 9;
        const int g = lane >> 4, c = (lane & 15) + 16 * ct;
        const double sc = ldexp(1.0, 46 - egm[par * OZ_COLS + c]);
        unsigned dig[OZ_SLICES];
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t) dig[t] = 0u;
        long long ysum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = oz_pair_of_slot(ks, 16 * g + 4 * d + q);
            long long Y = llrint(oz_gval(a, fg, im, c, p, par) * sc);
            ysum += Y;
#pragma unroll
            for (int t = OZ_SLICES - 1; t >= 0; --t) {
                const long long dd = ((Y + 128) & 255) - 128;
                Y = (Y - dd) >> 8;
                dig[t] |= ((unsigned)dd & 255u) << (8 * q);
            }
        }
        atomicAdd(&csum[par][c], (unsigned long long)ysum);
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t)
            dst[((size_t)((par * OZ_SLICES + t) * 2 + ct) * 64 + lane) * 4 + d] = dig[t];
    }
    __syncthreads();
    if (tid < 2 * OZ_COLS)
        a.gcolsum[((ubase + ks) * 2) * OZ_COLS + tid] = (double)(long long)csum[tid >> 5][tid & 31];
}

__global__ __launch_bounds__(256) void k_oz_gdig(const OzArgs a) {
    __shared__ unsigned long long csum[2][OZ_COLS];
    oz_gdig_body(a, blockIdx.x % a.nks, (blockIdx.x / a.nks) % a.nm, blockIdx.x / (a.nks * a.nm), csum);
}

// ---------------------------------------------------------------------------
// Analysis: one block of 4 waves per (degree segment, half of the pairs,
// field group).  Every lane runs the Q-form recurrence of one pair from the
// segment seeds and writes the 6 digit bytes of each degree to LDS as
// [parity][slice][row][pair] (row = degree pair index within 32 degrees);
// wave w then contracts parity w >> 1, column tile w & 1 over the block's 4
// ksteps of 64 pairs (the field operand stays in registers for the whole
// segment) and adds its rows to the FP64 output.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void anal_oz_body(const OzArgs &a, const double *__restrict__ asc,
                                             const double2 *__restrict__ qc, int h, int gs, int fg,
                                             unsigned char *Alds) {
    const int im = a.segm[gs], s = gs - a.segbase[im], m = im * a.mres;
    const int l0 = m + s * a.seg;
    const int lstop = min(min(l0 + a.seg, a.lmax + 1), a.ltr + 1);
    const int nst = lstop - l0;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int pbase = h * 256;
    if (nst <= 0 || pbase + 255 < a.pfirst[gs]) return;
    const int p = pbase + tid;
    const bool act = p < a.np;
    const double x = act ? a.x[p] : 0.0;
    const size_t q = (size_t)gs * a.np + (act ? p : 0);
    double prev = act ? a.seedq0[q] : 0.0, cur = act ? a.seedq1[q] : 0.0;
    int kk = act ? a.seedk[q] : 0;
    double live = kk == 0 ? 1.0 : 0.0;  // lanes still in the scaled range contribute zeros
    // read-only tables passed as restrict kernel arguments: per-degree
    // constants become wave-uniform scalar loads
    const double2 *qcm = qc + a.coff[im] - m;
    const double *ascm = asc + a.coff[im] - m;
    const int *eplm = a.epl + a.coff[im] - m;

    // field operand of this wave: parity wp, column tile wct, 4 ksteps x 6 slices
    const int wp = w >> 1, wct = w & 1;
    v4i Bw[4][OZ_SLICES];
    const size_t ubase = ((size_t)fg * a.nm + im) * a.nks;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int kg = 4 * h + ks;
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t)
            Bw[ks][t] = kg < a.nks ? ((const v4i *)a.gdig)[(ubase + kg) * OZ_CHUNK_V4 +
                                                           ((wp * OZ_SLICES + t) * 2 + wct) * 64 + lane]
                                   : v4i{0, 0, 0, 0};
    }
    const int c = (lane & 15) + 16 * wct;
    double csum = 0.0;
    for (int ks = 0; ks < 4; ++ks) {
        const int kg = 4 * h + ks;
        if (kg < a.nks) csum += a.gcolsum[((ubase + kg) * 2 + wp) * OZ_COLS + c];
    }
    const int egc = a.eg[(((size_t)fg * a.nm + im) * 2 + wp) * OZ_COLS + c];
    const int f = fg * OZ_FIELDS + (c >> 1);
    const int wpos = 4 * (tid & 63) + (tid >> 6);  // LDS position of this lane's pair (see k_oz_gdig)

    for (int t0 = 0; t0 < nst; t0 += 32) {
        // recurrence and digits of 32 degrees
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
                const int i = blk * 8 + qq;
                const int t = t0 + i;
                // wave-uniform degree: scalar loads of its per-degree constants; rows
                // beyond the segment are never stored, so they need no mask
                const int l = __builtin_amdgcn_readfirstlane(l0 + min(t, nst - 1));
                // U = 2^47 + round(P 2^(46-epl)) lands in the low 48 bits of the mantissa
                const unsigned long long bits =
                    (unsigned long long)__double_as_longlong(fma(cur * live, ascm[l], kMagic));
                const unsigned uh = (unsigned)(bits >> 32) ^ 0x8080u, ul = (unsigned)bits ^ 0x80808080u;
                unsigned char *row = Alds + ((i & 1) * OZ_SLICES * 16 + (i >> 1)) * OZ_ROWB + wpos;
                row[0 * 16 * OZ_ROWB] = (unsigned char)(uh >> 8);
                row[1 * 16 * OZ_ROWB] = (unsigned char)uh;
                row[2 * 16 * OZ_ROWB] = (unsigned char)(ul >> 24);
                row[3 * 16 * OZ_ROWB] = (unsigned char)(ul >> 16);
                row[4 * 16 * OZ_ROWB] = (unsigned char)(ul >> 8);
                row[5 * 16 * OZ_ROWB] = (unsigned char)ul;
                const double beta = qcm[l + 1].x;
                const double nx = fma(x, cur, -(beta * prev));
                prev = cur;
                cur = nx;
            }
            const int lchk = __builtin_amdgcn_readfirstlane(l0 + min(t0 + blk * 8 + 7, nst - 1));
            if (kk > 0 && fmax(fabs(cur), fabs(prev)) * qcm[lchk].y > kHiS) {
                cur *= kDownS;
                prev *= kDownS;
                --kk;
                live = kk == 0 ? 1.0 : 0.0;
            }
        }
        __syncthreads();
        v4i acc[OZ_SLICES];
#pragma unroll
        for (int v = 0; v < OZ_SLICES; ++v) acc[v] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            v4i Aop[OZ_SLICES];
#pragma unroll
            for (int sl = 0; sl < OZ_SLICES; ++sl)
                Aop[sl] = *(const v4i *)(Alds + ((wp * OZ_SLICES + sl) * 16 + (lane & 15)) * OZ_ROWB + 64 * ks +
                                         16 * (lane >> 4));
#pragma unroll
            for (int t = 0; t < OZ_SLICES; ++t)
#pragma unroll
                for (int sl = 0; sl + t < OZ_SLICES; ++sl)
                    acc[sl + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aop[sl], Bw[ks][t], acc[sl + t], 0, 0, 0);
        }
        // rows 4 (lane >> 4) + i are degree pairs, column c
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * (lane >> 4) + i;
            const int t = t0 + 2 * row + wp;
            if (t < nst && f < a.nf) {
                const int l = l0 + t;
                double E = kBias * csum;
#pragma unroll
                for (int v = 0; v < OZ_SLICES; ++v) E = fma((double)acc[v][i], ldexp(1.0, 80 - 8 * v), E);
                const double val = E * a.g * ldexp(1.0, eplm[l] - 46 + egc - 46);
                atomicAdd(a.spec_out + 2 * ((size_t)f * a.spec_fstride + (size_t)(a.lmax + 1) * m + l) + (c & 1),
                          val);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_anal_oz(const OzArgs a, const double *__restrict__ asc,
                                                  const double2 *__restrict__ qc) {
    // rows padded to 272 bytes: the 16 lanes of an A-operand read hit distinct banks
    __shared__ __align__(16) unsigned char Alds[2 * OZ_SLICES * 16 * OZ_ROWB];
    const int item = blockIdx.x;
    anal_oz_body(a, asc, qc, item % a.nhalf, (item / a.nhalf) % a.nseg_total, item / (a.nhalf * a.nseg_total), Alds);
}

// ---------------------------------------------------------------------------
// Batches of several configurations (uneven lmax).  One launch per stage
// covers every group: blocks of the per-order prologues are assigned through
// a prefix sum of the groups' block counts, and the two Legendre contractions
// run as persistent kernels that take work units (largest orders of the
// largest configurations first) from a device counter until none is left.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int oz_group_of(const OzGroups &G, int x) {
    int gi = 0;
    while (gi + 1 < G.ngroups && G.prefix[gi + 1] <= x) ++gi;
    return gi;
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bexp_grp(const OzGroups G) {
    __shared__ double cmax[256];
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    oz_bexp_body<IT>(a, b % a.nm, b / a.nm, cmax);
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bdig_grp(const OzGroups G) {
    __shared__ unsigned long long csum[2][OZ_COLS];
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    oz_bdig_body<IT>(a, b % a.ch_total, b / a.ch_total, csum);
}

__global__ __launch_bounds__(256) void k_oz_gexp_grp(const OzGroups G) {
    __shared__ double cmax[2][256];
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    oz_gexp_body(a, b % a.nm, b / a.nm, cmax);
}

__global__ __launch_bounds__(256) void k_oz_gdig_grp(const OzGroups G) {
    __shared__ unsigned long long csum[2][OZ_COLS];
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    oz_gdig_body(a, b % a.nks, (b / a.nks) % a.nm, b / (a.nks * a.nm), csum);
}

__global__ void k_oz_reset(int *counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

// Persistent synthesis: unit u of group gi is (pair block, order, field group)
// with the pair block fastest, as the single-configuration grid.
__global__ __launch_bounds__(256) void k_synth_oz_grp(const OzGroups G) {
    __shared__ double2 abch[4 * 32];
    __shared__ v4i bch[OZ_CHUNK_V4];
    __shared__ int s_unit;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_unit = atomicAdd(G.counter, 1);
        __syncthreads();
        const int u = s_unit;
        if (u >= G.prefix[G.ngroups]) break;
        const int gi = oz_group_of(G, u);
        const OzArgs &a = G.args[gi];
        const int lu = u - G.prefix[gi];
        const int gx = (a.np + 63) / 64;
        const int unit = lu / gx;
        synth_oz_body(a, lu % gx, unit % a.nm, unit / a.nm, abch, bch);
    }
}

// Persistent analysis: unit u of group gi is (pair half, segment, field group).
__global__ __launch_bounds__(256) void k_anal_oz_grp(const OzGroups G) {
    __shared__ __align__(16) unsigned char Alds[2 * OZ_SLICES * 16 * OZ_ROWB];
    __shared__ int s_unit;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_unit = atomicAdd(G.counter, 1);
        __syncthreads();
        const int u = s_unit;
        if (u >= G.prefix[G.ngroups]) break;
        const int gi = oz_group_of(G, u);
        const OzArgs &a = G.args[gi];
        const int lu = u - G.prefix[gi];
        anal_oz_body(a, a.asc, a.qc, lu % a.nhalf, (lu / a.nhalf) % a.nseg_total, lu / (a.nhalf * a.nseg_total),
                     Alds);
    }
}

}  // namespace

static_assert(sizeof(OzGroups) <= 4000, "grouped kernel arguments must fit the kernel argument segment");

hipError_t launch_oz_synth_grp(const OzGroups &bexp, const OzGroups &bdig, const OzGroups &units, int in_f32,
                               int persistent_blocks, hipStream_t st) {
    const unsigned nb1 = bexp.prefix[bexp.ngroups], nb2 = bdig.prefix[bdig.ngroups];
    if (in_f32) {
        hipLaunchKernelGGL(k_oz_bexp_grp<float>, dim3(nb1), dim3(256), 0, st, bexp);
        hipLaunchKernelGGL(k_oz_bdig_grp<float>, dim3(nb2), dim3(256), 0, st, bdig);
    } else {
        hipLaunchKernelGGL(k_oz_bexp_grp<double>, dim3(nb1), dim3(256), 0, st, bexp);
        hipLaunchKernelGGL(k_oz_bdig_grp<double>, dim3(nb2), dim3(256), 0, st, bdig);
    }
    hipLaunchKernelGGL(k_oz_reset, dim3(1), dim3(64), 0, st, units.counter);
    hipLaunchKernelGGL(k_synth_oz_grp, dim3(persistent_blocks), dim3(256), 0, st, units);
    return hipGetLastError();
}

hipError_t launch_oz_anal_grp(const OzGroups &gexp, const OzGroups &gdig, const OzGroups &units,
                              int persistent_blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_oz_gexp_grp, dim3((unsigned)gexp.prefix[gexp.ngroups]), dim3(256), 0, st, gexp);
    hipLaunchKernelGGL(k_oz_gdig_grp, dim3((unsigned)gdig.prefix[gdig.ngroups]), dim3(256), 0, st, gdig);
    hipLaunchKernelGGL(k_oz_reset, dim3(1), dim3(64), 0, st, units.counter);
    hipLaunchKernelGGL(k_anal_oz_grp, dim3(persistent_blocks), dim3(256), 0, st, units);
    return hipGetLastError();
}

hipError_t launch_oz_asc(const OzAscArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_oz_asc, dim3((a.n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_oz_anal(const OzArgs &a, hipStream_t st) {
    if (a.nks > OZ_MAX_KS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_oz_gexp, dim3((unsigned)a.nm * a.nfg), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_oz_gdig, dim3((unsigned)a.nks * a.nm * a.nfg), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_anal_oz, dim3((unsigned)a.nhalf * a.nseg_total * a.nfg), dim3(256), 0, st, a, a.asc, a.qc);
    return hipGetLastError();
}

hipError_t launch_oz_seed(const OzSeedArgs &a, hipStream_t st) {
    dim3 grid((a.np + 255) / 256, a.nm);
    hipLaunchKernelGGL(k_oz_seed, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_oz_synth(const OzArgs &a, hipStream_t st) {
    const dim3 g1((unsigned)a.nm * a.nfg), g2((unsigned)a.ch_total * a.nfg);
    if (a.in_f32) {
        hipLaunchKernelGGL(k_oz_bexp<float>, g1, dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_oz_bdig<float>, g2, dim3(256), 0, st, a);
    } else {
        hipLaunchKernelGGL(k_oz_bexp<double>, g1, dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_oz_bdig<double>, g2, dim3(256), 0, st, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int gx = (a.np + 63) / 64;
    hipLaunchKernelGGL(k_synth_oz, dim3(grid_for_units(a.xcd, gx, a.nm * a.nfg)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace shtk

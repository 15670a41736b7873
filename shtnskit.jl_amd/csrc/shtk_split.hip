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
#include <type_traits>

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
// 2^52 + 0x808080808080: the same for balanced base-256 digits (see k_oz_gprep)
constexpr double kMagicB = 0x1p52 + 141289400074368.0;

__device__ __forceinline__ int oz_nch(int lmax, int m) { return (lmax - m + 128) / 128; }

// Read-only tables read through the constant address space: wave-uniform reads
// become scalar loads (counted apart from the vector memory operations) even in
// kernels that also write global memory, so waiting for them never waits for
// the kernel's own output stores and atomics.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}

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

// byte b of v0..v3 packed into one word (v0 in the low byte)
__device__ __forceinline__ unsigned pack4(unsigned v0, unsigned v1, unsigned v2, unsigned v3, int b) {
    const unsigned sel = (unsigned)b | ((unsigned)(b + 4) << 8) | 0x0c0c0000u;
    const unsigned t01 = __builtin_amdgcn_perm(v1, v0, sel);
    const unsigned t23 = __builtin_amdgcn_perm(v3, v2, sel);
    return __builtin_amdgcn_perm(t23, t01, 0x05040100u);
}

// ---------------------------------------------------------------------------
// Field operand of the synthesis.  k_oz_bexp: one block per (order, field
// group); wave w reads the coefficient columns of fields 4w..4w+3 with lanes
// over consecutive degrees, finds the column exponents eb (|a| < 2^eb over the
// order's degrees) and clears the column sums.  k_oz_bdig: one block per
// (field group, chunk) stages the chunk's 16 fields x 4 quarters x 32 degrees
// in LDS with coalesced reads, then writes one 4-byte word of each of the 6
// digit planes per thread-unit, in the B-operand layout
//   [chunk][parity][slice t][column tile ct][lane][16 bytes],
// lane = c + 16 g (column c of the tile, quarter g), byte j <-> degree
// l = m + g Kq + 32 chunk + 2 j + parity, and adds its integer column sums.
// Digits come from the kMagicB FMA as in the analysis operand.
// ---------------------------------------------------------------------------
template <typename IT>
__device__ __forceinline__ void oz_bexp_body(const OzArgs &a, int im, int fg) {
    const int m = im * a.mres;
    const int lend = min(a.lmax, a.ltr);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int fl = 4 * w + k, f = fg * OZ_FIELDS + fl;
        double mr = 0.0, mi = 0.0;
        if (f < a.nf) {
            // one (re, im) vector load per coefficient, four in flight per lane
            using C2 = typename std::conditional<sizeof(IT) == 8, double2, float2>::type;
            const C2 *src = (const C2 *)a.spec_in + (size_t)f * a.spec_fstride + (size_t)(a.lmax + 1) * m;
            int l = m + lane;
            for (; l + 192 <= lend; l += 256) {
                C2 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = src[l + 64 * u];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    mr = fmax(mr, fabs((double)v[u].x));
                    mi = fmax(mi, fabs((double)v[u].y));
                }
            }
            for (; l <= lend; l += 64) {
                const C2 v = src[l];
                mr = fmax(mr, fabs((double)v.x));
                mi = fmax(mi, fabs((double)v.y));
            }
        }
#pragma unroll
        for (int msk = 32; msk >= 1; msk >>= 1) {
            mr = fmax(mr, __shfl_xor(mr, msk));
            mi = fmax(mi, __shfl_xor(mi, msk));
        }
        if (lane < 2) {
            const double v = lane ? mi : mr;
            a.eb[((size_t)fg * a.nm + im) * OZ_COLS + 2 * fl + lane] = v > 0.0 ? ilogb(v) + 1 : 0;
        }
    }
    if (tid < 2 * OZ_COLS) a.colsum[((size_t)fg * a.nm + im) * 2 * OZ_COLS + tid] = 0ull;
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bexp(const OzArgs a) {
    oz_bexp_body<IT>(a, blockIdx.x % a.nm, blockIdx.x / a.nm);
}

constexpr int OZ_BSTG = 129;  // LDS stride of one field's 4 x 32 staged values (+1 against bank conflicts)

template <typename IT>
__device__ __forceinline__ void oz_bdig_body(const OzArgs &a, int gc, int fg, double2 *stg,
                                             unsigned long long (*csum)[OZ_COLS]) {
    const int im = a.chm[gc], rho = gc - a.chbase[im];
    const int m = im * a.mres, lmax = a.lmax;
    const int Kq = 32 * oz_nch(lmax, m);
    const int lend = min(lmax, a.ltr);
    const int tid = threadIdx.x;
    if (tid < 2 * OZ_COLS) csum[tid >> 5][tid & 31] = 0ull;
    // stage: element e = (field, quarter, degree offset) with the degree fastest
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int e = tid + 256 * k;
        const int l32 = e & 31, g = (e >> 5) & 3, fl = e >> 7;
        const int f = fg * OZ_FIELDS + fl;
        const int l = m + g * Kq + 32 * rho + l32;
        double2 v = make_double2(0.0, 0.0);
        if (f < a.nf && l <= lend) {
            using C2 = typename std::conditional<sizeof(IT) == 8, double2, float2>::type;
            const C2 c = ((const C2 *)a.spec_in)[(size_t)f * a.spec_fstride + (size_t)(lmax + 1) * m + l];
            v = make_double2((double)c.x, (double)c.y);
        }
        stg[fl * OZ_BSTG + g * 32 + l32] = v;
    }
    __syncthreads();
    unsigned *dst = (unsigned *)(a.bdig + ((size_t)fg * a.ch_total + gc) * OZ_CHUNK_V4);
    const int *ebm = a.eb + ((size_t)fg * a.nm + im) * OZ_COLS;
    for (int u = tid; u < 1024; u += 256) {
        const int d = u & 3, lane = (u >> 2) & 63, ct = (u >> 8) & 1, par = u >> 9;
        const int g = lane >> 4, c = (lane & 15) + 16 * ct;
        const double sc = ldexp(1.0, 46 - ebm[c]);
        const double *sv = (const double *)(stg + (c >> 1) * OZ_BSTG + g * 32) + (c & 1);
        unsigned hw[4], lw[4];
        double ys = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l32 = 2 * (4 * d + q) + par;
            const double r = fma(sv[2 * l32], sc, kMagicB);
            ys += r - kMagicB;
            const unsigned long long bits = (unsigned long long)__double_as_longlong(r);
            hw[q] = (unsigned)(bits >> 32) ^ 0x8080u;
            lw[q] = (unsigned)bits ^ 0x80808080u;
        }
        atomicAdd(&csum[par][c], (unsigned long long)(long long)ys);
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t)
            dst[((size_t)((par * OZ_SLICES + t) * 2 + ct) * 64 + lane) * 4 + d] =
                t < 2 ? pack4(hw[0], hw[1], hw[2], hw[3], 1 - t) : pack4(lw[0], lw[1], lw[2], lw[3], 5 - t);
    }
    __syncthreads();
    if (tid < 2 * OZ_COLS)
        atomicAdd(a.colsum + ((size_t)fg * a.nm + im) * 2 * OZ_COLS + tid, csum[tid >> 5][tid & 31]);
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bdig(const OzArgs a) {
    __shared__ double2 stg[OZ_FIELDS * OZ_BSTG];
    __shared__ unsigned long long csum[2][OZ_COLS];
    oz_bdig_body<IT>(a, blockIdx.x % a.ch_total, blockIdx.x / a.ch_total, stg, csum);
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
        if (tid < 128) abch[(tid >> 5) * OZ_ABS + (tid & 31)] = abnext;
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
                const double2 co = abch[g * OZ_ABS + i];
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
    __shared__ double2 abch[4 * OZ_ABS];
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
// OZ_TPAD entries past the end of both tables are zero (the analysis reads
// whole 8- and 32-degree blocks unclamped).
__global__ __launch_bounds__(256) void k_oz_asc(const OzAscArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n + OZ_TPAD) return;
    if (i >= a.n) {
        a.epl[i] = 0;
        a.asc[i] = 0.0;
        return;
    }
    const int e = a.emaxl[i];
    const int ep = e < -201 ? -200 : min(e + 1, 64);
    a.epl[i] = ep;
    a.asc[i] = a.qc[i].y * ldexp(1.0, 46 - ep);
}

// ---------------------------------------------------------------------------
// Analysis field operand.  G+- = w (F_N +- F_S) per pair (even degrees use G+,
// odd G-), written as balanced signed base-256 digits in the B-operand layout
//   [kstep][parity][slice t][column tile ct][lane][16 bytes]
// with one column exponent per (pair block, parity, column) and the integer
// column sums of every kstep.  A pair block is the 256 NH latitude pairs that
// one analysis workgroup contracts (NH = 1 or 2 halves of 256 pairs).  Within
// a block, K slot q (kstep q >> 6, lane group (q >> 4) & 3, byte q & 15) holds
// local pair q: the analysis waves write the digit bytes of 4 consecutive
// pairs as one dword (lane quads exchange bytes through DPP).
//
// k_oz_gprep: one workgroup of 256 NH threads per (pair block, order, field
// group), taking the group's fields in four rounds of four (their north and
// south rows staged in LDS with coalesced reads, the next round's reads issued
// before the current round's digits).  In a round, thread (field, parity,
// re/im, jj = tid % 16 NH) holds the 16 slots of one B lane word of one column
// (kstep jj >> 2 of the block, lane group jj & 3), so each digit word is
// written once from registers; the column maximum is a shuffle reduction over
// the 16 NH threads of a column and the kstep column sums one over 4 threads.
//
// Digits: with kMagicB = 2^52 + 0x808080808080, fma(v, 2^(46-e), kMagicB)
// leaves Y + 0x808080808080 in the low 48 mantissa bits (Y = round(v 2^(46-e)),
// |Y| <= 2^46); its bytes are b_k + 128 for the balanced digits b_k of Y, so a
// XOR with 0x80 per byte gives the int8 digits, and Y itself is the exact
// difference of the FMA result and kMagicB.
// ---------------------------------------------------------------------------
constexpr int OZ_MAX_KS = 32;  // np <= 2048

template <int NH>
__device__ __forceinline__ int oz_pair_of_slot(int q) {
    return q;
}

// LDS of the analysis A digits: [parity][slice][row 16][256 NH pair bytes + 16],
// slices padded by 160 bytes (40 dwords): the four slices a lane quad writes at
// once then start 8 banks apart (modulo 32) and 24 apart (modulo 64), so the
// 16 quads x 4 slices of a wave's digit store hit every bank once, and the two
// high-word slices of the parity pair likewise.
template <int NH>
struct OzLds {
    static constexpr int ROWB = 256 * NH + 16;          // row stride (bytes)
    static constexpr int SLS = 16 * ROWB + 160;         // slice stride
    static constexpr int PS = OZ_SLICES * SLS;          // parity stride
    static constexpr int BYTES = 2 * PS;
};

// 4x4 byte transpose inside each lane quad: lane q ends with byte q of the
// words of lanes 4j .. 4j+3 (in lane order), in two halves, each a DPP
// exchange (lane ^ 1, then lane ^ 2) and a byte permute.
__device__ __forceinline__ unsigned qbt_first(unsigned v, unsigned sel1) {
    const unsigned x = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    return __builtin_amdgcn_perm(x, v, sel1);
}
__device__ __forceinline__ unsigned qbt_second(unsigned b, unsigned sel2) {
    const unsigned y = (unsigned)__builtin_amdgcn_mov_dpp((int)b, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    return __builtin_amdgcn_perm(y, b, sel2);
}

// LDS staging of the Fourier rows of 4 fields of a pair block (north, south),
// loaded with consecutive threads on consecutive latitudes, plus the block's
// latitude weights.  Pair pl sits at gp_slot(pl) = pl + pl / 16: the threads of
// a column read pairs 16 apart, which the padding spreads over all banks.
__device__ __forceinline__ int gp_slot(int pl) { return pl + (pl >> 4); }

template <int NH>
struct GprepLds {
    static constexpr int ROW = 256 * NH + 16 * NH;
    double2 F[4][2][ROW];
    double w[ROW];
};

template <int NH>
__device__ __forceinline__ void oz_gprep_body(const OzArgs &a, int pb, int im, int fg, GprepLds<NH> &L) {
    constexpr int TPF = 16 * NH, NP = 256 * NH;  // threads per column, pairs per block
    // four rounds of four fields; in each, thread (field fr of the round, parity
    // par, re/im ri, word jj) builds one B lane word of one column, while the
    // next round's rows are in flight
    const int tid = threadIdx.x, jj = tid % TPF, combo = tid / TPF;
    const int fr = combo >> 2, par = (combo >> 1) & 1, ri = combo & 1;
    const int ksl = jj >> 2, g = jj & 3;
    const int ks = pb * 4 * NH + ksl;
    const size_t ubase = ((size_t)fg * a.nm + im) * a.nks;
    v4i *dst = (v4i *)a.gdig + (ubase + ks) * OZ_CHUNK_V4;
    {
        const int p = pb * NP + tid;
        L.w[gp_slot(tid)] = p < a.np ? a.wlat[p] : 0.0;
    }
    double2 ld[8];
    auto issue = [&](int rd) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + NP * i, fe = e / (2 * NP), ns = (e / NP) & 1, pl = e % NP;
            const int ff = fg * OZ_FIELDS + 4 * rd + fe, p = pb * NP + pl;
            double2 val = make_double2(0.0, 0.0);
            if (ff < a.nf && p < a.np) {
                const int ps = a.nlat - 1 - p;
                if (!ns || ps != p) val = a.F0[((size_t)ff * a.nm + im) * a.nlat + (ns ? ps : p)];
            }
            ld[i] = val;
        }
    };
    issue(0);
    for (int rd = 0; rd < 4; ++rd) {
        __syncthreads();  // the previous round's rows are consumed
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + NP * i;
            L.F[e / (2 * NP)][(e / NP) & 1][gp_slot(e % NP)] = ld[i];
        }
        __syncthreads();
        if (rd < 3) issue(rd + 1);
        // G+- = w (F_N +- F_S) of 16 consecutive pairs: word d, byte j = pair 4 d + j
        double v[16], mx = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int pl = gp_slot(oz_pair_of_slot<NH>(64 * ksl + 16 * g + 4 * d + j));
                const double2 FN = L.F[fr][0][pl], FS = L.F[fr][1][pl];
                const double fn = ri ? FN.y : FN.x, fs = ri ? FS.y : FS.x;
                const double x = L.w[pl] * (par ? fn - fs : fn + fs);
                v[4 * j + d] = x;
                mx = fmax(mx, fabs(x));
            }
#pragma unroll
        for (int msk = 1; msk < TPF; msk <<= 1) mx = fmax(mx, __shfl_xor(mx, msk));
        const int e = mx > 0.0 ? ilogb(mx) + 1 : 0;
        const int c = 2 * (4 * rd + fr) + ri;
        if (jj == 0) a.eg[((((size_t)fg * a.nm + im) * a.npb + pb) * 2 + par) * OZ_COLS + c] = e;
        const double sc = ldexp(1.0, 46 - e);
        unsigned hw[16], lw[16];
        double ys = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double r = fma(v[i], sc, kMagicB);
            ys += r - kMagicB;
            const unsigned long long bits = (unsigned long long)__double_as_longlong(r);
            hw[i] = (unsigned)(bits >> 32) ^ 0x8080u;
            lw[i] = (unsigned)bits ^ 0x80808080u;
        }
        ys += __shfl_xor(ys, 1);
        ys += __shfl_xor(ys, 2);
        if (g == 0) a.gcolsum[((ubase + ks) * 2 + par) * OZ_COLS + c] = ys;
        // word d of the lane holds bytes j = 0..3 (j in byte j)
        v4i *row = dst + ((par * OZ_SLICES) * 2 + (c >> 4)) * 64 + (c & 15) + 16 * g;
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t) {
            v4i wd;
#pragma unroll
            for (int d = 0; d < 4; ++d)
                wd[d] = (int)(t < 2 ? pack4(hw[d], hw[4 + d], hw[8 + d], hw[12 + d], 1 - t)
                                    : pack4(lw[d], lw[4 + d], lw[8 + d], lw[12 + d], 5 - t));
            row[t * 2 * 64] = wd;
        }
    }
}

template <int NH>
__global__ __launch_bounds__(256 * NH) void k_oz_gprep(const OzArgs a) {
    __shared__ GprepLds<NH> L;
    oz_gprep_body<NH>(a, blockIdx.x % a.npb, (blockIdx.x / a.npb) % a.nm, blockIdx.x / (a.npb * a.nm), L);
}

// ---------------------------------------------------------------------------
// Analysis: one workgroup of 4 NH waves per (pair block, run of degree
// segments, field group).  Every lane runs the Q-form recurrence of one pair
// and writes the 6 digit bytes of each degree to LDS as
// [parity][slice][row][pair] (row = degree pair index within 32 degrees);
// wave w then contracts parity (w >> 1) & 1, column tile w & 1 over the K half
// w >> 2 of the block (the field operand of its 4 ksteps stays in registers
// for the whole run).  With two halves the waves of a parity / column tile
// exchange the integer sums of half their rows through LDS, so each output
// has one writer: plain stores when one block covers all pairs, atomics only
// when several pair blocks add to it (np > 512).
// ---------------------------------------------------------------------------
// Degree I of the chunk: digits of P_l 2^(46-epl) into the LDS rows, the Q-form
// step, and after every 8th degree the scale check of a chain still in the
// scaled range.  The quad byte transpose of a digit word is spread over the two
// following degrees (its first DPP exchange one degree after the word is
// formed, its second one degree later), so no DPP read follows the VALU write
// of its source directly and the wave needs no wait states for them; I = 32
// drains the last two degrees.
template <int I, int NH>
__device__ __forceinline__ void oz_digit_steps(const OzRec &R, double x, double &prev, double &cur, int &kk,
                                               double &live, unsigned char *aul, unsigned char *auh, unsigned sel1,
                                               unsigned sel2, unsigned &uh_even, unsigned &w1, unsigned &w1h,
                                               unsigned &b2, unsigned &b2h) {
    constexpr int ROWB = OzLds<NH>::ROWB, PS = OzLds<NH>::PS;
    // second half of the transposes of degree I - 2's words: store them
    if constexpr (I >= 2) {
        constexpr int D = I - 2;
        *(unsigned *)(aul + (D & 1) * PS + (D >> 1) * ROWB) = qbt_second(b2, sel2) ^ 0x80808080u;
        if constexpr (D & 1) *(unsigned *)(auh + (D >> 1) * ROWB) = qbt_second(b2h, sel2) ^ 0x80808080u;
    }
    // first half of the transposes of degree I - 1's words
    if constexpr (I >= 1 && I <= 32) {
        constexpr int D = I - 1;
        b2 = qbt_first(w1, sel1);
        if constexpr (D & 1) b2h = qbt_first(w1h, sel1);
    }
    if constexpr (I < 32) {
        constexpr int H = I >> 4, N = I & 15;
        // U = 2^47 + round(P 2^(46-epl)) lands in the low 48 bits of the mantissa
        double dm = kMagic;
        const double cl = cur * live;
        SHTK_FMAC_BC(dm, R.a[H], cl, N, "");
        const unsigned long long bits = (unsigned long long)__double_as_longlong(dm);
        const unsigned uh = (unsigned)(bits >> 32);
        // slices 2..5 (bytes 3..0 of the low word); the XOR 0x80 per byte at the
        // store gives balanced digits
        w1 = (unsigned)bits;
        if constexpr (I & 1) {
            // slices 0, 1 (bytes 1, 0 of the high word) of degrees I - 1 and I
            w1h = __builtin_amdgcn_perm(uh, uh_even, 0x05040100u);
        } else {
            uh_even = uh;
        }
        double nx = x * cur;
        SHTK_FMAC_BC(nx, R.b[H], prev, N, "-");
        prev = cur;
        cur = nx;
        if constexpr ((I & 7) == 7) {
            const double cq = readlane_f64(R.c[H], N);  // c_l of this degree
            if (kk > 0 && fmax(fabs(cur), fabs(prev)) * cq > kHiS) {
                cur *= kDownS;
                prev *= kDownS;
                --kk;
                live = kk == 0 ? 1.0 : 0.0;
            }
        }
        oz_digit_steps<I + 1, NH>(R, x, prev, cur, kk, live, aul, auh, sel1, sel2, uh_even, w1, w1h, b2, b2h);
    } else if constexpr (I == 32) {
        // degree 31: its second half right after the first
        oz_digit_steps<33, NH>(R, x, prev, cur, kk, live, aul, auh, sel1, sel2, uh_even, w1, w1h, b2, b2h);
    }
}

template <int NH>
__device__ __forceinline__ void anal_oz_body(const OzArgs &a, const double *__restrict__ asc,
                                             const double2 *__restrict__ qc, const int *__restrict__ epl,
                                             const int *__restrict__ pfirst, int pb, int run, int fg,
                                             unsigned char *Alds, int *Xch) {
    constexpr int ROWB = OzLds<NH>::ROWB, SLS = OzLds<NH>::SLS, PS = OzLds<NH>::PS, RPW = NH == 2 ? 2 : 4;
    const cptr<int> ru = as_const((const int *)a.run) + 4 * run;
    const int im = ru[0], gs0 = ru[1], nsr = ru[2], m = im * a.mres;
    const int s0 = gs0 - as_const(a.segbase)[im];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int pbase = pb * 256 * NH;
    const int p = pbase + tid;
    const bool act = p < a.np;
    const double x = act ? a.x[p] : 0.0;
    const long long co = as_const(a.coff)[im] - m;
    // per-degree tables of the order (padded past lmax + 1), read per lane into OzRec
    const double *ascl = asc + co, *qbl = a.qb + co, *qccl = a.qcc + co;
    const cptr<int> eplm = as_const(epl) + co;
    const cptr<int> pfc = as_const(pfirst);

    // contraction role of this wave: K half kh, parity wp, column tile wct
    const int kh = w >> 2, wp = (w >> 1) & 1, wct = w & 1;
    v4i Bw[4][OZ_SLICES];
    const size_t ubase = ((size_t)fg * a.nm + im) * a.nks;
    const int kb = pb * 4 * NH;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int t = 0; t < OZ_SLICES; ++t)
            Bw[ks][t] = ((const v4i *)a.gdig)[(ubase + kb + 4 * kh + ks) * OZ_CHUNK_V4 +
                                              ((wp * OZ_SLICES + t) * 2 + wct) * 64 + lane];
    const int c = (lane & 15) + 16 * wct;
    double csum = 0.0;
    for (int k = 0; k < 4 * NH; ++k) csum += a.gcolsum[((ubase + kb + k) * 2 + wp) * OZ_COLS + c];
    const int egc = a.eg[((((size_t)fg * a.nm + im) * a.npb + pb) * 2 + wp) * OZ_COLS + c];
    const int f = fg * OZ_FIELDS + (c >> 1);
    const bool fok = f < a.nf;
    double *outc = a.spec_out + 2 * ((size_t)(fok ? f : 0) * a.spec_fstride + (size_t)(a.lmax + 1) * m) + (c & 1);
    const bool atom = a.npb > 1;
    // digit stores: lane q of a quad writes slice 5 - q of the low word for the
    // quad's 4 pairs, and (every other degree) slice 1 - (q & 1) of degree
    // pair member q >> 1 of the high words
    const int q4 = lane & 3;
    const unsigned sel1 = (q4 & 1) ? 0x03070105u : 0x06020400u;
    const unsigned sel2 = (q4 & 2) ? 0x03020706u : 0x05040100u;
    const int dw = 4 * (16 * w + (lane >> 2));  // byte offset of the quad's dword in a row
    unsigned char *aul = Alds + (5 - q4) * SLS + dw;
    unsigned char *auh = Alds + (1 - (q4 & 1)) * SLS + (q4 >> 1) * PS + dw;

    // Segments of the run in which every pair of this block is still in the
    // scaled range contribute nothing (pfirst does not increase along an
    // order); the recurrence starts from the seeds of the first live segment
    // and continues through the rest of the run: at a segment boundary L the
    // Q-form of the next segment is the current one times c_{L-1} (the last
    // scale of the segment ending), so no further per-lane loads are needed.
    int sr0 = 0;
    while (sr0 < nsr && pbase + 256 * NH - 1 < pfc[gs0 + sr0]) ++sr0;
    if (!atom) oz_zero_holes(a, im, s0, nsr, m + (s0 + sr0) * a.seg, fg);
    if (sr0 == nsr) return;
    double prev, cur;
    int kk;
    {
        const size_t q = (size_t)(gs0 + sr0) * a.np + (act ? p : 0);
        prev = act ? a.seedq0[q] : 0.0;
        cur = act ? a.seedq1[q] : 0.0;
        kk = act ? a.seedk[q] : 0;
    }
    double live = kk == 0 ? 1.0 : 0.0;  // lanes still in the scaled range contribute zeros
    const int rn = lane & 15;
    auto load_rec = [&](OzRec &r, int lc) {
        r.a[0] = ascl[lc + rn];
        r.a[1] = ascl[lc + 16 + rn];
        r.b[0] = qbl[lc + rn + 1];
        r.b[1] = qbl[lc + 16 + rn + 1];
        r.c[0] = qccl[lc + rn];
        r.c[1] = qccl[lc + 16 + rn];
    };
    OzRec nrec;
    load_rec(nrec, m + (s0 + sr0) * a.seg);
    const int wpu = __builtin_amdgcn_readfirstlane(wp);
    const int r0u = __builtin_amdgcn_readfirstlane(NH == 2 ? 2 * kh : 0);  // first row index i of this wave
    const int grp = lane >> 4;
    for (int sr = sr0; sr < nsr; ++sr) {
        const int l0 = m + (s0 + sr) * a.seg;
        const int lstop = min(min(l0 + a.seg, a.lmax + 1), a.ltr + 1);
        const int nst = lstop - l0;
        if (nst <= 0) break;
        if (sr > sr0) {
            if ((a.seg & 31) == 0) {
                const double cb = as_const(qc)[co + l0 - 1].y;
                prev *= cb;
                cur *= cb;
            } else {
                // segments that end inside a 32-degree chunk ran the recurrence
                // past their end: start the next one from its seeds
                const size_t q = (size_t)(gs0 + sr) * a.np + (act ? p : 0);
                prev = act ? a.seedq0[q] : 0.0;
                cur = act ? a.seedq1[q] : 0.0;
                kk = act ? a.seedk[q] : 0;
                live = kk == 0 ? 1.0 : 0.0;
            }
        }
        for (int t0 = 0; t0 < nst; t0 += 32) {
            const int lc = __builtin_amdgcn_readfirstlane(l0 + t0);
            // per-degree exponents of the rows 4 g + r0 + r this wave emits; rows
            // past the segment read valid (padded) table entries and are never stored
            int ev[4][RPW];
#pragma unroll
            for (int gg = 0; gg < 4; ++gg)
#pragma unroll
                for (int r = 0; r < RPW; ++r) ev[gg][r] = eplm[lc + wpu + 2 * (4 * gg + r0u + r)];
            // Recurrence and digits of 32 degrees from this chunk's records; the
            // next chunk's (this segment's next, or the next segment's first) load
            // meanwhile.
            const OzRec rec = nrec;
            // (clamped to lmax: reads stay inside the order's padded tables; a chunk
            // that does not exist never uses them)
            load_rec(nrec, min(t0 + 32 < nst ? lc + 32 : m + (s0 + sr + 1) * a.seg, a.lmax));
            asm volatile("s_nop 4");  // wait states before DPP reads the refilled records
            unsigned uh_even = 0, w1 = 0, w1h = 0, b2 = 0, b2h = 0;
            oz_digit_steps<0, NH>(rec, x, prev, cur, kk, live, aul, auh, sel1, sel2, uh_even, w1, w1h, b2, b2h);
            __syncthreads();
            v4i acc[OZ_SLICES];
#pragma unroll
            for (int v = 0; v < OZ_SLICES; ++v) acc[v] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                v4i Aop[OZ_SLICES];
#pragma unroll
                for (int sl = 0; sl < OZ_SLICES; ++sl)
                    Aop[sl] = *(const v4i *)(Alds + wp * PS + sl * SLS + (lane & 15) * ROWB +
                                             64 * (4 * kh + ks) + 16 * (lane >> 4));
#pragma unroll
                for (int t = 0; t < OZ_SLICES; ++t)
#pragma unroll
                    for (int sl = 0; sl + t < OZ_SLICES; ++sl)
                        acc[sl + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aop[sl], Bw[ks][t], acc[sl + t], 0, 0, 0);
            }
            if constexpr (NH == 2) {
                // rows i = 2 kh, 2 kh + 1 stay with this wave; its sums of the other
                // two go to the partner of the other K half (wave w ^ 4)
                int *xo = Xch + w * (2 * OZ_SLICES * 64) + lane;
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int v = 0; v < OZ_SLICES; ++v) xo[(r * OZ_SLICES + v) * 64] = acc[v][2 * (1 - kh) + r];
            }
            __syncthreads();
            // rows 4 (lane >> 4) + i are degree pairs, column c
            const int *xi = Xch + (w ^ 4) * (2 * OZ_SLICES * 64) + lane;
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int i = r0u + r;
                const int row = 4 * grp + i;
                const int t = t0 + 2 * row + wp;
                const int e = grp == 0 ? ev[0][r] : grp == 1 ? ev[1][r] : grp == 2 ? ev[2][r] : ev[3][r];
                if (t < nst && fok) {
                    double E = kBias * csum;
#pragma unroll
                    for (int v = 0; v < OZ_SLICES; ++v) {
                        int s = acc[v][NH == 2 ? 2 * kh + r : r];
                        if constexpr (NH == 2) s += xi[(r * OZ_SLICES + v) * 64];
                        E = fma((double)s, ldexp(1.0, 80 - 8 * v), E);
                    }
                    const double val = E * a.g * ldexp(1.0, e - 46 + egc - 46);
                    double *dst = outc + 2 * (l0 + t);
                    if (atom)
                        atomicAdd(dst, val);
                    else
                        *dst = val;
                }
            }
        }
    }
}

template <int NH>
__global__ __launch_bounds__(256 * NH) void k_anal_oz(const OzArgs a, const double *__restrict__ asc,
                                                       const double2 *__restrict__ qc, const int *__restrict__ epl,
                                                       const int *__restrict__ pfirst) {
    // rows padded by 16 bytes: the 16 lanes of an A-operand read hit distinct banks
    __shared__ __align__(16) unsigned char Alds[OzLds<NH>::BYTES];
    __shared__ int Xch[NH == 2 ? 8 * 2 * OZ_SLICES * 64 : 1];
    const int item = blockIdx.x;
    anal_oz_body<NH>(a, asc, qc, epl, pfirst, item % a.npb, (item / a.npb) % a.nrun, item / (a.npb * a.nrun), Alds,
                     Xch);
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
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    oz_bexp_body<IT>(a, b % a.nm, b / a.nm);
}

template <typename IT>
__global__ __launch_bounds__(256) void k_oz_bdig_grp(const OzGroups G) {
    __shared__ double2 stg[OZ_FIELDS * OZ_BSTG];
    __shared__ unsigned long long csum[2][OZ_COLS];
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    oz_bdig_body<IT>(a, b % a.ch_total, b / a.ch_total, stg, csum);
}

template <int NH>
__global__ __launch_bounds__(256 * NH) void k_oz_gprep_grp(const OzGroups G) {
    const int gi = oz_group_of(G, blockIdx.x);
    const OzArgs &a = G.args[gi];
    const int b = blockIdx.x - G.prefix[gi];
    __shared__ GprepLds<NH> L;
    oz_gprep_body<NH>(a, b % a.npb, (b / a.npb) % a.nm, b / (a.npb * a.nm), L);
}

__global__ void k_oz_reset(int *counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

// Persistent synthesis: unit u of group gi is (pair block, order, field group)
// with the pair block fastest, as the single-configuration grid.
__global__ __launch_bounds__(256) void k_synth_oz_grp(const OzGroups G) {
    __shared__ double2 abch[4 * OZ_ABS];
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

// Persistent analysis: unit u of group gi is (pair block, run, field group);
// every group of one launch has the same pair-block size NH.
template <int NH>
__global__ __launch_bounds__(256 * NH) void k_anal_oz_grp(const OzGroups G) {
    __shared__ __align__(16) unsigned char Alds[OzLds<NH>::BYTES];
    __shared__ int Xch[NH == 2 ? 8 * 2 * OZ_SLICES * 64 : 1];
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
        anal_oz_body<NH>(a, a.asc, a.qc, a.epl, a.pfirst, lu % a.npb, (lu / a.npb) % a.nrun, lu / (a.npb * a.nrun),
                         Alds, Xch);
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

hipError_t launch_oz_anal_grp(const OzGroups *gprep, const OzGroups *units, int persistent_blocks, hipStream_t st) {
    // units[0]: groups with two-half pair blocks, units[1]: with one half
    for (int k = 0; k < 2; ++k) {
        if (units[k].ngroups == 0) continue;
        const dim3 g1((unsigned)gprep[k].prefix[gprep[k].ngroups]);
        if (k == 0)
            hipLaunchKernelGGL(k_oz_gprep_grp<2>, g1, dim3(512), 0, st, gprep[k]);
        else
            hipLaunchKernelGGL(k_oz_gprep_grp<1>, g1, dim3(256), 0, st, gprep[k]);
        hipLaunchKernelGGL(k_oz_reset, dim3(1), dim3(64), 0, st, units[k].counter);
        if (k == 0)
            hipLaunchKernelGGL(k_anal_oz_grp<2>, dim3(persistent_blocks), dim3(512), 0, st, units[k]);
        else
            hipLaunchKernelGGL(k_anal_oz_grp<1>, dim3(persistent_blocks), dim3(256), 0, st, units[k]);
    }
    return hipGetLastError();
}

hipError_t launch_oz_asc(const OzAscArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_oz_asc, dim3((a.n + OZ_TPAD + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_oz_anal(const OzArgs &a, hipStream_t st) {
    if (a.nks > OZ_MAX_KS || a.nks != 4 * a.nh * a.npb) return hipErrorInvalidValue;
    const dim3 g1((unsigned)a.npb * a.nm * a.nfg), g2((unsigned)a.npb * a.nrun * a.nfg);
    if (a.nh == 2) {
        hipLaunchKernelGGL(k_oz_gprep<2>, g1, dim3(512), 0, st, a);
        hipLaunchKernelGGL(k_anal_oz<2>, g2, dim3(512), 0, st, a, a.asc, a.qc, a.epl, a.pfirst);
    } else {
        hipLaunchKernelGGL(k_oz_gprep<1>, g1, dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_anal_oz<1>, g2, dim3(256), 0, st, a, a.asc, a.qc, a.epl, a.pfirst);
    }
    return hipGetLastError();
}

hipError_t launch_oz_seed(const OzSeedArgs &a, hipStream_t st) {
    dim3 grid((a.np + 255) / 256, a.nm);
    hipLaunchKernelGGL(k_oz_seed, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

namespace {

// Vector (spheroidal/toroidal) transforms through the scalar split contraction.
// With d_l = sqrt((l^2 - m^2)/(4 l^2 - 1)) the orthonormal rows satisfy
//     sin(theta) dP_l/dtheta = l d_{l+1} P_{l+1} - (l+1) d_l P_{l-1},
// so sin(theta) V_theta and sin(theta) V_phi of src/kernels.jl:155-169 are
// same-order scalar syntheses to degree lmax + 1:
//     sin V_theta = sum_l P_l [(l-1) d_l S_{l-1} - (l+2) d_{l+1} S_{l+1} - i m T_l]
//     sin V_phi   = sum_l P_l [i m S_l + (l-1) d_l T_{l-1} - (l+2) d_{l+1} T_{l+1}]
// and the analysis of src/kernels.jl:205-215 is two scalar analyses with the
// weights w/sin(theta), G_theta and G_phi (degrees to lmax + 1), followed by
//     S_l = [l d_{l+1} G_theta,l+1 - (l+1) d_l G_theta,l-1 - i m G_phi,l] / (l(l+1))
//     T_l = [i m G_theta,l + l d_{l+1} G_phi,l+1 - (l+1) d_l G_phi,l-1] / (l(l+1)).
__device__ __forceinline__ double vs_d(int l, int m) {
    return l > m ? sqrt((double)(l * l - m * m) / (4.0 * l * l - 1.0)) : 0.0;
}

// Canonical coefficient l of one order, zero outside [lo, hi].
template <class IT>
__device__ __forceinline__ double2 vs_coef(const IT *A, const double *sm, int divide, int l, int lo, int hi) {
    if (l < lo || l > hi) return make_double2(0.0, 0.0);
    double c = 1.0;
    if (sm) c = divide ? 1.0 / sm[l] : sm[l];
    return make_double2(c * (double)A[2 * l], c * (double)A[2 * l + 1]);
}

// (S, T) of nf fields -> coefficients of 2 nf scalar fields of degree lmax + 1
// ([theta fields][phi fields], column stride lmax + 2).  One block per
// (order, field); S = T = 0 for l = 0 and l > ltr (src/kernels.jl:155-169 sums
// l >= max(1, m)).
template <class IT>
__global__ __launch_bounds__(256) void k_vsplit_coef(const IT *S, const IT *T, double *C, const double *scale,
                                                     int divide, int lmax, int mres, int nm, int ltr,
                                                     long long in_fstride, long long out_fstride, int nf) {
    const int im = blockIdx.x % nm, f = blockIdx.x / nm, m = im * mres;
    const size_t col = (size_t)(lmax + 1) * m;
    const IT *Sm = S + 2 * ((size_t)f * in_fstride + col);
    const IT *Tm = T + 2 * ((size_t)f * in_fstride + col);
    const double *sm = scale ? scale + col : nullptr;
    double *Ct = C + 2 * ((size_t)f * out_fstride + (size_t)(lmax + 2) * m);
    double *Cp = C + 2 * ((size_t)(nf + f) * out_fstride + (size_t)(lmax + 2) * m);
    const int lo = max(1, m), hi = min(ltr, lmax);
    const double mm = (double)m;
    for (int l = m + (int)threadIdx.x; l <= lmax + 1; l += 256) {
        const double dm = (l - 1) * vs_d(l, m), dp = (l + 2) * vs_d(l + 1, m);
        const double2 s1 = vs_coef(Sm, sm, divide, l - 1, lo, hi), s0 = vs_coef(Sm, sm, divide, l, lo, hi);
        const double2 s2 = vs_coef(Sm, sm, divide, l + 1, lo, hi);
        const double2 t1 = vs_coef(Tm, sm, divide, l - 1, lo, hi), t0 = vs_coef(Tm, sm, divide, l, lo, hi);
        const double2 t2 = vs_coef(Tm, sm, divide, l + 1, lo, hi);
        *(double2 *)(Ct + 2 * l) = make_double2(dm * s1.x - dp * s2.x + mm * t0.y, dm * s1.y - dp * s2.y - mm * t0.x);
        *(double2 *)(Cp + 2 * l) = make_double2(dm * t1.x - dp * t2.x - mm * s0.y, dm * t1.y - dp * t2.y + mm * s0.x);
    }
}

// G of 2 nf scalar fields (degree lmax + 1) -> (S, T) in the configured
// convention.  One block per (column, field); columns that are not orders and
// rows l < max(1, m) or l > ltr are written as zero.
template <class OT>
__global__ __launch_bounds__(256) void k_vsplit_out(const double *G, OT *S, OT *T, const double *scale, int divide,
                                                    int lmax, int mmax, int mres, int ltr, long long g_fstride,
                                                    long long out_fstride, int nf) {
    const int mc = blockIdx.x % (mmax + 1), f = blockIdx.x / (mmax + 1), m = mc;
    const size_t col = (size_t)(lmax + 1) * mc;
    OT *So = S + 2 * ((size_t)f * out_fstride + col);
    OT *To = T + 2 * ((size_t)f * out_fstride + col);
    const double2 *Gt = (const double2 *)G + (size_t)f * g_fstride + (size_t)(lmax + 2) * mc;
    const double2 *Gp = (const double2 *)G + (size_t)(nf + f) * g_fstride + (size_t)(lmax + 2) * mc;
    const bool order = mc % mres == 0;
    const int lo = max(1, m), hi = min(ltr, lmax);
    const double mm = (double)m;
    for (int l = (int)threadIdx.x; l <= lmax; l += 256) {
        double sr = 0.0, si = 0.0, tr = 0.0, ti = 0.0;
        if (order && l >= lo && l <= hi) {
            const double k = 1.0 / ((double)l * (l + 1));
            const double dp = l * vs_d(l + 1, m), dm = (l + 1) * vs_d(l, m);
            const double2 gt2 = Gt[l + 1], gt0 = Gt[l], gp2 = Gp[l + 1], gp0 = Gp[l];
            const double2 gt1 = l - 1 >= m ? Gt[l - 1] : make_double2(0.0, 0.0);
            const double2 gp1 = l - 1 >= m ? Gp[l - 1] : make_double2(0.0, 0.0);
            double c = k;
            if (scale) c = divide ? k / scale[col + l] : k * scale[col + l];
            sr = c * (dp * gt2.x - dm * gt1.x + mm * gp0.y);
            si = c * (dp * gt2.y - dm * gt1.y - mm * gp0.x);
            tr = c * (dp * gp2.x - dm * gp1.x - mm * gt0.y);
            ti = c * (dp * gp2.y - dm * gp1.y + mm * gt0.x);
        }
        So[2 * l] = (OT)sr;
        So[2 * l + 1] = (OT)si;
        To[2 * l] = (OT)tr;
        To[2 * l + 1] = (OT)ti;
    }
}

}  // namespace

hipError_t launch_vsplit_coef(const void *S, const void *T, double *C, int in_f32, const double *scale, int divide,
                              int lmax, int mres, int nm, int ltr, long long in_fstride, long long out_fstride,
                              int nf, hipStream_t st) {
    const dim3 grid((unsigned)nm * nf);
    if (in_f32)
        hipLaunchKernelGGL(k_vsplit_coef<float>, grid, dim3(256), 0, st, (const float *)S, (const float *)T, C,
                           scale, divide, lmax, mres, nm, ltr, in_fstride, out_fstride, nf);
    else
        hipLaunchKernelGGL(k_vsplit_coef<double>, grid, dim3(256), 0, st, (const double *)S, (const double *)T, C,
                           scale, divide, lmax, mres, nm, ltr, in_fstride, out_fstride, nf);
    return hipGetLastError();
}

hipError_t launch_vsplit_out(const double *G, void *S, void *T, int out_f32, const double *scale, int divide,
                             int lmax, int mmax, int mres, int ltr, long long g_fstride, long long out_fstride,
                             int nf, hipStream_t st) {
    const dim3 grid((unsigned)(mmax + 1) * nf);
    if (out_f32)
        hipLaunchKernelGGL(k_vsplit_out<float>, grid, dim3(256), 0, st, G, (float *)S, (float *)T, scale, divide,
                           lmax, mmax, mres, ltr, g_fstride, out_fstride, nf);
    else
        hipLaunchKernelGGL(k_vsplit_out<double>, grid, dim3(256), 0, st, G, (double *)S, (double *)T, scale,
                           divide, lmax, mmax, mres, ltr, g_fstride, out_fstride, nf);
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

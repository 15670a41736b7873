// Batched Legendre contraction on the FP32 matrix cores of gfx950
// (SHTK_PREC_FP32): FP64 Legendre values, nodes, weights and cross-tile sums,
// the contraction itself as v_mfma_f32_16x16x4_f32 with FP32 accumulation over
// K-tiles of 64 products, each tile's sum added into FP64 accumulators.
//
// The contraction is that of the batch transforms (src/batch_transforms.jl:
// 400-425 for the analysis, 519-535 for the synthesis): one Legendre row per
// (order, latitude) shared by every field.  Results carry FP32-grade error
// (relative L2 ~1e-7 per field); the FP64 kernels and the int8 split
// (SHTK_PREC_SPLIT) give FP64-grade results.
//
// Synthesis: one block of 4 waves per (64 latitude pairs, order, field group of
// 16 fields = 32 real columns).  Wave w owns pairs 16w..16w+15 (MFMA rows); its
// lane group g runs the P-form recurrence over quarter g of the order's degrees
// from plan-time quarter seeds (the tables of the split contraction), so every
// lane produces its own A operand, P * 2^-ep (ep: the pair's exponent) rounded
// to FP32.  One MFMA per degree and column tile sums the 4 quarters at that
// degree offset (K = 4).  The coefficients are staged in LDS per chunk of 32
// degrees per quarter.
//
// Analysis: one workgroup of 4 NH waves per (pair block of 256 NH pairs, run of
// degree segments, field group).  Every lane runs the Q-form recurrence of one
// pair from the segment seeds and writes P_l * 2^-epl (epl: the degree's
// exponent over all pairs) as FP32 to LDS; wave w contracts parity (w >> 1) & 1,
// column tile w & 1 over K half w >> 2 of the block against the field operand
// w (F_N +- F_S) (FP32, built by k_f32_gprep, held in registers for the run).
// With two K halves the partner waves exchange half their rows, so every
// coefficient has one writer (atomics only when several pair blocks cover np > 512).
#include <type_traits>

#include "shtk_internal.h"

namespace shtk {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr double kHiS = 0x1p100;
constexpr double kDownS = 0x1p-512;

template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}

__device__ __forceinline__ int quarter_chunks(int lmax, int m) { return (lmax - m + 128) / 128; }

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Staged coefficients of a chunk: [field][quarter][degree] complex floats with
// strides chosen so that the B-operand reads (16 columns x 4 quarters of one
// degree) and the staging writes (consecutive degrees) hit distinct banks.
constexpr int F32_GS = 40;   // float2 per quarter row (32 degrees + 8)
constexpr int F32_FS = 161;  // float2 per field (4 quarter rows + 1)

template <typename IT>
__device__ __forceinline__ void synth_f32_body(const OzArgs &a, int pbk, int im, int fg, float2 *stg,
                                               double2 *abch) {
    const int m = im * a.mres, lmax = a.lmax;
    const int nch = quarter_chunks(lmax, m), Kq = 32 * nch;
    const int lend = min(lmax, a.ltr);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int p = pbk * 64 + w * 16 + r;
    const bool act = p < a.np;
    const double x = act ? a.x[p] : 0.0;
    const size_t sq = ((size_t)im * 4 + g) * a.np + (act ? p : 0);
    double prev = act ? a.q0[sq] : 0.0, cur = act ? a.q1[sq] : 0.0;
    int kk = act ? a.qk[sq] : 0;
    const double scl = ldexp(1.0, -(act ? a.ep[(size_t)im * a.np + p] : 0));  // |P| 2^-ep < 1
    const double2 *ab = a.ab + a.coff[im] - m;
    using C2 = typename std::conditional<sizeof(IT) == 8, double2, float2>::type;
    const C2 *spec = (const C2 *)a.spec_in + (size_t)(lmax + 1) * m;

    v4f acc[2][2];
    double accd[2][2][4];
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            acc[par][ct] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 4; ++e) accd[par][ct][e] = 0.0;
        }
    const float *stf = (const float *)stg;
    // chunk operands (coefficients, recurrence coefficients) are loaded one chunk
    // ahead into registers and staged in LDS at the chunk start
    C2 nv[8];
    double2 nab = make_double2(0.0, 0.0);
    auto load_chunk = [&](int rho) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = tid + 256 * k;
            const int l32 = e & 31, gg = (e >> 5) & 3, fl = e >> 7;
            const int f = fg * OZ_FIELDS + fl;
            const int l = m + gg * Kq + 32 * rho + l32;
            const bool ok = f < a.nf && l <= lend;
            nv[k] = spec[ok ? (size_t)f * a.spec_fstride + l : 0];
            if (!ok) nv[k].x = nv[k].y = 0;
        }
        if (tid < 128) {
            const int l = m + (tid >> 5) * Kq + 32 * rho + (tid & 31);
            nab = ab[min(l, lmax) + 1];
        }
    };
    load_chunk(0);
    for (int rho = 0; rho < nch; ++rho) {
        __syncthreads();  // the previous chunk's operands are consumed
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = tid + 256 * k;
            const int l32 = e & 31, gg = (e >> 5) & 3, fl = e >> 7;
            stg[fl * F32_FS + gg * F32_GS + l32] = make_float2((float)nv[k].x, (float)nv[k].y);
        }
        if (tid < 128) abch[(tid >> 5) * OZ_ABS + (tid & 31)] = nab;
        if (rho + 1 < nch) load_chunk(rho + 1);
        __syncthreads();
        // a lane still in the scaled range contributes zeros (scale 0); so does a
        // degree past lend: the recurrence keeps running there with the last
        // table entry and can overflow, and an infinite A times a zero B would be NaN
        double sl = kk == 0 ? scl : 0.0;
        const int lq = m + g * Kq + 32 * rho;  // degree of step 0 of this lane's quarter
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = blk * 8 + q;
                const float af = lq + i <= lend ? (float)(cur * sl) : 0.f;
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const int c = r + 16 * ct;  // B column of this lane
                    const float bv = stf[2 * ((c >> 1) * F32_FS + g * F32_GS + i) + (c & 1)];
                    acc[i & 1][ct] = mfma4(af, bv, acc[i & 1][ct]);
                }
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
        }
        // FP32 tiles of 16 degrees x 4 quarters end here; sums continue in FP64
#pragma unroll
        for (int par = 0; par < 2; ++par)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
                for (int e = 0; e < 4; ++e) accd[par][ct][e] += (double)acc[par][ct][e];
                acc[par][ct] = v4f{0.f, 0.f, 0.f, 0.f};
            }
    }

    // lane holds rows 4 (lane >> 4) + e (pairs), column (lane & 15) + 16 ct
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int c = (lane & 15) + 16 * ct;
        const int f = fg * OZ_FIELDS + (c >> 1);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int pp = pbk * 64 + w * 16 + 4 * g + e;
            const bool live = pp < a.np;
            const double sc = a.g * ldexp(1.0, live ? a.ep[(size_t)im * a.np + pp] : 0);
            const double E = accd[0][ct][e], O = accd[1][ct][e];
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

template <typename IT>
__global__ __launch_bounds__(256) void k_synth_f32(const OzArgs a) {
    __shared__ float2 stg[OZ_FIELDS * F32_FS];
    __shared__ double2 abch[4 * OZ_ABS];
    const int gx = (a.np + 63) / 64;
    int pbk, unit;
    if (!unit_of_block(a.xcd, gx, a.nm * a.nfg, &unit, &pbk)) return;
    synth_f32_body<IT>(a, pbk, unit % a.nm, unit / a.nm, stg, abch);
}

// ---------------------------------------------------------------------------
// Analysis field operand: G+- = w (F_N +- F_S) per pair as FP32, stored in the
// B-operand layout of the analysis waves
//   [field group][order][pair block][K half][parity][column tile][j][lane]
// with lane = c + 16 g for column c of the tile and pair 4 j + g of the K half.
// One thread per (pair, order, field group); pairs past np get zeros.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_f32_gprep(const OzArgs a) {
    const int npp = a.npb * 256 * a.nh;
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int im = blockIdx.y, fg = blockIdx.z;
    if (p >= npp) return;
    const int pb = p / (256 * a.nh), pl = p % (256 * a.nh), kh = pl >> 8, q = pl & 255;
    const int j = q >> 2, gq = q & 3;
    const bool act = p < a.np;
    const double wl = act ? a.wlat[p] : 0.0;
    const int ps = a.nlat - 1 - p;
    float *base = a.gf32 + ((((size_t)fg * a.nm + im) * a.npb + pb) * a.nh + kh) * 4 * 4096;
    // all 32 columns of this pair per parity in registers, then written as the
    // pair's two 16-column runs (consecutive threads: consecutive 64-byte runs)
    float v[2][OZ_COLS];
#pragma unroll
    for (int fl = 0; fl < OZ_FIELDS; ++fl) {
        const int f = fg * OZ_FIELDS + fl;
        double2 FN = make_double2(0.0, 0.0), FS = make_double2(0.0, 0.0);
        if (act && f < a.nf) {
            const double2 *Fi = a.F0 + ((size_t)f * a.nm + im) * a.nlat;
            FN = Fi[p];
            if (ps != p) FS = Fi[ps];
        }
        v[0][2 * fl] = (float)(wl * (FN.x + FS.x));
        v[0][2 * fl + 1] = (float)(wl * (FN.y + FS.y));
        v[1][2 * fl] = (float)(wl * (FN.x - FS.x));
        v[1][2 * fl + 1] = (float)(wl * (FN.y - FS.y));
    }
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            float4 *dst = (float4 *)(base + ((par * 2 + ct) * 64 + j) * 64 + 16 * gq);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                dst[k] = make_float4(v[par][16 * ct + 4 * k], v[par][16 * ct + 4 * k + 1], v[par][16 * ct + 4 * k + 2],
                                     v[par][16 * ct + 4 * k + 3]);
        }
}

// LDS position of pair q (0..255) of a K half: pairs 4 j + g are read four j at a
// time by lane group g, so they sit at g * 72 + j (72 = 64 + 8 against bank conflicts
// of the writes); a parity row holds NH halves of 4 * 72 and 4 more floats.
__device__ __forceinline__ int f32_pos(int q) { return (q & 3) * 72 + (q >> 2); }

// Degree I of a 32-degree chunk: P_l 2^-epl (FP32) into the LDS rows, the
// Q-form step, and after every 8th degree the scale check of a chain still in
// the scaled range.  lv46 = live x 2^-46 (live: 1 once the lane left the
// scaled range, else 0).
template <int I, int RSF>
__device__ __forceinline__ void f32_digit_steps(const OzRec &R, double x, double &prev, double &cur, int &kk,
                                                double &lv46, float *Alds, int wpos) {
    if constexpr (I < 32) {
        constexpr int H = I >> 4, N = I & 15;
        double v = 0.0;
        const double cl = cur * lv46;
        SHTK_FMAC_BC(v, R.a[H], cl, N, "");
        Alds[((I & 1) * 16 + (I >> 1)) * RSF + wpos] = (float)v;
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
                lv46 = kk == 0 ? 0x1p-46 : 0.0;
            }
        }
        f32_digit_steps<I + 1, RSF>(R, x, prev, cur, kk, lv46, Alds, wpos);
    }
}

template <int NH>
__device__ __forceinline__ void anal_f32_body(const OzArgs &a, const double *__restrict__ asc,
                                              const double2 *__restrict__ qc, const int *__restrict__ epl,
                                              const int *__restrict__ pfirst, int pb, int run, int fg, float *Alds,
                                              double *Xch) {
    constexpr int RSF = NH * 288 + 4, RPW = NH == 2 ? 2 : 4;
    const cptr<int> ru = as_const((const int *)a.run) + 4 * run;
    const int im = ru[0], gs0 = ru[1], nsr = ru[2], m = im * a.mres;
    const int s0 = gs0 - as_const(a.segbase)[im];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int pbase = pb * 256 * NH;
    const int p = pbase + tid;
    const bool act = p < a.np;
    const double x = act ? a.x[p] : 0.0;
    const long long co = as_const(a.coff)[im] - m;
    const double *ascl = asc + co, *qbl = a.qb + co, *qccl = a.qcc + co;  // padded past lmax + 1
    const cptr<int> eplm = as_const(epl) + co;
    const cptr<int> pfc = as_const(pfirst);

    // contraction role of this wave: K half kh, parity wp, column tile wct
    const int kh = w >> 2, wp = (w >> 1) & 1, wct = w & 1;
    float Bw[64];
    {
        const float *src = a.gf32 + (((((size_t)fg * a.nm + im) * a.npb + pb) * NH + kh) * 4 + wp * 2 + wct) * 4096;
#pragma unroll
        for (int j = 0; j < 64; ++j) Bw[j] = src[j * 64 + lane];
    }
    const int c = (lane & 15) + 16 * wct;
    const int f = fg * OZ_FIELDS + (c >> 1);
    const bool fok = f < a.nf;
    double *outc = a.spec_out + 2 * ((size_t)(fok ? f : 0) * a.spec_fstride + (size_t)(a.lmax + 1) * m) + (c & 1);
    const bool atom = a.npb > 1;
    const int wpos = (tid >> 8) * 288 + f32_pos(tid & 255);  // LDS column of this lane's pair

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
    double lv46 = kk == 0 ? 0x1p-46 : 0.0;  // lanes still in the scaled range contribute zeros
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
    const int grp = lane >> 4;
    const int r0 = NH == 2 ? 2 * kh : 0;  // first of the rows 4 grp + e this wave emits
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
                const size_t q = (size_t)(gs0 + sr) * a.np + (act ? p : 0);
                prev = act ? a.seedq0[q] : 0.0;
                cur = act ? a.seedq1[q] : 0.0;
                kk = act ? a.seedk[q] : 0;
                lv46 = kk == 0 ? 0x1p-46 : 0.0;
            }
        }
        for (int t0 = 0; t0 < nst; t0 += 32) {
            const int lc = __builtin_amdgcn_readfirstlane(l0 + t0);
            // this chunk's records; the next chunk's (clamped to lmax: reads stay
            // inside the padded tables, a chunk that does not exist never uses them)
            const OzRec rec = nrec;
            load_rec(nrec, min(t0 + 32 < nst ? lc + 32 : m + (s0 + sr + 1) * a.seg, a.lmax));
            asm volatile("s_nop 4");  // wait states before DPP reads the refilled records
            f32_digit_steps<0, RSF>(rec, x, prev, cur, kk, lv46, Alds, wpos);
            __syncthreads();
            // rows: the 16 degrees of parity wp; K: the 256 pairs of half kh
            double accd[4] = {0.0, 0.0, 0.0, 0.0};
            const float *arow = Alds + (wp * 16 + (lane & 15)) * RSF + kh * 288 + grp * 72;
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const float4 A4 = *(const float4 *)(arow + 16 * t4 + 4 * jj);
                    const int j0 = 16 * t4 + 4 * jj;
                    acc = mfma4(A4.x, Bw[j0], acc);
                    acc = mfma4(A4.y, Bw[j0 + 1], acc);
                    acc = mfma4(A4.z, Bw[j0 + 2], acc);
                    acc = mfma4(A4.w, Bw[j0 + 3], acc);
                }
                // FP32 tile of 64 pairs done; the sum over tiles is FP64
#pragma unroll
                for (int e = 0; e < 4; ++e) accd[e] += (double)acc[e];
            }
            if constexpr (NH == 2) {
                // rows 2 kh, 2 kh + 1 stay with this wave; the other two go to the
                // partner of the other K half (wave w ^ 4)
                double *xo = Xch + w * 2 * 64 + lane;
                xo[0] = accd[2 * (1 - kh)];
                xo[64] = accd[2 * (1 - kh) + 1];
            }
            __syncthreads();
            const double *xi = Xch + (w ^ 4) * 2 * 64 + lane;
#pragma unroll
            for (int rr = 0; rr < RPW; ++rr) {
                const int row = 4 * grp + r0 + rr;
                const int t = t0 + 2 * row + wp;
                if (t < nst && fok) {
                    double s = accd[r0 + rr];
                    if constexpr (NH == 2) s += xi[rr * 64];
                    const double val = s * a.g * ldexp(1.0, eplm[l0 + t]);
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
__global__ __launch_bounds__(256 * NH) void k_anal_f32(const OzArgs a, const double *__restrict__ asc,
                                                        const double2 *__restrict__ qc, const int *__restrict__ epl,
                                                        const int *__restrict__ pfirst) {
    __shared__ __align__(16) float Alds[2 * 16 * (NH * 288 + 4)];
    __shared__ double Xch[NH == 2 ? 8 * 2 * 64 : 1];
    const int item = blockIdx.x;
    anal_f32_body<NH>(a, asc, qc, epl, pfirst, item % a.npb, (item / a.npb) % a.nrun, item / (a.npb * a.nrun),
                      Alds, Xch);
}

}  // namespace

hipError_t launch_f32_synth(const OzArgs &a, hipStream_t st) {
    const int gx = (a.np + 63) / 64;
    const dim3 grid(grid_for_units(a.xcd, gx, a.nm * a.nfg));
    if (a.in_f32)
        hipLaunchKernelGGL(k_synth_f32<float>, grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_synth_f32<double>, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_f32_anal(const OzArgs &a, hipStream_t st) {
    if (a.nh != 1 && a.nh != 2) return hipErrorInvalidValue;
    const int npp = a.npb * 256 * a.nh;
    hipLaunchKernelGGL(k_f32_gprep, dim3((npp + 255) / 256, a.nm, a.nfg), dim3(256), 0, st, a);
    const dim3 g2((unsigned)a.npb * a.nrun * a.nfg);
    if (a.nh == 2)
        hipLaunchKernelGGL(k_anal_f32<2>, g2, dim3(512), 0, st, a, a.asc, a.qc, a.epl, a.pfirst);
    else
        hipLaunchKernelGGL(k_anal_f32<1>, g2, dim3(256), 0, st, a, a.asc, a.qc, a.epl, a.pfirst);
    return hipGetLastError();
}

}  // namespace shtk

// Legendre stage of the spherical harmonic transforms, FP64, for gfx950.
//
// Reference behaviour (SHTnsKit.jl):
//   rows       src/legendre.jl:233-322 (orthonormal + Condon-Shortley recurrence,
//              scaled exponents for seeds that underflow), 470-555 (d/dtheta and
//              1/sin(theta) rows)
//   synthesis  src/kernels.jl:57-65 and 155-169, orchestrated per m by
//              src/core_transforms.jl:749-869 and src/sphtor_transforms.jl:455-525
//   analysis   src/kernels.jl:33-40 and 205-215, orchestrated per m by
//              src/core_transforms.jl:655-746 and src/sphtor_transforms.jl:528-612
//
// Design (MI355X-first, not a translation of the per-(m, latitude) loops):
//   * Gauss nodes come in +/- x pairs; P_l^m(-x) = (-1)^(l-m) P_l^m(x), so one
//     recurrence chain serves a north/south latitude pair and the even/odd
//     degree sums give both hemispheres.
//   * Rows are cut into degree segments.  Within a segment the recurrence runs
//     on Q_l = P_l / c_l, Q_{l+1} = x Q_l - beta_{l+1} Q_{l-1} (one mul and one
//     fma per step); c_l is folded into the coefficients (synthesis) or into
//     the reduced sums (analysis).  A plan-time kernel stores each segment's
//     starting pair, so analysis work items are (order, segment) pairs of at
//     most `seg` degrees: balanced work, no cross-item reductions.
//   * Per-step coefficient records are staged once per segment in LDS and read
//     as broadcasts; the VALU only runs the recurrence and the contraction.
//   * Values that underflow FP64 carry a 2^(512k) scale; such chains run the
//     recurrence without contributing (physical value < 2^-359).  Chains that
//     stay below that level for a whole segment are skipped entirely (the
//     polar caps of high orders); blocks whose chains are all in range at the
//     segment start take a branch-free loop.
//   * Analysis: each lane sums its K chains in registers; per-degree sums over
//     lanes go through a wave-private LDS transpose every few degrees.
#include <algorithm>

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

template <typename T>
__device__ __forceinline__ void add_c(T *p, size_t idx, double re, double im) {
    atomicAdd(p + 2 * idx, (T)re);
    atomicAdd(p + 2 * idx + 1, (T)im);
}

__device__ __forceinline__ int nseg_of(int lmax, int m, int S) { return (lmax - m + S) / S; }

// Read-only tables read through the constant address space: wave-uniform
// reads become scalar loads, also in kernels that write global memory.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}


// ---------------------------------------------------------------------------
// Plan-time seeds: one thread per (order, latitude pair) runs the P-form
// recurrence with scaled exponents (src/legendre.jl:252-322) through the row
// and stores each segment's starting pair in the segment's Q normalisation.
// It also records, per segment, the first pair that is live anywhere in the
// segment (pfirst) and the end of the range of pairs still scaled at its start
// (pfast).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seed(const SeedArgs a) {
    const int im = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.np) return;
    const int m = im * a.mres;
    const int lmax = a.lmax, S = a.seg;
    const double x = a.x[p];
    const double2 *ab = a.ab + a.coff[im] - m;
    const size_t sidx = (size_t)im * a.np + p;
    double cur = a.sectv[sidx], prev = 0.0;
    int k = a.sectk[sidx];
    const int nseg = nseg_of(lmax, m, S);
    int l = m;
    for (int s = 0; s < nseg; ++s) {
        const int gs = a.segbase[im] + s;
        const int l0 = m + s * S;
        const size_t q = (size_t)gs * a.np + p;
        if (s == 0) {
            a.seedq0[q] = 0.0;
            a.seedq1[q] = cur;
        } else {
            a.seedq0[q] = prev;
            a.seedq1[q] = cur / ab[l0].x;
        }
        a.seedk[q] = k;
        if (k > 0) atomicMax(&a.pfast[gs], p + 1);
        const int lend = min(l0 + S, lmax + 1);
        for (; l < lend; ++l) {
            const double2 co = ab[l + 1];
            const double nx = fma(co.x * x, cur, -(co.y * prev));
            prev = cur;
            cur = nx;
            if (k > 0 && (((l - l0) & (SCALE_CHECK - 1)) == SCALE_CHECK - 1 || l == lend - 1) &&
                fmax(fabs(cur), fabs(prev)) > kHi) {
                cur *= kDown;
                prev *= kDown;
                --k;
            }
        }
        if (k == 0) atomicMin(&a.pfirst[gs], p);
    }
}

// ---------------------------------------------------------------------------
// Synthesis.  One workgroup of four waves per (pair block, order, field); the
// waves take the degree segments of the order round-robin, run the P-form
// recurrence (P_{l+1} = a x P_l - b P_{l-1}) from the segment seeds, and add
// their partial sums through LDS at the end in a fixed order.  Per-degree
// coefficients are wave-uniform and stream through scalar loads, prefetched
// one block of degrees ahead.
// ---------------------------------------------------------------------------
template <int K, int PAR, bool FAST>
__device__ __forceinline__ void sstep(const double4 &r, const double *x, double *prev, double *cur, const double *cm,
                                      double (*acc)[4]) {
    // r = (beta_{l+1}, c_l, Re c_l a_l, Im c_l a_l): the coefficients arrive
    // scaled by c_l (k_spec_qscale)
    const double cr = r.z, ci = r.w;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double v = FAST ? cur[j] : cur[j] * cm[j];
        acc[j][2 * PAR] = fma(v, cr, acc[j][2 * PAR]);
        acc[j][2 * PAR + 1] = fma(v, ci, acc[j][2 * PAR + 1]);
        const double nx = fma(x[j], cur[j], -(r.x * prev[j]));
        prev[j] = cur[j];
        cur[j] = nx;
    }
}

// c: the Q-form scale of the current degree (1 for the P form)
template <int K>
__device__ __forceinline__ void rescale(double *prev, double *cur, int *kk, double *cm, double c = 1.0) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (kk[j] > 0 && fmax(fabs(cur[j]), fabs(prev[j])) * c > kHi) {
            cur[j] *= kDown;
            prev[j] *= kDown;
            kk[j] -= 1;
            cm[j] = kk[j] == 0 ? 1.0 : 0.0;
        }
    }
}

// Degrees l0 .. lstop-1 of one segment; (l - l0) has the parity of (l - m).
// Q form: Q_{l+1} = x Q_l - beta_{l+1} Q_{l-1}, P_l = c_l Q_l, so the
// contraction weight of degree l is c_l a_l (pre-scaled by k_spec_qscale) and a
// chain step costs one multiply and three FMAs.
template <int K, bool FAST, typename IT>
__device__ __forceinline__ void synth_seg(cptr<double> qbm, cptr<double> qcm, const IT *__restrict__ al, int l0,
                                          int lstop, int lmax, const double *x, double *prev, double *cur,
                                          int *kk, double (*acc)[4]) {
    double cm[K];
#pragma unroll
    for (int j = 0; j < K; ++j) cm[j] = kk[j] == 0 ? 1.0 : 0.0;
    // beta_{l+1}, c_l, c_l a_l; the tables are padded, the coefficient column is not
    auto ld = [&](int l) {
        const int lc = min(l, lmax);
        const double c = qcm[l];
        return make_double4(qbm[l + 1], c, (double)al[2 * lc], (double)al[2 * lc + 1]);
    };
    double4 r0 = ld(l0), r1 = ld(l0 + 1), r2 = ld(l0 + 2), r3 = ld(l0 + 3);
    int l = l0;
    for (; l + 4 <= lstop; l += 4) {
        const double4 n0 = ld(l + 4), n1 = ld(l + 5), n2 = ld(l + 6), n3 = ld(l + 7);
        sstep<K, 0, FAST>(r0, x, prev, cur, cm, acc);
        sstep<K, 1, FAST>(r1, x, prev, cur, cm, acc);
        sstep<K, 0, FAST>(r2, x, prev, cur, cm, acc);
        sstep<K, 1, FAST>(r3, x, prev, cur, cm, acc);
        if (!FAST && ((l - l0) & (SCALE_CHECK - 1)) == SCALE_CHECK - 4) rescale<K>(prev, cur, kk, cm, r3.y);
        r0 = n0;
        r1 = n1;
        r2 = n2;
        r3 = n3;
    }
    if (l < lstop) sstep<K, 0, FAST>(r0, x, prev, cur, cm, acc);
    if (l + 1 < lstop) sstep<K, 1, FAST>(r1, x, prev, cur, cm, acc);
    if (l + 2 < lstop) sstep<K, 0, FAST>(r2, x, prev, cur, cm, acc);
}

template <int K, typename IT>
__global__ __launch_bounds__(256) void k_synth_s(const LegArgs a) {
    __shared__ double red[4][4 * K][64];
    int b, unit;
    if (!unit_of_block(a.xcd, a.gx, a.gy * a.gz, &unit, &b)) return;
    const int im = unit % a.gy, f = unit / a.gy;
    const int m = im * a.mres;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int S = a.seg, np = a.np, lmax = a.lmax, ltr = a.ltr;
    const int pb0 = b * 64 * K, pb1 = min(pb0 + 64 * K, np);
    double x[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = pb0 + lane + 64 * j;
        x[j] = p < np ? a.x[p] : 0.0;
    }
    const long long co = as_const(a.coff)[im] - m;
    const cptr<double> qbm = as_const(a.qb) + co, qcm = as_const(a.qcc) + co;
    const IT *__restrict__ al = (const IT *)a.in0 + 2 * ((size_t)f * a.spec_fstride + (size_t)(lmax + 1) * m);
    double acc[K][4];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.0;
    const int nseg = nseg_of(lmax, m, S);
    // segment seeds of the wave's next segment are loaded while the current one runs
    double nq0[K], nq1[K];
    int nqk[K];
    auto load_seeds = [&](int s) {
        const int gs = as_const(a.segbase)[im] + min(s, nseg - 1);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int p = pb0 + lane + 64 * j;
            const bool act = p < np;
            const size_t q = (size_t)gs * np + (act ? p : 0);
            nq0[j] = act ? a.seedq0[q] : 0.0;
            nq1[j] = act ? a.seedq1[q] : 0.0;
            nqk[j] = act ? a.seedk[q] : 0;
        }
    };
    if (w < nseg) load_seeds(w);
    for (int s = w; s < nseg; s += 4) {
        const int l0 = m + s * S;
        if (l0 > ltr) break;
        const int gs = as_const(a.segbase)[im] + s;
        double prev[K], cur[K];
        int kk[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            prev[j] = nq0[j];
            cur[j] = nq1[j];
            kk[j] = nqk[j];
        }
        if (s + 4 < nseg) load_seeds(s + 4);
        if (pb1 <= as_const(a.pfirst)[gs]) continue;
        const bool fast = pb0 >= as_const(a.pfast)[gs];
        const int lstop = min(l0 + S, ltr + 1);
        if (fast) synth_seg<K, true>(qbm, qcm, al, l0, lstop, lmax, x, prev, cur, kk, acc);
        else synth_seg<K, false>(qbm, qcm, al, l0, lstop, lmax, x, prev, cur, kk, acc);
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) red[w][4 * j + q][lane] = acc[j][q];
    __syncthreads();
    if (w != 0) return;
    const double gm = a.g * mfac_of(a.mfac_mode, m);
    double2 *Fo = a.F0 + ((size_t)f * a.nm + im) * a.nlat;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = pb0 + lane + 64 * j;
        if (p >= np) continue;
        double t[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            t[q] = ((red[0][4 * j + q][lane] + red[1][4 * j + q][lane]) + red[2][4 * j + q][lane]) +
                   red[3][4 * j + q][lane];
        const double sc = gm * (a.wlat ? a.wlat[p] : 1.0);
        Fo[p] = make_double2((t[0] + t[2]) * sc, (t[1] + t[3]) * sc);
        const int ps = a.nlat - 1 - p;
        if (ps != p) Fo[ps] = make_double2((t[0] - t[2]) * sc, (t[1] - t[3]) * sc);
    }
}

// ---------------------------------------------------------------------------
// Scalar synthesis with the per-degree operands in VGPRs (64-degree segments).
// Lane 16 r + n of a wave holds the record (beta_{l+1}, c_l a_l) of degree
// l0 + 16 g + n in record group g (the four 16-lane rows hold the same
// records), and each use broadcasts lane n of the row with DPP row_newbcast:n
// as the first operand of v_fmac_f64, a full-rate FP64 FMA: the degree loop has
// no scalar loads and no waits.  The records of a wave's next segment are
// loaded (vector loads) while the current one runs, like the seeds.
// ---------------------------------------------------------------------------
struct SRec {
    double b;   // beta_{l+1}
    double2 c;  // c_l a_l
};

// One degree for K chains: acc += (c_l a_l) u (u = Q_l, times 0 for chains that
// are still scaled) and Q_{l+1} = x Q_l - beta_{l+1} Q_{l-1}.  NOP puts five
// wait states before the first DPP read, for records just written by a VALU copy
// or an EXEC write by the rescale.  The asm statements are volatile so that the
// NOP stays in front of the others.
template <int K, int PAR, int N, bool NOP, bool FAST>
__device__ __forceinline__ void dstep(const SRec &r, const double *x, double *prev, double *cur, const double *cm,
                                      double (*acc)[4]) {
    double t[K], u[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        t[j] = x[j] * cur[j];
        u[j] = FAST ? cur[j] : cur[j] * cm[j];
    }
    // the contraction FMAs of all chains first, so that each product x Q_l has
    // 2K instructions to complete before its recurrence FMA
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (NOP && j == 0)
            asm volatile(
                "s_nop 4\n\t"
                "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
                : "+v"(acc[j][2 * PAR]), "+v"(acc[j][2 * PAR + 1])
                : "v"(r.c.x), "v"(r.c.y), "v"(u[j]), "i"(N));
        else
            asm volatile(
                "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
                : "+v"(acc[j][2 * PAR]), "+v"(acc[j][2 * PAR + 1])
                : "v"(r.c.x), "v"(r.c.y), "v"(u[j]), "i"(N));
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
        asm volatile("v_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(t[j])
                     : "v"(r.b), "v"(prev[j]), "i"(N));
#pragma unroll
    for (int j = 0; j < K; ++j) {
        prev[j] = cur[j];
        cur[j] = t[j];
    }
}

// Steps N = 0 .. 15 of one record group while l0g + N < lstop; (l - m) has
// the parity of N.  The scaled chains are checked every 8 degrees with
// c_l of the last degree (scalar load, unscaled blocks never take it).
template <int K, bool FAST, int N>
__device__ __forceinline__ bool dgroup_from(const SRec &r, cptr<double> qcm, int l0g, int nrem, const double *x,
                                            double *prev, double *cur, int *kk, double *cm, double (*acc)[4]) {
    if constexpr (N == 16) {
        return true;
    } else {
        if (N >= nrem) return false;
        dstep<K, N & 1, N, N == 0 || (!FAST && N == 8), FAST>(r, x, prev, cur, cm, acc);
        if (!FAST && (N & 7) == 7) rescale<K>(prev, cur, kk, cm, qcm[l0g + N]);
        return dgroup_from<K, FAST, N + 1>(r, qcm, l0g, nrem, x, prev, cur, kk, cm, acc);
    }
}

template <int K, bool FAST>
__device__ __forceinline__ void dsynth_seg(const SRec (&R)[4], cptr<double> qcm, int l0, int lstop, const double *x,
                                           double *prev, double *cur, int *kk, double (*acc)[4]) {
    double cm[K];
#pragma unroll
    for (int j = 0; j < K; ++j) cm[j] = kk[j] == 0 ? 1.0 : 0.0;
    // a test per step (scalar, beside the VALU work) ends the segment at lstop;
    // straight-line groups without it raise the VGPR count past 128
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int l0g = l0 + 16 * g;
        if (!dgroup_from<K, FAST, 0>(R[g], qcm, l0g, lstop - l0g, x, prev, cur, kk, cm, acc)) return;
    }
}

// waves per k_synth_d workgroup, taking the segments round-robin and adding
// their sums through LDS (2 waves time the same; for one field, 1 wave per
// workgroup leaves a long tail of low orders: 82 against 65 us)
constexpr int DSW = 4;

// NW = DSW: units order-major.  NW = 1 (launches with many units, i.e. batches
// of fields): one wave per (order, pair block, field) runs all its segments,
// units lowest order first for every field so the long columns start first.
template <int K, int NW>
__global__ __launch_bounds__(64 * NW) void k_synth_d(const LegArgs a) {
    __shared__ double red[NW > 1 ? NW : 1][NW > 1 ? 4 * K : 1][64];
    int b, unit;
    if (!unit_of_block(a.xcd, a.gx, a.gy * a.gz, &unit, &b)) return;
    const int im = NW > 1 ? unit % a.gy : unit / a.gz, f = NW > 1 ? unit / a.gy : unit % a.gz;
    const int m = im * a.mres;
    const int lane = threadIdx.x & 63;
    const int w = NW > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : 0;
    const int np = a.np, lmax = a.lmax, ltr = a.ltr;
    const int pb0 = b * 64 * K, pb1 = min(pb0 + 64 * K, np);
    double x[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = pb0 + lane + 64 * j;
        x[j] = p < np ? a.x[p] : 0.0;
    }
    const long long co = as_const(a.coff)[im] - m;
    const cptr<double> qcm = as_const(a.qcc) + co;
    const double *qbm = a.qb + co;  // padded: degrees to lmax + 64 are readable
    // c_l a_l (k_spec_qscale, or a_l times c_l here), one order column; reads
    // are clamped to lmax
    const double2 *cal = (const double2 *)a.in0 + (size_t)f * a.spec_fstride + (size_t)(lmax + 1) * m;
    const double *qcl = a.qcc + co;
    double acc[K][4];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.0;
    const int nseg = nseg_of(lmax, m, DSEG);
    // the next segment's seeds, records and pair bounds are loaded while the
    // current one runs (the bounds through scalar loads: the degree loop issues
    // none, so their wait at the next segment start finds them in)
    const int segb = as_const(a.segbase)[im];
    double nq0[K], nq1[K];
    int nqk[K], npf = 0, npa = 0;
    SRec nr[4];
    auto load_next = [&](int s) {
        const int sc = min(s, nseg - 1);
        const int gs = segb + sc;
        npf = as_const(a.pfirst)[gs];
        npa = as_const(a.pfast)[gs];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int p = pb0 + lane + 64 * j;
            const bool act = p < np;
            const size_t q = (size_t)gs * np + (act ? p : 0);
            nq0[j] = act ? a.seedq0[q] : 0.0;
            nq1[j] = act ? a.seedq1[q] : 0.0;
            nqk[j] = act ? a.seedk[q] : 0;
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int l = m + sc * DSEG + 16 * g + (lane & 15);
            nr[g].b = qbm[l + 1];
            const int lc = min(l, lmax);
            // no branch around the loads (it would complete them at its join)
            double2 c = cal[lc];
            const double qv = qcl[lc];  // the table is padded past lmax: always readable
            const double q = a.qscale_in ? qv : 1.0;
            c.x *= q;
            c.y *= q;
            nr[g].c = c;
        }
    };
    if (w < nseg) load_next(w);
    for (int s = w; s < nseg; s += NW) {
        const int l0 = m + s * DSEG;
        if (l0 > ltr) break;
        const int pf = npf, pfa = npa;
        double prev[K], cur[K];
        int kk[K];
        SRec R[4];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            prev[j] = nq0[j];
            cur[j] = nq1[j];
            kk[j] = nqk[j];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) R[g] = nr[g];
        if (s + NW < nseg) load_next(s + NW);
        if (pb1 <= pf) continue;
        const bool fast = pb0 >= pfa;
        const int lstop = min(l0 + DSEG, ltr + 1);
        if (fast) dsynth_seg<K, true>(R, qcm, l0, lstop, x, prev, cur, kk, acc);
        else dsynth_seg<K, false>(R, qcm, l0, lstop, x, prev, cur, kk, acc);
    }
    if constexpr (NW > 1) {
#pragma unroll
        for (int j = 0; j < K; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) red[w][4 * j + q][lane] = acc[j][q];
        __syncthreads();
        if (w != 0) return;
    }
    const double gm = a.g * mfac_of(a.mfac_mode, m);
    double2 *Fo = a.F0 + ((size_t)f * a.nm + im) * a.nlat;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = pb0 + lane + 64 * j;
        if (p >= np) continue;
        double t[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if constexpr (NW > 1) {
                double v = red[0][4 * j + q][lane];
#pragma unroll
                for (int ww = 1; ww < NW; ++ww) v += red[ww][4 * j + q][lane];
                t[q] = v;
            } else {
                t[q] = acc[j][q];
            }
        }
        const double sc = gm * (a.wlat ? a.wlat[p] : 1.0);
        Fo[p] = make_double2((t[0] + t[2]) * sc, (t[1] + t[3]) * sc);
        const int ps = a.nlat - 1 - p;
        if (ps != p) Fo[ps] = make_double2((t[0] - t[2]) * sc, (t[1] - t[3]) * sc);
    }
}

// ---------------------------------------------------------------------------
// Vector synthesis: s*Vt = sum_l [D_l S_l - i m P_l T_l], s*Vp = sum_l [i m P_l S_l + D_l T_l],
// D_l = s dP_l/dtheta = l x P_l - g_l P_{l-1}; output times wlat (1/s, or 1 in
// Robert form).  Parity: P(-x) = (-1)^(l-m) P(x), D(-x) = -(-1)^(l-m) D(x).
// Accumulators per parity: A = sum D S, B = sum P T, C = sum P S, E = sum D T.
// ---------------------------------------------------------------------------
template <int K, int PAR, bool FAST>
__device__ __forceinline__ void vstep(double beta, double dl, double dg, const double *R, const double *x,
                                      double *prev, double *cur, const double *cm, double (*acc)[2][8]) {
    // Q form: D_l = l c_l x Q_l - g_l c_{l-1} Q_{l-1} = dl (x Q_l) - dg Q_{l-1},
    // P_l = c_l Q_l; R = (S', T', c_l S', c_l T') of this degree
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double xc = x[j] * cur[j];
        double D = fma(dl, xc, -(dg * prev[j]));
        double Q = cur[j];
        if (!FAST) {
            D *= cm[j];
            Q *= cm[j];
        }
        double *A = acc[j][PAR];
        A[0] = fma(D, R[0], A[0]);
        A[1] = fma(D, R[1], A[1]);
        A[2] = fma(Q, R[6], A[2]);
        A[3] = fma(Q, R[7], A[3]);
        A[4] = fma(Q, R[4], A[4]);
        A[5] = fma(Q, R[5], A[5]);
        A[6] = fma(D, R[2], A[6]);
        A[7] = fma(D, R[3], A[7]);
        const double nx = fma(-beta, prev[j], xc);
        prev[j] = cur[j];
        cur[j] = nx;
    }
}

// One degree's uniform operands: beta_{l+1}, (l c_l, g_l c_{l-1}) and the
// prologue record (S', T', c_l S', c_l T'), all through scalar loads.
struct VDeg {
    double beta, dl, dg, R[8];
};

template <int K, bool FAST>
__device__ __forceinline__ void vsynth_seg(cptr<double> qbm, cptr<double> qdm, cptr<double> qcm, cptr<double> rec,
                                           int l0, int lstop, int lmax, const double *x, double *prev,
                                           double *cur, int *kk, double (*acc)[2][8]) {
    double cm[K];
#pragma unroll
    for (int j = 0; j < K; ++j) cm[j] = kk[j] == 0 ? 1.0 : 0.0;
    auto ld = [&](int l) {
        VDeg d;
        d.beta = qbm[l + 1];
        const int ld2 = 2 * min(l, lmax + 1);
        d.dl = qdm[ld2];
        d.dg = qdm[ld2 + 1];
        const int lc = min(l, lmax);
#pragma unroll
        for (int k = 0; k < 8; ++k) d.R[k] = rec[8 * lc + k];
        return d;
    };
    // two degrees (both parities) per iteration, the next two loaded meanwhile
    VDeg r0 = ld(l0), r1 = ld(l0 + 1);
    int l = l0;
    for (; l + 2 <= lstop; l += 2) {
        const VDeg n0 = ld(l + 2), n1 = ld(l + 3);
        vstep<K, 0, FAST>(r0.beta, r0.dl, r0.dg, r0.R, x, prev, cur, cm, acc);
        vstep<K, 1, FAST>(r1.beta, r1.dl, r1.dg, r1.R, x, prev, cur, cm, acc);
        if (!FAST && ((l - l0) & (SCALE_CHECK - 1)) == SCALE_CHECK - 2) rescale<K>(prev, cur, kk, cm, qcm[l + 2]);
        r0 = n0;
        r1 = n1;
    }
    if (l < lstop) vstep<K, 0, FAST>(r0.beta, r0.dl, r0.dg, r0.R, x, prev, cur, cm, acc);
}

constexpr int VS_CHAINS = 2;

template <int K>
__global__ __launch_bounds__(256) void k_synth_v(const LegArgs a) {
    __shared__ double red[4][16][64];
    int b, unit;
    if (!unit_of_block(a.xcd, a.gx, a.gy * a.gz, &unit, &b)) return;
    const int im = unit % a.gy, f = unit / a.gy;
    if (a.skip_m0 && im == 0) return;  // written by the zonal kernel
    const int m = im * a.mres;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int S = a.seg, np = a.np, lmax = a.lmax, ltr = a.ltr;
    const int pb0 = b * 64 * K, pb1 = min(pb0 + 64 * K, np);
    double x[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = pb0 + lane + 64 * j;
        x[j] = p < np ? a.x[p] : 0.0;
    }
    const long long co = as_const(a.coff)[im] - m;
    const cptr<double> qbm = as_const(a.qb) + co, qcm = as_const(a.qcc) + co;
    const cptr<double> qdm = as_const((const double *)a.qd) + 2 * co;
    // prologue records (k_vspec_prep), same element offsets as the spectral arrays
    const cptr<double> rec =
        as_const((const double *)a.in0) + 8 * ((size_t)f * a.spec_fstride + (size_t)(lmax + 1) * m);
    double acc[K][2][8];
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[j][q][r] = 0.0;
    const int nseg = nseg_of(lmax, m, S);
    for (int s = w; s < nseg; s += 4) {
        const int l0 = m + s * S;
        if (l0 > ltr) break;
        const int gs = as_const(a.segbase)[im] + s;
        if (pb1 <= as_const(a.pfirst)[gs]) continue;
        const bool fast = pb0 >= as_const(a.pfast)[gs];
        const int lstop = min(l0 + S, ltr + 1);
        double prev[K], cur[K];
        int kk[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int p = pb0 + lane + 64 * j;
            const bool act = p < np;
            const size_t q = (size_t)gs * np + (act ? p : 0);
            prev[j] = act ? a.seedq0[q] : 0.0;
            cur[j] = act ? a.seedq1[q] : 0.0;
            kk[j] = act ? a.seedk[q] : 0;
        }
        if (fast) vsynth_seg<K, true>(qbm, qdm, qcm, rec, l0, lstop, lmax, x, prev, cur, kk, acc);
        else vsynth_seg<K, false>(qbm, qdm, qcm, rec, l0, lstop, lmax, x, prev, cur, kk, acc);
    }
    const double mm = (double)m;
    const double gm = a.g * mfac_of(a.mfac_mode, m);
    const size_t fo = ((size_t)f * a.nm + im) * a.nlat;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j > 0) __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 8; ++r) red[w][8 * q + r][lane] = acc[j][q][r];
        __syncthreads();
        const int p = pb0 + lane + 64 * j;
        if (w == 0 && p < np) {
            double e[8], o[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                e[r] = ((red[0][r][lane] + red[1][r][lane]) + red[2][r][lane]) + red[3][r][lane];
                o[r] = ((red[0][8 + r][lane] + red[1][8 + r][lane]) + red[2][8 + r][lane]) + red[3][8 + r][lane];
            }
#pragma unroll
            for (int r = 2; r < 6; ++r) {
                e[r] *= mm;
                o[r] *= mm;
            }
            const double sc = gm * (a.wlat ? a.wlat[p] : 1.0);
            // [0..1] = A, [2..3] = m B, [4..5] = m C, [6..7] = E
            const double tNr = e[0] + o[0] + (e[3] + o[3]), tNi = e[1] + o[1] - (e[2] + o[2]);
            const double tSr = -e[0] + o[0] + (e[3] - o[3]), tSi = -e[1] + o[1] - (e[2] - o[2]);
            const double pNr = -(e[5] + o[5]) + e[6] + o[6], pNi = e[4] + o[4] + e[7] + o[7];
            const double pSr = -(e[5] - o[5]) - e[6] + o[6], pSi = (e[4] - o[4]) - e[7] + o[7];
            a.F0[fo + p] = make_double2(tNr * sc, tNi * sc);
            a.F1[fo + p] = make_double2(pNr * sc, pNi * sc);
            const int ps = a.nlat - 1 - p;
            if (ps != p) {
                a.F0[fo + ps] = make_double2(tSr * sc, tSi * sc);
                a.F1[fo + ps] = make_double2(pSr * sc, pSi * sc);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Vector synthesis with the per-degree operands in VGPRs (64-degree segments),
// the treatment of k_synth_d: lane 16 r + n holds the record of degree
// l0 + 16 g + n of record group g -- beta_{l+1}, l c_l, g_l c_{l-1} and the
// prologue record (S', T', c_l S', c_l T') -- and every use broadcasts lane n of
// the row (DPP row_newbcast:n) into v_fmac_f64, so the degree loop has no scalar
// loads and no waits.  The next group's records are loaded (vector loads)
// while the current group runs.  Per degree and chain: x Q_l, D_l = l c_l x Q_l
// - g_l c_{l-1} Q_{l-1} (two FMAs), eight contraction FMAs and the recurrence FMA.
// ---------------------------------------------------------------------------
struct VRec {
    double b, dl, dg;  // beta_{l+1}, l c_l, g_l c_{l-1}
    double R[8];       // S'r, S'i, T'r, T'i, c S'r, c S'i, c T'r, c T'i
};


template <int K, int PAR, int N, bool NOP, bool FAST>
__device__ __forceinline__ void vdstep(const VRec &r, const double *x, double *prev, double *cur, const double *cm,
                                       double (*acc)[2][8]) {
    double xc[K], D[K], Q[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        xc[j] = x[j] * cur[j];
        D[j] = 0.0;
    }
    // five wait states between a VALU write of the records (or of EXEC in a
    // rescale) and their first DPP read
    if (NOP) asm volatile("s_nop 4");
#pragma unroll
    for (int j = 0; j < K; ++j) SHTK_FMAC_BC(D[j], r.dl, xc[j], N, "");
#pragma unroll
    for (int j = 0; j < K; ++j) SHTK_FMAC_BC(D[j], r.dg, prev[j], N, "-");
#pragma unroll
    for (int j = 0; j < K; ++j) {
        Q[j] = cur[j];
        if (!FAST) {
            D[j] *= cm[j];
            Q[j] *= cm[j];
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double *A = acc[j][PAR];
        SHTK_FMAC_BC(A[0], r.R[0], D[j], N, "");
        SHTK_FMAC_BC(A[1], r.R[1], D[j], N, "");
        SHTK_FMAC_BC(A[2], r.R[6], Q[j], N, "");
        SHTK_FMAC_BC(A[3], r.R[7], Q[j], N, "");
        SHTK_FMAC_BC(A[4], r.R[4], Q[j], N, "");
        SHTK_FMAC_BC(A[5], r.R[5], Q[j], N, "");
        SHTK_FMAC_BC(A[6], r.R[2], D[j], N, "");
        SHTK_FMAC_BC(A[7], r.R[3], D[j], N, "");
    }
#pragma unroll
    for (int j = 0; j < K; ++j) SHTK_FMAC_BC(xc[j], r.b, prev[j], N, "-");
#pragma unroll
    for (int j = 0; j < K; ++j) {
        prev[j] = cur[j];
        cur[j] = xc[j];
    }
}

// Steps N = 0 .. 15 of one record group while N < nrem; (l - m) has the parity
// of N.  Scaled chains are checked every 8 degrees with c_l of the last degree.
template <int K, bool FAST, int N>
__device__ __forceinline__ bool vdgroup_from(const VRec &r, cptr<double> qcm, int l0g, int nrem, const double *x,
                                             double *prev, double *cur, int *kk, double *cm, double (*acc)[2][8]) {
    if constexpr (N == 16) {
        return true;
    } else {
        if (N >= nrem) return false;
        vdstep<K, N & 1, N, N == 0 || (!FAST && N == 8), FAST>(r, x, prev, cur, cm, acc);
        if (!FAST && (N & 7) == 7) rescale<K>(prev, cur, kk, cm, qcm[l0g + N]);
        return vdgroup_from<K, FAST, N + 1>(r, qcm, l0g, nrem, x, prev, cur, kk, cm, acc);
    }
}

// The accumulated sums of the pair block -> Fourier columns of both components
// (north and south rows by parity), as k_synth_v.
template <int K, int NW = 4>
__device__ __forceinline__ void vsynth_out(const LegArgs &a, double (*acc)[2][8], double (*red)[16][64], int w,
                                           int lane, int pb0, int im, int f) {
    const int m = im * a.mres, np = a.np;
    const double mm = (double)m;
    const double gm = a.g * mfac_of(a.mfac_mode, m);
    const size_t fo = ((size_t)f * a.nm + im) * a.nlat;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if constexpr (NW > 1) {
            if (j > 0) __syncthreads();
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int r = 0; r < 8; ++r) red[w][8 * q + r][lane] = acc[j][q][r];
            __syncthreads();
        }
        const int p = pb0 + lane + 64 * j;
        if (w == 0 && p < np) {
            double e[8], o[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                if constexpr (NW > 1) {
                    e[r] = ((red[0][r][lane] + red[1][r][lane]) + red[2][r][lane]) + red[3][r][lane];
                    o[r] = ((red[0][8 + r][lane] + red[1][8 + r][lane]) + red[2][8 + r][lane]) + red[3][8 + r][lane];
                } else {
                    e[r] = acc[j][0][r];
                    o[r] = acc[j][1][r];
                }
            }
#pragma unroll
            for (int r = 2; r < 6; ++r) {
                e[r] *= mm;
                o[r] *= mm;
            }
            const double sc = gm * (a.wlat ? a.wlat[p] : 1.0);
            // [0..1] = A, [2..3] = m B, [4..5] = m C, [6..7] = E
            const double tNr = e[0] + o[0] + (e[3] + o[3]), tNi = e[1] + o[1] - (e[2] + o[2]);
            const double tSr = -e[0] + o[0] + (e[3] - o[3]), tSi = -e[1] + o[1] - (e[2] - o[2]);
            const double pNr = -(e[5] + o[5]) + e[6] + o[6], pNi = e[4] + o[4] + e[7] + o[7];
            const double pSr = -(e[5] - o[5]) - e[6] + o[6], pSi = (e[4] - o[4]) - e[7] + o[7];
            a.F0[fo + p] = make_double2(tNr * sc, tNi * sc);
            a.F1[fo + p] = make_double2(pNr * sc, pNi * sc);
            const int ps = a.nlat - 1 - p;
            if (ps != p) {
                a.F0[fo + ps] = make_double2(tSr * sc, tSi * sc);
                a.F1[fo + ps] = make_double2(pSr * sc, pSi * sc);
            }
        }
    }
}

// VD: the records are built from the caller's canonical FP64 S and T as they
// are loaded (LegArgs.vdirect); a template parameter so that neither form
// branches around its loads (a branch would complete them at its join).
// NW = 4: the block's 4 waves take the order's segments round-robin and add
// their sums through LDS (units in order-major order); NW = 1: one wave runs all
// segments of its (order, pair block, field) with no barrier, units ordered
// lowest order first for every field (longest work first).
template <int K, bool VD, int NW>
__global__ __launch_bounds__(64 * NW) void k_synth_vd(const LegArgs a) {
    __shared__ double red[NW > 1 ? 4 : 1][16][64];
    int b, unit;
    if (!unit_of_block(a.xcd, a.gx, a.gy * a.gz, &unit, &b)) return;
    const int im = NW > 1 ? unit % a.gy : unit / a.gz, f = NW > 1 ? unit / a.gy : unit % a.gz;
    if (a.skip_m0 && im == 0) return;  // written by the zonal kernel
    const int m = im * a.mres;
    const int lane = threadIdx.x & 63;
    const int w = NW > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : 0;
    const int np = a.np, lmax = a.lmax, ltr = a.ltr;
    const int pb0 = b * 64 * K, pb1 = min(pb0 + 64 * K, np);
    double x[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = pb0 + lane + 64 * j;
        x[j] = p < np ? a.x[p] : 0.0;
    }
    const long long co = as_const(a.coff)[im] - m;
    const cptr<double> qcm = as_const(a.qcc) + co;
    const double *qbm = a.qb + co;                              // padded past lmax + 1
    const double2 *qdm = a.qd + co;                             // entries l = m .. lmax + 1
    const size_t cbase = (size_t)f * a.spec_fstride + (size_t)(lmax + 1) * m;
    const double4 *rec = (const double4 *)a.in0 + 2 * cbase;
    // vdirect: the caller's canonical FP64 S and T, records built as they are loaded
    const double2 *Sd = (const double2 *)a.in0 + cbase, *Td = (const double2 *)a.in1 + cbase;
    const double *qcl = a.qcc + co;
    double acc[K][2][8];
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[j][q][r] = 0.0;
    const int nseg = nseg_of(lmax, m, DSEG);
    const int segb = as_const(a.segbase)[im];
    // record group (segment s, group g) of this lane's degree
    auto load_rec = [&](VRec &d, int s, int g) {
        const int l = m + min(s, nseg - 1) * DSEG + 16 * g + (lane & 15);
        d.b = qbm[l + 1];
        const double2 dd = qdm[min(l, lmax + 1)];
        d.dl = dd.x;
        d.dg = dd.y;
        const int lc = min(l, lmax);
        if constexpr (VD) {
            // raw values only: the products with c_l are formed when the group
            // starts (a use here would wait for the loads at once)
            const double2 sv = Sd[lc], tv = Td[lc];
            d.R[0] = sv.x;
            d.R[1] = sv.y;
            d.R[2] = tv.x;
            d.R[3] = tv.y;
            d.R[4] = qcl[lc];
        } else {
            const double4 r0 = rec[2 * (size_t)lc], r1 = rec[2 * (size_t)lc + 1];
            d.R[0] = r0.x;
            d.R[1] = r0.y;
            d.R[2] = r0.z;
            d.R[3] = r0.w;
            d.R[4] = r1.x;
            d.R[5] = r1.y;
            d.R[6] = r1.z;
            d.R[7] = r1.w;
        }
    };
    VRec nr;
    if (w < nseg) load_rec(nr, w, 0);
    for (int s = w; s < nseg; s += NW) {
        const int l0 = m + s * DSEG;
        if (l0 > ltr) break;
        const int gs = segb + s;
        const int pf = as_const(a.pfirst)[gs], pfa = as_const(a.pfast)[gs];
        const int lstop = min(l0 + DSEG, ltr + 1);
        if (pb1 <= pf) {
            if (s + NW < nseg) load_rec(nr, s + NW, 0);
            continue;
        }
        double prev[K], cur[K];
        int kk[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int p = pb0 + lane + 64 * j;
            const bool act = p < np;
            const size_t q = (size_t)gs * np + (act ? p : 0);
            prev[j] = act ? a.seedq0[q] : 0.0;
            cur[j] = act ? a.seedq1[q] : 0.0;
            kk[j] = act ? a.seedk[q] : 0;
        }
        const bool fast = pb0 >= pfa;
        double cm[K];
#pragma unroll
        for (int j = 0; j < K; ++j) cm[j] = kk[j] == 0 ? 1.0 : 0.0;
        for (int g = 0; g < 4; ++g) {
            const int l0g = l0 + 16 * g;
            VRec r = nr;
            if constexpr (VD) {
                const double c = r.R[4];
                r.R[4] = c * r.R[0];
                r.R[5] = c * r.R[1];
                r.R[6] = c * r.R[2];
                r.R[7] = c * r.R[3];
            }
            // the next group: this segment's next, or the first of the wave's next segment
            if (g < 3 && l0g + 16 < lstop) load_rec(nr, s, g + 1);
            else if (s + NW < nseg) load_rec(nr, s + NW, 0);
            const bool full = fast ? vdgroup_from<K, true, 0>(r, qcm, l0g, lstop - l0g, x, prev, cur, kk, cm, acc)
                                   : vdgroup_from<K, false, 0>(r, qcm, l0g, lstop - l0g, x, prev, cur, kk, cm, acc);
            if (!full || l0g + 16 >= lstop) break;
        }
    }
    vsynth_out<K, NW>(a, acc, red, w, lane, pb0, im, f);
}

// Prologue of the vector synthesis: per (order, field) and l >= m the record
// (S', T', c_l S', c_l T') with S' = conv(l, m) S [/ l(l+1)], T' likewise.
template <typename IT>
__global__ __launch_bounds__(256) void k_vspec_prep(const IT *S, const IT *T, double *rec, const double *scale,
                                                    int divide, int ll1, const double *qcc,
                                                    const long long *coff, int lmax, int mres, int nm,
                                                    long long per_field) {
    const int im = blockIdx.x % nm, f = blockIdx.x / nm, m = im * mres;
    const double *cm = qcc + coff[im] - m;
    const long long base = (long long)f * per_field + (long long)(lmax + 1) * m;
    for (int l = m + (int)threadIdx.x; l <= lmax; l += 256) {
        const long long q = base + l;
        double sc = 1.0;
        if (scale) {
            const double cs = scale[q - (long long)f * per_field];
            sc = divide ? 1.0 / cs : cs;
        }
        if (ll1) sc = l > 0 ? sc / ((double)l * (l + 1)) : 0.0;
        const double c = cm[l];
        const double sr = sc * (double)S[2 * q], si = sc * (double)S[2 * q + 1];
        const double tr = sc * (double)T[2 * q], ti = sc * (double)T[2 * q + 1];
        double4 *r = (double4 *)(rec + 8 * q);
        r[0] = make_double4(sr, si, tr, ti);
        r[1] = make_double4(c * sr, c * si, c * tr, c * ti);
    }
}

// ---------------------------------------------------------------------------
// Vector transforms on grids with nodes at the poles (first pair = theta 0, pi).
// The Legendre kernels weight those rows by 1/sin(theta) := 0, which is the
// exact limit for m = 0 and m >= 2.  For m = 1 the pole limits are finite:
// dP_l/dtheta = P_l/sin(theta) = -h_l at the north pole and dP_l/dtheta =
// (-1)^(l+1) h_l, P_l/sin(theta) = (-1)^l h_l at the south pole, with
// h_l = sqrt((2l+1) l (l+1) / (4 pi)) / 2 (orthonormal, Condon-Shortley phase;
// _dPdtheta_at_pole / _P_over_sinth_at_pole, src/kernels.jl:72-99, and
// src/legendre.jl:405-423).  These kernels rewrite the pole rows of order 1
// (one block per field); in Robert form the synthesis rows are sin(theta) times
// the field, i.e. zero for every order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pole_h(int l) {
    return 0.5 * 0.28209479177387814 * sqrt((2.0 * l + 1.0) * (double)l * (double)(l + 1));
}

template <int NV>
__device__ __forceinline__ void block_sum256(double (*red)[256], double *v) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int k = 0; k < NV; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = red[k][0];
}

__global__ __launch_bounds__(256) void k_vpole_synth(const LegArgs a, int robert, double wscale) {
    __shared__ double red[8][256];
    const int f = blockIdx.x, nlat = a.nlat;
    if (robert) {
        for (int im = threadIdx.x; im < a.nm; im += 256) {
            const size_t fo = ((size_t)f * a.nm + im) * nlat;
            const double2 z = make_double2(0.0, 0.0);
            a.F0[fo] = z;
            a.F1[fo] = z;
            a.F0[fo + nlat - 1] = z;
            a.F1[fo + nlat - 1] = z;
        }
        return;
    }
    // order 1: prologue records (S', T', c S', c T'), or with vdirect the caller's S and T
    const size_t c1 = (size_t)f * a.spec_fstride + (size_t)(a.lmax + 1);
    const double *rec = (const double *)a.in0 + 8 * c1;
    const double2 *Sd = (const double2 *)a.in0 + c1, *Td = (const double2 *)a.in1 + c1;
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int lt = min(a.ltr, a.lmax);
    for (int l = 1 + (int)threadIdx.x; l <= lt; l += 256) {
        const double h = pole_h(l), hs = (l & 1) ? -h : h;
        double r[4];
        if (a.vdirect) {
            const double2 sv = Sd[l], tv = Td[l];
            r[0] = sv.x;
            r[1] = sv.y;
            r[2] = tv.x;
            r[3] = tv.y;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = rec[8 * (size_t)l + k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = fma(h, r[k], v[k]);
            v[4 + k] = fma(hs, r[k], v[4 + k]);
        }
    }
    block_sum256<8>(red, v);
    if (threadIdx.x == 0) {
        const double gm = a.g * mfac_of(a.mfac_mode, 1) * wscale;
        const size_t fo = ((size_t)f * a.nm + 1) * nlat;
        // north: Vt = -sum h S' + i sum h T', Vp = -i sum h S' - sum h T'
        a.F0[fo] = make_double2(gm * (-v[0] - v[3]), gm * (v[2] - v[1]));
        a.F1[fo] = make_double2(gm * (v[1] - v[2]), gm * (-v[0] - v[3]));
        // south (sums with (-1)^l h): Vt = -sum S' - i sum T', Vp = i sum S' - sum T'
        a.F0[fo + nlat - 1] = make_double2(gm * (v[7] - v[4]), gm * (-v[5] - v[6]));
        a.F1[fo + nlat - 1] = make_double2(gm * (-v[5] - v[6]), gm * (v[4] - v[7]));
    }
}

// Analysis: adds the weighted pole rows of order 1 to S_l1, T_l1 (after the
// Legendre analysis wrote them), S += w [dP Vt - i P/s Vp], T += w [i P/s Vt + dP Vp];
// conv (optional): the outputs are in the plan's convention, divide by it.
template <typename OT>
__global__ __launch_bounds__(256) void k_vpole_anal(const LegArgs a, double wpole, const double *conv) {
    const int f = blockIdx.x, nlat = a.nlat;
    const size_t fo = ((size_t)f * a.nm + 1) * nlat;
    const double2 tN = a.F0[fo], pN = a.F1[fo], tS = a.F0[fo + nlat - 1], pS = a.F1[fo + nlat - 1];
    const double gm = a.g * mfac_of(a.mfac_mode, 1) * wpole;
    const size_t co = 2 * ((size_t)f * a.spec_fstride + (size_t)(a.lmax + 1));
    OT *cs = (OT *)a.out0 + co, *ct = (OT *)a.out1 + co;
    const int lt = min(a.ltr, a.lmax);
    for (int l = 1 + (int)threadIdx.x; l <= lt; l += 256) {
        const double h = pole_h(l), hs = (l & 1) ? -h : h;
        const double dN = -h, qN = -h, dS = -hs, qS = hs;  // dP/dtheta and P/sin(theta)
        double sc = gm;
        if (a.ll1_mode) sc /= (double)l * (l + 1);
        if (conv) sc /= conv[(size_t)(a.lmax + 1) + l];
        const double sr = dN * tN.x + qN * pN.y + dS * tS.x + qS * pS.y;
        const double si = dN * tN.y - qN * pN.x + dS * tS.y - qS * pS.x;
        const double tr = -qN * tN.y + dN * pN.x - qS * tS.y + dS * pS.x;
        const double ti = qN * tN.x + dN * pN.y + qS * tS.x + dS * pS.y;
        cs[2 * l] = (OT)((double)cs[2 * l] + sc * sr);
        cs[2 * l + 1] = (OT)((double)cs[2 * l + 1] + sc * si);
        ct[2 * l] = (OT)((double)ct[2 * l] + sc * tr);
        ct[2 * l + 1] = (OT)((double)ct[2 * l + 1] + sc * ti);
    }
}

// ---------------------------------------------------------------------------
// Analysis work items: (global segment, pair block of 512, field).
// ---------------------------------------------------------------------------
struct Item {
    int gs, pbi, im, s, m, l0, lend, lstop, pb0, pb1, pstart, f;
    bool fast;
};

// Zero the outputs of an item that no reduction writes (l beyond ltr, l < m,
// orders between multiples of mres); skipped in accumulate mode (pre-zeroed).
template <typename OT>
__device__ void zero_fill(const LegArgs &a, const Item &it, OT *col, bool all_dead, int lane) {
    if (a.atomic || it.pbi != 0) return;
    const int lz0 = all_dead ? it.l0 : it.lstop;
    for (int l = lz0 + lane; l < it.lend; l += 64) st_c(col, l, 0.0, 0.0);
    if (it.s == 0) {
        for (int l = lane; l < it.m; l += 64) st_c(col, l, 0.0, 0.0);
        for (int c = 1; c < a.mres && it.m + c <= a.mmax; ++c)
            for (int l = lane; l <= a.lmax; l += 64) st_c(col + 2 * (size_t)c * (a.lmax + 1), l, 0.0, 0.0);
    }
}

// Scalar analysis: one wavefront per item (segment, pair block of 512,
// field), K <= 8 chains per lane (pairs pstart + lane + 64 j), Q-form
// recurrence.  Each output coefficient has one writer, so there is no
// inter-wave reduction: every 8 degrees the lanes' 16 partial sums (8 degrees
// x re/im) go through a wave-private LDS tile, lane (o = lane & 15, q =
// lane >> 4) adds the partials of output o from lanes 16q .. 16q+15, two
// shuffle stages finish, and 16 lanes store 8 consecutive coefficients.
// beta_{l+1} and c_l of each 8-degree block arrive in one batch of scalar
// loads from the padded tables.
constexpr int ANW_WAVES = 4;
constexpr int ANW_RS = 65;  // LDS row stride (doubles) of the reduction tile

__device__ __forceinline__ Item make_item_at(const LegArgs &a, int item, int pblk) {
    Item it;
    it.gs = item % a.gx;
    it.pbi = (item / a.gx) % a.gy;
    it.im = as_const(a.segm)[it.gs];
    it.s = it.gs - as_const(a.segbase)[it.im];
    it.m = it.im * a.mres;
    it.l0 = it.m + it.s * a.seg;
    it.lend = min(it.l0 + a.seg, a.lmax + 1);
    it.lstop = min(it.lend, a.ltr + 1);
    it.pb0 = it.pbi * pblk;
    it.pb1 = min(it.pb0 + pblk, a.np);
    it.pstart = max(it.pb0, as_const(a.pfirst)[it.gs]);
    it.fast = it.pstart >= as_const(a.pfast)[it.gs];
    it.f = item / (a.gx * a.gy);
    return it;
}

template <int K, bool FAST, typename OT>
__device__ __forceinline__ void anal_w_run(const LegArgs &a, const Item &it, double *red, OT *col, int lane) {
    const int np = a.np, nlat = a.nlat;
    const double2 *Fi = a.F0 + ((size_t)it.f * a.nm + it.im) * nlat;
    double x[K], prev[K], cur[K], gpr[K], gpi[K], gmr[K], gmi[K];
    int kk[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = it.pstart + lane + 64 * j;
        const bool act = p < it.pb1;
        const int pc = act ? p : 0;
        x[j] = a.x[pc];
        const size_t q = (size_t)it.gs * np + pc;
        prev[j] = act ? a.seedq0[q] : 0.0;
        cur[j] = act ? a.seedq1[q] : 0.0;
        kk[j] = act ? a.seedk[q] : 0;
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
    const double gm = a.g * mfac_of(a.mfac_mode, it.m);
    const int nst = it.lstop - it.l0;
    const long long co = as_const(a.coff)[it.im] - it.m;
    const cptr<double> qcm = as_const(a.qcc) + co;
    const int o = lane & 15, qg = lane >> 4;
    // Checked path: ke = -4096 x scale count, so ldexp(cur, ke) is the lane's
    // term -- the value itself when live, exactly 0 while scaled (|cur| < 2^1024)
    // -- in one instruction; slow has bit j set while chain j holds a scaled
    // lane, and only those chains take the rescale checks.
    unsigned slow = 0;
    if (!FAST) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            kk[j] *= -4096;
            if (__any(kk[j] < 0)) slow |= 1u << j;
        }
    }
    // beta_{l+1} of the item's degrees, staged in the wave's LDS slice and read
    // as broadcasts (no scalar loads, and so no scalar-load waits, in the loop)
    double *bl = red + 16 * ANW_RS;
    for (int i = lane; i < ((nst + 7) & ~7); i += 64) bl[i] = a.qb[co + it.l0 + i + 1];  // padded table
    __builtin_amdgcn_wave_barrier();
    for (int c0 = 0; c0 < nst; c0 += 8) {
        const int lc = __builtin_amdgcn_readfirstlane(it.l0 + c0);
        double bq[8];
#pragma unroll
        for (int st = 0; st < 8; ++st) bq[st] = bl[c0 + st];
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            double pr = 0.0, pi = 0.0;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const double v = FAST ? cur[j] : ldexp(cur[j], kk[j]);
                pr = fma(v, (st & 1) ? gmr[j] : gpr[j], pr);
                pi = fma(v, (st & 1) ? gmi[j] : gpi[j], pi);
                const double nx = fma(x[j], cur[j], -(bq[st] * prev[j]));
                prev[j] = cur[j];
                cur[j] = nx;
            }
            // partial sums go to the wave's LDS tile right away
            red[(2 * st) * ANW_RS + lane] = pr;
            red[(2 * st + 1) * ANW_RS + lane] = pi;
        }
        if (!FAST && slow) {
            const double cl = qcm[lc + 7];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (!(slow & (1u << j))) continue;
                if (kk[j] < 0 && fmax(fabs(cur[j]), fabs(prev[j])) * cl > kHi) {
                    cur[j] *= kDown;
                    prev[j] *= kDown;
                    kk[j] += 4096;
                }
                if (!__any(kk[j] < 0)) slow &= ~(1u << j);
            }
        }
        // wave-private transpose-reduce of the 16 partial sums
        __builtin_amdgcn_wave_barrier();
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) sum += red[o * ANW_RS + 16 * qg + i];
        __builtin_amdgcn_wave_barrier();
        sum += __shfl_xor(sum, 16);
        sum += __shfl_xor(sum, 32);
        const int t = c0 + (o >> 1);
        if (qg == 0 && t < nst) {
            const double v = gm * qcm[it.l0 + t] * sum;
            OT *dst = col + 2 * (size_t)(it.l0 + t) + (o & 1);
            if (a.atomic)
                atomicAdd(dst, (OT)v);
            else
                *dst = (OT)v;
        }
    }
}

template <int K, typename OT>
__device__ __forceinline__ void anal_w_dispatch(const LegArgs &a, const Item &it, double *red, OT *col, int lane) {
    if (it.fast)
        anal_w_run<K, true>(a, it, red, col, lane);
    else
        anal_w_run<K, false>(a, it, red, col, lane);
}

template <typename OT>
__global__ __launch_bounds__(64 * ANW_WAVES) void k_anal_w(const LegArgs a) {
    __shared__ double red_all[ANW_WAVES][16 * ANW_RS + SEG_MAX + 8];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int item = blockIdx.x * ANW_WAVES + w;
    if (item >= a.gx * a.gy * a.gz) return;
    const Item it = make_item_at(a, item, 512);
    OT *col = (OT *)a.out0 + 2 * ((size_t)it.f * a.spec_fstride + (size_t)(a.lmax + 1) * it.m);
    const bool dead = it.lstop <= it.l0 || it.pstart >= it.pb1;
    zero_fill(a, it, col, dead, lane);
    if (dead) return;
    double *red = red_all[w];
    const int keff = (it.pb1 - it.pstart + 63) / 64;
    switch (keff) {
        case 1: anal_w_dispatch<1>(a, it, red, col, lane); break;
        case 2: anal_w_dispatch<2>(a, it, red, col, lane); break;
        case 3: anal_w_dispatch<3>(a, it, red, col, lane); break;
        case 4: anal_w_dispatch<4>(a, it, red, col, lane); break;
        case 5: anal_w_dispatch<5>(a, it, red, col, lane); break;
        case 6: anal_w_dispatch<6>(a, it, red, col, lane); break;
        case 7: anal_w_dispatch<7>(a, it, red, col, lane); break;
        default: anal_w_dispatch<8>(a, it, red, col, lane); break;
    }
}

// Vector analysis, one wavefront per item (segment, pair block of 512,
// field) with up to 8 chains per lane, so no coefficient has two writers.
// Every 2 degrees the lanes' 16 partial sums (2 degrees x SD, SQ, TQ, TD x
// re/im) go through a wave-private LDS tile; lane (o = lane & 15, q = lane >> 4)
// adds 16 of them and two shuffles complete output o = 8 step + quantity; a
// third shuffle pairs each quantity with the one S or T needs beside it.
constexpr int AVW_WAVES = 4;
constexpr int AVW_RS = 65;     // LDS row stride (doubles) of the reduction tile
constexpr int AVW_STEPS = 2;   // degrees per reduction (16 partial sums per lane)

struct AVRec {
    double b, dl, dg;  // beta_{l+1}, l c_l, g_l c_{l-1}
};

// One degree of the vector analysis for K chains; N: the degree's lane in its
// record group (DPP broadcast source), PAR: its parity (l - m).
template <int K, int N, int PAR, bool NOP, bool FAST>
__device__ __forceinline__ void avw_step(const AVRec &r, const double *x, double *prev, double *cur, const int *kk,
                                         const double (*gt)[4], const double (*gp)[4], double *q8) {
    // even l - m: SD <- Gt-, SQ <- Gp+, TQ <- Gt+, TD <- Gp-; odd: the other signs
    constexpr int e = PAR ? 0 : 2, od = PAR ? 2 : 0;
    double xc[K], D[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        xc[j] = x[j] * cur[j];
        D[j] = 0.0;
    }
    if (NOP) asm volatile("s_nop 4");
#pragma unroll
    for (int j = 0; j < K; ++j) SHTK_FMAC_BC(D[j], r.dl, xc[j], N, "");
#pragma unroll
    for (int j = 0; j < K; ++j) SHTK_FMAC_BC(D[j], r.dg, prev[j], N, "-");
    double sdr = 0, sdi = 0, sqr = 0, sqi = 0, tqr = 0, tqi = 0, tdr = 0, tdi = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double Dj = D[j], Q = cur[j];
        if (!FAST) {
            const double cmj = kk[j] == 0 ? 1.0 : 0.0;
            Dj *= cmj;
            Q *= cmj;
        }
        sdr = fma(Dj, gt[j][e], sdr);
        sdi = fma(Dj, gt[j][e + 1], sdi);
        sqr = fma(Q, gp[j][od], sqr);
        sqi = fma(Q, gp[j][od + 1], sqi);
        tqr = fma(Q, gt[j][od], tqr);
        tqi = fma(Q, gt[j][od + 1], tqi);
        tdr = fma(Dj, gp[j][e], tdr);
        tdi = fma(Dj, gp[j][e + 1], tdi);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) SHTK_FMAC_BC(xc[j], r.b, prev[j], N, "-");
#pragma unroll
    for (int j = 0; j < K; ++j) {
        prev[j] = cur[j];
        cur[j] = xc[j];
    }
    q8[0] = sdr;
    q8[1] = sdi;
    q8[2] = sqr;
    q8[3] = sqi;
    q8[4] = tqr;
    q8[5] = tqi;
    q8[6] = tdr;
    q8[7] = tdi;
}

// Degrees c16 + N, c16 + N + 1 (N even) of a record group, their reduction and
// stores, then the next pair; stops at nst.  Scaled chains are checked every 8
// degrees with c_l of the last degree.
template <int K, bool FAST, int N, typename OT>
__device__ __forceinline__ void avw_pairs(const LegArgs &a, const Item &it, const AVRec &r, cptr<double> qcm,
                                          double *red, OT *cs, OT *ct, int lane, int c16, int nst, const double *x,
                                          double *prev, double *cur, int *kk, const double (*gt)[4],
                                          const double (*gp)[4], double gm, double mm) {
    if constexpr (N < 16) {
        const int c0 = c16 + N;
        if (c0 >= nst) return;
        double q8[8];
        avw_step<K, N, 0, N == 0 || (!FAST && N == 8), FAST>(r, x, prev, cur, kk, gt, gp, q8);
        double *rw = red + lane;
#pragma unroll
        for (int q = 0; q < 8; ++q) rw[q * AVW_RS] = q8[q];
        avw_step<K, N + 1, 1, false, FAST>(r, x, prev, cur, kk, gt, gp, q8);
#pragma unroll
        for (int q = 0; q < 8; ++q) rw[(8 + q) * AVW_RS] = q8[q];
        if (!FAST && ((N + 1) & 7) == 7) {
            const double cl = qcm[it.l0 + c0 + 1];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (kk[j] > 0 && fmax(fabs(cur[j]), fabs(prev[j])) * cl > kHi) {
                    cur[j] *= kDown;
                    prev[j] *= kDown;
                    kk[j] -= 1;
                }
            }
        }
        // lane (o, qg) adds 16 of the 64 partials of output o = 8 step + quantity
        const int o = lane & 15, qg = lane >> 4, kq = lane & 7, so = o >> 3;
        __builtin_amdgcn_wave_barrier();
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) sum += red[o * AVW_RS + 16 * qg + i];
        __builtin_amdgcn_wave_barrier();
        sum += __shfl_xor(sum, 16);
        sum += __shfl_xor(sum, 32);
        // S = SD - i m c SQ, T = i m c TQ + TD: quantity k pairs with k ^ 3
        const double other = __shfl_xor(sum, 3);
        const int t = c0 + so;
        if (qg == 0 && t < nst && (kq < 2 || kq >= 6)) {
            const int l = it.l0 + t;
            double sc = gm;
            if (a.ll1_mode) sc *= l > 0 ? 1.0 / ((double)l * (l + 1)) : 0.0;
            const double mc = mm * qcm[l];
            const double sg = (kq == 0 || kq == 7) ? 1.0 : -1.0;
            const double v = sc * fma(sg * mc, other, sum);
            OT *dst = (kq < 2 ? cs : ct) + 2 * (size_t)l + (kq & 1);
            if (a.atomic)
                atomicAdd(dst, (OT)v);
            else
                *dst = (OT)v;
        }
        avw_pairs<K, FAST, N + 2>(a, it, r, qcm, red, cs, ct, lane, c16, nst, x, prev, cur, kk, gt, gp, gm, mm);
    }
}

template <int K, bool FAST, typename OT>
__device__ __forceinline__ void anal_vw_run(const LegArgs &a, const Item &it, double *red, OT *cs, OT *ct,
                                            int lane) {
    const int np = a.np, nlat = a.nlat;
    const size_t fo = ((size_t)it.f * a.nm + it.im) * nlat;
    double x[K], prev[K], cur[K];
    double gt[K][4], gp[K][4];  // (+re, +im, -re, -im)
    int kk[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int p = it.pstart + lane + 64 * j;
        const bool act = p < it.pb1;
        const int pc = act ? p : 0;
        x[j] = a.x[pc];
        const size_t q = (size_t)it.gs * np + pc;
        prev[j] = act ? a.seedq0[q] : 0.0;
        cur[j] = act ? a.seedq1[q] : 0.0;
        kk[j] = act ? a.seedk[q] : 0;
        double2 tN = make_double2(0.0, 0.0), tS = tN, pN = tN, pS = tN;
        double wl = 0.0;
        if (act) {
            tN = a.F0[fo + p];
            pN = a.F1[fo + p];
            const int ps = nlat - 1 - p;
            if (ps != p) {
                tS = a.F0[fo + ps];
                pS = a.F1[fo + ps];
            }
            wl = a.wlat ? a.wlat[p] : 1.0;
        }
        gt[j][0] = wl * (tN.x + tS.x);
        gt[j][1] = wl * (tN.y + tS.y);
        gt[j][2] = wl * (tN.x - tS.x);
        gt[j][3] = wl * (tN.y - tS.y);
        gp[j][0] = wl * (pN.x + pS.x);
        gp[j][1] = wl * (pN.y + pS.y);
        gp[j][2] = wl * (pN.x - pS.x);
        gp[j][3] = wl * (pN.y - pS.y);
    }
    const double gm = a.g * mfac_of(a.mfac_mode, it.m);
    const double mm = (double)it.m;
    const int nst = it.lstop - it.l0;
    const long long co = as_const(a.coff)[it.im] - it.m;
    const cptr<double> qcm = as_const(a.qcc) + co;
    // per-degree records in VGPRs, broadcast by DPP (no scalar loads in the loop):
    // lane 16 r + n holds (beta_{l+1}, l c_l, g_l c_{l-1}) of degree l0 + c16 + n
    const double *qbv = a.qb + co;                       // padded past lmax + 1
    const double2 *qdv = a.qd + co;                      // entries l = m .. lmax + 1
    AVRec nr;
    auto load_rec = [&](AVRec &d, int c16) {
        const int l = min(it.l0 + c16 + (lane & 15), a.lmax + 1);
        d.b = qbv[l + 1];
        const double2 dd = qdv[l];
        d.dl = dd.x;
        d.dg = dd.y;
    };
    load_rec(nr, 0);
    for (int c16 = 0; c16 < nst; c16 += 16) {
        const AVRec r = nr;
        if (c16 + 16 < nst) load_rec(nr, c16 + 16);
        avw_pairs<K, FAST, 0>(a, it, r, qcm, red, cs, ct, lane, c16, nst, x, prev, cur, kk, gt, gp, gm, mm);
    }
}

template <int K, typename OT>
__device__ __forceinline__ void anal_vw_dispatch(const LegArgs &a, const Item &it, double *red, OT *cs, OT *ct,
                                                 int lane) {
    if (it.fast)
        anal_vw_run<K, true>(a, it, red, cs, ct, lane);
    else
        anal_vw_run<K, false>(a, it, red, cs, ct, lane);
}

// KMAX: the most chains any item of the launch needs (np / 64 up to 8); a
// kernel compiled for fewer chains allocates fewer VGPRs and runs more waves.
template <typename OT, int KMAX>
__global__ __launch_bounds__(64 * AVW_WAVES, KMAX <= 4 ? 3 : 1) void k_anal_vw(const LegArgs a) {
    __shared__ double red_all[AVW_WAVES][8 * AVW_STEPS * AVW_RS];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int item = blockIdx.x * AVW_WAVES + w;
    if (item >= a.gx * a.gy * a.gz) return;
    const Item it = make_item_at(a, item, a.apblk);
    const size_t co = 2 * ((size_t)it.f * a.spec_fstride + (size_t)(a.lmax + 1) * it.m);
    OT *cs = (OT *)a.out0 + co;
    OT *ct = (OT *)a.out1 + co;
    const bool dead = it.lstop <= it.l0 || it.pstart >= it.pb1;
    zero_fill(a, it, cs, dead, lane);
    zero_fill(a, it, ct, dead, lane);
    if (dead || (a.skip_m0 && it.im == 0)) return;  // order 0: written by the zonal kernel
    double *red = red_all[w];
    const int keff = (it.pb1 - it.pstart + 63) / 64;
    if (KMAX <= 2 || keff <= 2) anal_vw_dispatch<2>(a, it, red, cs, ct, lane);
    else if (KMAX <= 4 || keff <= 4) anal_vw_dispatch<4>(a, it, red, cs, ct, lane);
    else anal_vw_dispatch<8>(a, it, red, cs, ct, lane);
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

// Coefficients of the scalar synthesis in the Q form: out = c_l(m) * conv(l, m)
// * in for l >= m of every order of the plan (other entries are not read by the
// synthesis and are left untouched).  One block per (order, field).
template <typename IT>
__global__ __launch_bounds__(256) void k_spec_qscale(const IT *in, double *out, const double *scale, int divide,
                                                     const double *qcc, const long long *coff, int lmax, int mres,
                                                     int nm, long long per_field) {
    const int im = blockIdx.x % nm, f = blockIdx.x / nm, m = im * mres;
    const double *cm = qcc + coff[im] - m;
    const long long base = (long long)f * per_field + (long long)(lmax + 1) * m;
    auto factor = [&](int l, long long q) {
        double s = cm[l];
        if (scale) {
            const double cs = scale[q - (long long)f * per_field];
            s = divide ? s / cs : s * cs;
        }
        return s;
    };
    // the first 4 x 256 degrees: every load issued before the first store
    constexpr int U = 4;
    double vr[U], vi[U], sc[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const int l = m + (int)threadIdx.x + 256 * k;
        const long long q = base + l;
        const bool ok = l <= lmax;
        vr[k] = ok ? (double)in[2 * q] : 0.0;
        vi[k] = ok ? (double)in[2 * q + 1] : 0.0;
        sc[k] = ok ? factor(l, q) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const int l = m + (int)threadIdx.x + 256 * k;
        const long long q = base + l;
        if (l <= lmax) {
            out[2 * q] = sc[k] * vr[k];
            out[2 * q + 1] = sc[k] * vi[k];
        }
    }
    for (int l = m + (int)threadIdx.x + 256 * U; l <= lmax; l += 256) {
        const long long q = base + l;
        const double s = factor(l, q);
        out[2 * q] = s * (double)in[2 * q];
        out[2 * q + 1] = s * (double)in[2 * q + 1];
    }
}

__global__ void k_zero(unsigned long long *p, size_t n8) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n8) p[q] = 0ull;
}

}  // namespace

hipError_t launch_seed(const SeedArgs &a, hipStream_t st) {
    dim3 grid((a.np + 255) / 256, a.nm);
    hipLaunchKernelGGL(k_seed, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int K>
static void launch_synth_k(LegArgs a, int nfields, int vector, hipStream_t st) {
    a.gx = (a.np + 64 * K - 1) / (64 * K);
    a.gy = a.nm_run > 0 ? min(a.nm_run, a.nm) : a.nm;
    a.gz = nfields;
    const dim3 grid(grid_for_units(a.xcd, a.gx, a.gy * a.gz));
    if (vector) {
        // inputs are the FP64 prologue records of k_vspec_prep
        if (a.seg == DSEG && a.vwaves == 1) {
            if (a.vdirect) hipLaunchKernelGGL((k_synth_vd<K, true, 1>), grid, dim3(64), 0, st, a);
            else hipLaunchKernelGGL((k_synth_vd<K, false, 1>), grid, dim3(64), 0, st, a);
        } else if (a.seg == DSEG && a.vdirect)
            hipLaunchKernelGGL((k_synth_vd<K, true, 4>), grid, dim3(256), 0, st, a);
        else if (a.seg == DSEG)
            hipLaunchKernelGGL((k_synth_vd<K, false, 4>), grid, dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL((k_synth_v<K>), grid, dim3(256), 0, st, a);
    } else {
        if (a.seg == DSEG && (long long)a.gx * a.gy * a.gz >= a.s1w_units)
            hipLaunchKernelGGL((k_synth_d<K, 1>), grid, dim3(64), 0, st, a);
        else if (a.seg == DSEG)
            hipLaunchKernelGGL((k_synth_d<K, DSW>), grid, dim3(64 * DSW), 0, st, a);
        else if (a.io_f32)
            hipLaunchKernelGGL((k_synth_s<K, float>), grid, dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL((k_synth_s<K, double>), grid, dim3(256), 0, st, a);
    }
}

hipError_t launch_leg_synth(const LegArgs &a, int nfields, int vector, int chains, hipStream_t st) {
    if (a.seg > SEG_MAX || (a.seg & 7)) return hipErrorInvalidValue;
    if (vector) {
        if (chains == 1) launch_synth_k<1>(a, nfields, 1, st);
        else launch_synth_k<VS_CHAINS>(a, nfields, 1, st);
    }
    else if (chains <= 1) launch_synth_k<1>(a, nfields, 0, st);
    else if (chains == 2) launch_synth_k<2>(a, nfields, 0, st);
    else launch_synth_k<4>(a, nfields, 0, st);
    return hipGetLastError();
}

hipError_t launch_leg_anal(const LegArgs &a0, int nfields, int vector, hipStream_t st) {
    if (a0.seg > SEG_MAX || (a0.seg & 7)) return hipErrorInvalidValue;
    LegArgs a = a0;
    if (!vector || a.apblk != 256) a.apblk = 512;  // pairs per work item
    const int pblk = a.apblk;
    a.gx = a.nseg_total;
    a.gy = (a.np + pblk - 1) / pblk;
    a.gz = nfields;
    const dim3 grid((unsigned)a.gx * a.gy * nfields);
    if (vector) {
        const dim3 gw((unsigned)(((long long)a.gx * a.gy * nfields + AVW_WAVES - 1) / AVW_WAVES));
        const int kmax = (std::min(a.np, pblk) + 63) / 64;
        if (kmax <= 2) {
            if (a.io_f32) hipLaunchKernelGGL((k_anal_vw<float, 2>), gw, dim3(64 * AVW_WAVES), 0, st, a);
            else hipLaunchKernelGGL((k_anal_vw<double, 2>), gw, dim3(64 * AVW_WAVES), 0, st, a);
        } else if (kmax <= 4) {
            if (a.io_f32) hipLaunchKernelGGL((k_anal_vw<float, 4>), gw, dim3(64 * AVW_WAVES), 0, st, a);
            else hipLaunchKernelGGL((k_anal_vw<double, 4>), gw, dim3(64 * AVW_WAVES), 0, st, a);
        } else {
            if (a.io_f32) hipLaunchKernelGGL((k_anal_vw<float, 8>), gw, dim3(64 * AVW_WAVES), 0, st, a);
            else hipLaunchKernelGGL((k_anal_vw<double, 8>), gw, dim3(64 * AVW_WAVES), 0, st, a);
        }
    } else {
        const dim3 gw((unsigned)(((long long)a.gx * a.gy * nfields + ANW_WAVES - 1) / ANW_WAVES));
        if (a.io_f32)
            hipLaunchKernelGGL(k_anal_w<float>, gw, dim3(64 * ANW_WAVES), 0, st, a);
        else
            hipLaunchKernelGGL(k_anal_w<double>, gw, dim3(64 * ANW_WAVES), 0, st, a);
    }
    return hipGetLastError();
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

hipError_t launch_spec_qscale(const void *in, double *out, int in_f32, const double *scale, int divide,
                              const double *qcc, const long long *coff, int lmax, int mres, int nm,
                              long long per_field, int nfields, hipStream_t st) {
    const dim3 grid((unsigned)nm * nfields);
    if (in_f32)
        hipLaunchKernelGGL(k_spec_qscale<float>, grid, dim3(256), 0, st, (const float *)in, out, scale, divide, qcc,
                           coff, lmax, mres, nm, per_field);
    else
        hipLaunchKernelGGL(k_spec_qscale<double>, grid, dim3(256), 0, st, (const double *)in, out, scale, divide,
                           qcc, coff, lmax, mres, nm, per_field);
    return hipGetLastError();
}

hipError_t launch_vspec_prep(const void *S, const void *T, double *rec, int in_f32, const double *scale, int divide,
                             int ll1, const double *qcc, const long long *coff, int lmax, int mres, int nm,
                             long long per_field, int nfields, hipStream_t st) {
    const dim3 grid((unsigned)nm * nfields);
    if (in_f32)
        hipLaunchKernelGGL(k_vspec_prep<float>, grid, dim3(256), 0, st, (const float *)S, (const float *)T, rec, scale,
                           divide, ll1, qcc, coff, lmax, mres, nm, per_field);
    else
        hipLaunchKernelGGL(k_vspec_prep<double>, grid, dim3(256), 0, st, (const double *)S, (const double *)T, rec,
                           scale, divide, ll1, qcc, coff, lmax, mres, nm, per_field);
    return hipGetLastError();
}

hipError_t launch_vpole_synth(const LegArgs &a, int nfields, int robert, double wscale, hipStream_t st) {
    hipLaunchKernelGGL(k_vpole_synth, dim3((unsigned)nfields), dim3(256), 0, st, a, robert, wscale);
    return hipGetLastError();
}

hipError_t launch_vpole_anal(const LegArgs &a, int nfields, double wpole, const double *conv, hipStream_t st) {
    if (a.io_f32)
        hipLaunchKernelGGL(k_vpole_anal<float>, dim3((unsigned)nfields), dim3(256), 0, st, a, wpole, conv);
    else
        hipLaunchKernelGGL(k_vpole_anal<double>, dim3((unsigned)nfields), dim3(256), 0, st, a, wpole, conv);
    return hipGetLastError();
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t st) {
    const size_t n8 = bytes / 8;
    dim3 grid((unsigned)((n8 + 255) / 256));
    hipLaunchKernelGGL(k_zero, grid, dim3(256), 0, st, (unsigned long long *)p, n8);
    return hipGetLastError();
}

}  // namespace shtk

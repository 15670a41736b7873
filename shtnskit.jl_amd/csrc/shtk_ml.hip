// Single-order transforms: the Legendre sums of one stored order between a
// coefficient column (degrees m..ltr) and one Fourier column (nlat latitudes),
// with no phi transform -- analysis_packed_ml / synthesis_packed_ml
// (src/transforms.jl:351-437) and analysis_sphtor_ml / synthesis_sphtor_ml
// (src/sphtor_transforms.jl:1188-1322).
//
// One lane per latitude pair (north p, south nlat-1-p).  Each lane runs the
// orthonormal P-form recurrence P_l = a_l x P_{l-1} - b_l P_{l-2} from the
// plan's scaled sectoral seed (value * 2^(-512 k), live at k = 0, rescaled by
// 2^-512 when it passes 2^100 every 8 degrees), and D_l = s dP_l/dtheta =
// l x P_l - g_l P_{l-1}.  South values follow from the parity (-1)^(l-m).
// Vector rows use dP/dtheta = D/s and P/s; at exact poles (s = 0) order 1 takes
// the limits -h_l (north), h_l = sqrt((2l+1) l (l+1) / (4 pi)) / 2, and other
// orders 0 (src/kernels.jl:72-99).  The per-degree operands come through LDS in
// chunks of 64 degrees.  Analysis sums reduce over the wavefront and are stored
// as per-block FP64 partials; a second kernel adds them, applies the convention
// map (and 1/(l(l+1)) for the vector pair) and writes the output dtype.
#include "shtk_internal.h"

namespace shtk {

namespace {

constexpr double kHiM = 0x1p100;
constexpr double kDownM = 0x1p-512;

__device__ __forceinline__ double ml_pole_h(int l) {
    return 0.5 * 0.28209479177387814 * sqrt((2.0 * l + 1.0) * (double)l * (double)(l + 1));
}

// Recurrence state of one lane; step() moves from degree l to l + 1 with the
// coefficients co = (a_{l+1}, b_{l+1}).
struct MlRec {
    double x, cur, prev;
    int k;
    __device__ __forceinline__ void step(double2 co, int l, int m) {
        const double nx = fma(co.x * x, cur, -(co.y * prev));
        prev = cur;
        cur = nx;
        if (k > 0 && ((l - m) & (SCALE_CHECK - 1)) == SCALE_CHECK - 1 && fmax(fabs(cur), fabs(prev)) > kHiM) {
            cur *= kDownM;
            prev *= kDownM;
            --k;
        }
    }
};

template <typename T>
__device__ __forceinline__ double2 ld_c(const T *p, size_t i) {
    return make_double2((double)p[2 * i], (double)p[2 * i + 1]);
}

template <typename T>
__device__ __forceinline__ void st_c2(T *p, size_t i, double re, double im) {
    p[2 * i] = (T)re;
    p[2 * i + 1] = (T)im;
}

// Order 0: dP_l/dtheta from the differentiated recurrence d_l = sqrt((2l+1)/(2l-1))
// (x d_{l-1} - l s P_{l-1}) (_zonal_dtheta_row!, src/legendre.jl:331-338), which
// keeps its accuracy near the poles where l x P_l - g_l P_{l-1} cancels.  With
// g_l = l sqrt((2l+1)/(2l-1)) of order 0: d_l = (g_l / l) x d_{l-1} - g_l s P_{l-1};
// gl = g_l / l comes staged with the chunk.
__device__ __forceinline__ double ml_zonal_step(double g, double gl, double x, double s, double Pm1, double d) {
    return fma(gl * x, d, -(g * s * Pm1));
}

// (dn, qn): dP/dtheta and P/sin(theta) at the north node of the lane (Robert
// form synthesis: D and P themselves, zero at the poles); dz: the order-0
// derivative of this degree; g: g_l of the order.
__device__ __forceinline__ void ml_vrow(int m, int l, double g, double P, double Pm1, double x, double s,
                                        double invs, double dz, bool pole, bool robert_syn, double &dn,
                                        double &qn) {
    const double D = m == 0 ? s * dz : fma((double)l * x, P, -(g * Pm1));
    if (m == 0 && !robert_syn) {
        dn = dz;
        qn = P * invs;
        return;
    }
    if (pole) {
        const double v = (m == 1 && !robert_syn) ? -ml_pole_h(l) : 0.0;
        dn = v;
        qn = v;
    } else if (robert_syn) {
        dn = D;
        qn = P;
    } else {
        dn = D * invs;
        qn = P * invs;
    }
}

// Per-degree operands of a chunk of ML_CH degrees, staged in LDS by the block's
// wave (one load per lane instead of a dependent global load per degree).
constexpr int ML_CH = 64;
struct MlStage {
    double2 ab[ML_CH];  // (a_{l+1}, b_{l+1})
    double g[ML_CH];    // g_l
    double gl[ML_CH];   // g_l / l (order 0)
    double c[ML_CH];    // convention factor (synthesis)
    double2 q0[ML_CH], q1[ML_CH];  // coefficients (synthesis)
};

template <typename IT>
__device__ __forceinline__ void ml_stage(const MlArgs &a, MlStage &S, int c0, int n, const IT *in0, const IT *in1,
                                         bool coefs) {
    const int t = threadIdx.x;
    if (t < n) {
        const int l = c0 + t, m = a.m;
        S.ab[t] = a.ab[l + 1 - m];  // l + 1 <= lmax + 1: inside the order's table
        S.g[t] = a.gtab[l - m];
        S.gl[t] = l > 0 ? S.g[t] / (double)l : 0.0;
        if (coefs) {
            S.c[t] = a.conv ? a.conv[l] : 1.0;
            S.q0[t] = ld_c(in0, (size_t)(l - m));
            if (a.vector) S.q1[t] = ld_c(in1, (size_t)(l - m));
        }
    }
}

}  // namespace

template <typename IT, typename OT>
__global__ __launch_bounds__(64) void k_ml_synth(const MlArgs a) {
    __shared__ MlStage S;
    const int p0 = blockIdx.x * 64 + threadIdx.x, f = blockIdx.y;
    const bool act = p0 < a.np;
    const int p = act ? p0 : 0;
    const int m = a.m;
    MlRec r{a.x[p], act ? a.sectv[p] : 0.0, 0.0, act ? a.sectk[p] : 0};
    const double s = a.s[p];
    const bool pole = s == 0.0;
    const double invs = pole ? 0.0 : 1.0 / s;
    const IT *in0 = (const IT *)a.in0 + 2 * (size_t)f * a.ld_spec;
    const IT *in1 = a.vector ? (const IT *)a.in1 + 2 * (size_t)f * a.ld_spec : nullptr;
    double2 n0 = make_double2(0.0, 0.0), s0 = n0, n1 = n0, s1 = n0;
    const double mm = (double)m;
    double dz = 0.0;
    for (int c0 = m; c0 <= a.ltr; c0 += ML_CH) {
        const int n = min(ML_CH, a.ltr + 1 - c0);
        __syncthreads();
        ml_stage(a, S, c0, n, in0, in1, true);
        __syncthreads();
#pragma unroll 4
        for (int t = 0; t < n; ++t) {
            const int l = c0 + t;
            const double live = r.k == 0 ? 1.0 : 0.0;
            const double P = r.cur * live, Pm1 = r.prev * live;
            const double g = S.g[t];
            if (m == 0 && l > 0) dz = ml_zonal_step(g, S.gl[t], r.x, s, Pm1, dz);
            const double c = S.c[t];
            const double sg = ((l - m) & 1) ? -1.0 : 1.0;
            if (!a.vector) {
                const double2 q = S.q0[t];
                const double v = c * P;
                n0.x = fma(v, q.x, n0.x);
                n0.y = fma(v, q.y, n0.y);
                s0.x = fma(sg * v, q.x, s0.x);
                s0.y = fma(sg * v, q.y, s0.y);
            } else if (l >= 1) {
                double dn, qn;
                ml_vrow(m, l, g, P, Pm1, r.x, s, invs, dz, pole, a.robert, dn, qn);
                const double2 Sq = S.q0[t], Tq = S.q1[t];
                const double Sr = c * Sq.x, Si = c * Sq.y, Tr = c * Tq.x, Ti = c * Tq.y;
                const double mq = mm * qn;
                // Vt = d S - i m q T, Vp = i m q S + d T; south: d -> -sg d, q -> sg q
                n0.x += dn * Sr + mq * Ti;
                n0.y += dn * Si - mq * Tr;
                n1.x += -mq * Si + dn * Tr;
                n1.y += mq * Sr + dn * Ti;
                s0.x += sg * (-dn * Sr + mq * Ti);
                s0.y += sg * (-dn * Si - mq * Tr);
                s1.x += sg * (-mq * Si - dn * Tr);
                s1.y += sg * (mq * Sr - dn * Ti);
            }
            if (l < a.ltr) r.step(S.ab[t], l, m);
        }
    }
    if (!act) return;
    const double g = a.g;  // Robert form: the sums above already carry sin(theta)
    OT *o0 = (OT *)a.out0 + 2 * (size_t)f * a.ld_grid;
    OT *o1 = a.vector ? (OT *)a.out1 + 2 * (size_t)f * a.ld_grid : nullptr;
    const int ps = a.nlat - 1 - p;
    st_c2(o0, p, g * n0.x, g * n0.y);
    if (a.vector) st_c2(o1, p, g * n1.x, g * n1.y);
    if (ps != p && ps < a.nrow) {
        st_c2(o0, ps, g * s0.x, g * s0.y);
        if (a.vector) st_c2(o1, ps, g * s1.x, g * s1.y);
    }
}

// Analysis: the lanes' terms of 16 degrees go to an LDS tile [4 * degree + q][lane];
// lane r then adds row r and stores it as this block's partial, acc[field][block]
// [l - m][4]; k_ml_fin adds the blocks' partials.
template <typename IT>
__global__ __launch_bounds__(64) void k_ml_anal(const MlArgs a) {
    __shared__ MlStage S;
    __shared__ double red[64][65];
    const int p0 = blockIdx.x * 64 + threadIdx.x, f = blockIdx.y;
    const bool act = p0 < a.np;
    const int p = act ? p0 : 0;
    const int m = a.m, nl = a.ltr - m + 1;
    MlRec r{a.x[p], act ? a.sectv[p] : 0.0, 0.0, act ? a.sectk[p] : 0};
    const double s = a.s[p];
    const bool pole = s == 0.0;
    const double invs = pole ? 0.0 : 1.0 / s;
    const int ps = a.nlat - 1 - p;
    const double wl = act ? a.w[p] : 0.0;
    // weighted inputs at the north and south rows (Robert form: divided by s off the poles)
    const double rs = (a.vector && a.robert && !pole) ? invs : 1.0;
    double2 vN0 = make_double2(0.0, 0.0), vS0 = vN0, vN1 = vN0, vS1 = vN0;
    {
        const IT *in0 = (const IT *)a.in0 + 2 * (size_t)f * a.ld_grid;
        const bool south = ps != p && ps < a.nrow;
        vN0 = ld_c(in0, p);
        if (south) vS0 = ld_c(in0, ps);
        if (a.vector) {
            const IT *in1 = (const IT *)a.in1 + 2 * (size_t)f * a.ld_grid;
            vN1 = ld_c(in1, p);
            if (south) vS1 = ld_c(in1, ps);
        }
        const double sc = wl * rs;
        vN0.x *= sc; vN0.y *= sc; vS0.x *= sc; vS0.y *= sc;
        vN1.x *= sc; vN1.y *= sc; vS1.x *= sc; vS1.y *= sc;
    }
    double *acc = a.acc + 4 * ((size_t)f * gridDim.x + blockIdx.x) * nl;
    const double mm = (double)m;
    double dz = 0.0;
    int cfirst = m, clast = a.ltr;
    if (a.zseed) {
        // order 0 in segments: this block's segment blockIdx.z from its start values
        const int sgm = blockIdx.z;
        const size_t zq = (size_t)sgm * a.np + p;
        r.prev = a.zseed[zq];
        r.cur = a.zseed[(size_t)a.zsegs * a.np + zq];
        dz = a.zseed[(size_t)2 * a.zsegs * a.np + zq];
        cfirst = sgm * ZSEG;
        clast = min(a.ltr, cfirst + ZSEG - 1);
    }
    for (int c0 = cfirst; c0 <= clast; c0 += ML_CH) {
        const int n = min(ML_CH, clast + 1 - c0);
        __syncthreads();
        ml_stage<IT>(a, S, c0, n, nullptr, nullptr, false);
        __syncthreads();
#pragma unroll 4
        for (int t = 0; t < n; ++t) {
            const int l = c0 + t;
            const double live = r.k == 0 ? 1.0 : 0.0;
            const double P = r.cur * live, Pm1 = r.prev * live;
            const double g = S.g[t];
            if (m == 0 && l > 0) dz = ml_zonal_step(g, S.gl[t], r.x, s, Pm1, dz);
            const double sg = ((l - m) & 1) ? -1.0 : 1.0;
            double v[4];
            if (!a.vector) {
                v[0] = P * (vN0.x + sg * vS0.x);
                v[1] = P * (vN0.y + sg * vS0.y);
                v[2] = v[3] = 0.0;
            } else {
                double dn = 0.0, qn = 0.0;
                if (l >= 1) ml_vrow(m, l, g, P, Pm1, r.x, s, invs, dz, pole, false, dn, qn);
                // north (d, q) and south (-sg d, sg q) rows:
                // S += d Vt - i m q Vp,  T += i m q Vt + d Vp
                const double mq = mm * qn;
                const double tdx = vN0.x - sg * vS0.x, tdy = vN0.y - sg * vS0.y;  // Vt under d
                const double tqx = vN0.x + sg * vS0.x, tqy = vN0.y + sg * vS0.y;  // Vt under q
                const double pdx = vN1.x - sg * vS1.x, pdy = vN1.y - sg * vS1.y;  // Vp under d
                const double pqx = vN1.x + sg * vS1.x, pqy = vN1.y + sg * vS1.y;  // Vp under q
                v[0] = dn * tdx + mq * pqy;
                v[1] = dn * tdy - mq * pqx;
                v[2] = -mq * tqy + dn * pdx;
                v[3] = mq * tqx + dn * pdy;
            }
            const int row = 4 * (t & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) red[row + q][threadIdx.x] = v[q];
            if ((t & 15) == 15 || t == n - 1) {
                __syncthreads();
                const int cnt = 4 * ((t & 15) + 1), lfirst = l - (t & 15);
                if ((int)threadIdx.x < cnt) {
                    double u = 0.0;
#pragma unroll 8
                    for (int i = 0; i < 64; ++i) u += red[threadIdx.x][i];
                    acc[4 * (size_t)(lfirst - m) + threadIdx.x] = u;
                }
                __syncthreads();
            }
            if (l < a.ltr) r.step(S.ab[t], l, m);
        }
    }
}

// ---------------------------------------------------------------------------
// Order 0 of the vector pair in parallel segments of ZSEG degrees.  Order 0 is
// never in the scaled range, so a segment needs only P_{l0-1}, P_{l0} and the
// zonal derivative dP_{l0-1}/dtheta at its start (k_zonal_seed, once per plan);
// from there the kernels run the same two recurrences as k_ml_synth /
// k_ml_anal over the segment alone.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_zonal_seed(const MlArgs a, double *out, int nseg) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.np) return;
    const double x = a.x[p], s = a.s[p];
    double prev = 0.0, cur = a.sectv[p], dz = 0.0;  // order 0: scale count 0
    for (int l = 0; l <= a.lmax; ++l) {
        if (l % ZSEG == 0) {
            const int sg = l / ZSEG;
            out[((size_t)0 * nseg + sg) * a.np + p] = prev;
            out[((size_t)1 * nseg + sg) * a.np + p] = cur;
            out[((size_t)2 * nseg + sg) * a.np + p] = dz;
        }
        const double g = a.gtab[l];
        if (l > 0) dz = ml_zonal_step(g, g / (double)l, x, s, prev, dz);
        if (l < a.lmax) {
            const double2 co = a.ab[l + 1];
            const double nx = fma(co.x * x, cur, -(co.y * prev));
            prev = cur;
            cur = nx;
        }
    }
}

// Synthesis: block of ZW waves per (64 pairs, field); wave w runs segments
// w, w + ZW, ... (operands staged in its own LDS slice), the waves' sums are
// added through LDS at the end.
constexpr int ZW = 8;
template <typename IT, typename OT>
__global__ __launch_bounds__(64 * ZW) void k_zonal_synth(const MlArgs a) {
    __shared__ MlStage Sw[ZW];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p0 = blockIdx.x * 64 + lane, f = blockIdx.y;
    const bool act = p0 < a.np;
    const int p = act ? p0 : 0;
    const double x = a.x[p], s = a.s[p];
    const IT *in0 = (const IT *)a.in0 + 2 * (size_t)f * a.ld_spec;
    const IT *in1 = (const IT *)a.in1 + 2 * (size_t)f * a.ld_spec;
    double2 n0 = make_double2(0.0, 0.0), s0 = n0, n1 = n0, s1 = n0;
    const int nseg = a.ltr / ZSEG + 1;
    MlStage &S = Sw[w];
    for (int sgm = w; sgm < nseg; sgm += ZW) {
        const int l0 = sgm * ZSEG, n = min(ZSEG, a.ltr + 1 - l0);
        const size_t zq = (size_t)sgm * a.np + p;
        double prev = a.zseed[zq], cur = a.zseed[(size_t)a.zsegs * a.np + zq];
        double dz = a.zseed[(size_t)2 * a.zsegs * a.np + zq];
        if (lane < n) {
            const int l = l0 + lane;
            S.ab[lane] = a.ab[l + 1];
            S.g[lane] = a.gtab[l];
            S.gl[lane] = l > 0 ? S.g[lane] / (double)l : 0.0;
            S.c[lane] = a.conv ? a.conv[l] : 1.0;
            S.q0[lane] = ld_c(in0, (size_t)l);
            S.q1[lane] = ld_c(in1, (size_t)l);
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // this wave's staging is in
#pragma unroll 4
        for (int t = 0; t < n; ++t) {
            const int l = l0 + t;
            if (l > 0) dz = ml_zonal_step(S.g[t], S.gl[t], x, s, prev, dz);
            if (l >= 1) {
                // order 0: V_theta = dP/dtheta S, V_phi = dP/dtheta T (Robert form: times
                // sin(theta)); south rows take the parity (-1)^l with the opposite sign
                const double dn = a.robert ? s * dz : dz;
                const double c = S.c[t];
                const double2 Sq = S.q0[t], Tq = S.q1[t];
                const double Sr = c * Sq.x, Si = c * Sq.y, Tr = c * Tq.x, Ti = c * Tq.y;
                const double sg = (l & 1) ? -1.0 : 1.0;
                n0.x += dn * Sr;
                n0.y += dn * Si;
                n1.x += dn * Tr;
                n1.y += dn * Ti;
                s0.x += sg * (-dn * Sr);
                s0.y += sg * (-dn * Si);
                s1.x += sg * (-dn * Tr);
                s1.y += sg * (-dn * Ti);
            }
            const double2 co = S.ab[t];
            const double nx = fma(co.x * x, cur, -(co.y * prev));
            prev = cur;
            cur = nx;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // reads done before restaging
    }
    __syncthreads();
    double *red = (double *)Sw;  // [ZW][8][64] doubles, within the staging slices
    const double v8[8] = {n0.x, n0.y, s0.x, s0.y, n1.x, n1.y, s1.x, s1.y};
#pragma unroll
    for (int q = 0; q < 8; ++q) red[(w * 8 + q) * 64 + lane] = v8[q];
    __syncthreads();
    if (w != 0 || !act) return;
    double t8[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double u = 0.0;
#pragma unroll
        for (int ww = 0; ww < ZW; ++ww) u += red[(ww * 8 + q) * 64 + lane];
        t8[q] = u;
    }
    const double g = a.g;
    OT *o0 = (OT *)a.out0 + 2 * (size_t)f * a.ld_grid;
    OT *o1 = (OT *)a.out1 + 2 * (size_t)f * a.ld_grid;
    const int ps = a.nlat - 1 - p;
    st_c2(o0, p, g * t8[0], g * t8[1]);
    st_c2(o1, p, g * t8[4], g * t8[5]);
    if (ps != p && ps < a.nrow) {
        st_c2(o0, ps, g * t8[2], g * t8[3]);
        st_c2(o1, ps, g * t8[6], g * t8[7]);
    }
}

template <typename OT>
__global__ __launch_bounds__(256) void k_ml_fin(const MlArgs a) {
    const int f = blockIdx.x, m = a.m, nl = a.ltr - m + 1, nb = (a.np + 63) / 64;
    const double *acc = a.acc + 4 * (size_t)f * nb * nl;
    OT *o0 = (OT *)a.out0 + 2 * (size_t)f * a.ld_spec;
    OT *o1 = a.vector ? (OT *)a.out1 + 2 * (size_t)f * a.ld_spec : nullptr;
    for (int t = threadIdx.x; t < nl; t += 256) {
        const int l = m + t;
        double sc = a.g;
        if (a.conv) sc /= a.conv[l];
        if (a.vector) sc = l > 0 ? sc / ((double)l * (l + 1)) : 0.0;
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] += acc[4 * ((size_t)b * nl + t) + q];
        st_c2(o0, t, sc * v[0], sc * v[1]);
        if (a.vector) st_c2(o1, t, sc * v[2], sc * v[3]);
    }
}

hipError_t launch_zonal_seed(const MlArgs &a, double *out, int nseg, hipStream_t st) {
    if (a.m != 0 || nseg != a.lmax / ZSEG + 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zonal_seed, dim3((unsigned)((a.np + 255) / 256)), dim3(256), 0, st, a, out, nseg);
    return hipGetLastError();
}

hipError_t launch_ml_synth(const MlArgs &a, int nfields, hipStream_t st) {
    const dim3 grid((unsigned)((a.np + 63) / 64), (unsigned)nfields);
    if (a.zseed && a.vector && a.m == 0 && a.fourier_f64) {
        if (a.ltr / ZSEG + 1 > a.zsegs) return hipErrorInvalidValue;
        if (a.io_f32) hipLaunchKernelGGL((k_zonal_synth<float, double>), grid, dim3(64 * ZW), 0, st, a);
        else hipLaunchKernelGGL((k_zonal_synth<double, double>), grid, dim3(64 * ZW), 0, st, a);
        return hipGetLastError();
    }
    if (a.io_f32 && a.fourier_f64)
        hipLaunchKernelGGL((k_ml_synth<float, double>), grid, dim3(64), 0, st, a);
    else if (a.fourier_f64)
        hipLaunchKernelGGL((k_ml_synth<double, double>), grid, dim3(64), 0, st, a);
    else if (a.io_f32)
        hipLaunchKernelGGL((k_ml_synth<float, float>), grid, dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL((k_ml_synth<double, double>), grid, dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_ml_anal(const MlArgs &a, int nfields, hipStream_t st) {
    const bool zs = a.zseed && a.vector && a.m == 0;
    if (zs && a.ltr / ZSEG + 1 > a.zsegs) return hipErrorInvalidValue;
    if (a.zseed && !zs) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((a.np + 63) / 64), (unsigned)nfields, zs ? (unsigned)(a.ltr / ZSEG + 1) : 1u);
    if (a.fourier_f64) {
        hipLaunchKernelGGL(k_ml_anal<double>, grid, dim3(64), 0, st, a);
        if (hipPeekAtLastError() != hipSuccess) return hipGetLastError();
        if (a.io_f32)
            hipLaunchKernelGGL(k_ml_fin<float>, dim3((unsigned)nfields), dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(k_ml_fin<double>, dim3((unsigned)nfields), dim3(256), 0, st, a);
    } else if (a.io_f32) {
        hipLaunchKernelGGL(k_ml_anal<float>, grid, dim3(64), 0, st, a);
        if (hipPeekAtLastError() != hipSuccess) return hipGetLastError();
        hipLaunchKernelGGL(k_ml_fin<float>, dim3((unsigned)nfields), dim3(256), 0, st, a);
    } else {
        hipLaunchKernelGGL(k_ml_anal<double>, grid, dim3(64), 0, st, a);
        if (hipPeekAtLastError() != hipSuccess) return hipGetLastError();
        hipLaunchKernelGGL(k_ml_fin<double>, dim3((unsigned)nfields), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace shtk

// Host side of libshtk_mi355x.so: plan construction and the C ABI of include/shtk.h.
//
// A plan is the device-resident counterpart of SHTnsKit's SHTConfig
// (src/config.jl:170-206, create_gauss_config :633-655) for one Gauss grid:
// nodes and weights, the FP64 recurrence coefficients of every order in the
// segmented Q form, per-segment seeds (computed once by a device kernel), the
// coefficient convention map (src/normalization.jl:142-153) and the FFT
// twiddles.  Transforms enqueue the kernels of shtk_legendre.hip and
// shtk_fft.hip on the caller's stream.
#include <algorithm>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "shtk_internal.h"

extern "C" {
static int plan_tables(shtk_plan *p);  // defined with the entry points
}

// Workspace of shtk_synthesis_static: a file-scope device array, one copy per
// GPU at its own address, held until the process ends.  In host code the
// array's name is only a stand-in for the symbol; its device address on each
// GPU comes from hipGetSymbolAddress (static_scratch_setup, run once per device
// by shtk_plan_create with that device current) and only that address is ever
// passed to kernels.  Every call on a GPU shares that GPU's array.
constexpr size_t kStaticScratchBytes = SHTK_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kStaticScratchBytes];

namespace {

thread_local std::string g_err;
thread_local int g_errcode = SHTK_OK;

int fail(int code, const std::string &msg) {
    g_err = msg;
    g_errcode = code;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? SHTK_ENOMEM : SHTK_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                              \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

#define RC_TRY(expr)             \
    do {                         \
        int rc_ = (expr);        \
        if (rc_) return rc_;     \
    } while (0)

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // device address of `scratch` per device (nullptr: not set up)

int static_scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return SHTK_OK;
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "static scratch address");
    g_scratch_ptr[device] = ptr;
    return SHTK_OK;
}

void *static_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

template <typename T>
int upload(T **dst, const std::vector<T> &src, const char *what) {
    if (src.empty()) {
        *dst = nullptr;
        return SHTK_OK;
    }
    HIP_TRY(hipMalloc((void **)dst, sizeof(T) * src.size()), what);
    HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice), what);
    return SHTK_OK;
}

// Gauss-Legendre nodes (ascending) and weights: Newton iteration on P_n from the
// Abramowitz-Stegun initial guess, tolerance 1e-15, at most 50 steps, weights
// 2/((1-z^2) P_n'(z)^2), symmetric fill -- the scheme of src/legendre.jl:956-1003.
void gauss_nodes(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const int half = (n + 1) / 2;
    for (int k = 1; k <= half; ++k) {
        double z = std::cos(M_PI * (k - 0.25) / (n + 0.5));
        auto legendre = [n](double zz, double &pn, double &pnm1) {
            pnm1 = 1.0;
            pn = zz;
            for (int l = 2; l <= n; ++l) {
                const double pnp1 = ((2.0 * l - 1.0) * zz * pn - (l - 1.0) * pnm1) / l;
                pnm1 = pn;
                pn = pnp1;
            }
        };
        double pn, pnm1, pd;
        for (int it = 0; it < 50; ++it) {
            legendre(z, pn, pnm1);
            pd = n * (z * pn - pnm1) / (z * z - 1.0);
            const double z1 = z;
            z -= pn / pd;
            if (std::fabs(z - z1) < 1e-15) break;
        }
        legendre(z, pn, pnm1);
        pd = n * (z * pn - pnm1) / (z * z - 1.0);
        x[k - 1] = -z;
        x[n - k] = z;
        const double wk = 2.0 / ((1.0 - z * z) * pd * pd);
        w[k - 1] = wk;
        w[n - k] = wk;
    }
}

// Fejer's first rule on the midpoint nodes theta_j = (j + 1/2) pi / n (weights in
// x = cos(theta)): w_j = (2/n) (1 - 2 sum_{k=1}^{floor(n/2)} cos(2 k theta_j) / (4k^2 - 1)),
// the formula of _fejer1_weights (src/legendre.jl:1017-1032).  Nodes north first.
void fejer1_nodes(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        // node angle as create_regular_config (src/config.jl:1296), weight angle as the rule
        x[j] = std::cos((j + 0.5) * (M_PI / n));
        const double th = (j + 0.5) * M_PI * (1.0 / n);
        double series = 0.0;
        for (int k = 1; k <= n / 2; ++k) series += std::cos(2.0 * k * th) / (4.0 * k * k - 1.0);
        w[j] = (2.0 / n) * (1.0 - 2.0 * series);
    }
}

// Clenshaw-Curtis rule on theta_j = j pi / N, N = n - 1, both poles included,
// with the parity-dependent endpoint term of _clenshaw_curtis_weights
// (src/legendre.jl:1044-1075).  Nodes north first; the exact poles get x = +-1.
void clenshaw_curtis_nodes(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const int N = n - 1;
    const bool even = (N % 2) == 0;
    const double endpoint = even ? 1.0 / ((double)N * N - 1.0) : 1.0 / ((double)N * N);
    for (int j = 0; j <= N; ++j) {
        const double th = j * M_PI / N;
        x[j] = j == 0 ? 1.0 : (j == N ? -1.0 : std::cos(j * (M_PI / N)));
        if (j == 0 || j == N) {
            w[j] = endpoint;
            continue;
        }
        double v = 1.0;
        const int kmax = even ? N / 2 - 1 : (N - 1) / 2;
        for (int k = 1; k <= kmax; ++k) v -= 2.0 * std::cos(2.0 * k * th) / (4.0 * k * k - 1.0);
        if (even) v -= std::cos(N * th) / ((double)N * N - 1.0);
        w[j] = 2.0 * v / N;
    }
}

// Driscoll-Healy nodes theta_j = pi j / n (j < n) and weights (sqrt(8)/n sqrt(2))
// sin(pi j / n) sum_{l=0}^{n/2-1} sin((2l+1) pi j / n) / (2l+1), the formula of
// driscoll_healy_weights (src/mathutils.jl:149-183) with the nodes of
// create_regular_config (src/config.jl:1267-1273).  Node n, the south pole, is
// appended with weight 0 so that the node set is symmetric about the equator.
void driscoll_healy_nodes(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n + 1, 0.0);
    w.assign(n + 1, 0.0);
    const double nf = std::sqrt(8.0) / n * std::sqrt(2.0);
    for (int j = 0; j < n; ++j) {
        x[j] = std::cos(M_PI * j / n);
        double sum = 0.0;
        for (int l = 0; l <= n / 2 - 1; ++l) sum += std::sin((2.0 * l + 1.0) * M_PI * j / n) / (2.0 * l + 1.0);
        w[j] = nf * std::sin(M_PI * j / n) * sum;
    }
    x[n] = -1.0;
}

double coef_scale(const shtk_desc &d, int l, int m) {
    double nrm = 1.0;
    if (d.norm == SHTK_NORM_FOURPI) nrm = std::sqrt(4.0 * M_PI);
    else if (d.norm == SHTK_NORM_SCHMIDT) nrm = std::sqrt(4.0 * M_PI / (2.0 * l + 1.0));
    const double rs = (d.real_norm && m > 0) ? 1.0 / std::sqrt(2.0) : 1.0;
    const double ph = d.cs_phase ? 1.0 : ((m & 1) ? -1.0 : 1.0);
    return nrm * rs * ph;
}

// Pass radices of one FFT length: 8, 4, 2, 3, 5 first, then the other primes
// up to FFT_GEN_MAX (generic passes, whose sum is returned in *gen_sum).
bool factor_len(int n, std::vector<int> &r, int *gen_sum) {
    r.clear();
    *gen_sum = 0;
    for (int c : {8, 4, 2, 3, 5})
        while (n % c == 0) {
            r.push_back(c);
            n /= c;
        }
    for (int q = 7; n > 1 && q <= shtk::FFT_GEN_MAX; q += 2)
        while (n % q == 0) {
            r.push_back(q);
            *gen_sum += q;
            n /= q;
        }
    return n == 1 && (int)r.size() <= shtk::FFT_MAX_PASSES;
}

// The longitude FFT of length n: direct mixed-radix passes when n's prime
// factors above 5 are few and small, otherwise Bluestein's algorithm with the
// shortest 2^a 3^b 5^c convolution length >= 2n - 1.  FFTW (the reference's
// FFT, src/fftutils.jl:195-238) accepts every length; so does this plan up to
// FFT_MAX_LEN (direct) and FFT_MAX_LEN / 2 (Bluestein).
shtk::FftPlan make_fft_plan(int n) {
    shtk::FftPlan f;
    if (n < 1 || n > shtk::FFT_MAX_LEN) return f;
    int gs = 0;
    if (factor_len(n, f.radix, &gs) && gs <= shtk::FFT_GEN_BUDGET) {
        f.ok = true;
        f.nfft = n;
        return f;
    }
    for (int m = 2 * n - 1; m <= shtk::FFT_MAX_LEN; ++m) {
        int t = m;
        for (int c : {2, 3, 5})
            while (t % c == 0) t /= c;
        if (t != 1 || !factor_len(m, f.radix, &gs)) continue;
        f.ok = f.blue = true;
        f.nfft = m;
        return f;
    }
    f.radix.clear();
    return f;
}

void free_plan(shtk_plan *p) {
    if (!p) return;
    void *ptrs[] = {p->x,      p->s,      p->wscaled, p->winvs,  p->winvs2, p->invs,  p->qc,  p->ab, p->gtab,
                    p->qd,     p->coff,   p->segbase, p->seedq0, p->seedq1, p->seedk, p->pfirst,
                    p->pfast,  p->segm,   p->convscale, p->tw, p->chirp, p->bhat, p->ws,     p->sectv, p->sectk, p->oz_q0,
                    p->oz_q1,  p->oz_qk,  p->oz_ep,     p->oz_chbase, p->oz_ws, p->oz_emaxl, p->oz_epl,
                    p->oz_asc, p->oz_chm, p->oz_run, p->qb, p->qcc, p->zseed};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (p->order_ev) (void)hipEventDestroy(p->order_ev);
    if (p->side_fork) (void)hipEventDestroy(p->side_fork);
    if (p->side_join) (void)hipEventDestroy(p->side_join);
    if (p->side) (void)hipStreamDestroy(p->side);
    free_plan(p->vshadow);
    delete p;
}

int validate(const shtk_desc *d) {
    if (!d) return fail(SHTK_EINVAL, "desc is NULL");
    if (d->lmax < 0) return fail(SHTK_EINVAL, "lmax must be >= 0");
    if (d->mmax < 0 || d->mmax > d->lmax) return fail(SHTK_EINVAL, "mmax must satisfy 0 <= mmax <= lmax");
    if (d->mres < 1) return fail(SHTK_EINVAL, "mres must be >= 1");
    if (d->grid < SHTK_GRID_GAUSS || d->grid > SHTK_GRID_DRISCOLL_HEALY)
        return fail(SHTK_EINVAL, "unknown grid type (gauss, regular, regular with poles, Driscoll-Healy)");
    if (d->grid == SHTK_GRID_DRISCOLL_HEALY && (d->nlat % 2 != 0 || d->nlat != 2 * (d->lmax + 1)))
        return fail(SHTK_EINVAL, "DH weights require nlat=2*(lmax+1)=" + std::to_string(2 * (d->lmax + 1)) +
                                     ", got nlat=" + std::to_string(d->nlat));
    if (d->grid == SHTK_GRID_GAUSS && d->nlat < d->lmax + 1)
        return fail(SHTK_EINVAL, "Gauss-Legendre quadrature requires nlat >= lmax+1");
    if (d->grid == SHTK_GRID_REGULAR && d->nlat < d->lmax + 2)
        return fail(SHTK_EINVAL, "nlat must be >= lmax+2 for regular grids");
    if (d->grid == SHTK_GRID_REGULAR_POLES && d->nlat < std::max(2, d->lmax + 1))
        return fail(SHTK_EINVAL, "nlat must be >= max(2, lmax+1) for regular grids with poles");
    if (d->nlon < 2 * d->mmax + 1) return fail(SHTK_EINVAL, "nlon must be >= 2*mmax+1");
    if (d->norm < 0 || d->norm > 2) return fail(SHTK_EINVAL, "unknown normalization");
    if (d->phi_scale != SHTK_PHI_DFT && d->phi_scale != SHTK_PHI_QUAD)
        return fail(SHTK_EINVAL, "unknown phi scaling mode");
    if (d->nlat + (d->grid == SHTK_GRID_DRISCOLL_HEALY ? 1 : 0) > 8192)
        return fail(SHTK_EUNSUPPORTED, "nlat > 8192 (8191 for Driscoll-Healy grids) is not supported by this backend");
    return SHTK_OK;
}

struct DeviceGuard {
    int prev = -1;
    bool switched = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) switched = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (switched) (void)hipSetDevice(prev);
    }
};

// Workspace layout: [F0][F1][spec0][spec1], each region 256-byte aligned.
struct WsLayout {
    size_t f_bytes, spec_bytes, total;
    size_t zonal_off;  // vector: FP64 sums of the order-0 rewrite of the analysis
};

WsLayout ws_layout(const shtk_plan *p, int nf, int vector) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    WsLayout L;
    L.f_bytes = al(sizeof(double2) * (size_t)nf * p->nm * p->d.nlat);
    // spectral staging: canonical-convention copies (non-canonical plans) and the
    // Q-form coefficients of the scalar synthesis
    L.spec_bytes = al(sizeof(double2) * (size_t)nf * (p->d.lmax + 1) * (p->d.mmax + 1));
    L.total = (vector ? 2 : 1) * (L.f_bytes + L.spec_bytes);
    // vector synthesis records: 8 doubles per coefficient (k_vspec_prep)
    if (vector) L.total = std::max(L.total, 2 * L.f_bytes + 4 * L.spec_bytes);
    L.zonal_off = L.total;
    if (vector) L.total += al(sizeof(double) * 4 * (size_t)nf * ((p->np + 63) / 64) * (p->d.lmax + 1));
    return L;
}

int ensure_ws(shtk_plan *p, size_t bytes) {
    if (p->ws_bytes >= bytes) return SHTK_OK;
    if (p->ws) {
        HIP_TRY(hipDeviceSynchronize(), "workspace resize sync");
        HIP_TRY(hipFree(p->ws), "workspace free");
        p->ws = nullptr;
        p->ws_bytes = 0;
    }
    HIP_TRY(hipMalloc(&p->ws, bytes), "workspace allocation");
    p->ws_bytes = bytes;
    ++p->ws_gen;
    return SHTK_OK;
}

hipEvent_t pool_event(shtk_plan *p) {
    if (!p->event_pool.empty()) {
        hipEvent_t e = p->event_pool.back();
        p->event_pool.pop_back();
        return e;
    }
    // timing markers only: no system-scope fence (it adds microseconds around every kernel)
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

// Runs one kernel launcher; with profiling on, brackets it with events.
template <typename F>
hipError_t timed(shtk_plan *p, const char *name, hipStream_t st, F &&launch) {
    if (!p->profiling) return launch();
    hipEvent_t a = pool_event(p), b = pool_event(p);
    if (a) (void)hipEventRecord(a, st);
    hipError_t e = launch();
    if (b) (void)hipEventRecord(b, st);
    if (a && b) p->pending.push_back({name, a, b});
    return e;
}

// Fork/join of the plan's second stream: fork() returns a stream on which work
// runs after everything enqueued on st so far (the side stream, or st itself
// when it cannot be made, e.g. st is being captured before the side stream
// exists); join() makes st wait for what was enqueued there.  Under graph
// capture the event record/wait pair becomes a fork and a join of the graph.
struct SideStream {
    shtk_plan *p;
    hipStream_t st, s;
    SideStream(shtk_plan *p_, hipStream_t st_) : p(p_), st(st_), s(st_) {
        if (!p->side) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
            hipStream_t ns = nullptr;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (hipStreamCreateWithFlags(&ns, hipStreamNonBlocking) != hipSuccess) return;
            if (hipEventCreateWithFlags(&e0, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess ||
                hipEventCreateWithFlags(&e1, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
                if (e0) (void)hipEventDestroy(e0);
                (void)hipStreamDestroy(ns);
                return;
            }
            p->side = ns;
            p->side_fork = e0;
            p->side_join = e1;
        }
        if (hipEventRecord(p->side_fork, st) != hipSuccess) return;
        if (hipStreamWaitEvent(p->side, p->side_fork, 0) != hipSuccess) return;
        s = p->side;
    }
    bool forked() const { return s != st; }
    hipError_t join() {
        if (s == st) return hipSuccess;
        hipError_t e = hipEventRecord(p->side_join, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, p->side_join, 0);
        s = st;
        return e;
    }
    ~SideStream() { (void)join(); }
};

// Calls on one plan share its workspaces.  The host mutex orders their
// enqueueing; on the device, a call issued on another stream than the previous
// call first waits for the event recorded at the end of that call, so kernels
// of two streams never overlap on the workspaces.  A plan used from one stream
// only records no events (each record is a queue marker between launches): the
// first call from a second stream synchronises the device once, and from then
// on every call records.  Work being captured into a graph skips all of this
// (the graph's owner orders its launches).
struct StreamOrder {
    shtk_plan *p;
    hipStream_t st;
    bool active = false;
    StreamOrder(shtk_plan *p_, hipStream_t s) : p(p_), st(s) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
        active = true;
        if (p->order_used && p->order_st != st) {
            if (p->order_multi) {
                (void)hipStreamWaitEvent(st, p->order_ev, 0);
            } else {
                (void)hipDeviceSynchronize();
                p->order_multi = true;
            }
        }
    }
    ~StreamOrder() {
        if (!active) return;
        p->order_st = st;
        p->order_used = true;
        if (p->order_multi && p->order_ev) (void)hipEventRecord(p->order_ev, st);
    }
    StreamOrder(const StreamOrder &) = delete;
    StreamOrder &operator=(const StreamOrder &) = delete;
};

shtk::LegArgs base_args(const shtk_plan *p, int ltr) {
    shtk::LegArgs a{};
    a.s1w_units = p->ssynth_1w_units;
    a.lmax = p->d.lmax;
    a.mmax = p->d.mmax;
    a.mres = p->d.mres;
    a.nlat = p->d.nlat;
    a.np = p->np;
    a.nm = p->nm;
    a.ltr = ltr;
    a.seg = p->seg;
    a.x = p->x;
    a.qc = p->qc;
    a.qd = p->qd;
    a.qb = p->qb;
    a.qcc = p->qcc;
    a.ab = p->ab;
    a.gtab = p->gtab;
    a.coff = p->coff;
    a.segbase = p->segbase;
    a.segm = p->segm;
    a.nseg_total = p->nseg_total;
    a.seedq0 = p->seedq0;
    a.seedq1 = p->seedq1;
    a.seedk = p->seedk;
    a.pfirst = p->pfirst;
    a.pfast = p->pfast;
    a.spec_fstride = (long long)(p->d.lmax + 1) * (p->d.mmax + 1);
    a.xcd = p->xcd_map;
    return a;
}

shtk::FftArgs fft_args(const shtk_plan *p) {
    shtk::FftArgs f{};
    f.nlat = p->d.nlat;
    f.nlon = p->d.nlon;
    f.np = p->np;
    f.nm = p->nm;
    f.mres = p->d.mres;
    f.mmax = p->d.mmax;
    f.nrad = (int)p->fft.radix.size();
    f.radix = 0;
    for (size_t q = 0; q < p->fft.radix.size(); ++q)
        f.radix |= (unsigned long long)p->fft.radix[q] << (shtk::FFT_RADIX_BITS * q);
    f.tw = p->tw;
    f.nfft = p->fft.nfft;
    f.blue = p->fft.blue ? 1 : 0;
    f.chirp = p->chirp;
    f.bhat = p->bhat;
    f.s = p->s;
    f.tile_pairs = p->fft_tile;
    f.nrow = p->nrow;
    return f;
}

int check_common(shtk_plan *p, int nfields, int dtype, std::initializer_list<const void *> ptrs) {
    if (!p) return fail(SHTK_EINVAL, "plan is NULL");
    if (nfields < 1) return fail(SHTK_EINVAL, "nfields must be >= 1");
    if (dtype != SHTK_F64 && dtype != SHTK_F32) return fail(SHTK_EINVAL, "dtype must be SHTK_F64 or SHTK_F32");
    for (const void *q : ptrs)
        if (!q) return fail(SHTK_EINVAL, "array pointer is NULL");
    if (!p->fft.ok)
        return fail(SHTK_EUNSUPPORTED, "nlon=" + std::to_string(p->d.nlon) + " exceeds the LDS FFT's length limit (" +
                                           std::to_string(shtk::FFT_MAX_LEN) + " with prime factors <= " +
                                           std::to_string(shtk::FFT_GEN_MAX) + ", " +
                                           std::to_string(shtk::FFT_MAX_LEN / 2) + " otherwise)");
    return SHTK_OK;
}

// Contraction precision of the *_ex entry points; the matrix-core paths share the
// split tables, which cover lmax <= 4095 (4094 for the vector forms, which run
// the scalar contraction to degree lmax + 1).
int check_precision(const shtk_plan *p, int precision, int lmax_limit) {
    if (precision != SHTK_PREC_FP64 && precision != SHTK_PREC_SPLIT && precision != SHTK_PREC_FP32)
        return fail(SHTK_EINVAL, "precision must be SHTK_PREC_FP64, SHTK_PREC_SPLIT or SHTK_PREC_FP32");
    if (precision != SHTK_PREC_FP64 && p && p->d.lmax > lmax_limit)
        return fail(SHTK_EUNSUPPORTED, "matrix-core (split or FP32) transforms support lmax <= " +
                                           std::to_string(lmax_limit));
    return SHTK_OK;
}

int norm_ltr(const shtk_plan *p, int ltr, int *out) {
    if (ltr < 0) ltr = p->d.lmax;
    if (ltr > p->d.lmax) return fail(SHTK_EINVAL, "ltr must be <= lmax");
    *out = ltr;
    return SHTK_OK;
}

// Split-precision tables: quarter seeds and per-pair exponents (one kernel,
// first use only) and the chunk offsets of every order.
int ensure_oz_tables(shtk_plan *p, hipStream_t st) {
    if (p->oz_q0) return SHTK_OK;
    const int np = p->np, nm = p->nm, lmax = p->d.lmax;
    std::vector<int> chbase(nm), chm;
    int total = 0;
    for (int im = 0; im < nm; ++im) {
        chbase[im] = total;
        const int n = (lmax - im * p->d.mres + 128) / 128;
        total += n;
        chm.insert(chm.end(), n, im);
    }
    const size_t nq = (size_t)nm * 4 * np;
    HIP_TRY(hipMalloc((void **)&p->oz_q0, sizeof(double) * nq), "split tables");
    HIP_TRY(hipMalloc((void **)&p->oz_q1, sizeof(double) * nq), "split tables");
    HIP_TRY(hipMalloc((void **)&p->oz_qk, sizeof(int) * nq), "split tables");
    HIP_TRY(hipMalloc((void **)&p->oz_ep, sizeof(int) * (size_t)nm * np), "split tables");
    RC_TRY(upload(&p->oz_chbase, chbase, "split tables"));
    RC_TRY(upload(&p->oz_chm, chm, "split tables"));
    p->oz_ch_total = total;
    RC_TRY(upload(&p->oz_emaxl, std::vector<int>(p->qc_len, -100000), "split tables"));
    const size_t npad = p->qc_len + shtk::OZ_TPAD;
    HIP_TRY(hipMalloc((void **)&p->oz_epl, sizeof(int) * npad), "split tables");
    HIP_TRY(hipMalloc((void **)&p->oz_asc, sizeof(double) * npad), "split tables");
    // analysis work runs: up to oz_run_len consecutive segments of one order,
    // orders of equal run count split evenly, lowest order (longest) first
    std::vector<int4> runs;
    for (int im = 0, gs = 0; im < nm; ++im) {
        const int nseg = (lmax - im * p->d.mres + p->seg) / p->seg;
        const int nr = (nseg + p->oz_run_len - 1) / p->oz_run_len;
        for (int r = 0; r < nr; ++r) {
            const int s0 = r * nseg / nr, s1 = (r + 1) * nseg / nr;
            runs.push_back(make_int4(im, gs + s0, s1 - s0, 0));
        }
        gs += nseg;
    }
    RC_TRY(upload(&p->oz_run, runs, "split tables"));
    p->oz_nrun = (int)runs.size();
    shtk::OzSeedArgs sa{};
    sa.lmax = lmax;
    sa.mres = p->d.mres;
    sa.np = np;
    sa.nm = nm;
    sa.x = p->x;
    sa.ab = p->ab;
    sa.coff = p->coff;
    sa.sectv = p->sectv;
    sa.sectk = p->sectk;
    sa.q0 = p->oz_q0;
    sa.q1 = p->oz_q1;
    sa.qk = p->oz_qk;
    sa.ep = p->oz_ep;
    sa.emaxl = p->oz_emaxl;
    HIP_TRY(shtk::launch_oz_seed(sa, st), "split seed kernel");
    shtk::OzAscArgs aa{};
    aa.n = (int)p->qc_len;
    aa.qc = p->qc;
    aa.emaxl = p->oz_emaxl;
    aa.asc = p->oz_asc;
    aa.epl = p->oz_epl;
    HIP_TRY(shtk::launch_oz_asc(aa, st), "split scale kernel");
    HIP_TRY(hipStreamSynchronize(st), "split seed kernel");
    return SHTK_OK;
}

struct OzWs {
    size_t bdig, colsum, eb, total;
};

// Split analysis pair blocks: one workgroup contracts 256 nh pairs (nh = 2
// when np > 256, so that lmax <= 1023 needs one block per order and no
// atomics); ksteps of 64 pairs cover whole blocks, so every LDS position of a
// block has a field operand (zero beyond np).
// pair halves per split-analysis block: 512 pairs (two halves) for grids of more than
// 256 pairs unless SHTK_OZ_NH (1 or 2) chose otherwise at plan creation
int oz_nh(const shtk_plan *p) { return p->oz_nh_force ? p->oz_nh_force : (p->np > 256 ? 2 : 1); }
int oz_npb(const shtk_plan *p) { return (p->np + 256 * oz_nh(p) - 1) / (256 * oz_nh(p)); }
int oz_nks(const shtk_plan *p) { return 4 * oz_nh(p) * oz_npb(p); }

// synthesis: B operands per (field group, chunk); analysis: per (field group,
// order, kstep of 64 pairs) plus an FP64 canonical accumulator when the output
// is not canonical FP64
OzWs oz_ws_layout(const shtk_plan *p, int nfg, bool analysis) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    OzWs L;
    const size_t nks = oz_nks(p);
    if (!analysis) {
        L.bdig = al((size_t)nfg * p->oz_ch_total * shtk::OZ_CHUNK_V4 * 16);
        L.colsum = al(sizeof(unsigned long long) * (size_t)nfg * p->nm * 2 * shtk::OZ_COLS);
        L.eb = al(sizeof(int) * (size_t)nfg * p->nm * shtk::OZ_COLS);
    } else {
        L.bdig = al((size_t)nfg * p->nm * nks * shtk::OZ_CHUNK_V4 * 16);
        L.colsum = al(sizeof(double) * (size_t)nfg * p->nm * nks * 2 * shtk::OZ_COLS);
        L.eb = al(sizeof(int) * (size_t)nfg * p->nm * oz_npb(p) * 2 * shtk::OZ_COLS);
    }
    L.total = L.bdig + L.colsum + L.eb;
    return L;
}

int ensure_oz_ws(shtk_plan *p, size_t bytes) {
    if (p->oz_ws_bytes >= bytes) return SHTK_OK;
    if (p->oz_ws) {
        HIP_TRY(hipDeviceSynchronize(), "split workspace resize sync");
        HIP_TRY(hipFree(p->oz_ws), "split workspace free");
        p->oz_ws = nullptr;
        p->oz_ws_bytes = 0;
    }
    HIP_TRY(hipMalloc(&p->oz_ws, bytes), "split workspace allocation");
    p->oz_ws_bytes = bytes;
    ++p->ws_gen;
    return SHTK_OK;
}

// Split-precision argument blocks (tables and workspaces must be in place).
shtk::OzArgs oz_common_args(const shtk_plan *p, int nf, int ltr, double g) {
    shtk::OzArgs o{};
    o.lmax = p->d.lmax;
    o.mmax = p->d.mmax;
    o.mres = p->d.mres;
    o.nlat = p->d.nlat;
    o.np = p->np;
    o.nm = p->nm;
    o.ltr = ltr;
    o.nf = nf;
    o.nfg = (nf + shtk::OZ_FIELDS - 1) / shtk::OZ_FIELDS;
    o.x = p->x;
    o.ab = p->ab;
    o.coff = p->coff;
    o.spec_fstride = (long long)(p->d.lmax + 1) * (p->d.mmax + 1);
    o.g = g;
    o.xcd = p->xcd_map;
    return o;
}

shtk::OzArgs oz_synth_args(const shtk_plan *p, const void *spec_in, int in_f32, int nf, int ltr, double g,
                           double2 *F0) {
    shtk::OzArgs o = oz_common_args(p, nf, ltr, g);
    o.q0 = p->oz_q0;
    o.q1 = p->oz_q1;
    o.qk = p->oz_qk;
    o.ep = p->oz_ep;
    o.chbase = p->oz_chbase;
    o.chm = p->oz_chm;
    o.ch_total = p->oz_ch_total;
    o.spec_in = spec_in;
    o.in_f32 = in_f32;
    const OzWs W = oz_ws_layout(p, o.nfg, false);
    char *ow = (char *)p->oz_ws;
    o.bdig = (int4 *)ow;
    o.colsum = (unsigned long long *)(ow + W.bdig);
    o.eb = (int *)(ow + W.bdig + W.colsum);
    o.F0 = F0;
    return o;
}

shtk::OzArgs oz_anal_args(const shtk_plan *p, double2 *F0, double *acc, int nf, int ltr, const double *wlat,
                          double g) {
    shtk::OzArgs o = oz_common_args(p, nf, ltr, g);
    o.F0 = F0;
    o.qc = p->qc;
    o.asc = p->oz_asc;
    o.epl = p->oz_epl;
    o.qb = p->qb;
    o.qcc = p->qcc;
    o.seedq0 = p->seedq0;
    o.seedq1 = p->seedq1;
    o.seedk = p->seedk;
    o.segbase = p->segbase;
    o.segm = p->segm;
    o.pfirst = p->pfirst;
    o.seg = p->seg;
    o.nseg_total = p->nseg_total;
    o.nks = oz_nks(p);
    o.nh = oz_nh(p);
    o.npb = oz_npb(p);
    o.run = p->oz_run;
    o.nrun = p->oz_nrun;
    o.wlat = wlat;
    o.spec_out = acc;
    const OzWs W = oz_ws_layout(p, o.nfg, true);
    char *ow = (char *)p->oz_ws;
    o.gdig = (int4 *)ow;
    o.gcolsum = (double *)(ow + W.bdig);
    o.eg = (int *)(ow + W.bdig + W.colsum);
    return o;
}

// Bytes of split workspace for nf fields in both directions (analysis adds an
// FP64 canonical accumulator when the output is not canonical FP64).
size_t oz_ws_bytes(const shtk_plan *p, int nf, bool with_acc) {
    const int nfg = (nf + shtk::OZ_FIELDS - 1) / shtk::OZ_FIELDS;
    const size_t acc = with_acc ? (size_t)nf * (p->d.lmax + 1) * (p->d.mmax + 1) * 16 : 0;
    return std::max(oz_ws_layout(p, nfg, false).total, oz_ws_layout(p, nfg, true).total + acc);
}

// FP32 contraction workspace: the analysis field operand (FP32, k_f32_gprep
// layout: per field group, order, pair block and K half, 4 x 4096 floats) and
// an FP64 canonical accumulator when the output is not canonical FP64.
size_t f32_gbytes(const shtk_plan *p, int nf) {
    const size_t nfg = (nf + shtk::OZ_FIELDS - 1) / shtk::OZ_FIELDS;
    return (nfg * p->nm * oz_npb(p) * oz_nh(p) * 4 * 4096 * sizeof(float) + 255) / 256 * 256;
}

size_t f32_ws_bytes(const shtk_plan *p, int nf, bool with_acc) {
    const size_t acc = with_acc ? (size_t)nf * (p->d.lmax + 1) * (p->d.mmax + 1) * 16 : 0;
    return f32_gbytes(p, nf) + acc;
}

shtk::OzArgs f32_synth_args(const shtk_plan *p, const void *spec_in, int in_f32, int nf, int ltr, double g,
                            double2 *F0) {
    shtk::OzArgs o = oz_common_args(p, nf, ltr, g);
    o.q0 = p->oz_q0;
    o.q1 = p->oz_q1;
    o.qk = p->oz_qk;
    o.ep = p->oz_ep;
    o.spec_in = spec_in;
    o.in_f32 = in_f32;
    o.F0 = F0;
    return o;
}

shtk::OzArgs f32_anal_args(const shtk_plan *p, double2 *F0, double *acc, int nf, int ltr, const double *wlat,
                           double g) {
    shtk::OzArgs o = oz_common_args(p, nf, ltr, g);
    o.F0 = F0;
    o.qc = p->qc;
    o.asc = p->oz_asc;
    o.epl = p->oz_epl;
    o.qb = p->qb;
    o.qcc = p->qcc;
    o.seedq0 = p->seedq0;
    o.seedq1 = p->seedq1;
    o.seedk = p->seedk;
    o.segbase = p->segbase;
    o.segm = p->segm;
    o.pfirst = p->pfirst;
    o.seg = p->seg;
    o.nseg_total = p->nseg_total;
    o.nh = oz_nh(p);
    o.npb = oz_npb(p);
    o.run = p->oz_run;
    o.nrun = p->oz_nrun;
    o.wlat = wlat;
    o.spec_out = acc;
    o.gf32 = (float *)p->oz_ws;
    return o;
}

// One transform direction with its scalings (see the entry points below).
struct Job {
    int vector;     // 0 scalar, 1 spheroidal/toroidal pair
    int conv_div;   // non-canonical plans: spectral data divided (1) or multiplied (0) by the convention map
    const double *wlat;
    double g;
    int mfac_mode, ll1_mode, fft_mode, ltr;
    const char *leg_name, *fft_name;
    int prec = SHTK_PREC_FP64;
    int cplx_out = 0;  // scalar synthesis to a complex grid (real_output = false)
    int cplx_in = 0;   // scalar analysis of a complex grid
    // vector transforms on grids with poles, order-1 pole limits: 1 forward synthesis or
    // analysis (quadrature weight in the analysis), 2 adjoint direction (the weight moves to
    // the synthesis side: adjoint analysis; none on the analysis side: adjoint synthesis)
    int pole_fix = 0;
};

// Order 0 of the FP64 vector transforms, rewritten by the single-order kernels:
// near the poles D_l = l x P_l - g_l P_{l-1} cancels for m = 0 (the terms are ~l P,
// the difference ~sin^2), so dP_l/dtheta of order 0 comes from the differentiated
// recurrence of _zonal_dtheta_row! (src/legendre.jl:331-338) instead.  Columns are
// the plan's Fourier rows of order 0 (field stride nm * nlat) and the order-0
// coefficients (field stride spec_fstride).
shtk::MlArgs zonal_args(const shtk_plan *p, int ltr, long long spec_fstride) {
    shtk::MlArgs A{};
    A.lmax = p->d.lmax;
    A.m = 0;
    A.ltr = ltr;
    A.nlat = p->d.nlat;
    A.np = p->np;
    A.vector = 1;
    A.robert = p->d.robert_form;
    A.x = p->x;
    A.s = p->s;
    A.w = p->wscaled;
    A.ab = p->ab + p->hcoff[0];
    A.gtab = p->gtab + p->hcoff[0];
    A.sectv = p->sectv;
    A.sectk = p->sectk;
    A.ld_spec = spec_fstride;
    A.ld_grid = (long long)p->nm * p->d.nlat;
    A.nrow = p->d.nlat;
    A.fourier_f64 = 1;
    A.zseed = p->zseed;
    A.zsegs = p->zsegs;
    return A;
}

// Order-0 segment starts of the vector pair (shtk::launch_zonal_seed), made on
// the first vector call; without them the order-0 kernels run one serial pass.
int ensure_zseed(shtk_plan *p, hipStream_t st) {
    if (p->zseed) return SHTK_OK;
    const int nseg = p->d.lmax / shtk::ZSEG + 1;
    double *z = nullptr;
    HIP_TRY(hipMalloc((void **)&z, sizeof(double) * 3 * (size_t)nseg * p->np), "order-0 seeds");
    p->zseed = z;  // freed with the plan even if the launch fails
    shtk::MlArgs A = zonal_args(p, p->d.lmax, 0);
    HIP_TRY(shtk::launch_zonal_seed(A, z, nseg, st), "order-0 seed kernel");
    p->zsegs = nseg;
    return SHTK_OK;
}

// spectral -> grid: [convert] -> Legendre synthesis -> inverse phi FFT
// ext_ws: a caller-provided workspace of ext_bytes (nothing is allocated), or
// nullptr for the plan's own (grown on demand).
int run_to_grid(shtk_plan *p, const Job &j, const void *s0, const void *s1, void *g0, void *g1, int nf, int dtype,
                void *stream, void *ext_ws = nullptr, size_t ext_bytes = 0) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    DeviceGuard dg(p->device);
    hipStream_t st = (hipStream_t)stream;
    StreamOrder so(p, st);
    const WsLayout L = ws_layout(p, nf, j.vector);
    if (ext_ws) {
        if (L.total > ext_bytes)
            return fail(SHTK_EINVAL, "the call needs " + std::to_string(L.total) + " bytes of workspace; the static "
                                     "scratch holds " + std::to_string(ext_bytes));
    } else {
        RC_TRY(ensure_ws(p, L.total));
    }
    char *ws = ext_ws ? (char *)ext_ws : (char *)p->ws;
    double2 *F0 = (double2 *)ws, *F1 = (double2 *)(ws + L.f_bytes);
    char *spec = ws + (j.vector ? 2 : 1) * L.f_bytes;
    shtk::LegArgs a = base_args(p, j.ltr);
    a.in0 = s0;
    a.in1 = s1;
    a.io_f32 = dtype == SHTK_F32;
    a.vwaves = p->vsynth_waves;
    if (!j.vector && j.prec == SHTK_PREC_FP64 && p->canonical && !a.io_f32 && p->seg == shtk::DSEG) {
        // scalar FP64 synthesis of canonical FP64 coefficients: the kernel applies
        // c_l (Q form) as it loads them
        a.qscale_in = 1;
    } else if (!j.vector && j.prec == SHTK_PREC_FP64) {
        // scalar FP64 synthesis: coefficients times c_l (Q form), converted to the
        // canonical convention on the way
        double *dst = (double *)spec;
        HIP_TRY(shtk::launch_spec_qscale(s0, dst, a.io_f32, p->canonical ? nullptr : p->convscale, j.conv_div,
                                         p->qcc, p->coff, p->d.lmax, p->d.mres, p->nm, a.spec_fstride, nf, st),
                "coefficient scaling");
        a.in0 = dst;
        a.io_f32 = 0;
    } else if (j.vector && j.prec == SHTK_PREC_FP64 && p->canonical && !a.io_f32 && !j.ll1_mode &&
               p->seg == shtk::DSEG) {
        // vector FP64 synthesis of canonical FP64 S, T: the kernel builds its records
        // (S, T, c_l S, c_l T) as it loads them
        a.vdirect = 1;
    } else if (j.vector && j.prec == SHTK_PREC_FP64) {
        // vector FP64 synthesis: records (S', T', c_l S', c_l T') with the convention
        // map and the l(l+1) factor applied
        double *dst = (double *)spec;
        HIP_TRY(shtk::launch_vspec_prep(s0, s1, dst, a.io_f32, p->canonical ? nullptr : p->convscale, j.conv_div,
                                        j.ll1_mode, p->qcc, p->coff, p->d.lmax, p->d.mres, p->nm, a.spec_fstride,
                                        nf, st),
                "coefficient scaling");
        a.in0 = dst;
        a.in1 = nullptr;
        a.io_f32 = 0;
    } else if (!p->canonical) {
        const void *src[2] = {s0, s1};
        for (int q = 0; q <= j.vector; ++q) {
            void *dst = spec + q * L.spec_bytes;
            HIP_TRY(shtk::launch_spec_scale(src[q], dst, a.io_f32, 0, p->convscale, j.conv_div, p->d.lmax,
                                            p->d.mmax, p->d.mres, a.spec_fstride, nf, st),
                    "coefficient conversion");
            (q == 0 ? a.in0 : a.in1) = dst;
        }
        a.io_f32 = 0;
    }
    a.F0 = F0;
    a.F1 = F1;
    a.wlat = j.wlat;
    a.g = j.g;
    a.mfac_mode = j.mfac_mode;
    a.ll1_mode = j.ll1_mode;
    if (j.prec == SHTK_PREC_SPLIT) {
        RC_TRY(ensure_oz_tables(p, st));
        RC_TRY(ensure_oz_ws(p, oz_ws_bytes(p, nf, !p->canonical || dtype != SHTK_F64)));
        const shtk::OzArgs o = oz_synth_args(p, a.in0, a.io_f32, nf, j.ltr, j.g, F0);
        HIP_TRY(timed(p, "oz_synth", st, [&] { return shtk::launch_oz_synth(o, st); }),
                "split synthesis launch");
    } else if (j.prec == SHTK_PREC_FP32) {
        RC_TRY(ensure_oz_tables(p, st));
        const shtk::OzArgs o = f32_synth_args(p, a.in0, a.io_f32, nf, j.ltr, j.g, F0);
        HIP_TRY(timed(p, "f32_synth", st, [&] { return shtk::launch_f32_synth(o, st); }),
                "FP32 synthesis launch");
    } else {
        // vector: the order-0 rows come from the zonal kernel, on the side stream
        // beside the main kernel (which skips order 0); disjoint output columns
        const bool zonal = j.vector && j.pole_fix == 1 && j.ltr >= 1;
        a.skip_m0 = zonal;
        if (zonal) {
            RC_TRY(ensure_zseed(p, st));
            shtk::MlArgs z = zonal_args(p, j.ltr, a.spec_fstride);
            z.io_f32 = dtype == SHTK_F32;
            z.conv = p->canonical ? nullptr : p->convscale;  // order 0: index l
            z.g = j.g;
            z.in0 = s0;
            z.in1 = s1;
            z.out0 = F0;
            z.out1 = F1;
            SideStream side(p, st);
            HIP_TRY(timed(p, "leg_synth_sphtor_m0", side.s, [&] { return shtk::launch_ml_synth(z, nf, side.s); }),
                    "order-0 rows launch");
            HIP_TRY(timed(p, j.leg_name, st,
                          [&] { return shtk::launch_leg_synth(a, nf, 1, p->vsynth_chains, st); }),
                    "Legendre synthesis launch");
            HIP_TRY(side.join(), "order-0 rows join");
        } else {
            HIP_TRY(timed(p, j.leg_name, st,
                          [&] { return shtk::launch_leg_synth(a, nf, j.vector, j.vector ? p->vsynth_chains : p->synth_chains, st); }),
                    "Legendre synthesis launch");
        }
        const bool zero_rows = j.pole_fix == 1 && p->d.robert_form;  // Robert rows are sin(theta) V
        if (j.pole_fix && p->poles && (zero_rows || (p->d.mres == 1 && p->d.mmax >= 1)))
            HIP_TRY(shtk::launch_vpole_synth(a, nf, zero_rows, j.pole_fix == 2 ? p->hw[0] * p->scale_phi : 1.0, st),
                    "pole rows launch");
    }
    void *grids[2] = {g0, g1};
    for (int q = 0; q <= j.vector; ++q) {
        shtk::FftArgs fa = fft_args(p);
        fa.F = q == 0 ? F0 : F1;
        fa.grid = grids[q];
        fa.io_f32 = dtype == SHTK_F32;
        fa.mode = j.fft_mode;
        if (j.cplx_out)
            HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft_cplx(fa, nf, 1, st); }),
                    "phi synthesis launch");
        else
            HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft(fa, nf, 1, st); }),
                    "phi synthesis launch");
    }
    return SHTK_OK;
}

// grid -> spectral: forward phi FFT -> Legendre analysis -> [convert]
int run_to_spec(shtk_plan *p, const Job &j, const void *g0, const void *g1, void *s0, void *s1, int nf, int dtype,
                void *stream) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    DeviceGuard dg(p->device);
    hipStream_t st = (hipStream_t)stream;
    StreamOrder so(p, st);
    const WsLayout L = ws_layout(p, nf, j.vector);
    RC_TRY(ensure_ws(p, L.total));
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws, *F1 = (double2 *)(ws + L.f_bytes);
    char *spec = ws + (j.vector ? 2 : 1) * L.f_bytes;
    const void *grids[2] = {g0, g1};
    for (int q = 0; q <= j.vector; ++q) {
        shtk::FftArgs fa = fft_args(p);
        fa.F = q == 0 ? F0 : F1;
        fa.grid = (void *)grids[q];
        fa.io_f32 = dtype == SHTK_F32;
        fa.mode = j.fft_mode;
        if (j.cplx_in)
            HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft_cplx(fa, nf, 0, st); }),
                    "phi analysis launch");
        else
            HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft(fa, nf, 0, st); }),
                    "phi analysis launch");
    }
    shtk::LegArgs a = base_args(p, j.ltr);
    a.F0 = F0;
    a.F1 = F1;
    a.wlat = j.wlat;
    a.g = j.g;
    a.mfac_mode = j.mfac_mode;
    a.ll1_mode = j.ll1_mode;
    a.out0 = s0;
    a.out1 = s1;
    a.io_f32 = dtype == SHTK_F32;
    if (!p->canonical) {
        a.out0 = spec;
        a.out1 = spec + L.spec_bytes;
        a.io_f32 = 0;
    }
    if (j.prec == SHTK_PREC_FP32) {
        RC_TRY(ensure_oz_tables(p, st));
        const bool direct = p->canonical && dtype == SHTK_F64;
        RC_TRY(ensure_oz_ws(p, f32_ws_bytes(p, nf, !direct)));
        double *acc = direct ? (double *)s0 : (double *)((char *)p->oz_ws + f32_gbytes(p, nf));
        if (oz_npb(p) > 1)  // several pair blocks accumulate; one block stores every coefficient
            HIP_TRY(shtk::launch_zero(acc, (size_t)nf * a.spec_fstride * 16, st), "output clear");
        const shtk::OzArgs o = f32_anal_args(p, F0, acc, nf, j.ltr, j.wlat, j.g);
        HIP_TRY(timed(p, "f32_anal", st, [&] { return shtk::launch_f32_anal(o, st); }), "FP32 analysis launch");
        if (!direct)
            HIP_TRY(shtk::launch_spec_scale(acc, s0, 0, dtype == SHTK_F32, p->canonical ? nullptr : p->convscale,
                                            j.conv_div, p->d.lmax, p->d.mmax, p->d.mres, a.spec_fstride, nf, st),
                    "coefficient conversion");
        return SHTK_OK;
    }
    if (j.prec == SHTK_PREC_SPLIT) {
        RC_TRY(ensure_oz_tables(p, st));
        const bool direct = p->canonical && dtype == SHTK_F64;
        RC_TRY(ensure_oz_ws(p, oz_ws_bytes(p, nf, !direct)));
        const int nfg = (nf + shtk::OZ_FIELDS - 1) / shtk::OZ_FIELDS;
        double *acc = direct ? (double *)s0 : (double *)((char *)p->oz_ws + oz_ws_layout(p, nfg, true).total);
        if (oz_npb(p) > 1)  // several pair blocks accumulate; one block stores every coefficient
            HIP_TRY(shtk::launch_zero(acc, (size_t)nf * a.spec_fstride * 16, st), "output clear");
        const shtk::OzArgs o = oz_anal_args(p, F0, acc, nf, j.ltr, j.wlat, j.g);
        HIP_TRY(timed(p, "oz_anal", st, [&] { return shtk::launch_oz_anal(o, st); }), "split analysis launch");
        if (!direct)
            HIP_TRY(shtk::launch_spec_scale(acc, s0, 0, dtype == SHTK_F32, p->canonical ? nullptr : p->convscale,
                                            j.conv_div, p->d.lmax, p->d.mmax, p->d.mres, a.spec_fstride, nf, st),
                    "coefficient conversion");
        return SHTK_OK;
    }
    const int pblk = j.vector ? p->vanal_pblk : 512;  // pairs per analysis work item (launch_leg_anal)
    a.apblk = pblk;
    a.atomic = p->np > pblk;
    if (a.atomic) {
        const size_t bytes = (size_t)nf * a.spec_fstride * (a.io_f32 ? 8 : 16);
        void *outs[2] = {a.out0, a.out1};
        for (int q = 0; q <= j.vector; ++q)
            HIP_TRY(shtk::launch_zero(outs[q], bytes, st), "output clear");
    }
    const bool zonal = j.vector && j.pole_fix == 1 && j.ltr >= 1;
    if (zonal) {
        // order 0 from the zonal kernel on the side stream, beside the main kernel
        // (which skips it); the convention map of order 0 follows with the rest
        // (non-canonical plans)
        a.skip_m0 = 1;
        RC_TRY(ensure_zseed(p, st));
        shtk::MlArgs z = zonal_args(p, j.ltr, a.spec_fstride);
        z.io_f32 = a.io_f32;
        z.g = j.g;
        z.in0 = F0;
        z.in1 = F1;
        z.out0 = a.out0;
        z.out1 = a.out1;
        z.acc = (double *)(ws + L.zonal_off);
        SideStream side(p, st);
        HIP_TRY(timed(p, "leg_anal_sphtor_m0", side.s, [&] { return shtk::launch_ml_anal(z, nf, side.s); }),
                "order-0 rows launch");
        HIP_TRY(timed(p, j.leg_name, st, [&] { return shtk::launch_leg_anal(a, nf, 1, st); }),
                "Legendre analysis launch");
        HIP_TRY(side.join(), "order-0 rows join");
    } else {
        HIP_TRY(timed(p, j.leg_name, st, [&] { return shtk::launch_leg_anal(a, nf, j.vector, st); }),
                "Legendre analysis launch");
    }
    if (j.pole_fix && p->poles && p->d.mres == 1 && p->d.mmax >= 1)
        HIP_TRY(shtk::launch_vpole_anal(a, nf, j.pole_fix == 2 ? 1.0 : p->hw[0] * p->scale_phi, nullptr, st),
                "pole rows launch");
    if (!p->canonical) {
        void *dst[2] = {s0, s1};
        for (int q = 0; q <= j.vector; ++q)
            HIP_TRY(shtk::launch_spec_scale(spec + q * L.spec_bytes, dst[q], 0, dtype == SHTK_F32, p->convscale,
                                            j.conv_div, p->d.lmax, p->d.mmax, p->d.mres, a.spec_fstride, nf, st),
                    "coefficient conversion");
    }
    return SHTK_OK;
}

// Companion plan of degree lmax + 1 on the same nodes, canonical convention:
// the split-precision vector transforms run as scalar split transforms on it.
int ensure_vshadow(shtk_plan *p) {
    if (p->vshadow) return SHTK_OK;
    shtk_plan *s = new (std::nothrow) shtk_plan();
    if (!s) return fail(SHTK_ENOMEM, "host allocation failed");
    s->d = p->d;
    s->d.lmax = p->d.lmax + 1;
    s->d.norm = SHTK_NORM_ORTHONORMAL;
    s->d.cs_phase = 1;
    s->d.real_norm = 0;
    s->d.robert_form = 0;
    s->device = p->device;
    s->np = p->np;
    s->nm = p->nm;
    s->seg = p->seg;
    s->xcd_map = p->xcd_map;
    s->synth_chains = p->synth_chains;
    s->vsynth_chains = p->vsynth_chains;
    s->vanal_pblk = p->vanal_pblk;
    s->vsynth_waves = p->vsynth_waves;
    s->ssynth_1w_units = p->ssynth_1w_units;
    s->fft_tile = p->fft_tile;
    s->oz_run_len = p->oz_run_len;
    s->oz_nh_force = p->oz_nh_force;
    s->canonical = true;
    s->inv_scale = p->inv_scale;
    s->scale_phi = p->scale_phi;
    s->hx = p->hx;
    s->hw = p->hw;
    s->nrow = p->nrow;
    s->poles = p->poles;
    s->pole_weights = p->pole_weights;
    const int rc = plan_tables(s);
    if (rc) {
        free_plan(s);
        return rc;
    }
    p->vshadow = s;
    return SHTK_OK;
}

// Split-precision vector transforms (launch_vsplit_coef / launch_vsplit_out):
// sin(theta) V_theta and sin(theta) V_phi are scalar syntheses of degree
// lmax + 1, and the analysis is two scalar analyses with the weights
// w/sin(theta) followed by a three-term combination, so both directions run
// the companion plan's split contraction on 2 nf fields whose Fourier data are
// the contiguous [V_theta fields][V_phi fields] of the workspace.
int run_vsplit_to_grid(shtk_plan *p, const Job &j, const void *S, const void *T, void *Vt, void *Vp, int nf,
                       int dtype, void *stream) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    DeviceGuard dg(p->device);
    hipStream_t st = (hipStream_t)stream;
    StreamOrder so(p, st);
    const WsLayout L = ws_layout(p, nf, 1);
    RC_TRY(ensure_ws(p, L.total));
    RC_TRY(ensure_vshadow(p));
    shtk_plan *s = p->vshadow;
    RC_TRY(ensure_oz_tables(s, st));
    RC_TRY(ensure_oz_ws(s, j.prec == SHTK_PREC_FP32 ? f32_ws_bytes(s, 2 * nf, false)
                                                   : oz_ws_bytes(s, 2 * nf, false)));
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws, *F1 = F0 + (size_t)nf * p->nm * p->d.nlat;
    double *C = (double *)(ws + 2 * L.f_bytes);  // 2 nf (lmax+2)(mmax+1) complex <= the 4 spectral regions
    const long long in_fs = (long long)(p->d.lmax + 1) * (p->d.mmax + 1);
    const long long out_fs = (long long)(p->d.lmax + 2) * (p->d.mmax + 1);
    HIP_TRY(shtk::launch_vsplit_coef(S, T, C, dtype == SHTK_F32, p->canonical ? nullptr : p->convscale, j.conv_div,
                                     p->d.lmax, p->d.mres, p->nm, j.ltr, in_fs, out_fs, nf, st),
            "vector coefficient combination");
    if (j.prec == SHTK_PREC_FP32) {
        const shtk::OzArgs o = f32_synth_args(s, C, 0, 2 * nf, j.ltr + 1, j.g, F0);
        HIP_TRY(timed(p, j.leg_name, st, [&] { return shtk::launch_f32_synth(o, st); }), "FP32 synthesis launch");
    } else {
        const shtk::OzArgs o = oz_synth_args(s, C, 0, 2 * nf, j.ltr + 1, j.g, F0);
        HIP_TRY(timed(p, j.leg_name, st, [&] { return shtk::launch_oz_synth(o, st); }), "split synthesis launch");
    }
    // Low orders: near the poles sin(theta) V is far smaller than its terms there
    // (for m = 0 it vanishes like sin^2), so the division by sin(theta) would lift
    // the split rounding of those rows above 1e-12; the FP64 kernel recomputes the
    // orders m < max(16, (lmax+1)/32) (about 6 % of the work) over the split
    // result, after the split prologues have consumed C.
    const int m_lo = std::max(16, (p->d.lmax + 1) / 32);
    const int nm_lo = std::min(p->nm, (m_lo + p->d.mres - 1) / p->d.mres);
    HIP_TRY(shtk::launch_vspec_prep(S, T, C, dtype == SHTK_F32, p->canonical ? nullptr : p->convscale, j.conv_div,
                                    0, p->qcc, p->coff, p->d.lmax, p->d.mres, nm_lo, in_fs, nf, st),
            "coefficient scaling");
    shtk::LegArgs a = base_args(p, j.ltr);
    a.in0 = C;
    a.F0 = F0;
    a.F1 = F1;
    a.wlat = nullptr;  // sin(theta) V, as the split result
    a.g = j.g;
    a.nm_run = nm_lo;
    HIP_TRY(timed(p, "leg_synth_sphtor_low", st, [&] { return shtk::launch_leg_synth(a, nf, 1, p->vsynth_chains, st); }),
            "Legendre synthesis launch");
    // grids with poles: order-1 pole rows from the limits (the phi transform divides
    // those rows by 1), Robert form: every pole row zero
    if (p->poles && (p->d.robert_form || (p->d.mres == 1 && p->d.mmax >= 1)))
        HIP_TRY(shtk::launch_vpole_synth(a, nf, p->d.robert_form, 1.0, st), "pole rows launch");
    void *grids[2] = {Vt, Vp};
    for (int q = 0; q < 2; ++q) {
        shtk::FftArgs fa = fft_args(p);
        fa.F = q == 0 ? F0 : F1;
        fa.grid = grids[q];
        fa.io_f32 = dtype == SHTK_F32;
        fa.mode = j.fft_mode;
        if (j.cplx_out)
            HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft_cplx(fa, nf, 1, st); }),
                    "phi synthesis launch");
        else
            HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft(fa, nf, 1, st); }),
                    "phi synthesis launch");
    }
    return SHTK_OK;
}

int run_vsplit_to_spec(shtk_plan *p, const Job &j, const void *Vt, const void *Vp, void *S, void *T, int nf,
                       int dtype, void *stream) {
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    DeviceGuard dg(p->device);
    hipStream_t st = (hipStream_t)stream;
    StreamOrder so(p, st);
    const WsLayout L = ws_layout(p, nf, 1);
    RC_TRY(ensure_ws(p, L.total));
    RC_TRY(ensure_vshadow(p));
    shtk_plan *s = p->vshadow;
    RC_TRY(ensure_oz_tables(s, st));
    RC_TRY(ensure_oz_ws(s, j.prec == SHTK_PREC_FP32 ? f32_ws_bytes(s, 2 * nf, false)
                                                   : oz_ws_bytes(s, 2 * nf, false)));
    char *ws = (char *)p->ws;
    double2 *F0 = (double2 *)ws, *F1 = F0 + (size_t)nf * p->nm * p->d.nlat;
    const void *grids[2] = {Vt, Vp};
    for (int q = 0; q < 2; ++q) {
        shtk::FftArgs fa = fft_args(p);
        fa.F = q == 0 ? F0 : F1;
        fa.grid = (void *)grids[q];
        fa.io_f32 = dtype == SHTK_F32;
        fa.mode = j.fft_mode;
        HIP_TRY(timed(p, j.fft_name, st, [&] { return shtk::launch_fft(fa, nf, 0, st); }), "phi analysis launch");
    }
    double *G = (double *)(ws + 2 * L.f_bytes);
    const long long out_fs = (long long)(p->d.lmax + 1) * (p->d.mmax + 1);
    const long long g_fs = (long long)(p->d.lmax + 2) * (p->d.mmax + 1);
    if (oz_npb(s) > 1) HIP_TRY(shtk::launch_zero(G, (size_t)2 * nf * g_fs * 16, st), "analysis sums clear");
    if (j.prec == SHTK_PREC_FP32) {
        const shtk::OzArgs o = f32_anal_args(s, F0, G, 2 * nf, j.ltr + 1, j.wlat, j.g);
        HIP_TRY(timed(p, j.leg_name, st, [&] { return shtk::launch_f32_anal(o, st); }), "FP32 analysis launch");
    } else {
        const shtk::OzArgs o = oz_anal_args(s, F0, G, 2 * nf, j.ltr + 1, j.wlat, j.g);
        HIP_TRY(timed(p, j.leg_name, st, [&] { return shtk::launch_oz_anal(o, st); }), "split analysis launch");
    }
    HIP_TRY(shtk::launch_vsplit_out(G, S, T, dtype == SHTK_F32, p->canonical ? nullptr : p->convscale, j.conv_div,
                                    p->d.lmax, p->d.mmax, p->d.mres, j.ltr, g_fs, out_fs, nf, st),
            "vector coefficient combination");
    // grids with poles: the weights w/sin(theta) dropped the pole rows; order 1 takes
    // their limit terms, added to the outputs in the plan's convention
    if (p->poles && p->d.mres == 1 && p->d.mmax >= 1) {
        shtk::LegArgs a = base_args(p, j.ltr);
        a.F0 = F0;
        a.F1 = F1;
        a.g = j.g;
        a.ll1_mode = j.ll1_mode;
        a.mfac_mode = j.mfac_mode;
        a.out0 = S;
        a.out1 = T;
        a.io_f32 = dtype == SHTK_F32;
        a.spec_fstride = out_fs;
        HIP_TRY(shtk::launch_vpole_anal(a, nf, p->hw[0] * p->scale_phi, p->canonical ? nullptr : p->convscale, st),
                "pole rows launch");
    }
    return SHTK_OK;
}

int segment_length() {
    const char *e = std::getenv("SHTK_SEG");
    int s = e ? std::atoi(e) : 64;
    if (s < 8 || s > shtk::SEG_MAX || (s & 7)) s = 64;
    return s;
}

// Launch-shape knobs read once per plan: SHTK_XCD=1 keeps the pair blocks of
// an order on one XCD in the synthesis kernels, SHTK_SYNTH_CHAINS (1, 2 or 4)
// sets the latitude pairs per lane of the scalar synthesis.
int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// ---------------------------------------------------------------------------
// Uneven batches: several configurations (one plan each), nf[g] fields per
// configuration.  With SHTK_PREC_SPLIT every stage is one launch for all
// groups (the Legendre contractions as persistent work-queue kernels); with
// graphs on, the whole launch sequence of a call is captured once per
// (direction, arrays, dtype, stream) and replayed with one hipGraphLaunch.
// ---------------------------------------------------------------------------
struct BatchGraph {
    int dir, dtype;
    std::vector<const void *> ptrs;
    hipStream_t stream;
    hipGraphExec_t exec;
    std::vector<unsigned long long> gens;  // workspace generation of every plan at capture
};

}  // namespace

struct shtk_batch {
    int ngroups = 0, precision = SHTK_PREC_FP64, use_graph = 0;
    std::vector<shtk_plan *> plans;   // in caller order
    std::vector<int> nf;
    std::vector<int> order;           // groups by descending lmax (work order of the persistent kernels)
    int *counter = nullptr;           // device work counters (synthesis, analysis x 2 pair-block sizes)
    int blocks_synth = 0, blocks_anal = 0;  // persistent kernels' grid sizes (0: two blocks per CU)
    int device = 0, cus = 256;
    std::vector<BatchGraph> graphs;
    std::mutex mu;
};

namespace {

int batch_blocks(const shtk_batch *b, int dir) {
    const int n = dir == 0 ? b->blocks_synth : b->blocks_anal;
    return n > 0 ? n : 2 * b->cus;
}

// The grouped split launches of one direction (tables and workspaces ready).
int batch_split(shtk_batch *b, int dir, const void *const *in, void *const *out, int dtype, hipStream_t st) {
    shtk::OzGroups g1{}, g2{}, g3{};
    const int n = b->ngroups;
    g1.ngroups = g2.ngroups = g3.ngroups = n;
    g3.counter = b->counter;
    shtk::OzGroups gprep[2]{}, units[2]{};
    units[0].counter = b->counter + 1;
    units[1].counter = b->counter + 2;
    std::vector<double *> accs(n);
    if (dir == 0) {
        for (int q = 0; q < n; ++q) {
            const int gi = b->order[q];
            shtk_plan *p = b->plans[gi];
            const int nf = b->nf[gi];
            const WsLayout L = ws_layout(p, nf, 0);
            double2 *F0 = (double2 *)p->ws;
            const void *spec = in[gi];
            int in_f32 = dtype == SHTK_F32;
            if (!p->canonical) {
                void *dst = (char *)p->ws + L.f_bytes;
                HIP_TRY(shtk::launch_spec_scale(spec, dst, in_f32, 0, p->convscale, 0, p->d.lmax, p->d.mmax,
                                                p->d.mres, (long long)(p->d.lmax + 1) * (p->d.mmax + 1), nf, st),
                        "coefficient conversion");
                spec = dst;
                in_f32 = 0;
            }
            const shtk::OzArgs o = oz_synth_args(p, spec, in_f32, nf, p->d.lmax, p->inv_scale / p->d.nlon, F0);
            g1.args[q] = g2.args[q] = g3.args[q] = o;
            g1.prefix[q + 1] = g1.prefix[q] + o.nm * o.nfg;
            g2.prefix[q + 1] = g2.prefix[q] + o.ch_total * o.nfg;
            g3.prefix[q + 1] = g3.prefix[q] + ((o.np + 63) / 64) * o.nm * o.nfg;
        }
        // input dtype is the batch's for every group; converted groups read FP64
        bool any_f32 = false;
        for (int q = 0; q < n; ++q) any_f32 |= g1.args[q].in_f32 != 0;
        if (any_f32)
            for (int q = 0; q < n; ++q)
                if (!g1.args[q].in_f32) return fail(SHTK_EUNSUPPORTED, "split batch: mixed input precisions");
        HIP_TRY(shtk::launch_oz_synth_grp(g1, g2, g3, any_f32, batch_blocks(b, 0), st), "split batch synthesis");
        for (int gi = 0; gi < n; ++gi) {
            shtk_plan *p = b->plans[gi];
            shtk::FftArgs fa = fft_args(p);
            fa.F = (double2 *)p->ws;
            fa.grid = out[gi];
            fa.io_f32 = dtype == SHTK_F32;
            HIP_TRY(shtk::launch_fft(fa, b->nf[gi], 1, st), "phi synthesis launch");
        }
        return SHTK_OK;
    }
    for (int gi = 0; gi < n; ++gi) {
        shtk_plan *p = b->plans[gi];
        shtk::FftArgs fa = fft_args(p);
        fa.F = (double2 *)p->ws;
        fa.grid = (void *)in[gi];
        fa.io_f32 = dtype == SHTK_F32;
        HIP_TRY(shtk::launch_fft(fa, b->nf[gi], 0, st), "phi analysis launch");
    }
    for (int q = 0; q < n; ++q) {
        const int gi = b->order[q];
        shtk_plan *p = b->plans[gi];
        const int nf = b->nf[gi];
        const int nfg = (nf + shtk::OZ_FIELDS - 1) / shtk::OZ_FIELDS;
        const bool direct = p->canonical && dtype == SHTK_F64;
        double *acc = direct ? (double *)out[gi] : (double *)((char *)p->oz_ws + oz_ws_layout(p, nfg, true).total);
        accs[gi] = acc;
        if (oz_npb(p) > 1)
            HIP_TRY(shtk::launch_zero(acc, (size_t)nf * (p->d.lmax + 1) * (p->d.mmax + 1) * 16, st), "output clear");
        const shtk::OzArgs o = oz_anal_args(p, (double2 *)p->ws, acc, nf, p->d.lmax, p->wscaled, 1.0);
        // one launch pair per pair-block size: [0] two halves, [1] one half
        const int k = o.nh == 2 ? 0 : 1;
        shtk::OzGroups &gp = gprep[k], &gu = units[k];
        const int at = gu.ngroups++;
        gp.ngroups = gu.ngroups;
        gp.args[at] = gu.args[at] = o;
        gp.prefix[at + 1] = gp.prefix[at] + o.npb * o.nm * o.nfg;
        gu.prefix[at + 1] = gu.prefix[at] + o.npb * o.nrun * o.nfg;
    }
    HIP_TRY(shtk::launch_oz_anal_grp(gprep, units, batch_blocks(b, 1), st), "split batch analysis");
    for (int gi = 0; gi < n; ++gi) {
        shtk_plan *p = b->plans[gi];
        if (p->canonical && dtype == SHTK_F64) continue;
        HIP_TRY(shtk::launch_spec_scale(accs[gi], out[gi], 0, dtype == SHTK_F32, p->canonical ? nullptr : p->convscale,
                                        1, p->d.lmax, p->d.mmax, p->d.mres,
                                        (long long)(p->d.lmax + 1) * (p->d.mmax + 1), b->nf[gi], st),
                "coefficient conversion");
    }
    return SHTK_OK;
}

// One direction of a batch, eagerly (no graph).
int batch_issue(shtk_batch *b, int dir, const void *const *in, void *const *out, int dtype, hipStream_t st) {
    if (b->precision == SHTK_PREC_SPLIT) return batch_split(b, dir, in, out, dtype, st);
    // FP64 and FP32: one configuration at a time, largest first (a captured graph
    // replays the whole sequence with one launch)
    for (int q = 0; q < b->ngroups; ++q) {
        const int gi = b->order[q];
        const int rc = dir == 0 ? shtk_synthesis_ex(b->plans[gi], in[gi], out[gi], b->nf[gi], dtype, 1, -1,
                                                    b->precision, st)
                                : shtk_analysis_ex(b->plans[gi], in[gi], out[gi], b->nf[gi], dtype, -1, b->precision,
                                                   st);
        if (rc) return rc;
    }
    return SHTK_OK;
}

// Everything a call may allocate or build, done before any capture.
int batch_prepare(shtk_batch *b, int dtype, hipStream_t st) {
    for (int gi = 0; gi < b->ngroups; ++gi) {
        shtk_plan *p = b->plans[gi];
        std::lock_guard<std::recursive_mutex> lk(p->mu);
        RC_TRY(ensure_ws(p, ws_layout(p, b->nf[gi], 0).total));
        if (b->precision == SHTK_PREC_SPLIT) {
            RC_TRY(ensure_oz_tables(p, st));
            RC_TRY(ensure_oz_ws(p, oz_ws_bytes(p, b->nf[gi], !p->canonical || dtype != SHTK_F64)));
        } else if (b->precision == SHTK_PREC_FP32) {
            RC_TRY(ensure_oz_tables(p, st));
            RC_TRY(ensure_oz_ws(p, f32_ws_bytes(p, b->nf[gi], !p->canonical || dtype != SHTK_F64)));
        }
    }
    return SHTK_OK;
}

void batch_drop_graphs(shtk_batch *b) {
    for (BatchGraph &gr : b->graphs) (void)hipGraphExecDestroy(gr.exec);
    b->graphs.clear();
}

int batch_run(shtk_batch *b, int dir, const void *const *in, void *const *out, int dtype, void *stream) {
    if (!b) return fail(SHTK_EINVAL, "batch is NULL");
    if (!in || !out) return fail(SHTK_EINVAL, "array list is NULL");
    if (dtype != SHTK_F64 && dtype != SHTK_F32) return fail(SHTK_EINVAL, "dtype must be SHTK_F64 or SHTK_F32");
    for (int gi = 0; gi < b->ngroups; ++gi)
        if (!in[gi] || !out[gi]) return fail(SHTK_EINVAL, "array pointer is NULL");
    std::lock_guard<std::mutex> lk(b->mu);
    // the group plans stay locked for the whole call (their workspaces are in use,
    // also by a plan handed out through shtk_batch_plan)
    std::vector<std::unique_lock<std::recursive_mutex>> plan_locks;
    for (shtk_plan *p : b->plans) plan_locks.emplace_back(p->mu);
    DeviceGuard dg(b->device);
    hipStream_t st = (hipStream_t)stream;
    if (b->use_graph && !st) return fail(SHTK_EINVAL, "graph capture needs a non-default stream");
    RC_TRY(batch_prepare(b, dtype, st));
    std::vector<std::unique_ptr<StreamOrder>> order;
    for (shtk_plan *p : b->plans) order.emplace_back(new StreamOrder(p, st));
    if (!b->use_graph) return batch_issue(b, dir, in, out, dtype, st);
    std::vector<const void *> key;
    for (int gi = 0; gi < b->ngroups; ++gi) {
        key.push_back(in[gi]);
        key.push_back(out[gi]);
    }
    std::vector<unsigned long long> gens;
    for (shtk_plan *p : b->plans) gens.push_back(p->ws_gen);
    // a workspace moved since a graph was captured: every cached graph holds stale addresses
    for (const BatchGraph &gr : b->graphs)
        if (gr.gens != gens) {
            batch_drop_graphs(b);
            break;
        }
    for (const BatchGraph &gr : b->graphs)
        if (gr.dir == dir && gr.dtype == dtype && gr.stream == st && gr.ptrs == key) {
            HIP_TRY(hipGraphLaunch(gr.exec, st), "graph launch");
            return SHTK_OK;
        }
    for (shtk_plan *p : b->plans) p->profiling = false;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "graph capture begin");
    const int rc = batch_issue(b, dir, in, out, dtype, st);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    HIP_TRY(ec, "graph capture end");
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(ei, "graph instantiation");
    b->graphs.push_back({dir, dtype, key, st, exec, gens});
    HIP_TRY(hipGraphLaunch(exec, st), "graph launch");
    return SHTK_OK;
}

}  // namespace

extern "C" {

int shtk_batch_create(const shtk_desc *descs, const int *nfields, int ngroups, int precision, shtk_batch **out) {
    if (!out) return fail(SHTK_EINVAL, "batch output pointer is NULL");
    *out = nullptr;
    if (!descs || !nfields) return fail(SHTK_EINVAL, "descs and nfields must not be NULL");
    if (ngroups < 1 || ngroups > shtk::OZ_MAX_GROUPS)
        return fail(SHTK_EINVAL, "ngroups must be between 1 and " + std::to_string(shtk::OZ_MAX_GROUPS));
    if (precision != SHTK_PREC_FP64 && precision != SHTK_PREC_SPLIT && precision != SHTK_PREC_FP32)
        return fail(SHTK_EINVAL, "precision must be SHTK_PREC_FP64, SHTK_PREC_SPLIT or SHTK_PREC_FP32");
    for (int g = 0; g < ngroups; ++g) {
        if (nfields[g] < 1) return fail(SHTK_EINVAL, "every group needs at least one field");
        if (precision != SHTK_PREC_FP64 && descs[g].lmax > 4095)
            return fail(SHTK_EUNSUPPORTED, "matrix-core (split or FP32) transforms support lmax <= 4095");
    }
    shtk_batch *b = new (std::nothrow) shtk_batch();
    if (!b) return fail(SHTK_ENOMEM, "host allocation failed");
    b->ngroups = ngroups;
    b->precision = precision;
    (void)hipGetDevice(&b->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    for (int g = 0; g < ngroups; ++g) {
        shtk_plan *p = nullptr;
        const int rc = shtk_plan_create(&descs[g], &p);
        if (rc) {
            shtk_batch_destroy(b);
            return rc;
        }
        b->plans.push_back(p);
        b->nf.push_back(nfields[g]);
        b->order.push_back(g);
    }
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int x, int y) { return b->plans[x]->d.lmax > b->plans[y]->d.lmax; });
    if (hipMalloc((void **)&b->counter, 3 * sizeof(int)) != hipSuccess) {
        shtk_batch_destroy(b);
        return fail(SHTK_ENOMEM, "batch counter allocation failed");
    }
    *out = b;
    return SHTK_OK;
}

int shtk_batch_destroy(shtk_batch *b) {
    if (!b) return SHTK_OK;
    {
        DeviceGuard dg(b->device);
        (void)hipDeviceSynchronize();
        batch_drop_graphs(b);
        if (b->counter) (void)hipFree(b->counter);
    }
    for (shtk_plan *p : b->plans) shtk_plan_destroy(p);
    delete b;
    return SHTK_OK;
}

int shtk_batch_plan(shtk_batch *b, int group, shtk_plan **plan) {
    if (!b || !plan || group < 0 || group >= b->ngroups) return fail(SHTK_EINVAL, "bad batch group");
    *plan = b->plans[group];
    return SHTK_OK;
}

int shtk_batch_set_graph(shtk_batch *b, int on) {
    if (!b) return fail(SHTK_EINVAL, "batch is NULL");
    std::lock_guard<std::mutex> lk(b->mu);
    if (!on) batch_drop_graphs(b);
    b->use_graph = on ? 1 : 0;
    return SHTK_OK;
}

int shtk_batch_set_tiles(shtk_batch *b, int synth_blocks, int anal_blocks) {
    if (!b) return fail(SHTK_EINVAL, "batch is NULL");
    if (synth_blocks < 0 || anal_blocks < 0) return fail(SHTK_EINVAL, "block counts must be >= 0");
    std::lock_guard<std::mutex> lk(b->mu);
    batch_drop_graphs(b);
    b->blocks_synth = synth_blocks;
    b->blocks_anal = anal_blocks;
    return SHTK_OK;
}

int shtk_batch_synthesis(shtk_batch *b, const void *const *alm, void *const *f, int dtype, void *stream) {
    return batch_run(b, 0, alm, f, dtype, stream);
}

int shtk_batch_analysis(shtk_batch *b, const void *const *f, void *const *alm, int dtype, void *stream) {
    return batch_run(b, 1, f, alm, dtype, stream);
}

// Autotuner: times every candidate launch shape on the caller's arrays (round
// trips: synthesis into f, analysis back into alm) and keeps the fastest.
int shtk_batch_autotune(shtk_batch *b, void *const *alm, void *const *f, int dtype, int reps, void *stream,
                        int *best_synth_blocks, int *best_anal_blocks, double *best_ms) {
    if (!b || !alm || !f) return fail(SHTK_EINVAL, "batch and arrays must not be NULL");
    if (reps < 1) reps = 1;
    hipStream_t st = (hipStream_t)stream;
    const int graph_was = b->use_graph;
    RC_TRY(shtk_batch_set_graph(b, 0));
    // candidate grids of the two persistent kernels: 1..4 blocks per CU each
    std::vector<int> blocks = {0};
    if (b->precision == SHTK_PREC_SPLIT) blocks = {b->cus, 2 * b->cus, 3 * b->cus, 4 * b->cus};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0), "autotune events");
    HIP_TRY(hipEventCreate(&e1), "autotune events");
    double best = 1e300;
    int bb = 0, bp = 0;
    for (int nb : blocks)
        for (int fp : blocks) {
            RC_TRY(shtk_batch_set_tiles(b, nb, fp));
            // one untimed round trip, then reps timed ones
            RC_TRY(shtk_batch_synthesis(b, (const void *const *)alm, f, dtype, st));
            RC_TRY(shtk_batch_analysis(b, (const void *const *)f, alm, dtype, st));
            HIP_TRY(hipEventRecord(e0, st), "autotune timing");
            for (int r = 0; r < reps; ++r) {
                RC_TRY(shtk_batch_synthesis(b, (const void *const *)alm, f, dtype, st));
                RC_TRY(shtk_batch_analysis(b, (const void *const *)f, alm, dtype, st));
            }
            HIP_TRY(hipEventRecord(e1, st), "autotune timing");
            HIP_TRY(hipEventSynchronize(e1), "autotune timing");
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "autotune timing");
            if (ms / reps < best) {
                best = ms / reps;
                bb = nb;
                bp = fp;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    RC_TRY(shtk_batch_set_tiles(b, bb, bp));
    RC_TRY(shtk_batch_set_graph(b, graph_was));
    if (best_synth_blocks) *best_synth_blocks = bb;
    if (best_anal_blocks) *best_anal_blocks = bp;
    if (best_ms) *best_ms = best;
    return SHTK_OK;
}


const char *shtk_version(void) { return "0.2.0"; }

int shtk_last_error(char *buf, size_t len) {
    if (buf && len) std::snprintf(buf, len, "%s", g_err.c_str());
    return g_errcode;
}

int shtk_stream_sync(void *stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "stream synchronize");
    return SHTK_OK;
}

static int plan_tables(shtk_plan *p) {
    const shtk_desc &d = p->d;
    const int np = p->np, nm = p->nm, lmax = d.lmax, S = p->seg;
    std::vector<double> x(np), s(np), ws(np), winvs(np), winvs2(np), invs(np);
    for (int q = 0; q < np; ++q) {
        x[q] = p->hx[q];
        const double v = (1.0 - x[q]) * (1.0 + x[q]);
        s[q] = std::sqrt(v > 0 ? v : 0.0);
        const double w = p->hw[q];
        ws[q] = w * p->scale_phi;
        invs[q] = s[q] > 0 ? 1.0 / s[q] : 0.0;
        winvs[q] = ws[q] * invs[q];
        winvs2[q] = winvs[q] * invs[q];
    }

    // Per-order coefficients for l = m .. lmax+1: P-form (a_l, b_l) as
    // src/legendre.jl:242, 277-278 (seed kernel), and the segmented Q form.
    std::vector<double2> ab, qc, qd;
    std::vector<double> gt;
    std::vector<long long> coff(nm);
    std::vector<int> segbase(nm), segm;
    int nseg_total = 0;
    for (int im = 0; im < nm; ++im) {
        const int m = im * d.mres;
        coff[im] = (long long)ab.size();
        segbase[im] = nseg_total;
        const int nseg = (lmax - m + S) / S;
        for (int q = 0; q < nseg; ++q) segm.push_back(im);
        nseg_total += nseg;
        std::vector<double> a(lmax + 2 - m + 1, 0.0), b(lmax + 2 - m + 1, 0.0);
        for (int l = m; l <= lmax + 1; ++l) {
            double al = 0.0, bl = 0.0;
            if (l == m + 1) {
                al = std::sqrt(2.0 * m + 3.0);
            } else if (l >= m + 2) {
                al = std::sqrt(((2.0 * l - 1.0) * (2.0 * l + 1.0)) / ((double)(l - m) * (double)(l + m)));
                bl = std::sqrt(((2.0 * l + 1.0) * (double)(l - 1 - m) * (double)(l - 1 + m)) /
                               ((2.0 * l - 3.0) * (double)(l - m) * (double)(l + m)));
            }
            a[l - m] = al;
            b[l - m] = bl;
            ab.push_back(make_double2(al, bl));
        }
        double c = 1.0;
        for (int l = m; l <= lmax + 1; ++l) {
            const int t = l - m;
            // segment base: m for the first segment, l0-1 for the others
            if (l == m) c = 1.0;
            else if (t % S == 0) c = a[t];  // l = l0 of a new segment: c_{l0-1} = 1, c_{l0} = a_{l0}
            else c *= a[t];
            const double beta = (l >= m + 2) ? b[t] / (a[t] * a[t - 1]) : 0.0;
            qc.push_back(make_double2(beta, c));
            const double cprev = (l == m) ? 0.0 : c / a[t];
            const double gl = (l > 0 && l > m)
                                  ? std::sqrt((2.0 * l + 1.0) * ((double)l * l - (double)m * m) / (2.0 * l - 1.0))
                                  : 0.0;
            qd.push_back(make_double2((double)l * c, gl * cprev));
            gt.push_back(gl);
        }
    }
    p->nseg_total = nseg_total;

    // Sectoral seeds P_m^m(x_p) = (-1)^m sqrt(1/4pi) prod_k sqrt((2k+1)/(2k)) sin(theta)
    // (src/legendre.jl:252-271), carried as mantissa * 2^e (exact rescaling) and
    // stored in the scaled representation of the kernels.
    std::vector<double> sectv((size_t)nm * np);
    std::vector<int> sectk((size_t)nm * np);
    for (int q = 0; q < np; ++q) {
        double mant = 0.28209479177387814;
        int e = 0;
        for (int m = 0; m <= d.mmax; ++m) {
            if (m > 0) {
                mant = (-std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s[q]) * mant;
                int de = 0;
                mant = std::frexp(mant, &de);
                e += de;
            }
            if (m % d.mres) continue;
            const size_t idx = (size_t)(m / d.mres) * np + q;
            if (mant == 0.0) {
                sectv[idx] = 0.0;
                sectk[idx] = 0;
            } else if (e > shtk::SCALE_LIVE_MIN) {
                sectv[idx] = std::ldexp(mant, e);
                sectk[idx] = 0;
            } else {
                const int k = (shtk::SCALE_LIVE_MIN - e) / shtk::SCALE_SHIFT + 1;
                sectv[idx] = std::ldexp(mant, e + k * shtk::SCALE_SHIFT);
                sectk[idx] = k;
            }
        }
    }

    std::vector<double> conv;
    if (!p->canonical) {
        conv.assign((size_t)(lmax + 1) * (d.mmax + 1), 1.0);
        for (int m = 0; m <= d.mmax; ++m)
            for (int l = m; l <= lmax; ++l) conv[(size_t)l + (size_t)(lmax + 1) * m] = coef_scale(d, l, m);
    }
    p->fft = make_fft_plan(d.nlon);
    const long double PI_L = 3.141592653589793238462643383279502884L;
    std::vector<double2> tw, chirp, bhat;
    if (p->fft.ok) {
        const int M = p->fft.nfft, N = d.nlon;
        tw.resize(M);
        for (int t = 0; t < M; ++t) {
            const long double ang = 2.0L * PI_L * t / M;
            tw[t] = make_double2((double)cosl(ang), -(double)sinl(ang));
        }
        if (p->fft.blue) {
            // Bluestein: X_k = conj(b_k) sum_n (x_n conj(b_n)) b_{k-n}, b_n = e^{i pi n^2 / N};
            // the convolution kernel B (b_m at m and M - m) is even, so its spectrum is
            // b_0 + 2 sum_{m>=1} b_m cos(2 pi m k / M); the inverse DFT uses conj(B).
            chirp.resize(N);
            std::vector<long double> br(N), bi(N);
            for (int t = 0; t < N; ++t) {
                const long long q = ((long long)t * t) % (2LL * N);
                const long double ang = PI_L * q / N;
                br[t] = cosl(ang);
                bi[t] = sinl(ang);
                chirp[t] = make_double2((double)br[t], -(double)bi[t]);
            }
            std::vector<long double> cosM(M);
            for (int t = 0; t < M; ++t) cosM[t] = cosl(2.0L * PI_L * t / M);
            bhat.resize(2 * (size_t)M);
            for (int k = 0; k < M; ++k) {
                long double re = br[0], im = bi[0];
                long long idx = 0;
                for (int m = 1; m < N; ++m) {
                    idx += k;
                    if (idx >= M) idx %= M;
                    re += 2.0L * br[m] * cosM[idx];
                    im += 2.0L * bi[m] * cosM[idx];
                }
                bhat[k] = make_double2((double)(re / M), (double)(im / M));
                bhat[M + k] = make_double2((double)(re / M), -(double)(im / M));
            }
        }
    }

    RC_TRY(upload(&p->x, x, "plan tables"));
    RC_TRY(upload(&p->s, s, "plan tables"));
    RC_TRY(upload(&p->wscaled, ws, "plan tables"));
    RC_TRY(upload(&p->winvs, winvs, "plan tables"));
    RC_TRY(upload(&p->winvs2, winvs2, "plan tables"));
    RC_TRY(upload(&p->invs, invs, "plan tables"));
    RC_TRY(upload(&p->qc, qc, "plan tables"));
    p->qc_len = qc.size();
    {
        // contiguous, padded copies of the Q-form columns: kernels read whole
        // blocks of 8 (or 32) degrees through scalar loads without clamping
        std::vector<double> qb(qc.size() + shtk::OZ_TPAD, 0.0), qcc(qc.size() + shtk::OZ_TPAD, 0.0);
        for (size_t i = 0; i < qc.size(); ++i) {
            qb[i] = qc[i].x;
            qcc[i] = qc[i].y;
        }
        RC_TRY(upload(&p->qb, qb, "plan tables"));
        RC_TRY(upload(&p->qcc, qcc, "plan tables"));
    }
    RC_TRY(upload(&p->qd, qd, "plan tables"));
    RC_TRY(upload(&p->ab, ab, "plan tables"));
    RC_TRY(upload(&p->gtab, gt, "plan tables"));
    RC_TRY(upload(&p->coff, coff, "plan tables"));
    p->hcoff = coff;
    RC_TRY(upload(&p->segbase, segbase, "plan tables"));
    RC_TRY(upload(&p->segm, segm, "plan tables"));
    RC_TRY(upload(&p->convscale, conv, "plan tables"));
    RC_TRY(upload(&p->tw, tw, "plan tables"));
    RC_TRY(upload(&p->chirp, chirp, "plan tables"));
    RC_TRY(upload(&p->bhat, bhat, "plan tables"));

    // Segment seeds, computed once on the device.
    const size_t nseed = (size_t)nseg_total * np;
    HIP_TRY(hipMalloc((void **)&p->seedq0, sizeof(double) * nseed), "seed tables");
    HIP_TRY(hipMalloc((void **)&p->seedq1, sizeof(double) * nseed), "seed tables");
    HIP_TRY(hipMalloc((void **)&p->seedk, sizeof(int) * nseed), "seed tables");
    RC_TRY(upload(&p->pfirst, std::vector<int>(nseg_total, np), "seed tables"));
    RC_TRY(upload(&p->pfast, std::vector<int>(nseg_total, 0), "seed tables"));
    double *dsectv = nullptr;
    int *dsectk = nullptr;
    int rc = upload(&dsectv, sectv, "seed inputs");
    if (!rc) rc = upload(&dsectk, sectk, "seed inputs");
    if (!rc) {
        shtk::SeedArgs sa{};
        sa.lmax = lmax;
        sa.mmax = d.mmax;
        sa.mres = d.mres;
        sa.np = np;
        sa.nm = nm;
        sa.seg = S;
        sa.x = p->x;
        sa.ab = p->ab;
        sa.coff = p->coff;
        sa.segbase = p->segbase;
        sa.sectv = dsectv;
        sa.sectk = dsectk;
        sa.seedq0 = p->seedq0;
        sa.seedq1 = p->seedq1;
        sa.seedk = p->seedk;
        sa.pfirst = p->pfirst;
        sa.pfast = p->pfast;
        hipError_t e = shtk::launch_seed(sa, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(e, "seed kernel");
    }
    // the sectoral seeds stay with the plan for the split-precision tables
    p->sectv = dsectv;
    p->sectk = dsectk;
    return rc;
}

int shtk_plan_create(const shtk_desc *desc, shtk_plan **out) {
    if (!out) return fail(SHTK_EINVAL, "plan output pointer is NULL");
    *out = nullptr;
    int rc = validate(desc);
    if (rc) return rc;
    shtk_plan *p = new (std::nothrow) shtk_plan();
    if (!p) return fail(SHTK_ENOMEM, "host allocation failed");
    p->d = *desc;
    p->nrow = desc->nlat;
    if (desc->grid == SHTK_GRID_DRISCOLL_HEALY) p->d.nlat = desc->nlat + 1;  // plus the south pole
    const shtk_desc &d = p->d;
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return fail(SHTK_EHIP, "hipGetDevice failed (no ROCm device)");
    }
    if (hipEventCreateWithFlags(&p->order_ev, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
        delete p;
        return fail(SHTK_EHIP, "hipEventCreateWithFlags failed");
    }
    {
        const int rc_s = static_scratch_setup(p->device);
        if (rc_s) {
            (void)hipEventDestroy(p->order_ev);
            delete p;
            return rc_s;
        }
    }
    p->np = (d.nlat + 1) / 2;
    p->nm = d.mmax / d.mres + 1;
    p->seg = segment_length();
    p->xcd_map = env_int("SHTK_XCD", 0) != 0;
    // latitude pairs per lane of the scalar synthesis: 2 for the 64-degree VGPR-record
    // kernel (4 waves per SIMD), 4 for the scalar-load kernel of other segment lengths
    p->synth_chains = env_int("SHTK_SYNTH_CHAINS", p->seg == 64 ? 2 : 4);
    p->vsynth_chains = env_int("SHTK_VSYNTH_CHAINS", 1) == 2 ? 2 : 1;
    // vector analysis items of 256 pairs (chains <= 4 per lane, 3 waves/SIMD; outputs added
    // atomically when np > 256) instead of 512 (8 chains, 256 VGPRs, 1 wave/SIMD): 2.9 -> 2.3 ms
    // for 8 fields at lmax 1023
    p->vanal_pblk = env_int("SHTK_VANAL_PBLK", 256) == 512 ? 512 : 256;
    p->vsynth_waves = env_int("SHTK_VSYNTH_WAVES", 1) == 4 ? 4 : 1;
    // scalar FP64 synthesis blocks of one wave once a launch has 8192 units or more
    // (2 fields at lmax 1023: 126 -> 119 us; 16 fields: 943 -> 821 us; one field
    // keeps 4 waves per block, 65 against 72 us)
    p->ssynth_1w_units = std::max(1, env_int("SHTK_SSYNTH_1W_UNITS", 8192));
    p->fft_tile = env_int("SHTK_FFT_PAIRS", 0);
    p->oz_run_len = std::max(1, env_int("SHTK_OZ_RUN", 8));
    {
        const int nh = env_int("SHTK_OZ_NH", 0);
        p->oz_nh_force = (nh == 1 || nh == 2) ? nh : 0;
    }
    p->canonical = d.norm == SHTK_NORM_ORTHONORMAL && !d.real_norm && d.cs_phase;
    p->inv_scale = d.phi_scale == SHTK_PHI_QUAD ? d.nlon / (2.0 * M_PI) : (double)d.nlon;
    p->scale_phi = (2.0 * M_PI / d.nlon) * d.nlon / p->inv_scale;
    std::vector<double> xa, wa;
    if (d.grid == SHTK_GRID_GAUSS) {
        gauss_nodes(d.nlat, xa, wa);
        p->hx.assign(xa.rbegin(), xa.rend());  // north first (src/config.jl:639)
        p->hw.assign(wa.rbegin(), wa.rend());
    } else {
        if (d.grid == SHTK_GRID_REGULAR) fejer1_nodes(d.nlat, xa, wa);
        else if (d.grid == SHTK_GRID_REGULAR_POLES) clenshaw_curtis_nodes(d.nlat, xa, wa);
        else driscoll_healy_nodes(p->nrow, xa, wa);
        p->hx = xa;
        p->hw = wa;
    }
    p->poles = d.grid == SHTK_GRID_REGULAR_POLES || d.grid == SHTK_GRID_DRISCOLL_HEALY;
    p->pole_weights = d.grid == SHTK_GRID_REGULAR_POLES;
    rc = plan_tables(p);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return SHTK_OK;
}

int shtk_profile_enable(shtk_plan *plan, int on) {
    if (!plan) return fail(SHTK_EINVAL, "plan is NULL");
    std::lock_guard<std::recursive_mutex> lk(plan->mu);
    plan->profiling = on != 0;
    return SHTK_OK;
}

int shtk_profile_read(shtk_plan *plan, int max, char *names, double *total_ms, int *count, int *n_out) {
    if (!plan) return fail(SHTK_EINVAL, "plan is NULL");
    std::lock_guard<std::recursive_mutex> lk(plan->mu);
    DeviceGuard g(plan->device);
    for (auto &pe : plan->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventSynchronize(pe.stop), "profile event wait");
        HIP_TRY(hipEventElapsedTime(&ms, pe.start, pe.stop), "profile event time");
        bool found = false;
        for (auto &t : plan->prof_totals)
            if (t.first == pe.name) {
                t.second.first += ms;
                t.second.second += 1;
                found = true;
            }
        if (!found) plan->prof_totals.push_back({pe.name, {(double)ms, 1}});
        plan->event_pool.push_back(pe.start);
        plan->event_pool.push_back(pe.stop);
    }
    plan->pending.clear();
    int n = 0;
    for (auto &t : plan->prof_totals) {
        if (n >= max) break;
        if (names) std::snprintf(names + 48 * n, 48, "%s", t.first.c_str());
        if (total_ms) total_ms[n] = t.second.first;
        if (count) count[n] = t.second.second;
        ++n;
    }
    if (n_out) *n_out = n;
    plan->prof_totals.clear();
    return SHTK_OK;
}

int shtk_plan_destroy(shtk_plan *plan) {
    if (!plan) return SHTK_OK;
    DeviceGuard g(plan->device);
    (void)hipDeviceSynchronize();
    for (auto &pe : plan->pending) {
        (void)hipEventDestroy(pe.start);
        (void)hipEventDestroy(pe.stop);
    }
    for (hipEvent_t e : plan->event_pool) (void)hipEventDestroy(e);
    free_plan(plan);
    return SHTK_OK;
}

int shtk_plan_nodes(const shtk_plan *plan, double *x, double *w) {
    if (!plan) return fail(SHTK_EINVAL, "plan is NULL");
    for (int i = 0; i < plan->nrow; ++i) {
        if (x) x[i] = plan->hx[i];
        if (w) w[i] = plan->hw[i];
    }
    return SHTK_OK;
}

size_t shtk_workspace_size(const shtk_plan *plan, int nfields, int vector) {
    if (!plan || nfields < 1) return 0;
    return ws_layout(plan, nfields, vector ? 1 : 0).total;
}

// --- scalar ------------------------------------------------------------------
int shtk_synthesis(shtk_plan *p, const void *alm, void *f, int nfields, int dtype, int real_output, int ltr,
                   void *stream) {
    return shtk_synthesis_ex(p, alm, f, nfields, dtype, real_output, ltr, SHTK_PREC_FP64, stream);
}

// shtk_synthesis with the static scratch as its workspace: the same kernels,
// no device allocation (the plan's tables exist since shtk_plan_create).
int shtk_synthesis_static(shtk_plan *p, const void *alm, void *f, int nfields, int dtype, int real_output, int ltr,
                          void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {alm, f}));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    void *ws = static_scratch(p->device);
    if (!ws) return fail(SHTK_EHIP, "static scratch is not set up on this device");
    Job j{0, 0, nullptr, p->inv_scale / p->d.nlon, 0, 0, 0, ltr, "leg_synth", "fft_synth"};
    j.cplx_out = real_output != 1;
    return run_to_grid(p, j, alm, nullptr, f, nullptr, nfields, dtype, stream, ws, kStaticScratchBytes);
}

int shtk_synthesis_ex(shtk_plan *p, const void *alm, void *f, int nfields, int dtype, int real_output, int ltr,
                      int precision, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {alm, f}));
    RC_TRY(check_precision(p, precision, 4095));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    // real output: f = Re ifft(phi_inv_scale * F) with the Hermitian mirror, the 1/nlon
    // of the inverse FFT folded in g; complex output (src/core_transforms.jl:294-344):
    // the bins m >= 0 alone, f = ifft(phi_inv_scale * F)
    Job j{0, 0, nullptr, p->inv_scale / p->d.nlon, 0, 0, 0, ltr, "leg_synth", "fft_synth"};
    j.prec = precision;
    j.cplx_out = real_output != 1;
    return run_to_grid(p, j, alm, nullptr, f, nullptr, nfields, dtype, stream);
}

int shtk_analysis_ex(shtk_plan *p, const void *f, void *alm, int nfields, int dtype, int ltr, int precision,
                     void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {alm, f}));
    RC_TRY(check_precision(p, precision, 4095));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    Job j{0, 1, p->wscaled, 1.0, 0, 0, 0, ltr, "leg_anal", "fft_anal"};
    j.prec = precision;
    return run_to_spec(p, j, f, nullptr, alm, nullptr, nfields, dtype, stream);
}

int shtk_analysis_cplx(shtk_plan *p, const void *f, void *alm, int nfields, int dtype, int ltr, int precision,
                       void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {alm, f}));
    RC_TRY(check_precision(p, precision, 4095));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    Job j{0, 1, p->wscaled, 1.0, 0, 0, 0, ltr, "leg_anal", "fft_anal"};
    j.prec = precision;
    j.cplx_in = 1;
    return run_to_spec(p, j, f, nullptr, alm, nullptr, nfields, dtype, stream);
}

int shtk_analysis(shtk_plan *p, const void *f, void *alm, int nfields, int dtype, int ltr, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {alm, f}));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    Job j{0, 1, p->wscaled, 1.0, 0, 0, 0, ltr, "leg_anal", "fft_anal"};
    return run_to_spec(p, j, f, nullptr, alm, nullptr, nfields, dtype, stream);
}

int shtk_adjoint_synthesis(shtk_plan *p, const void *fbar, void *almbar, int nfields, int dtype, int real_output,
                           void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {fbar, almbar}));
    // src/core_transforms.jl:477-519: wm * phi_inv_scale/nlon * sum_i P FFT(fbar), no weights
    Job j{0, 0, nullptr, p->inv_scale / p->d.nlon, real_output ? 1 : 0, 0, 0, p->d.lmax, "leg_anal", "fft_anal"};
    return run_to_spec(p, j, fbar, nullptr, almbar, nullptr, nfields, dtype, stream);
}

int shtk_adjoint_analysis(shtk_plan *p, const void *almbar, void *fbar, int nfields, int dtype, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {fbar, almbar}));
    // src/core_transforms.jl:611-652: F = 2pi w_i sum_l P abar, ifft over m >= 0;
    // the real part is Re(F_0)/N + 2 Re(sum_{m>0} (F_m/2N) e^{im phi})
    Job j{0, 1, p->wscaled, 1.0, 2, 0, 0, p->d.lmax, "leg_synth", "fft_synth"};
    return run_to_grid(p, j, almbar, nullptr, fbar, nullptr, nfields, dtype, stream);
}

int shtk_adjoint_analysis_cplx(shtk_plan *p, const void *almbar, void *fbar, int nfields, int dtype, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {fbar, almbar}));
    // the whole of src/core_transforms.jl:611-652: F = 2pi w_i sum_l P abar over the
    // bins m >= 0 (the negative ones stay zero), ifft to a complex grid
    Job j{0, 1, p->wscaled, 1.0, 0, 0, 0, p->d.lmax, "leg_synth", "fft_synth"};
    j.cplx_out = 1;
    return run_to_grid(p, j, almbar, nullptr, fbar, nullptr, nfields, dtype, stream);
}

// --- vector ------------------------------------------------------------------
int shtk_synthesis_sphtor(shtk_plan *p, const void *S, const void *T, void *Vt, void *Vp, int nfields, int dtype,
                          int real_output, int ltr, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {S, T, Vt, Vp}));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    // kernel sums are sin(theta) V; Robert form keeps them, otherwise divide by sin(theta);
    // real_output = 0: complex grids of the bins m >= 0 (src/sphtor_transforms.jl:162-241)
    Job j{1, 0, p->d.robert_form ? nullptr : p->invs, p->inv_scale / p->d.nlon, 0, 0, 0, ltr,
          "leg_synth_sphtor", "fft_synth"};
    j.cplx_out = real_output != 1;
    j.pole_fix = 1;
    return run_to_grid(p, j, S, T, Vt, Vp, nfields, dtype, stream);
}

int shtk_analysis_sphtor(shtk_plan *p, const void *Vt, const void *Vp, void *S, void *T, int nfields, int dtype,
                         int ltr, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {S, T, Vt, Vp}));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    // Robert-form rows cannot be divided by sin(theta) at weighted pole nodes when
    // order 1 is stored (_validate_robert_analysis, src/config.jl:608-624)
    if (p->pole_weights && p->d.robert_form && p->d.mres == 1 && p->d.mmax >= 1 && ltr >= 1)
        return fail(SHTK_EINVAL, "Robert-form components with m=1 cannot be unweighted at nonzero-weight pole "
                                 "nodes; use a Gauss or pole-free regular grid");
    Job j{1, 1, p->d.robert_form ? p->winvs2 : p->winvs, 1.0, 0, 1, 0, ltr, "leg_anal_sphtor", "fft_anal"};
    j.pole_fix = 1;
    return run_to_spec(p, j, Vt, Vp, S, T, nfields, dtype, stream);
}

int shtk_synthesis_sphtor_ex(shtk_plan *p, const void *S, const void *T, void *Vt, void *Vp, int nfields,
                             int dtype, int real_output, int ltr, int precision, void *stream) {
    if (precision == SHTK_PREC_FP64)
        return shtk_synthesis_sphtor(p, S, T, Vt, Vp, nfields, dtype, real_output, ltr, stream);
    RC_TRY(check_precision(p, precision, 4094));
    RC_TRY(check_common(p, nfields, dtype, {S, T, Vt, Vp}));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    // the contraction gives sin(theta) V; the phi FFT divides by sin(theta) unless Robert form
    Job j{1, 0, nullptr, p->inv_scale / p->d.nlon, 0, 0, p->d.robert_form ? 0 : 1, ltr, "oz_synth_sphtor",
          "fft_synth"};
    j.prec = precision;
    j.cplx_out = real_output != 1;
    return run_vsplit_to_grid(p, j, S, T, Vt, Vp, nfields, dtype, stream);
}

int shtk_analysis_sphtor_ex(shtk_plan *p, const void *Vt, const void *Vp, void *S, void *T, int nfields, int dtype,
                            int ltr, int precision, void *stream) {
    if (precision == SHTK_PREC_FP64) return shtk_analysis_sphtor(p, Vt, Vp, S, T, nfields, dtype, ltr, stream);
    RC_TRY(check_precision(p, precision, 4094));
    RC_TRY(check_common(p, nfields, dtype, {S, T, Vt, Vp}));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    if (p->pole_weights && p->d.robert_form && p->d.mres == 1 && p->d.mmax >= 1 && ltr >= 1)
        return fail(SHTK_EINVAL, "Robert-form components with m=1 cannot be unweighted at nonzero-weight pole "
                                 "nodes; use a Gauss or pole-free regular grid");
    Job j{1, 1, p->d.robert_form ? p->winvs2 : p->winvs, 1.0, 0, 1, 0, ltr, "oz_anal_sphtor", "fft_anal"};
    j.prec = precision;
    return run_vsplit_to_spec(p, j, Vt, Vp, S, T, nfields, dtype, stream);
}

// --- single order ------------------------------------------------------------

// Checks of _validate_stored_order (src/sphtor_transforms.jl:1156-1169) and
// synthesis_packed_ml (src/transforms.jl:394-404); fills the order's tables.
static int ml_args(shtk_plan *p, int im, int ltr, int nfields, int dtype, int vector, shtk::MlArgs *a) {
    if (!p) return fail(SHTK_EINVAL, "plan is NULL");
    if (nfields < 1) return fail(SHTK_EINVAL, "nfields must be >= 1");
    if (dtype != SHTK_F64 && dtype != SHTK_F32) return fail(SHTK_EINVAL, "dtype must be SHTK_F64 or SHTK_F32");
    if (im < 0 || im > p->d.mmax / p->d.mres)
        return fail(SHTK_EINVAL, "im must satisfy 0 <= im <= mmax/mres = " + std::to_string(p->d.mmax / p->d.mres));
    RC_TRY(norm_ltr(p, ltr, &ltr));
    const int m = im * p->d.mres;
    if (ltr < m) return fail(SHTK_EINVAL, "ltr must satisfy ltr >= im*mres = " + std::to_string(m));
    shtk::MlArgs &A = *a;
    A = shtk::MlArgs{};
    A.lmax = p->d.lmax;
    A.m = m;
    A.ltr = ltr;
    A.nlat = p->d.nlat;
    A.np = p->np;
    A.vector = vector;
    A.robert = p->d.robert_form;
    A.io_f32 = dtype == SHTK_F32;
    A.x = p->x;
    A.s = p->s;
    A.w = p->wscaled;
    A.ab = p->ab + p->hcoff[im];
    A.gtab = p->gtab + p->hcoff[im];
    A.sectv = p->sectv + (size_t)im * p->np;
    A.sectk = p->sectk + (size_t)im * p->np;
    A.conv = p->canonical ? nullptr : p->convscale + (size_t)(p->d.lmax + 1) * m;
    A.ld_spec = ltr - m + 1;
    A.ld_grid = p->nrow;
    A.nrow = p->nrow;
    return SHTK_OK;
}

static int ml_synth(shtk_plan *p, int im, const void *S, const void *T, void *V0, void *V1, int nfields, int dtype, int ltr,
             int vector, void *stream) {
    shtk::MlArgs a;
    RC_TRY(ml_args(p, im, ltr, nfields, dtype, vector, &a));
    if (!S || !V0 || (vector && (!T || !V1))) return fail(SHTK_EINVAL, "NULL array argument");
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    DeviceGuard dg(p->device);
    StreamOrder so(p, (hipStream_t)stream);
    a.g = p->inv_scale;
    a.in0 = S;
    a.in1 = T;
    a.out0 = V0;
    a.out1 = V1;
    HIP_TRY(shtk::launch_ml_synth(a, nfields, (hipStream_t)stream), "single-order synthesis launch");
    return SHTK_OK;
}

static int ml_anal(shtk_plan *p, int im, const void *V0, const void *V1, void *S, void *T, int nfields, int dtype, int ltr,
            int vector, void *stream) {
    shtk::MlArgs a;
    RC_TRY(ml_args(p, im, ltr, nfields, dtype, vector, &a));
    if (!S || !V0 || (vector && (!T || !V1))) return fail(SHTK_EINVAL, "NULL array argument");
    // Robert-form rows cannot be divided by sin(theta) at weighted pole nodes for
    // order 1 (_validate_robert_analysis with m, src/config.jl:608-624)
    if (vector && p->pole_weights && p->d.robert_form && a.m == 1 && a.ltr >= 1)
        return fail(SHTK_EINVAL, "Robert-form components with m=1 cannot be unweighted at nonzero-weight pole "
                                 "nodes; use a Gauss or pole-free regular grid");
    std::lock_guard<std::recursive_mutex> lk(p->mu);
    DeviceGuard dg(p->device);
    hipStream_t st = (hipStream_t)stream;
    StreamOrder so(p, st);
    const size_t acc_bytes = sizeof(double) * 4 * (size_t)nfields * ((p->np + 63) / 64) * a.ld_spec;
    RC_TRY(ensure_ws(p, acc_bytes));
    a.g = 1.0;
    a.in0 = V0;
    a.in1 = V1;
    a.out0 = S;
    a.out1 = T;
    a.acc = (double *)p->ws;
    HIP_TRY(shtk::launch_ml_anal(a, nfields, st), "single-order analysis launch");
    return SHTK_OK;
}

int shtk_synthesis_ml(shtk_plan *p, int im, const void *Ql, void *Vm, int nfields, int dtype, int ltr,
                      void *stream) {
    return ml_synth(p, im, Ql, nullptr, Vm, nullptr, nfields, dtype, ltr, 0, stream);
}

int shtk_analysis_ml(shtk_plan *p, int im, const void *Vm, void *Ql, int nfields, int dtype, int ltr, void *stream) {
    return ml_anal(p, im, Vm, nullptr, Ql, nullptr, nfields, dtype, ltr, 0, stream);
}

int shtk_synthesis_sphtor_ml(shtk_plan *p, int im, const void *Sl, const void *Tl, void *Vt, void *Vp, int nfields,
                             int dtype, int ltr, void *stream) {
    return ml_synth(p, im, Sl, Tl, Vt, Vp, nfields, dtype, ltr, 1, stream);
}

int shtk_analysis_sphtor_ml(shtk_plan *p, int im, const void *Vt, const void *Vp, void *Sl, void *Tl, int nfields,
                            int dtype, int ltr, void *stream) {
    return ml_anal(p, im, Vt, Vp, Sl, Tl, nfields, dtype, ltr, 1, stream);
}

int shtk_adjoint_synthesis_sphtor(shtk_plan *p, const void *Vtb, const void *Vpb, void *Sb, void *Tb, int nfields,
                                  int dtype, int real_output, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {Vtb, Vpb, Sb, Tb}));
    Job j{1, 0, p->invs, p->inv_scale / p->d.nlon, real_output ? 1 : 0, 0, p->d.robert_form ? 2 : 0, p->d.lmax,
          "leg_anal_sphtor", "fft_anal"};
    j.pole_fix = 2;
    return run_to_spec(p, j, Vtb, Vpb, Sb, Tb, nfields, dtype, stream);
}

int shtk_adjoint_analysis_sphtor(shtk_plan *p, const void *Sb, const void *Tb, void *Vtb, void *Vpb, int nfields,
                                 int dtype, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {Vtb, Vpb, Sb, Tb}));
    if (p->pole_weights && p->d.robert_form && p->d.mres == 1 && p->d.mmax >= 1)
        return fail(SHTK_EINVAL, "Robert-form components with m=1 cannot be unweighted at nonzero-weight pole "
                                 "nodes; use a Gauss or pole-free regular grid");
    Job j{1, 1, p->winvs, 1.0, 2, 1, p->d.robert_form ? 1 : 0, p->d.lmax, "leg_synth_sphtor", "fft_synth"};
    j.pole_fix = 2;
    return run_to_grid(p, j, Sb, Tb, Vtb, Vpb, nfields, dtype, stream);
}

int shtk_adjoint_analysis_sphtor_cplx(shtk_plan *p, const void *Sb, const void *Tb, void *Vtb, void *Vpb,
                                      int nfields, int dtype, void *stream) {
    RC_TRY(check_common(p, nfields, dtype, {Vtb, Vpb, Sb, Tb}));
    if (p->pole_weights && p->d.robert_form && p->d.mres == 1 && p->d.mmax >= 1)
        return fail(SHTK_EINVAL, "Robert-form components with m=1 cannot be unweighted at nonzero-weight pole "
                                 "nodes; use a Gauss or pole-free regular grid");
    Job j{1, 1, p->winvs, 1.0, 0, 1, p->d.robert_form ? 1 : 0, p->d.lmax, "leg_synth_sphtor", "fft_synth"};
    j.pole_fix = 2;
    j.cplx_out = 1;
    return run_to_grid(p, j, Sb, Tb, Vtb, Vpb, nfields, dtype, stream);
}

}  // extern "C"

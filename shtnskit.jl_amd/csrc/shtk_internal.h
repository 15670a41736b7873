// Internal declarations shared by the host plan code and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/shtk.h"

namespace shtk {

// Scaled-exponent representation of Legendre values that underflow FP64
// (the role of src/legendre.jl:197-322).  A chain in "scaled" state holds
// p * 2^(SCALE_SHIFT * k) with k > 0; its physical value is below 2^-359 and
// its contributions are dropped.  When |p| exceeds 2^SCALE_HI the chain is
// shifted down by 2^-SCALE_SHIFT and k decreases; at k == 0 it is live.
constexpr int SCALE_SHIFT = 512;
constexpr int SCALE_HI = 100;
constexpr int SCALE_LIVE_MIN = -412;  // seeds at or above 2^-412 start live
constexpr int SCALE_CHECK = 8;        // recurrence steps between checks

// Degree segments.  The rows of order m are cut into segments of `seg`
// degrees starting at l0 = m + s*seg.  Inside a segment the recurrence runs on
// Q_l = P_l / c_l with c = 1 at the segment base (m for s = 0, l0-1 otherwise)
// and c_l = c_{l-1} a_l, which turns P_l = a_l x P_{l-1} - b_l P_{l-2} into
// Q_{l+1} = x Q_l - beta_{l+1} Q_{l-1},  beta_l = b_l / (a_l a_{l-1}).
// A plan-time kernel stores every segment's starting pair (Q_{l0-1}, Q_{l0}).
// dst += (NEG) bc[lane N of this lane's 16-lane row] * v: v_fmac_f64 with the
// first operand broadcast by DPP row_newbcast (full FP64 rate).  A VALU write
// of bc needs wait states before this reads it: callers put "s_nop 4" after
// refilling bc registers.
#define SHTK_FMAC_BC(dst, bc, v, N, NEG)                                                             \
    asm volatile("v_fmac_f64_dpp %0, " NEG "%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf" \
                 : "+v"(dst)                                                                    \
                 : "v"(bc), "v"(v), "i"(N))

constexpr int SEG_MAX = 256;
constexpr int DSEG = 64;  // segment length of the register-record Legendre kernels

// Coefficient tables, one block per stored order (offset coff[im], entries
// l = m .. lmax+1):
//   qc[l] = (beta_l, c_l)                  scalar and vector recurrences
//   qd[l] = (l c_l, g_l c_{l-1})           vector: s dP_l/dtheta = l x P_l - g_l P_{l-1},
//                                          g_l = sqrt((2l+1)(l^2-m^2)/(2l-1))
constexpr int OZ_TPAD = 64;  // entries past the end of the padded per-degree tables

struct LegArgs {
    int lmax, mmax, mres, nlat, np, nm, ltr, seg;
    const double *x;    // [np] cos(theta) of the north node of each pair
    const double *wlat; // [np] per-latitude factor, or nullptr for 1
    const double2 *qc;
    const double2 *qd;
    const double *qb, *qcc; // beta_l and c_l of qc as separate arrays, padded by OZ_TPAD entries
    const double2 *ab;    // P-form (a_l, b_l)
    const double *gtab;   // g_l = sqrt((2l+1)(l^2-m^2)/(2l-1))
    const long long *coff;
    const int *segbase;   // [nm] first global segment index of each order
    const int *segm;      // [nseg_total] order index im of each global segment
    int nseg_total;
    const double *seedq0; // [nseg_total][np] Q_{l0-1}
    const double *seedq1; // [nseg_total][np] Q_{l0}
    const int *seedk;     // [nseg_total][np] scale count at l0
    const int *pfirst;    // [nseg_total] first pair live anywhere in the segment (np: none)
    const int *pfast;     // [nseg_total] pairs >= pfast are live at the segment start
    double g;             // global factor
    int mfac_mode;        // 0: 1 for all m; 1: 2 for m >= 1; 2: 1/2 for m >= 1
    int io_f32;           // spectral arrays are complex float
    int ll1_mode;         // vector: 1 divides analysis outputs / synthesis inputs by l(l+1)
    int atomic;           // analysis: accumulate into a zeroed output (several pair blocks)
    const void *in0, *in1;  // spectral inputs (synthesis)
    void *out0, *out1;      // spectral outputs (analysis)
    long long spec_fstride; // complex elements per field in spectral arrays
    double2 *F0, *F1;       // Fourier workspace [field][im][nlat]
    int gx, gy, gz;         // work items form a (gx, gy, gz = nfields) box, x fastest
    int xcd;                // synthesis: 1 keeps the pair blocks of an order on one XCD
    int nm_run;             // synthesis: only orders im < nm_run (0: all)
    int skip_m0;            // vector: order 0 is computed by the zonal kernel (shtk_ml.hip)
    int qscale_in;          // scalar synthesis (64-degree segments): in0 holds canonical FP64
                            // coefficients, multiplied by c_l as they are loaded
    int vwaves;             // vector synthesis (64-degree segments): waves per block, 1 or 4
    int s1w_units;          // scalar synthesis (64-degree segments): one wave per block when the
                            // launch has at least this many (order, pair block, field) units
    int apblk;              // analysis: latitude pairs per work item (512, or 256 for the vector
                            // analysis when it lets the kernel run 3 waves per SIMD)
    int vdirect;            // vector synthesis (64-degree segments): in0 / in1 hold canonical FP64
                            // S and T, the records (S, T, c_l S, c_l T) built as they are loaded
};

// Work distribution over the 8 XCDs (each with its own 4 MB L2).  The
// hardware deals consecutive blocks of a 1-D grid round-robin over the XCDs.
// A "unit" is a group of `per_unit` blocks that read the same order's
// coefficients and spectra.  With xcd = 1 the grid holds 8 * per_unit *
// ceil(nunits/8) blocks and every unit runs entirely on XCD (unit mod 8), so
// its blocks share one L2 while consecutive orders still spread over all XCDs;
// blocks beyond the last unit return false.  With xcd = 0 block b is simply
// (unit b / per_unit, index b % per_unit).
constexpr int kXcds = 8;
__device__ inline bool unit_of_block(int xcd, int per_unit, int nunits, int *unit, int *index) {
    const int id = blockIdx.x;
    if (!xcd) {
        *unit = id / per_unit;
        *index = id % per_unit;
        return *unit < nunits;
    }
    const int x = id % kXcds, k = id / kXcds;
    *unit = x + kXcds * (k / per_unit);
    *index = k % per_unit;
    return *unit < nunits;
}
inline unsigned grid_for_units(int xcd, int per_unit, int nunits) {
    if (!xcd) return (unsigned)per_unit * nunits;
    return (unsigned)kXcds * per_unit * ((nunits + kXcds - 1) / kXcds);
}

struct SeedArgs {
    int lmax, mmax, mres, np, nm, seg;
    const double *x;
    const double2 *ab;  // P-form (a_l, b_l), same offsets as qc
    const long long *coff;
    const int *segbase;
    const double *sectv;  // [nm][np] scaled sectoral P_m^m
    const int *sectk;
    double *seedq0, *seedq1;
    int *seedk;
    int *pfirst, *pfast;
};

// FFT factorisations: FFT_RADIX_BITS bits per pass, first pass in the low bits.
// Radices 2, 3, 4, 5 and 8 have dedicated butterflies; other primes up to
// FFT_GEN_MAX run a generic pass (R table twiddles per output).  Lengths whose
// other prime factors would cost more take Bluestein's algorithm: a circular
// convolution of length nfft >= 2 nlon - 1 (2^a 3^b 5^c) through two FFTs.
constexpr int FFT_RADIX_BITS = 6;
constexpr int FFT_MAX_PASSES = 10;
constexpr int FFT_GEN_MAX = 61;     // largest prime of the generic pass
constexpr int FFT_GEN_BUDGET = 48;  // largest sum of generic primes before Bluestein is used
constexpr int FFT_MAX_LEN = 8192;   // LDS FFT length limit (8 values per thread, 1024 threads)

struct FftPlan {
    bool ok = false;    // nlon is supported
    bool blue = false;  // Bluestein
    int nfft = 0;       // LDS transform length: nlon, or the convolution length
    std::vector<int> radix;   // passes of the nlon (direct) or nfft (Bluestein) transform
};

struct FftArgs {
    int nlat, nlon, np, nm, mres, mmax;
    int nrad;
    unsigned long long radix; // FFT_RADIX_BITS bits per pass (of the length-nfft transform)
    const double2 *tw;        // e^{-2 pi i t / nfft}
    int nfft;                 // LDS row length: nlon, or the Bluestein convolution length
    int blue;                 // Bluestein: chirp[nlon] = e^{-pi i t^2 / nlon}, bhat[2][nfft] = kernel spectra / nfft
    const double2 *chirp, *bhat;
    double2 *F;               // Fourier workspace [field][im][nlat]
    void *grid;               // grid array (nrow, nlon, nfields)
    int nrow;                 // grid rows: nlat, or nlat - 1 when the last node (south pole) has no row
    int io_f32;
    int mode;          // grid row factor: 0 none, 1 1/sin(theta), 2 sin(theta) (Robert form)
    const double *s;   // [np] sin(theta) for Robert scaling
    int tile_pairs;    // packed rows (latitude pairs' worth) per tile (0: chosen from the sizes)
};

}  // namespace shtk

struct shtk_plan {
    shtk_desc d{};
    int np = 0, nm = 0, seg = 128, nseg_total = 0;
    int xcd_map = 0, synth_chains = 4, fft_tile = 0;  // launch-shape knobs
    int vsynth_chains = 1;
    int vsynth_waves = 1;  // waves per vector synthesis block (SHTK_VSYNTH_WAVES: 1 or 4)
    int ssynth_1w_units = 8192;  // scalar synthesis: units from which blocks hold one wave
                                 // (SHTK_SSYNTH_1W_UNITS)
    int vanal_pblk = 256;  // pairs per vector analysis work item (SHTK_VANAL_PBLK: 256 or 512)  // latitude pairs per lane of the vector synthesis (1 or 2)
    int oz_nh_force = 0;  // split analysis pair halves per block (0: by grid size)
    bool canonical = true;
    bool poles = false;  // nodes at theta = 0 and pi (the first pair)
    bool pole_weights = false;  // the pole nodes carry nonzero quadrature weights
    // Grid rows of the user arrays.  The Driscoll-Healy grid runs as the symmetric
    // grid of nlat + 1 nodes theta = i pi / nlat (d.nlat = nlat + 1 inside the plan)
    // whose last node, the south pole, has weight 0 and no row.
    int nrow = 0;
    double inv_scale = 0, scale_phi = 0;
    std::vector<double> hx, hw;  // north-first nodes and weights (nlat)
    std::vector<long long> hcoff;  // host copy of coff: first table entry of each order
    int device = 0;

    // device tables
    double *x = nullptr, *s = nullptr;   // [np]
    double *wscaled = nullptr;   // [np] w * scale_phi
    double *winvs = nullptr;     // [np] w * scale_phi / sin(theta)
    double *winvs2 = nullptr;    // [np] w * scale_phi / sin(theta)^2   (Robert-form vector analysis)
    double *invs = nullptr;      // [np] 1 / sin(theta)
    double2 *qc = nullptr, *qd = nullptr, *ab = nullptr;
    double *qb = nullptr, *qcc = nullptr;  // beta_l, c_l of qc, padded by OZ_TPAD zeros
    double *gtab = nullptr;
    long long *coff = nullptr;
    int *segbase = nullptr;
    double *seedq0 = nullptr, *seedq1 = nullptr;
    int *seedk = nullptr, *pfirst = nullptr, *pfast = nullptr, *segm = nullptr;
    double *convscale = nullptr;  // [(lmax+1)*(mmax+1)] configured -> canonical multiplier (non-canonical only)
    double2 *tw = nullptr;     // e^{-2 pi i t / nfft}, t < nfft
    double2 *chirp = nullptr;  // Bluestein chirp and kernel spectra (see FftArgs)
    double2 *bhat = nullptr;
    shtk::FftPlan fft;

    // benchmark instrumentation (shtk_profile_enable / shtk_profile_read)
    bool profiling = false;
    struct Pending {
        const char *name;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<std::string, std::pair<double, int>>> prof_totals;

    // lazily grown workspace
    void *ws = nullptr;
    size_t ws_bytes = 0;

    // split-precision tables (built on first use) and their workspace
    double *sectv = nullptr;  // [nm][np] scaled sectoral seeds
    int *sectk = nullptr;
    double *oz_q0 = nullptr, *oz_q1 = nullptr;  // [nm][4][np]
    int *oz_qk = nullptr, *oz_ep = nullptr, *oz_chbase = nullptr, *oz_chm = nullptr;
    int *oz_emaxl = nullptr, *oz_epl = nullptr;  // per (order, degree), qc indexing
    double *oz_asc = nullptr;
    int4 *oz_run = nullptr;  // analysis work runs: (order, first segment, segments, 0)
    int oz_nrun = 0, oz_run_len = 8;
    size_t qc_len = 0;
    int oz_ch_total = 0;
    void *oz_ws = nullptr;
    size_t oz_ws_bytes = 0;
    // companion plan of degree lmax + 1 on the same nodes (split-precision vector transforms)
    shtk_plan *vshadow = nullptr;
    // Serialises calls on this plan (host side); recursive because the uneven
    // batches hold it across calls of the single-configuration entry points.
    std::recursive_mutex mu;
    // Device-side ordering of calls that share the workspaces: the last call's
    // stream and an event recorded at its end (a call on another stream waits
    // for it first).
    hipEvent_t order_ev = nullptr;
    hipStream_t order_st = nullptr;
    bool order_used = false;
    bool order_multi = false;  // calls came from more than one stream: events are recorded
    // Second stream of the plan (made on first use, outside graph capture) and
    // its fork/join events: the order-0 rows of the vector transforms run on it
    // beside the main Legendre kernel.
    double *zseed = nullptr;  // order-0 segment starts of the vector pair (made on first use)
    int zsegs = 0;
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, side_join = nullptr;
    // Incremented whenever ws or oz_ws is reallocated: captured graphs that hold
    // the old addresses are stale once it changes.
    unsigned long long ws_gen = 0;
};

namespace shtk {

// Batched split-precision contraction (shtk_split.hip).  Fields are taken 16
// at a time ("field groups", 32 real columns).  Per order the degrees m..lmax
// are cut into 4 equal quarters of Kq = 32 * nch(m) degrees, nch(m) =
// ceil((lmax-m+1)/128); lane group g of a wavefront runs the recurrence over
// quarter g, 32 degrees per chunk.  The field operand is stored as balanced
// base-256 digits in the exact register layout of the int8 MFMA B operand.
constexpr int OZ_SLICES = 6;
constexpr int OZ_COLS = 32;
constexpr int OZ_FIELDS = OZ_COLS / 2;
// LDS row stride (double2) of the matrix-core synthesis kernels' staged recurrence
// coefficients, one row of 32 degrees per lane group: 33 puts the four rows a wave
// reads at once (one broadcast address per lane group) on different banks.
constexpr int OZ_ABS = 33;
constexpr int OZ_CHUNK_V4 = 2 * OZ_SLICES * 2 * 64;  // 16-byte B operands per chunk (1536 = 24 KB)

struct OzArgs {
    int lmax, mmax, mres, nlat, np, nm, ltr, nf, nfg;
    const double *x;
    const double2 *ab;
    const long long *coff;
    const double *q0, *q1;  // [nm][4][np] P_{lq-1}, P_{lq} at the quarter starts (scaled representation)
    const int *qk;          // [nm][4][np] scale count
    const int *ep;          // [nm][np] |P| < 2^ep over the order's degrees
    const int *chbase;      // [nm] first chunk of each order
    const int *chm;         // [ch_total] order of each chunk
    int ch_total;
    const void *spec_in;    // canonical coefficients (synthesis)
    int in_f32;
    long long spec_fstride;
    int4 *bdig;             // [nfg][ch_total][1536] B operands
    unsigned long long *colsum;  // [nfg][nm][2][32] sum over degrees of the integer field values, per parity
    int *eb;                // [nfg][nm][32] column exponents
    double2 *F0;            // Fourier data [field][im][nlat] (synthesis output, analysis input)
    double g;
    int xcd;
    // analysis
    const double2 *qc;      // Q-form (beta_l, c_l) per order (offset coff[im])
    const double *asc;      // c_l 2^(46 - epl_l) per order: digit scale of degree l
    const int *epl;         // |P_l| < 2^epl over all pairs, per order
    const double *qb, *qcc; // beta_l and c_l of qc (padded by OZ_TPAD)
    const double *seedq0, *seedq1;
    const int *seedk, *segbase, *segm, *pfirst;
    int seg, nseg_total, nks;
    int nh, npb;            // analysis pair blocks: nh halves of 256 pairs each, npb blocks
    const double *wlat;
    double *spec_out;       // FP64 canonical coefficients (pre-zeroed when npb > 1: accumulated)
    const int4 *run;        // [nrun] (order, first segment, segment count, 0)
    int nrun;
    int4 *gdig;             // [nfg][nm][nks][2][6][2][64] B operands (pairs as k)
    double *gcolsum;        // [nfg][nm][nks][2][32]
    int *eg;                // [nfg][nm][npb][2][32] column exponents per pair block
    float *gf32;            // FP32 contraction: analysis field operand (k_f32_gprep layout)
};

// Analysis with one pair block (npb == 1) stores every coefficient once and
// the output is not cleared beforehand: the work item of (run, field group)
// also stores the zeros of the coefficients no contraction of it writes -- the
// run's leading segments in which no pair is live, degrees past ltr, and, on
// the run that starts its order, l < m and the orders up to the next multiple
// of mres.  lr_live: first degree of the first live segment (the run's end if
// none).  Every thread of the block takes part.
__device__ inline void oz_zero_holes(const OzArgs &a, int im, int s0, int nsr, int lr_live, int fg) {
    const int m = im * a.mres;
    const int lr0 = m + s0 * a.seg, lr1 = min(m + (s0 + nsr) * a.seg, a.lmax + 1);
    const int ll = min(lr_live, lr1), le = max(ll, min(lr1, a.ltr + 1));
    const int nfl = min(OZ_FIELDS, a.nf - fg * OZ_FIELDS);
    double2 *base = (double2 *)a.spec_out + (size_t)fg * OZ_FIELDS * a.spec_fstride;
    auto zr = [&](int mm, int la, int lb) {  // degrees [la, lb) of order mm, every field of the group
        const int n = lb - la;
        if (n <= 0) return;
        for (int i = threadIdx.x; i < nfl * n; i += blockDim.x) {
            const int fi = i / n, l = la + (i - fi * n);
            base[(size_t)fi * a.spec_fstride + (size_t)(a.lmax + 1) * mm + l] = make_double2(0.0, 0.0);
        }
    };
    zr(m, lr0, ll);
    zr(m, le, lr1);
    if (s0 == 0) {
        zr(m, 0, m);
        for (int c = 1; c < a.mres && m + c <= a.mmax; ++c) zr(m + c, 0, a.lmax + 1);
    }
}

// Per-degree constants of a 32-degree chunk of the matrix-core analyses in
// VGPRs: degree 16 h + n of the chunk sits in lane n of each 16-lane row of
// a[h] (c_l 2^(46 - epl_l)), b[h] (beta_{l+1}) and c[h] (c_l), and reaches the
// FMAs as a DPP row_newbcast operand, so the digit loops have no scalar loads
// (whose waits also drain the LDS stores in flight).  The next chunk's are
// loaded while the current one runs.
struct OzRec {
    double a[2], b[2], c[2];
};

__device__ inline double readlane_f64(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

struct OzSeedArgs {
    int lmax, mres, np, nm;
    const double *x;
    const double2 *ab;
    const long long *coff;
    const double *sectv;
    const int *sectk;
    double *q0, *q1;
    int *qk, *ep;
    int *emaxl;             // per order and degree: max over pairs of ilogb|P_l| (atomicMax)
};

struct OzAscArgs {
    int n;                  // table length (all orders)
    const double2 *qc;
    const int *emaxl;
    double *asc;
    int *epl;
};

// Several configurations in one launch per stage, passed by value as kernel
// arguments (so a captured graph keeps them): prefix[g] is the running count
// of blocks (prologues) or work units (persistent contractions) before group
// g, prefix[ngroups] the total.
constexpr int OZ_MAX_GROUPS = 8;
struct OzGroups {
    OzArgs args[OZ_MAX_GROUPS];
    int prefix[OZ_MAX_GROUPS + 1];
    int ngroups;
    int *counter;  // work counter of the persistent kernels (reset in the stream before each launch)
};

hipError_t launch_oz_seed(const OzSeedArgs &a, hipStream_t st);
hipError_t launch_oz_synth_grp(const OzGroups &bexp, const OzGroups &bdig, const OzGroups &units, int in_f32,
                               int persistent_blocks, hipStream_t st);
hipError_t launch_oz_anal_grp(const OzGroups *gprep, const OzGroups *units, int persistent_blocks, hipStream_t st);
hipError_t launch_oz_synth(const OzArgs &a, hipStream_t st);
hipError_t launch_oz_asc(const OzAscArgs &a, hipStream_t st);
hipError_t launch_oz_anal(const OzArgs &a, hipStream_t st);
// FP32 matrix-core contraction (shtk_mfma32.hip): same arguments and outputs as
// the split kernels (synthesis reads the canonical coefficients directly; the
// analysis builds its FP32 field operand in a.gf32 and writes a.spec_out).
hipError_t launch_f32_synth(const OzArgs &a, hipStream_t st);
hipError_t launch_f32_anal(const OzArgs &a, hipStream_t st);

// Kernel launchers (shtk_legendre.hip, shtk_fft.hip).
hipError_t launch_seed(const SeedArgs &a, hipStream_t st);
hipError_t launch_leg_synth(const LegArgs &a, int nfields, int vector, int chains, hipStream_t st);
hipError_t launch_leg_anal(const LegArgs &a, int nfields, int vector, hipStream_t st);
hipError_t launch_fft(const FftArgs &a, int nfields, int inverse, hipStream_t st);
hipError_t launch_spec_scale(const void *in, void *out, int in_f32, int out_f32, const double *scale, int divide,
                             int lmax, int mmax, int mres, long long nelem_per_field, int nfields,
                             hipStream_t st);
hipError_t launch_spec_qscale(const void *in, double *out, int in_f32, const double *scale, int divide,
                              const double *qcc, const long long *coff, int lmax, int mres, int nm,
                              long long per_field, int nfields, hipStream_t st);
hipError_t launch_vspec_prep(const void *S, const void *T, double *rec, int in_f32, const double *scale, int divide,
                             int ll1, const double *qcc, const long long *coff, int lmax, int mres, int nm,
                             long long per_field, int nfields, hipStream_t st);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t st);

// Single-order transforms (shtk_ml.hip): one stored order, coefficient columns
// [field][l - m] (ld_spec = ltr - m + 1) and Fourier columns [field][lat] (ld_grid = nlat).
struct MlArgs {
    int lmax, m, ltr, nlat, np;
    int vector;          // 1: spheroidal/toroidal pair
    int robert;          // vector rows in Robert form
    int io_f32;          // complex float inputs and outputs
    const double *x;     // [np] cos(theta) of the north nodes
    const double *s;     // [np] sin(theta) of the north nodes
    const double *w;     // [np] weights times the phi scale (analysis)
    const double2 *ab;   // (a_l, b_l) of this order at index l - m
    const double *gtab;  // g_l of this order at index l - m
    const double *sectv; // [np] scaled sectoral seeds of this order
    const int *sectk;
    const double *conv;  // convention factors indexed by l, or nullptr (canonical)
    double g;            // synthesis: phi inverse scale; analysis: 1
    const void *in0, *in1;
    void *out0, *out1;
    double *acc;         // analysis: FP64 partials [field][pair block of 64][l - m][4]
    int nrow;            // rows of the Fourier columns (nlat, or nlat - 1: no south-pole row)
    long long ld_spec, ld_grid;
    int fourier_f64;     // the Fourier columns are FP64 whatever io_f32 says (plan workspace)
    // order 0 of the vector pair in segments of ZSEG degrees run in parallel (nullptr:
    // one serial pass): [3][zsegs][np] = P_{l0-1}, P_{l0}, dP_{l0-1}/dtheta at l0 = ZSEG s
    const double *zseed;
    int zsegs;
};
constexpr int ZSEG = 64;  // degrees per order-0 segment (the staging chunk of the kernels)
hipError_t launch_ml_synth(const MlArgs &a, int nfields, hipStream_t st);
hipError_t launch_ml_anal(const MlArgs &a, int nfields, hipStream_t st);
// Plan-time order-0 segment starts for MlArgs.zseed (one thread per pair runs the
// zonal recurrences once): out [3][nseg][np].
hipError_t launch_zonal_seed(const MlArgs &a, double *out, int nseg, hipStream_t st);
// Pole rows of order 1 for the vector transforms on grids with nodes at the
// poles (one block per field): synthesis rewrites them from the prologue
// records in a.in0 times wscale (robert = 1: zeros every order's pole rows),
// analysis adds the pole terms times wpole to the order-1 outputs in a.out0 / a.out1
// (conv: the outputs are in the plan's convention map, or nullptr).
hipError_t launch_vpole_synth(const LegArgs &a, int nfields, int robert, double wscale, hipStream_t st);
hipError_t launch_vpole_anal(const LegArgs &a, int nfields, double wpole, const double *conv, hipStream_t st);
// FFTs of complex grid rows: synthesis (inverse = 1) turns the bins m = 0 .. mmax
// of F into complex rows; analysis (inverse = 0) keeps the bins m = 0 .. mmax of
// each complex row's forward transform.
hipError_t launch_fft_cplx(const FftArgs &a, int nfields, int inverse, hipStream_t st);
// Split-precision vector transforms: (S, T) -> coefficients of 2 nf scalar
// fields of degree lmax + 1, and the analysis sums of those back to (S, T).
hipError_t launch_vsplit_coef(const void *S, const void *T, double *C, int in_f32, const double *scale, int divide,
                              int lmax, int mres, int nm, int ltr, long long in_fstride, long long out_fstride,
                              int nf, hipStream_t st);
hipError_t launch_vsplit_out(const double *G, void *S, void *T, int out_f32, const double *scale, int divide,
                             int lmax, int mmax, int mres, int ltr, long long g_fstride, long long out_fstride,
                             int nf, hipStream_t st);

}  // namespace shtk

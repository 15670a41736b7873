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
constexpr int SEG_MAX = 256;

// Coefficient tables, one block per stored order (offset coff[im], entries
// l = m .. lmax+1):
//   qc[l] = (beta_l, c_l)                  scalar and vector recurrences
//   qd[l] = (l c_l, g_l c_{l-1})           vector: s dP_l/dtheta = l x P_l - g_l P_{l-1},
//                                          g_l = sqrt((2l+1)(l^2-m^2)/(2l-1))
struct LegArgs {
    int lmax, mmax, mres, nlat, np, nm, ltr, seg;
    const double *x;    // [np] cos(theta) of the north node of each pair
    const double *wlat; // [np] per-latitude factor, or nullptr for 1
    const double2 *qc;
    const double2 *qd;
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
};

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

struct FftArgs {
    int nlat, nlon, np, nm, mres, mmax;
    int nrad;
    unsigned long long radix; // 4 bits per pass, first pass in the low bits
    const double2 *tw;        // e^{-2 pi i t / nlon}
    double2 *F;               // Fourier workspace [field][im][nlat]
    void *grid;               // grid array (nlat, nlon, nfields)
    int io_f32;
    int mode;          // grid row factor: 0 none, 1 1/sin(theta), 2 sin(theta) (Robert form)
    const double *s;   // [np] sin(theta) for Robert scaling
};

}  // namespace shtk

struct shtk_plan {
    shtk_desc d{};
    int np = 0, nm = 0, seg = 128, nseg_total = 0;
    bool canonical = true;
    double inv_scale = 0, scale_phi = 0;
    std::vector<double> hx, hw;  // north-first nodes and weights (nlat)
    int device = 0;

    // device tables
    double *x = nullptr, *s = nullptr;   // [np]
    double *wscaled = nullptr;   // [np] w * scale_phi
    double *winvs = nullptr;     // [np] w * scale_phi / sin(theta)
    double *winvs2 = nullptr;    // [np] w * scale_phi / sin(theta)^2   (Robert-form vector analysis)
    double *invs = nullptr;      // [np] 1 / sin(theta)
    double2 *qc = nullptr, *qd = nullptr, *ab = nullptr;
    double *gtab = nullptr;
    long long *coff = nullptr;
    int *segbase = nullptr;
    double *seedq0 = nullptr, *seedq1 = nullptr;
    int *seedk = nullptr, *pfirst = nullptr, *pfast = nullptr, *segm = nullptr;
    double *convscale = nullptr;  // [(lmax+1)*(mmax+1)] configured -> canonical multiplier (non-canonical only)
    double2 *tw = nullptr;
    std::vector<int> radix;  // FFT factorization of nlon (empty: unsupported length)

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
    std::mutex mu;
};

namespace shtk {

// Kernel launchers (shtk_legendre.hip, shtk_fft.hip).
hipError_t launch_seed(const SeedArgs &a, hipStream_t st);
hipError_t launch_leg_synth(const LegArgs &a, int nfields, int vector, hipStream_t st);
hipError_t launch_leg_anal(const LegArgs &a, int nfields, int vector, hipStream_t st);
hipError_t launch_fft(const FftArgs &a, int nfields, int inverse, hipStream_t st);
hipError_t launch_spec_scale(const void *in, void *out, int in_f32, int out_f32, const double *scale, int divide,
                             int lmax, int mmax, int mres, long long nelem_per_field, int nfields,
                             hipStream_t st);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t st);

}  // namespace shtk

// Internal declarations shared by the host plan code and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
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

// Arguments of the Legendre-stage kernels.  The Fourier intermediate F is
// [field][im][i] complex double (im = m / mres, i = 0..nlat-1, north first).
struct LegArgs {
    int lmax, mmax, mres, nlat, np, nm, ltr;
    const double *x;    // [np] cos(theta) of the north node of each pair
    const double *s;    // [np] sin(theta)
    const double *wlat; // [np] per-latitude factor, or nullptr for 1
    const double2 *ab;  // per order: (a_l, b_l) for l = m .. lmax+1
    const long long *aboff;
    const double *seedv; // [nm][np] scaled sectoral seed
    const int *seedk;    // [nm][np] scale count k
    const double2 *dd;   // per order: (l*d_{l+1,m}, (l+1)*d_{l,m}) for l = m .. lmax (vector transforms)
    double g;            // global factor
    int mfac_mode;       // 0: 1 for all m; 1: 2 for m >= 1; 2: 1/2 for m >= 1
    int io_f32;          // spectral arrays are complex float
    const void *in0, *in1; // spectral inputs (synthesis) or unused
    void *out0, *out1;     // spectral outputs (analysis) or unused
    long long spec_fstride; // complex elements per field in spectral arrays
    double2 *F0, *F1;       // Fourier workspace (F1 for vector transforms)
    const double *ll1inv;   // [lmax+1] 1/(l(l+1)) (l >= 1), 0 at l = 0; vector analysis only
};

struct FftArgs {
    int nlat, nlon, np, nm, mres, mmax;
    int nrad;
    int radix[24];
    const double2 *tw; // e^{-2 pi i t / nlon}
    double2 *F;        // Fourier workspace [field][im][nlat]
    void *grid;        // grid array (nlat, nlon, nfields)
    int io_f32;
    int mode;          // grid row factor: 0 none, 1 1/sin(theta), 2 sin(theta) (Robert form)
    const double *s;   // [np] sin(theta) for Robert scaling
    int xcd_groups;    // pairs per XCD group for block remapping
};

}  // namespace shtk

struct shtk_plan {
    shtk_desc d{};
    int np = 0, nm = 0;
    bool canonical = true;
    double inv_scale = 0, scale_phi = 0;
    std::vector<double> hx, hw;  // north-first nodes and weights (nlat)
    int device = 0;

    // device tables
    double *x = nullptr, *s = nullptr, *w = nullptr;   // [np]
    double *wscaled = nullptr;                         // [np] w * scale_phi
    double *wadj = nullptr;                            // [np] w * nlon * scale_phi / nlon (adjoint analysis)
    double2 *ab = nullptr;
    long long *aboff = nullptr;
    double2 *dd = nullptr;
    double *seedv = nullptr;
    int *seedk = nullptr;
    double *convscale = nullptr;  // [(lmax+1)*(mmax+1)] configured -> canonical multiplier (non-canonical only)
    double *ll1inv = nullptr;
    double2 *tw = nullptr;
    std::vector<int> radix;  // FFT factorization of nlon (empty: unsupported length)

    // lazily grown workspace
    void *ws = nullptr;
    size_t ws_bytes = 0;
    std::mutex mu;
};

namespace shtk {

// Kernel launchers (shtk_legendre.hip, shtk_fft.hip).
hipError_t launch_leg_synth(const LegArgs &a, int nfields, int vector, hipStream_t st);
hipError_t launch_leg_anal(const LegArgs &a, int nfields, int vector, hipStream_t st);
hipError_t launch_fft(const FftArgs &a, int nfields, int inverse, hipStream_t st);
hipError_t launch_spec_scale(const void *in, void *out, int in_f32, int out_f32, const double *scale, int divide,
                             int lmax, int mmax, int mres, long long nelem_per_field, int nfields,
                             hipStream_t st);

}  // namespace shtk

/*
 * shtk.h -- C ABI of libshtk_mi355x.so, the MI355X (gfx950) spherical harmonic
 * transform backend for SHTnsKit.jl.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are HIP
 * device allocations (ROCArray / torch tensor storage); `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  Work is enqueued on
 * that stream and the call returns without synchronising, except where noted;
 * results are ready once the stream reaches them.  The Julia adapter calls
 * AMDGPU.synchronize() after each call, as ext/SHTnsKitAMDGPUExt.jl:696,746,786
 * does, so SHTnsKit's "results ready on return" contract holds.
 *
 * Memory layouts are SHTnsKit's (Julia column-major):
 *   grid      f[i + nlat*j + nlat*nlon*k]            i latitude (north first), j longitude, k field
 *   spectrum  a[l + (lmax+1)*m + (lmax+1)*(mmax+1)*k] complex as interleaved (re, im)
 * Entries with l < m (and orders m that are not multiples of mres) are written
 * as zero on output and ignored on input.
 *
 * dtype selects the element type of the caller's arrays (SHTK_F64: double /
 * complex double, SHTK_F32: float / complex float).  All Legendre recurrences,
 * reductions and FFTs run in FP64 whatever the I/O type.
 *
 * Return value: SHTK_OK (0) or an error code; shtk_last_error() gives the
 * message of the last failure on the calling thread.
 */
#ifndef SHTK_H
#define SHTK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum shtk_status {
    SHTK_OK = 0,
    SHTK_EINVAL = 1,      /* bad size, pointer or keyword (DimensionMismatch / ArgumentError) */
    SHTK_EUNSUPPORTED = 2, /* valid request this backend does not implement (BackendUnavailableError) */
    SHTK_ENOMEM = 3,      /* device allocation failed */
    SHTK_EHIP = 4         /* HIP runtime error */
};

enum shtk_norm { SHTK_NORM_ORTHONORMAL = 0, SHTK_NORM_FOURPI = 1, SHTK_NORM_SCHMIDT = 2 };
enum shtk_phi_scale { SHTK_PHI_DFT = 0, SHTK_PHI_QUAD = 1 };
/* Latitude grids of create_gauss_config and create_regular_config (src/config.jl:633-655,
 * 1235-1330): Gauss-Legendre nodes; Fejer-1 midpoints theta = (i + 1/2) pi / nlat
 * (include_poles = false); Clenshaw-Curtis nodes theta = i pi / (nlat - 1) with both
 * poles (include_poles = true); Driscoll-Healy nodes theta = i pi / nlat with the
 * weights of driscoll_healy_weights (src/mathutils.jl:149-183; use_dh_weights = true):
 * the north pole is a node, the south pole is not. */
enum shtk_grid { SHTK_GRID_GAUSS = 0, SHTK_GRID_REGULAR = 1, SHTK_GRID_REGULAR_POLES = 2, SHTK_GRID_DRISCOLL_HEALY = 3 };
enum shtk_dtype { SHTK_F64 = 0, SHTK_F32 = 1 };

/* Contraction precision of the batched entry points (*_ex).
 * SHTK_PREC_FP64: the FP64 recurrence-and-contraction kernels (the plain path).
 * SHTK_PREC_SPLIT: FP64 Legendre values and tables, contraction as an Ozaki
 *   split on int8 matrix cores (6 slices, exact int32 accumulation, FP64
 *   cross-slice sums); FP64-grade results (relative L2 error ~1e-12).
 * SHTK_PREC_FP32: FP64 Legendre values and tables, contraction on the FP32
 *   matrix cores (v_mfma_f32_16x16x4_f32) with FP32 accumulation over K-tiles of
 *   64 products and FP64 sums across tiles; FP32-grade results (relative L2
 *   error ~1e-7 per field). */
enum shtk_precision { SHTK_PREC_FP64 = 0, SHTK_PREC_SPLIT = 1, SHTK_PREC_FP32 = 2 };

/* Grid and spectral descriptor.  Mirrors the keyword set of
 * create_gauss_config(lmax, nlat; mmax, mres, nlon, norm, cs_phase, real_norm,
 * robert_form) (src/config.jl:633-655) and of create_regular_config(...;
 * include_poles) (src/config.jl:1235-1330) plus the phi scaling mode
 * (src/SHTnsKit.jl:141-177).  Constraints as src/config.jl:465-481 and the
 * grid minimums of _min_nlat (src/config.jl:452-458): nlat >= lmax+1 (Gauss,
 * regular with poles, and at least 2 there), nlat >= lmax+2 (regular),
 * nlat = 2(lmax+1) (Driscoll-Healy), nlon >= 2*mmax+1, 0 <= mmax <= lmax, mres >= 1. */
typedef struct shtk_desc {
    int lmax, mmax, mres, nlat, nlon;
    int norm;        /* enum shtk_norm */
    int cs_phase;    /* 1: Condon-Shortley phase */
    int real_norm;   /* 1: SHTns REAL_NORM convention */
    int robert_form; /* 1: vector fields in Robert form (multiplied by sin(theta)) */
    int phi_scale;   /* enum shtk_phi_scale */
    int grid;        /* enum shtk_grid */
} shtk_desc;

typedef struct shtk_plan shtk_plan;

/* Builds a plan: nodes and weights (Gauss-Legendre src/legendre.jl:956-1003,
 * north first as src/config.jl:639; Fejer-1 and Clenshaw-Curtis
 * src/legendre.jl:1017-1075 as create_regular_config), FP64 recurrence coefficients, sectoral
 * seeds with scaled exponents, FFT twiddles.  Replaces create_gauss_config
 * (src/config.jl:633-655) and the device table builders of
 * ext/SHTnsKitAMDGPUExt.jl:495-557 (which hold a full nlat x (lmax+1) x (mmax+1)
 * table; this plan holds O(lmax^2 + nlat*mmax) data).  Uses the current HIP
 * device; allocates device memory. */
int shtk_plan_create(const shtk_desc *desc, shtk_plan **plan);

/* Single-order transforms of stored order im (physical m = im * mres) for degrees
 * m..ltr (ltr < 0: lmax; ltr >= m), with no phi transform.  Coefficient columns
 * Ql (ltr - m + 1, nfields) and Fourier columns Vm (nlat, nfields) are complex
 * (interleaved re/im of dtype).  Synthesis: Vm = phi_inv_scale * sum_l Ql' P_l^m;
 * analysis: Ql = (the convention map of) phi_scale * sum_i w_i Vm_i P_l^m(x_i).
 * Replace synthesis_packed_ml / analysis_packed_ml (src/transforms.jl:351-437). */
int shtk_synthesis_ml(shtk_plan *plan, int im, const void *Ql, void *Vm, int nfields, int dtype, int ltr,
                      void *stream);
int shtk_analysis_ml(shtk_plan *plan, int im, const void *Vm, void *Ql, int nfields, int dtype, int ltr,
                     void *stream);
/* Vector pair of one order: Vt = dS/dtheta - (i m / sin) T, Vp = (i m / sin) S + dT/dtheta
 * (times sin(theta) in Robert form), and the analysis that inverts it, with the order-1
 * pole limits on grids with poles.  Replace synthesis_sphtor_ml / analysis_sphtor_ml
 * (src/sphtor_transforms.jl:1188-1322); synthesis_sph_ml, synthesis_tor_ml,
 * synthesis_grad_ml and the QST _ml pair (src/qst_transforms.jl:386-425) compose them. */
int shtk_synthesis_sphtor_ml(shtk_plan *plan, int im, const void *Sl, const void *Tl, void *Vt, void *Vp,
                             int nfields, int dtype, int ltr, void *stream);
int shtk_analysis_sphtor_ml(shtk_plan *plan, int im, const void *Vt, const void *Vp, void *Sl, void *Tl,
                            int nfields, int dtype, int ltr, void *stream);

/* Frees the plan and its workspace (gpu_clear_cache!, ext/GPUCommon.jl:202-441). */
int shtk_plan_destroy(shtk_plan *plan);

/* Host copies of cfg.x (cos theta) and cfg.w, north first, length nlat each. */
int shtk_plan_nodes(const shtk_plan *plan, double *x, double *w);

/* Bytes of device workspace a call with `nfields` fields uses (grown lazily). */
size_t shtk_workspace_size(const shtk_plan *plan, int nfields, int vector);

/* Scalar synthesis: alm (lmax+1, mmax+1, nfields) complex -> f (nlat, nlon, nfields).
 * ltr < 0 means lmax; 0 <= ltr <= lmax truncates the degree sum.  real_output = 1:
 * the real grid (Hermitian mirror of the bins m > 0); real_output = 0: the complex
 * grid (interleaved re/im of dtype) of the bins m >= 0 alone, as synthesis_cplx
 * and _synthesis with real_output = false (src/core_transforms.jl:289-344).
 * Replaces synthesis / synthesis_batch (src/core_transforms.jl:271-344,
 * src/batch_transforms.jl:455-655) behind _gpu_adapter_synthesis
 * (src/device_utils.jl:91-93). */
int shtk_synthesis(shtk_plan *plan, const void *alm, void *f, int nfields, int dtype, int real_output,
                   int ltr, void *stream);

/* Scalar synthesis that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as shtk_synthesis (FP64 contraction), with a
 * fixed workspace: a device array of SHTK_STATIC_SCRATCH_BYTES per GPU, held
 * until the process ends (set up by shtk_plan_create).  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other streams, other
 * plans) race; a call whose workspace, shtk_workspace_size(plan, nfields, 0),
 * exceeds the array returns SHTK_EINVAL (one field at lmax 1023 needs 33.6 MB). */
#define SHTK_STATIC_SCRATCH_BYTES (64u << 20)
int shtk_synthesis_static(shtk_plan *plan, const void *alm, void *f, int nfields, int dtype, int real_output,
                          int ltr, void *stream);

/* Scalar analysis: real f (nlat, nlon, nfields) -> alm.  Replaces analysis /
 * analysis_batch (src/core_transforms.jl:169-207, src/batch_transforms.jl:312-429)
 * behind _gpu_adapter_analysis (src/device_utils.jl:87-89). */
int shtk_analysis(shtk_plan *plan, const void *f, void *alm, int nfields, int dtype, int ltr, void *stream);

/* Batched scalar synthesis with a contraction precision (enum shtk_precision):
 * the batch transforms of src/batch_transforms.jl:455-655, which share one
 * Legendre row per (m, latitude) across all fields (:519-535).  Same layouts and
 * arguments as shtk_synthesis; SHTK_PREC_SPLIT needs a scalar canonical-or-
 * convertible plan with lmax <= 4095. */
int shtk_synthesis_ex(shtk_plan *plan, const void *alm, void *f, int nfields, int dtype, int real_output,
                      int ltr, int precision, void *stream);

/* Batched scalar analysis with a contraction precision: the batch analysis of
 * src/batch_transforms.jl:337-429 (one Legendre row per (m, latitude) shared by
 * all fields, :403-418).  SHTK_PREC_SPLIT stores every coefficient once when
 * nlat <= 1024 (bit-reproducible); larger grids add the partial sums of several
 * latitude blocks with device atomics, so their last bits may differ from run to run. */
int shtk_analysis_ex(shtk_plan *plan, const void *f, void *alm, int nfields, int dtype, int ltr, int precision,
                     void *stream);

/* Scalar analysis of complex fields: f (nlat, nlon, nfields) complex (interleaved
 * re/im of dtype) -> alm.  analysis with a complex eltype (src/core_transforms.jl:
 * 169-207): the m-loop reads the bins m >= 0 of the complex FFT (no Hermitian
 * fold).  precision as shtk_analysis_ex. */
int shtk_analysis_cplx(shtk_plan *plan, const void *f, void *alm, int nfields, int dtype, int ltr, int precision,
                       void *stream);

/* Vector synthesis (S, T) -> (Vt, Vp): synthesis_sphtor / synthesis_sphtor_batch
 * (src/sphtor_transforms.jl:162-241, src/batch_transforms.jl:745-815), behind
 * _gpu_adapter_synthesis_sphtor (src/sphtor_transforms.jl:304-311). */
int shtk_synthesis_sphtor(shtk_plan *plan, const void *S, const void *T, void *Vt, void *Vp, int nfields,
                          int dtype, int real_output, int ltr, void *stream);

/* Vector analysis (Vt, Vp) -> (S, T): analysis_sphtor / analysis_sphtor_batch
 * (src/sphtor_transforms.jl:252-289, src/batch_transforms.jl:689-725), behind
 * _gpu_adapter_analysis_sphtor (src/sphtor_transforms.jl:298-302).  On grids with
 * nodes at the poles the order-1 rows use the pole limits (src/kernels.jl:72-99);
 * Robert form with order 1 stored and nonzero pole weights (Clenshaw-Curtis) is
 * SHTK_EINVAL, as _validate_robert_analysis (src/config.jl:608-624). */
int shtk_analysis_sphtor(shtk_plan *plan, const void *Vt, const void *Vp, void *S, void *T, int nfields,
                         int dtype, int ltr, void *stream);

/* Batched vector synthesis with a contraction precision: synthesis_sphtor_batch
 * (src/batch_transforms.jl:745-815) and synthesis_sphtor (src/sphtor_transforms.jl:162-241).
 * Same layouts and arguments as shtk_synthesis_sphtor.  SHTK_PREC_SPLIT runs
 * sin(theta) V as a scalar split synthesis of degree lmax + 1 (the derivative
 * rows written as l d_{l+1} P_{l+1} - (l+1) d_l P_{l-1}); needs lmax <= 4094. */
int shtk_synthesis_sphtor_ex(shtk_plan *plan, const void *S, const void *T, void *Vt, void *Vp, int nfields,
                             int dtype, int real_output, int ltr, int precision, void *stream);

/* Batched vector analysis with a contraction precision: analysis_sphtor_batch
 * (src/batch_transforms.jl:689-725) and analysis_sphtor (src/sphtor_transforms.jl:252-289).
 * SHTK_PREC_SPLIT runs two scalar split analyses of degree lmax + 1 per field
 * with the weights w/sin(theta) and combines them in FP64; needs lmax <= 4094. */
int shtk_analysis_sphtor_ex(shtk_plan *plan, const void *Vt, const void *Vp, void *S, void *T, int nfields,
                            int dtype, int ltr, int precision, void *stream);

/* Adjoint of synthesis: real cotangent fbar (nlat, nlon, nfields) -> almbar.
 * _adjoint_synthesis, src/core_transforms.jl:477-519 (pullback of the
 * synthesis rrule, ext/SHTnsKitAdvancedADExt.jl:76-89, and of the batch rrule
 * :113-133). */
int shtk_adjoint_synthesis(shtk_plan *plan, const void *fbar, void *almbar, int nfields, int dtype,
                           int real_output, void *stream);

/* Adjoint of analysis: almbar -> real fbar (nlat, nlon, nfields), the real part
 * of _adjoint_analysis (src/core_transforms.jl:611-652) that ProjectTo(f) keeps
 * for a real primal (ext/SHTnsKitAdvancedADExt.jl:55-67). */
int shtk_adjoint_analysis(shtk_plan *plan, const void *almbar, void *fbar, int nfields, int dtype,
                          void *stream);

/* The whole (complex) adjoint of analysis, _adjoint_analysis (src/core_transforms.jl:611-652):
 * almbar -> complex fbar (nlat, nlon, nfields), interleaved re/im of dtype: the inverse
 * FFT of the bins m >= 0 (the negative bins are zero).  What the analysis rrule pulls
 * back for a complex primal (ext/SHTnsKitAdvancedADExt.jl:55-67). */
int shtk_adjoint_analysis_cplx(shtk_plan *plan, const void *almbar, void *fbar, int nfields, int dtype,
                               void *stream);

/* Vector adjoints: _adjoint_synthesis_sphtor (src/core_transforms.jl:537-597)
 * and the real part of _adjoint_analysis_sphtor (src/sphtor_transforms.jl:385-452). */
int shtk_adjoint_synthesis_sphtor(shtk_plan *plan, const void *Vtb, const void *Vpb, void *Sb, void *Tb,
                                  int nfields, int dtype, int real_output, void *stream);
int shtk_adjoint_analysis_sphtor(shtk_plan *plan, const void *Sb, const void *Tb, void *Vtb, void *Vpb,
                                 int nfields, int dtype, void *stream);
/* The whole (complex) _adjoint_analysis_sphtor (src/sphtor_transforms.jl:385-452):
 * complex Vtb, Vpb (interleaved re/im of dtype). */
int shtk_adjoint_analysis_sphtor_cplx(shtk_plan *plan, const void *Sb, const void *Tb, void *Vtb, void *Vpb,
                                      int nfields, int dtype, void *stream);

/* Per-kernel timing for benchmarks; no counterpart in the reference.  While
 * enabled, every kernel this plan launches is bracketed by HIP events on the
 * caller's stream.  shtk_profile_read() waits for the recorded events and
 * returns, per kernel name (up to `max` entries, names NUL-terminated in
 * 48-byte slots), the accumulated milliseconds and launch counts, then resets. */
/* ------------------------------------------------------------------------
 * Uneven batches: ngroups configurations (descs[g], nfields[g] fields each)
 * transformed together, e.g. many fields of different lmax.  The reference
 * handles such a set as one transform per field, or one batch call per
 * configuration (src/batch_transforms.jl:312-655); here every stage is one
 * launch for all groups: with SHTK_PREC_SPLIT the Legendre contractions run as
 * persistent kernels over a work queue ordered largest-first.  Arrays are
 * passed per group (alm[g] of shape (lmax+1, mmax+1, nfields[g]), f[g] of
 * shape (nlat, nlon, nfields[g])); ngroups <= 8. */
typedef struct shtk_batch shtk_batch;
int shtk_batch_create(const shtk_desc *descs, const int *nfields, int ngroups, int precision, shtk_batch **batch);
int shtk_batch_destroy(shtk_batch *batch);
/* The plan of group g (owned by the batch). */
int shtk_batch_plan(shtk_batch *batch, int group, shtk_plan **plan);
/* on = 1: each call captures its launch sequence in a hipGraph on first use
 * (per direction, arrays, dtype and stream; the stream must not be the
 * default stream) and replays it with a single graph launch afterwards. */
int shtk_batch_set_graph(shtk_batch *batch, int on);
/* Launch shapes: grid sizes of the persistent synthesis and analysis kernels
 * (0: two blocks per CU). */
int shtk_batch_set_tiles(shtk_batch *batch, int synth_blocks, int anal_blocks);
int shtk_batch_synthesis(shtk_batch *batch, const void *const *alm, void *const *f, int dtype, void *stream);
int shtk_batch_analysis(shtk_batch *batch, const void *const *f, void *const *alm, int dtype, void *stream);
/* Autotuner: times round trips (synthesis into f, analysis back into alm, which
 * therefore ends up holding its own projection) for every candidate launch
 * shape and keeps the fastest; reports it if the pointers are not NULL. */
int shtk_batch_autotune(shtk_batch *batch, void *const *alm, void *const *f, int dtype, int reps, void *stream,
                        int *best_synth_blocks, int *best_anal_blocks, double *best_ms);

int shtk_profile_enable(shtk_plan *plan, int on);
int shtk_profile_read(shtk_plan *plan, int max, char *names, double *total_ms, int *count, int *n_out);

/* Blocks until `stream` is idle. */
int shtk_stream_sync(void *stream);

/* Copies the last error message of the calling thread into buf (NUL-terminated);
 * returns the status code of that error. */
int shtk_last_error(char *buf, size_t len);

/* Library version string, e.g. "0.1.0". */
const char *shtk_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SHTK_H */

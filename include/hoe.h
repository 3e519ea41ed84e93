/*
 * hoe.h -- C ABI of libhoe, the MI355X (gfx950) Explicit higher-order MRI
 * encoding operator.
 *
 * The operator evaluates, for sample s and masked voxel v,
 *
 *     phi(s,v) = sum_l k_l(s) * b_l(v) + times[s] * fieldmap[v]      (cycles)
 *     E[(c,s), v] = exp(+2*pi*i*phi(s,v)) * csm[v,c] / sqrt(n_vox)
 *
 * with b_l the 9 or 16 unnormalised real solid harmonics (plus 4 concomitant
 * columns z^2, x^2+y^2, xz, yz when concomitant coefficients are given).
 * Reference model: HighOrderMRI.jl src/EncodingOperator/Explicit/
 * array_operator.jl:241-348 and docs/src/theory/encoding-model.md:77-123.
 *
 * Conventions shared by every entry point:
 *   - complex numbers are interleaved float pairs (re, im), the memory layout
 *     of Julia ComplexF32 and numpy complex64;
 *   - image vectors span the full grid (n_x*n_y*n_z, x fastest); voxels
 *     outside the mask are ignored on input and written as zero on output;
 *   - signal vectors are (n_sam, n_cha) with samples fastest, i.e. entry
 *     s + c*n_sam (array_operator.jl:271, encoding-model.md:316-337);
 *   - forward and adjoint include 1/sqrt(n_vox); the normal includes 1/n_vox;
 *   - flags: HOE_DEVICE_PTRS marks vector arguments as device pointers on the
 *     handle's GPU (otherwise host pointers, copied in and out); HOE_ASYNC
 *     returns without synchronising the handle's stream (device pointers only);
 *   - a handle is not reentrant: one thread at a time per handle; distinct
 *     handles may be used from distinct threads (explicit_gpu_operator.jl:66-67);
 *   - every call returns a hoe_status; hoe_last_error() gives the message.
 */
#ifndef HOE_H
#define HOE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hoe_handle hoe_handle;

/* Status codes. The host bindings map them to the reference's exceptions:
 * EINVAL -> ArgumentError, EDIM -> DimensionMismatch, EASSERT ->
 * AssertionError, ENOMEM -> OutOfMemoryError, ECLOSED -> ArgumentError
 * ("operator is closed", explicit_gpu_operator.jl:40-44). */
typedef enum {
    HOE_OK = 0,
    HOE_EINVAL = 1,
    HOE_EDIM = 2,
    HOE_EASSERT = 3,
    HOE_ENOMEM = 4,
    HOE_EHIP = 5,
    HOE_ECLOSED = 6,
    HOE_ENODEV = 7
} hoe_status;

/* Execution backends for one handle (switchable with hoe_set_backend). */
typedef enum {
    /* FP64 phase, FP32 sincos, FP32 VALU contraction with all coils of a
     * batch in registers (replaces explicit_forward_kernel! /
     * explicit_adjoint_kernel!, CUDA_kernel.jl:91-671). */
    HOE_BACKEND_DIRECT = 0,
    /* FP64 phase, FP32 sincos, exp(i*phi) tiles split hi+lo in FP16 and
     * contracted on MFMA (v_mfma_f32_32x32x16_f16) with FP32 accumulation
     * (replaces the phase-reuse GEMM path, explicit_gpu_operator.jl:117-297). */
    HOE_BACKEND_MIXED_FP16 = 1,
    /* As MIXED_FP16 with BF16 hi+lo tiles (v_mfma_f32_32x32x16_bf16). */
    HOE_BACKEND_MIXED_BF16 = 2,
    /* Float64 operator (the reference's ComplexF64 operators, test_explicit_gpu.jl:51):
     * FP64 model, FP64 phase, FP64 sincos (device math library sincospi after an exact
     * reduction modulo one cycle), FP64 VALU contraction and reductions. Handles of this
     * backend come from hoe_create_f64 and take ComplexF64 vectors; they cannot switch
     * to another backend, nor another handle to this one. */
    HOE_BACKEND_FP64 = 3
} hoe_backend;

enum {
    HOE_DEVICE_PTRS = 1,
    HOE_ASYNC = 2
};

/* n_d value for hoe_normal / hoe_cg (with D == NULL): reuse the weights
 * stored by the last hoe_set_weights / hoe_normal / hoe_cg call. */
#define HOE_CURRENT_WEIGHTS (-1)

/* Model description for hoe_create. Every pointer is a HOST pointer; the
 * library copies what it needs to the device. Mirrors the keyword arguments
 * of HighOrderExplicitOp(grid, kspha, times; fieldmap, csm, mask,
 * concomitant, recon_terms, k_nominal, gpus, backend)
 * (kernel_operator.jl:199-221, array_operator.jl:83-97). */
typedef struct {
    int32_t n_x, n_y, n_z;       /* grid.matrixSize */
    const float *x, *y, *z;      /* grid.x/.y/.z: n_grid coordinates in metres, x fastest (Grid.jl:65-96) */
    const uint8_t *mask;         /* n_grid entries, nonzero = inside; NULL = all voxels */
    int32_t n_term;              /* rows of kspha: 9 or 16 */
    int32_t n_sam;               /* samples */
    const float *kspha;          /* n_term x n_sam, term fastest; cycles/m^p */
    const float *times;          /* n_sam, seconds */
    const float *fieldmap;       /* n_grid, Hz; NULL = zeros */
    int32_t n_cha;               /* coils */
    const float *csm;            /* complex n_grid x n_cha, grid fastest; NULL = ones (n_cha must be 1) */
    const float *concomitant;    /* 4 x n_sam (z^2, x^2+y^2, xz, yz), cycles/m^2; NULL = none (prep_kspha.jl:78-98) */
    const char *recon_terms;     /* binary order flags "1111"/"111"; NULL = all (prep_kspha.jl:1-56) */
    const float *k_nominal;      /* 3 x n_sam first-order rows used when the first-order flag is 0; NULL = kspha rows 2..4 */
    double dwell;                /* seconds between coefficient samples, for delays (hoe_set_coefficients); 0 = unknown */
    int32_t backend;             /* hoe_backend */
    int32_t device;              /* HIP device ordinal ("gpus=[id]") */
} hoe_model_desc;

/* Model description for hoe_create_f64: the fields of hoe_model_desc with Float64
 * data (Grid{Float64}, Float64 kspha/times/fieldmap, ComplexF64 csm), for the
 * FP64 backend. No dwell (delays are not provided in FP64) and no backend field. */
typedef struct {
    int32_t n_x, n_y, n_z;
    const double *x, *y, *z;
    const uint8_t *mask;
    int32_t n_term;
    int32_t n_sam;
    const double *kspha;         /* n_term x n_sam, term fastest */
    const double *times;
    const double *fieldmap;      /* NULL = zeros */
    int32_t n_cha;
    const double *csm;           /* complex n_grid x n_cha interleaved (re, im), grid fastest; NULL = ones */
    const double *concomitant;   /* 4 x n_sam; NULL = none */
    const char *recon_terms;
    const double *k_nominal;     /* 3 x n_sam; NULL = kspha rows 2..4 */
    int32_t device;
} hoe_model_desc_f64;

typedef struct {
    int64_t n_grid;   /* prod(matrixSize): length of image vectors */
    int64_t n_vox;    /* voxels inside the mask */
    int64_t n_sam;
    int64_t n_cha;
    int32_t n_term;   /* effective phase terms (n_term, +4 with concomitant) */
    int32_t backend;
    int32_t device;
    int32_t closed;
    int32_t horner;   /* phase evaluation: bit 0 = Horner runs in the forward kernels, bit 1 = in the MFMA adjoint,
                         bit 2 = the MFMA kernels evaluate anchored phases (FP64 anchor, FP32 offsets) on them
                         for the coefficients last set (the fieldmap offsets and the Horner offsets within
                         bounds; reading it waits for the handle's queued work) */
    int32_t n_runs;   /* runs of consecutive masked voxels sharing (y, z) */
} hoe_info_t;

/* Library version string. */
const char *hoe_version(void);

/* Message of the most recent failure of hoe_create on this thread, or of
 * the most recent failing call on h (h != NULL). */
const char *hoe_last_error(const hoe_handle *h);

/* Number of visible GPUs (0 when no GPU or no HIP runtime). Never fails. */
int hoe_device_count(void);

/* Construct the operator: validate, prepare the coefficients
 * (prep_kspha / prep_concomitant), evaluate the FP64 basis and upload the
 * masked model. Replaces CUDAExplicitOp / ExplicitGPUOp
 * construction (kernel_operator.jl:199-315, explicit_gpu_operator.jl:299-392).
 * The FP64 basis is evaluated on the host from the FP32 coordinates. */
int hoe_create(const hoe_model_desc *desc, hoe_handle **out);

/* The FP64 operator (HOE_BACKEND_FP64): same validation, coefficient preparation,
 * basis (FP64 from the FP64 coordinates) and masking as hoe_create. On its handle
 * hoe_forward, hoe_adjoint, hoe_normal and hoe_cg take ComplexF64 vectors (interleaved
 * double pairs, same layouts); weights passed as float to hoe_normal / hoe_cg /
 * hoe_set_weights are widened, hoe_set_weights_f64 takes them in FP64. hoe_grad,
 * hoe_set_coefficients, hoe_set_kspha, hoe_set_kspha_dt and hoe_normal_noalloc return
 * HOE_EINVAL. Replaces the reference's Float64 operators (HighOrderExplicitOp with
 * Grid{Float64}, array_operator.jl:83-181; test_explicit_gpu.jl:51, 2e-12). */
int hoe_create_f64(const hoe_model_desc_f64 *desc, hoe_handle **out);

int hoe_get_info(const hoe_handle *h, hoe_info_t *info);

/* Switch the execution backend of an open handle. */
int hoe_set_backend(hoe_handle *h, int32_t backend);

/* y = E x. x: n_grid complex, y: n_sam*n_cha complex.
 * Replaces prod! (array_operator.jl:154-158 -> explicit_array_forward :241-288;
 * explicit_gpu_forward!, explicit_gpu_operator.jl:226-244). */
int hoe_forward(hoe_handle *h, const void *x, void *y, int flags);

/* x = E^H y. Replaces ctprod! (array_operator.jl:167-171 ->
 * explicit_array_adjoint :294-348; explicit_gpu_adjoint!,
 * explicit_gpu_operator.jl:246-270). */
int hoe_adjoint(hoe_handle *h, const void *y, void *x, int flags);

/* out = E^H D E x / with D real, nonnegative, finite and ALREADY squared
 * (W^H W); n_d is 0 (D = NULL, unit weights), n_sam (repeated over coils)
 * or n_sam*n_cha. Invalid D -> HOE_EINVAL, wrong n_d -> HOE_EDIM.
 * Replaces normalOperator(::ExplicitGPUOp, weights) + explicit_gpu_normal!
 * (explicit_gpu_operator.jl:272-297, 399-416). D is always a host pointer. */
int hoe_normal(hoe_handle *h, const float *D, int64_t n_d, const void *x, void *out, int flags);

/* out = E^H D E x / n_vox with the weights stored by the last hoe_set_weights /
 * hoe_normal / hoe_cg call, computed like hoe_normal but without allocating: no
 * hipMalloc, hipFree or other allocating call. Has no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation). x and out are device
 * pointers on the handle's GPU (flags must contain HOE_DEVICE_PTRS; HOE_ASYNC as
 * for hoe_normal). The workspaces (partial sums, MFMA fragments, the weighted
 * signal, the masked vectors, the fragment scales) come from one static 512 MiB
 * device array per GPU, which each GPU the library runs on holds until the process
 * ends. An operator that needs more returns HOE_EINVAL with the required size (the
 * 200 x 200, 32-coil, 16 000-sample class fits on every backend). Every call on
 * one GPU shares that GPU's array and leaves its contents behind: calls that overlap
 * on one GPU (queued on different handles' streams) race. */
int hoe_normal_noalloc(hoe_handle *h, const void *x, void *out, int flags);

/* Store normal-operator weights (rules as for hoe_normal) for later calls
 * with D == NULL, n_d == HOE_CURRENT_WEIGHTS; D == NULL, n_d == 0 stores unit
 * weights. The stored weights play the role of the diagonal captured by
 * normalOperator(op, D) (explicit_gpu_operator.jl:399-416). */
int hoe_set_weights(hoe_handle *h, const float *D, int64_t n_d);

/* hoe_set_weights with FP64 weights, for FP64 handles (same rules; on other handles the
 * weights are rounded to FP32). */
int hoe_set_weights_f64(hoe_handle *h, const double *D, int64_t n_d);

/* Replace the field-coefficient traces and the global field/data delay tau
 * (seconds). kspha holds n_term_eff x n_sam values (term fastest) that are
 * used as-is (no recon_terms processing; concomitant rows last when the
 * model has them). The effective coefficients are the traces resampled at
 * sample index s + tau/dwell with uniform Catmull-Rom interpolation (clamped
 * ends); tau = 0 uses them unchanged. kspha == NULL keeps the current
 * traces. The t*fieldmap term is not delayed, as in the reference's delay
 * model (FindDelay.jl:179-201, array_operator.jl:187-235). kspha is a host
 * pointer, or a device pointer on the handle's GPU with HOE_DEVICE_PTRS (then
 * HOE_ASYNC returns without waiting); the records are rebuilt on the device,
 * behind the work already queued on the handle. */
int hoe_set_coefficients(hoe_handle *h, const float *kspha, double tau, int flags);

/* Replace kspha (n_term x n_sam, term fastest, the layout of hoe_model_desc)
 * and process it exactly as hoe_create does: the handle's recon_terms flags
 * (prep_kspha.jl:1-56), with k_nominal (3 x n_sam, NULL = kspha's own rows
 * 2..4, the default of array_operator.jl:92). Concomitant rows, times, maps
 * and the backend are kept; the delay tau is reset to 0. Replaces building a
 * new operator for each trial delay (FindDelay.jl:184-187). Host pointers. */
int hoe_set_kspha(hoe_handle *h, const float *kspha, const float *k_nominal);

/* Set (kspha_dt != NULL) or clear (NULL) the trajectory-derivative
 * coefficients dk/dt: n_term x n_sam floats (term fastest, the kspha input
 * layout, not processed by recon_terms). While set, hoe_forward computes the
 * derivative product B x = 1/sqrt(nVox) sum_v 2 pi i (sum_l kdt_l(s) b_l(v))
 * exp(2 pi i phi) csm x (explicit_array_derivative, array_operator.jl:187-235;
 * delay estimation, FindDelay.jl:139-214) on the handle's backend (the
 * mixed backends at their own accuracy: 3e-5 FP16, 1e-4 BF16), hoe_adjoint
 * stays E^H (the reference's ctprod, array_operator.jl:167-171), and
 * hoe_normal / hoe_cg / hoe_grad return HOE_EINVAL. A model with
 * concomitant terms returns HOE_EINVAL: the reference builds the derivative
 * operator without them (array_operator.jl:139). Host pointer. */
int hoe_set_kspha_dt(hoe_handle *h, const float *kspha_dt);

/* Gradient of L(h,tau) = ||E(h,tau) x - y||^2.
 * x: n_grid complex, y: n_sam*n_cha complex (flags as for hoe_forward).
 * g_h: n_term_eff*n_sam floats (term fastest), dL/dh: host memory, or device
 * memory on the handle's GPU with HOE_DEVICE_PTRS (written in place, no host
 * copy); g_tau: dL/dtau (seconds^-1) and loss: L, host scalars. Reductions over
 * voxels, samples and coils run in FP64. Any of g_h, g_tau, loss may be NULL.
 * The call returns when the results are written. No reference counterpart
 * exists; the nearest reference code is the tau-sensitivity product
 * (array_operator.jl:187-235). */
int hoe_grad(hoe_handle *h, const void *x, const void *y, float *g_h, double *g_tau,
             double *loss, int flags);

/* Device-resident CGNR on (E^H D E + lambda I) x = E^H D y, x0 = 0, with FP64
 * inner products and residual norms. y: n_sam*n_cha complex data; D as in
 * hoe_normal (already squared); stops after iters iterations or when
 * ||r||/||r0|| <= rel_tol. x_out: n_grid complex. resid_log (host, may be
 * NULL): iters+1 doubles, ||r_k||/||r_0|| (0 after an early stop).
 * Replaces recon_HOOp's solve! with CGNR (recon_HOOp.jl:73-82). */
int hoe_cg(hoe_handle *h, const void *y, const float *D, int64_t n_d, int32_t iters,
           double lambda, double rel_tol, void *x_out, double *resid_log, int flags);

/* Release device resources. Idempotent; later calls on h return
 * HOE_ECLOSED (explicit_gpu_operator.jl:435-446). */
int hoe_close(hoe_handle *h);

/* hoe_close, then free the handle. NULL is accepted. */
void hoe_destroy(hoe_handle *h);

/* Synchronise the handle's stream (after HOE_ASYNC calls). */
int hoe_sync(hoe_handle *h);

/* Device timing of the most recent hoe_forward / hoe_adjoint / hoe_normal /
 * hoe_grad call made without HOE_ASYNC: milliseconds spent in that call's
 * dominant kernel launches (per launch, averaged) measured with HIP events on
 * the handle's stream. Calls with HOE_ASYNC record no events.
 * which: 0 = forward contraction kernel, 1 = adjoint contraction kernel,
 * 2 = whole call. Returns a negative value when unavailable. */
double hoe_last_kernel_ms(const hoe_handle *h, int which);

/* ---- field-model inputs upstream of kspha ---- */

/* GIRF prediction: apply the frequency-domain gradient response Hw to n_other
 * independent 3-channel nominal gradient waveforms (apply_girf,
 * src/GIRF/apply_girf.jl:39-150, with _apply_girf_core :111-150).
 * g:   3 x n_point x n_other reals, channel fastest (Julia (3, nPoint, nOther)).
 * hw:  complex n_freq x 3 x n_out, frequency fastest, DC centred (fftshift
 *      order), interleaved (re, im).
 * out: n_out x n_point x n_other reals, channel fastest.
 * The FFT length is nextprod([2,3,5], max(n_point, n_freq)); waveforms are
 * zero padded to it (circular convolution); a shorter response is expanded in
 * the time domain exactly as the reference does (apply_girf.jl:72-90, which
 * leaves out time sample n_freq/2+1, 1-based); rbw < 1 clears
 * round(N*(1-rbw)/2) bins at each end of the centred spectrum. FP32
 * arithmetic (hipFFT C2C) for hoe_apply_girf, FP64 (Z2Z) for _f64. Runs on
 * the given device; pointers are host memory, or device memory on that device
 * with HOE_DEVICE_PTRS. No handle: errors are read with hoe_last_error(NULL). */
int hoe_apply_girf(const float *g, int64_t n_point, int64_t n_other, const float *hw, int32_t n_freq,
                   int32_t n_out, double rbw, int32_t device, float *out, int flags);
int hoe_apply_girf_f64(const double *g, int64_t n_point, int64_t n_other, const double *hw, int32_t n_freq,
                       int32_t n_out, double rbw, int32_t device, double *out, int flags);

/* nextprod([2,3,5], n): the FFT length apply_girf uses for n = max(nPoint, nFreq)
 * (apply_girf.jl:69). */
int64_t hoe_girf_fft_length(int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* HOE_H */

// Internal state of a libhoe handle and the launchers shared by the C ABI
// (hoe_api.cpp) and the kernels (hoe_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hoe.h"

namespace hoe {

// Work decomposition constants.
constexpr int kWG = 256;        // threads per workgroup (4 wave64)
constexpr int kRowPad = 256;    // samples and voxels are padded to whole workgroups
constexpr int kKStep = 8;       // voxels (forward) or samples (adjoint) per MFMA K step
constexpr int kMaxTerms = 20;   // 16 solid harmonics + 4 concomitant columns

struct DeviceBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// K unit (8 consecutive voxels) of the MFMA Horner forward. For a unit inside one run,
// x_a (an FP32 value: the middle of its voxels' x range, rounded) and fm_a (the fieldmap
// of its first voxel) anchor the phase: voxel j carries u_j = x_j - x_a and
// dfm_j = fieldmap_j - fm_a in FP32 (HoeUnitVox), so that its phase is
// P(x_a) + t * fm_a (FP64, reduced modulo one cycle) plus a cubic in u_j and t * dfm_j
// (FP32). run = -1 marks a unit that crosses a run boundary.
struct alignas(16) HoeUnit {
    int32_t run;
    float xa;
    float fma;
    float pad;
};
// Largest fieldmap offset max|t| * max|fieldmap - anchor fieldmap|, in cycles, for which the
// MFMA Horner kernels evaluate anchored phases: its FP32 rounding (about 3 * 2^-24 of it,
// 1.4e-6 cycles at the limit) stays far below the FP16 hi/lo split's 2^-22.
constexpr double kAnchorMaxCycles = 8.0;
// Largest Horner offset |P(x_a + u) - P(x_a)| over the samples and the anchor groups, in
// cycles, for which they stay on: the device bounds it after each record build by
// max_s sum_i |A_s[i]| D_i, D_i the largest change of record entry i's basis polynomial
// within a K unit or adjoint group (2-12 cycles on trajectories that reach the grid's
// Nyquist radius; the FP32 rounding of the offsets grows in proportion). With both limits
// an anchored phase is at most 1 + 32 + 8 cycles, inside the range where v_sin_f32 /
// v_cos_f32 take it without a prior reduction (measured: the same 1.2e-7 error on
// +-64 cycles as on one).
constexpr double kAnchorMaxOffsetCycles = 32.0;

struct alignas(64) HoeUnitVox {
    float ud[8][2];   // (u_j, dfm_j)
};
// Anchors of a 64-voxel wave of the MFMA Horner adjoint: group c = 2 g + h holds the wave's
// voxels of run r_lo + g in wave half h (rows 32 h .. 32 h + 31); x_c is the middle of their
// x range (rounded to FP32), fm_c the fieldmap of the group's first voxel.
struct alignas(32) HoeAdjAnchor {
    float xa[4];
    float fm[4];
};

// Stream-K schedule of one contraction (see build_schedule in hoe_kernels.hip), or of one
// pass of it: a pass covers the stationary tiles [t_lo, t_hi) over the whole K range.
struct Sched {
    bool built = false;
    int tiles = 0;      // 256-row tiles of the stationary index
    int t_lo = 0, t_hi = 0;   // tiles of this pass
    int n_ks = 0;       // K units (8 streamed rows each)
    int cpx = 0;        // chunks per XCD part
    int n_chunks = 0;   // workgroups launched (8 * cpx)
    int n_seg = 0;      // segments = partial-sum blocks
    int max_tile_segs = 0;   // most segments of one tile
    DeviceBuf seg;        // int4 [n_seg]: tile, k0, k1, 0
    DeviceBuf chunk_off;  // int [n_chunks + 1]: first segment of each chunk
    DeviceBuf tile_off;   // int [tiles + 1]: CSR offsets into tile_list
    DeviceBuf tile_list;  // int [n_seg]: segments of each tile, in segment order
};

struct Handle {
    // ---- model sizes ----
    int64_t n_grid = 0, n_vox = 0, n_sam = 0, n_cha = 0;
    int64_t nv_pad = 0, ns_pad = 0;   // padded to kRowPad
    int32_t L = 0;                    // effective terms (b_1 = 1 included)
    int32_t LK = 0;                   // kernel record length: active terms + fieldmap, padded to a bucket
    std::vector<int> active;          // 0-based term indices (>= 1) with a not identically zero basis
    bool horner = false;              // forward kernels evaluate the phase per run by Horner in x
    bool horner_adj = false;          // MFMA adjoint by Horner (per wave, when its 64 voxels span <= 2 runs)
    bool anchored = false;            // MFMA Horner kernels evaluate anchored phases (FP32 offsets bounded)
    double anchor_fm_cycles = 0.0;    // max |t| * max |fieldmap - anchor fieldmap| over the anchor groups
    int32_t n_runs = 0;               // runs: consecutive masked voxels sharing (y, z)
    int32_t t_index = 0;              // index of t in the sample record (= number of active terms + 1)
    int32_t n_term_in = 0;            // 9 or 16
    int32_t ncs = 0;                  // coil stride of [row][coil] buffers: 32*ceil(n_cha/32)
    int32_t sb = 0;                   // doubles per voxel record (>= LK)
    int32_t sk = 0;                   // doubles per sample record (>= LK + 1)
    int32_t backend = HOE_BACKEND_DIRECT;
    bool f64 = false;                 // FP64 backend: ComplexF64 vectors, FP64 model and arithmetic
    int32_t cbytes = 8;               // bytes per complex entry of vectors and workspaces (8 or 16)
    int32_t device = 0;
    int32_t cu_count = 256;
    bool closed = false;
    bool weights_unit = false;        // dvec holds unit weights
    bool dt_on = false;               // hoe_forward computes the derivative product B x (kspha_dt set)
    double dt_sigma = 1.0;            // power of two with 2 pi dt_sigma max|psi| <= 1 (MFMA derivative forward)
    double dt_w = 0.0;                // 2 pi dt_sigma
    bool has_conc = false;            // the model carries the 4 concomitant terms
    bool order[4] = {true, true, true, true};   // recon_terms flags: orders 0..3 kept
    float csm_max = 0.0f;             // max |Re|,|Im| bound of the masked coil maps: max |csm|
    double dwell = 0.0, tau = 0.0;
    std::string err;

    // ---- host copies needed to rebuild coefficients ----
    std::vector<float> traces;        // L x n_sam, term fastest (prepared coefficients)
    std::vector<double> traces64;     // the same for the FP64 backend
    bool traces_stale = false;        // traces_d was written from device memory; the host copy is older
    std::vector<float> times;         // n_sam
    std::vector<int32_t> mask_idx_h;  // n_vox

    // ---- device model ----
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    DeviceBuf mask_idx;   // int32 [n_vox]
    DeviceBuf basis;      // double [nv_pad][sb]: b of the active terms, fieldmap
    DeviceBuf coef;       // double [ns_pad][sk]: k_1, k of the active terms, t
    DeviceBuf dcoef;      // double [ns_pad][sk] d k / d tau (for the delay gradient)
    DeviceBuf dtrec;      // double [ns_pad][sk] kspha_dt records (derivative product), t slot 0
    DeviceBuf csm;        // float2 [nv_pad][ncs]
    DeviceBuf hrec;       // double2 [nv_pad]: (x, fieldmap) of each voxel, for the Horner forward
    DeviceBuf units;      // HoeUnit [nv_pad / 8]: K-unit anchors of the MFMA Horner forward
    DeviceBuf uvox;       // HoeUnitVox [nv_pad / 8]: voxel offsets from the anchors
    DeviceBuf aanc;       // HoeAdjAnchor [nv_pad / 64]: group anchors of the MFMA Horner adjoint
    DeviceBuf aud;        // float2 [nv_pad]: (x - x_c, fieldmap - fm_c) against the voxel's adjoint group
    DeviceBuf abound;     // double [32] D_i per record entry, then (byte 256) the float bits of the
                          // Horner offset bound of the current records (kAnchorMaxOffsetCycles)
    DeviceBuf vrun;       // int [nv_pad]: run of each voxel
    DeviceBuf beta;       // double [n_runs][LK][4]: x-polynomial coefficients of each record entry

    // ---- workspaces ----
    DeviceBuf xm;         // float2 [nv_pad]        masked image
    DeviceBuf zbuf;       // float2 [nv_pad][ncs]   csm .* x (direct path)
    DeviceBuf zfrag;      // MFMA B fragments of csm .* x (mixed path)
    DeviceBuf ybuf;       // float2 [ns_pad][ncs]   weighted signal, coil fastest
    DeviceBuf yfrag;      // MFMA B fragments of the weighted signal (mixed path)
    DeviceBuf part_f;     // float2 [n_seg][256][ncs]   forward partial sums
    DeviceBuf part_a;     // float2 [n_seg][groups][256] adjoint partial sums, coils combined per group
    DeviceBuf part_g;     // double [n_seg][batches][256][L] gradient partials
    DeviceBuf dvec;       // float  [ns_pad][ncs] normal weights (coil fastest)
    DeviceBuf io_x;       // float2 [n_grid]  staging for host vectors
    DeviceBuf io_x2;      // float2 [n_grid]
    DeviceBuf io_y;       // float2 [n_sam*n_cha]
    DeviceBuf io_y2;      // float2 [n_sam*n_cha]
    DeviceBuf scal;       // small scalars: scales, CG dot products
    DeviceBuf cg_vec;     // float2 [4][nv_pad] CG vectors x, r, p, q
    DeviceBuf red;        // double [4096] block partial sums
    DeviceBuf gk;         // double [L][n_sam] dL/dk reduced
    DeviceBuf cg_log;     // double [iters+1] CG relative residuals
    // mixed-backend gradient (32 coils): MFMA fragments and partials
    DeviceBuf gzfrag;     // A fragments of csm .* x: [nv_pad/32][ncs/8][64 lanes][4] uint4
    DeviceBuf grfrag;     // B fragments of the residual: [ns_pad/32][ncs/8][64][2] uint4
    DeviceBuf gbfrag;     // B fragments of the basis columns: [nv_pad/32][2][64][2] uint4
    DeviceBuf gbscale;    // float [32]: power-of-two scale of each basis column
    DeviceBuf part_gm;    // float [n_seg][256][32] gradient partials
    DeviceBuf ghbuf;      // float [n_sam][L] dL/dh (term fastest)
    DeviceBuf growmap;    // int [L]: kernel row of each original term, -1 if its basis is zero
    DeviceBuf traces_d;   // float [n_sam][L]: the coefficient traces (term fastest), source of coef/dcoef
    DeviceBuf times_d;    // float [n_sam]
    DeviceBuf act_d;      // int [active terms]: 0-based term index of each kernel record entry 1..
    std::vector<float> gbscale_h;
    std::vector<double> rec_bmax;     // max |b| over the masked voxels of each record entry (entry 0: 1)
    std::vector<double> adj_tile_w;   // stream-K work weight of each 256-voxel tile in the MFMA adjoint
    std::vector<double> fwd_k_w;      // stream-K work weight of each K unit (8 voxels) in the Horner forward
    int gb_type = -1;     // backend whose element type gbfrag holds

    // schedules per backend (index = hoe_backend), as passes over tile ranges whose
    // partial sums share part_f / part_a, and for the gradient kernels
    std::vector<Sched> sched_f[4], sched_a[4];
    Sched sched_g, sched_gm;
    int occ_f[4] = {0, 0, 0, 0}, occ_a[4] = {0, 0, 0, 0};   // resident workgroups per CU

    // timing of the last call (ms)
    double last_ms[3] = {-1, -1, -1};
};

// ---- coil batching and schedules (hoe_kernels.hip) ----
int direct_cb(const Handle *h, bool fwd);   // coils per direct-kernel batch: 8, 16 or 32
int direct_batches(const Handle *h, bool fwd);
int mixed_nb(const Handle *h);         // 32-coil groups per wave in the MFMA kernels: 1 or 2
int mixed_groups(const Handle *h);
// Coil groups (grid rows) of a backend's adjoint contraction: one partial sum per voxel,
// segment and group.
int adj_groups(const Handle *h, int backend);
hipError_t build_schedule(Sched &sc, int tiles, int n_ks, int slots, int kmax,
                          const std::vector<double> *tile_w = nullptr, const std::vector<double> *k_w = nullptr);
// Build the forward/adjoint schedules of a backend and size the partial buffers.
hipError_t prepare_backend(Handle *h, int backend);
// Free the device arrays of the passes and empty the list.
void drop_passes(std::vector<Sched> &passes);
hipError_t prepare_grad(Handle *h);
// True when the current backend computes the gradient on the matrix cores.
bool grad_on_mfma(const Handle *h);
// Rebuild the per-(run, sample) table of the Horner adjoint from the coefficient records.

// ---- launchers (hoe_kernels.hip) ----
// All run on h->stream and return hipSuccess or the first HIP error.
hipError_t launch_gather(Handle *h, const void *x_full, void *xm);
hipError_t launch_scatter(Handle *h, const void *xm, void *x_full, double scale);
// Forward of x_full (device, n_grid complex) into the user signal layout (complex [c][s])
// with scale: contraction into part_f and its reduction, pass by pass.
hipError_t run_forward_user(Handle *h, const void *x_full, void *y_out, double scale, bool time_it);
// Forward of x_full into ybuf (coil fastest) with scale, optional weights dvec
// (weights == true) and optional subtraction of data (y_data != nullptr, user layout).
hipError_t run_forward_ybuf(Handle *h, const void *x_full, double scale, bool weights, const void *y_data,
                            bool time_it);
// Load a user signal (complex [c][s]) into ybuf (coil fastest), optionally weighted.
hipError_t launch_load_ybuf(Handle *h, const void *y_user, bool weights);
// Adjoint contraction from ybuf -> part_a and coil reduction into xm (masked) with scale,
// pass by pass. Entries at nv_write and beyond are not written (-1: all nv_pad, the
// padding as zeros); without a mask a caller's full-grid output takes nv_write = n_vox.
hipError_t run_adjoint_contract(Handle *h, double scale, void *xm_out, bool time_it, int nv_write = -1);
// Normal operator on masked vectors: out_m = scale * E^H D E x_m (+ lambda x_m).
// have_max: slot 0 already holds max|xm_in| and the other slots are clear (launch from a
// gather that computed it).
hipError_t run_normal_masked(Handle *h, const void *xm_in, void *xm_out, double scale, double lambda,
                             bool time_it, bool have_max = false, int nv_write = -1);
// Gather of a full-grid x (with max|x| for the fragment scale) and the normal operator.
hipError_t run_normal_full(Handle *h, const void *x_full, void *xm_out, double scale, bool time_it,
                           int nv_write = -1);
// Gradient partial kernel: uses zbuf (csm .* x) and ybuf (residual, coil fastest) -> gk.
hipError_t run_grad_contract(Handle *h, bool time_it);
// zbuf = csm .* xm
hipError_t launch_build_zbuf(Handle *h, const void *xm);
// CG helper kernels.
hipError_t cg_run(Handle *h, const void *rhs_m, int iters, double lambda, double rel_tol,
                  void *x_m, double *resid_log_dev);
// squared norm of a complex vector (n entries) into a device double.
hipError_t launch_norm2(Handle *h, const void *v, int64_t n, double *out_dev);
// dL/dh into ghbuf and dL/dtau into *gtau_dev from gk (after run_grad_contract);
// shift, m, w: the delay resampling k(s) = sum_t w[t] h(clamp(s + m - 1 + t)).
// gh_out: device array for dL/dh ([n_sam][L]); nullptr writes the handle's ghbuf.
hipError_t launch_grad_chain(Handle *h, double *gtau_dev, bool shift, int m, const double w[4], float *gh_out);
// coef and dcoef from traces_d (k_build_records): resampling shift m and Catmull-Rom weights
// w, dw (shift == 0: the traces unchanged, dcoef from dw at offset 0).
hipError_t launch_build_records(Handle *h, int shift, int m, const double w[4], const double dw[4]);

// GIRF prediction (hoe_girf.hip): FFT length nextprod([2,3,5], n), and the whole
// prediction of n_other 3-channel waveforms; on failure err holds the message.
int64_t girf_fft_length(int64_t n);
template <typename R>
int girf_run(const R *g, int64_t n_point, int64_t n_other, const R *hw, int32_t n_freq, int32_t n_out, double rbw,
             int32_t device, R *out, int flags, std::string &err);

}  // namespace hoe

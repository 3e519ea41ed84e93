// C ABI of libhoe: validation, model preparation and the entry points
// declared in include/hoe.h.
//
// Model preparation follows HighOrderMRI.jl:
//   - order selection prep_kspha (src/EncodingOperator/prep_kspha.jl:1-56),
//   - concomitant columns prep_concomitant (prep_kspha.jl:78-98),
//   - the 16 unnormalised solid harmonics basisfunc_spha
//     (src/SphericalHarmonics/basisfunction.jl:47-78), here evaluated in FP64
//     from the FP32 grid coordinates,
//   - column-major mask order and csm/fieldmap masking
//     (src/EncodingOperator/Explicit/array_operator.jl:121-126),
//   - normal-operator weight rules (explicit_gpu_operator.jl:399-416).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "hoe_internal.h"

using hoe::Handle;

namespace {

thread_local std::string g_err;

const int kCurrentWeights = HOE_CURRENT_WEIGHTS;

struct Fail {
    int code;
};

int set_err(Handle *h, int code, const std::string &msg) {
    if (h) h->err = msg;
    g_err = msg;
    return code;
}

#define HOE_TRY_HIP(h, expr)                                                                      \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return set_err((h), _e == hipErrorOutOfMemory ? HOE_ENOMEM : HOE_EHIP,                \
                           std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr);    \
    } while (0)

hipError_t dalloc(hoe::DeviceBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    b.bytes = bytes;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        b.bytes = 0;
        return e;
    }
    return hipMemset(b.p, 0, bytes);
}

void dfree(hoe::DeviceBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

inline int64_t round_up(int64_t a, int64_t m) { return (a + m - 1) / m * m; }

// Real solid harmonics 2..16 in the order of basisfunc_spha (basisfunction.jl:47-78),
// FP64 from FP32 coordinates. out[0..14] = b_2..b_16.
void solid_harmonics(double x, double y, double z, double *out) {
    const double r2 = x * x + y * y + z * z;
    out[0] = x;                                   // 2
    out[1] = y;                                   // 3
    out[2] = z;                                   // 4
    out[3] = x * y;                               // 5
    out[4] = z * y;                               // 6
    out[5] = 3.0 * z * z - r2;                    // 7
    out[6] = x * z;                               // 8
    out[7] = x * x - y * y;                       // 9
    out[8] = 3.0 * y * x * x - y * y * y;         // 10
    out[9] = x * z * y;                           // 11
    out[10] = (5.0 * z * z - r2) * y;             // 12
    out[11] = 5.0 * z * z * z - 3.0 * z * r2;     // 13
    out[12] = (5.0 * z * z - r2) * x;             // 14
    out[13] = x * x * z - y * y * z;              // 15
    out[14] = x * x * x - 3.0 * x * y * y;        // 16
}

// Coefficients of basis term l (0-based: 1..15 solid harmonics 2..16, n_term..n_term+3
// the concomitant columns z^2, x^2+y^2, xz, yz) as a cubic in x at fixed (y, z):
// b_l(x, y, z) = p[0] + p[1] x + p[2] x^2 + p[3] x^3.
void x_poly(int l, int n_term, double y, double z, double p[4]) {
    p[0] = p[1] = p[2] = p[3] = 0.0;
    if (l >= n_term) {
        switch (l - n_term) {
            case 0: p[0] = z * z; break;
            case 1: p[0] = y * y; p[2] = 1.0; break;
            case 2: p[1] = z; break;
            default: p[0] = y * z; break;
        }
        return;
    }
    switch (l + 1) {
        case 2: p[1] = 1.0; break;
        case 3: p[0] = y; break;
        case 4: p[0] = z; break;
        case 5: p[1] = y; break;
        case 6: p[0] = z * y; break;
        case 7: p[0] = 2.0 * z * z - y * y; p[2] = -1.0; break;
        case 8: p[1] = z; break;
        case 9: p[0] = -y * y; p[2] = 1.0; break;
        case 10: p[0] = -y * y * y; p[2] = 3.0 * y; break;
        case 11: p[1] = y * z; break;
        case 12: p[0] = (4.0 * z * z - y * y) * y; p[2] = -y; break;
        case 13: p[0] = 2.0 * z * z * z - 3.0 * z * y * y; p[2] = -3.0 * z; break;
        case 14: p[1] = 4.0 * z * z - y * y; p[3] = -1.0; break;
        case 15: p[0] = -y * y * z; p[2] = z; break;
        default: p[1] = -3.0 * y * y; p[3] = 1.0; break;   // 16: x^3 - 3 x y^2
    }
}

// Uniform Catmull-Rom weights at fractional offset f in [0,1) for taps
// -1, 0, 1, 2, and their derivatives with respect to f.
void catmull_rom(double f, double w[4], double dw[4]) {
    const double f2 = f * f, f3 = f2 * f;
    w[0] = 0.5 * (-f3 + 2.0 * f2 - f);
    w[1] = 0.5 * (3.0 * f3 - 5.0 * f2 + 2.0);
    w[2] = 0.5 * (-3.0 * f3 + 4.0 * f2 + f);
    w[3] = 0.5 * (f3 - f2);
    dw[0] = 0.5 * (-3.0 * f2 + 4.0 * f - 1.0);
    dw[1] = 0.5 * (9.0 * f2 - 10.0 * f);
    dw[2] = 0.5 * (-9.0 * f2 + 8.0 * f + 1.0);
    dw[3] = 0.5 * (3.0 * f2 - 2.0 * f);
}

// prep_kspha (prep_kspha.jl:1-56) into the first n_term rows of the traces: a cleared
// order-0/2/3 flag zeroes its rows, a cleared order-1 flag replaces rows 2-4 with
// k_nominal (3 x n_sam, term fastest; NULL keeps kspha's own rows, the reference's
// default k_nominal = kspha[2:4, :]). Concomitant rows are left as they are.
template <typename T>
void prepare_traces_t(const Handle *h, std::vector<T> &traces, const T *kspha, const T *k_nominal) {
    const int L = h->L, n_term = h->n_term_in;
    const bool *t = h->order;
    for (int64_t s = 0; s < h->n_sam; ++s) {
        T *r = &traces[(size_t)s * L];
        for (int l = 0; l < n_term; ++l) r[l] = kspha[(size_t)s * n_term + l];
        if (!t[0]) r[0] = T(0);
        if (!t[1] && k_nominal)
            for (int l = 0; l < 3; ++l) r[1 + l] = k_nominal[(size_t)s * 3 + l];
        if (!t[2])
            for (int l = 4; l < 9; ++l) r[l] = T(0);
        if (!t[3] && n_term == 16)
            for (int l = 9; l < 16; ++l) r[l] = T(0);
    }
}

void prepare_traces(Handle *h, const float *kspha, const float *k_nominal) {
    prepare_traces_t(h, h->traces, kspha, k_nominal);
}

// Sample records of the FP64 backend from its traces (no delay): [k_1, k of the active
// terms, t, 0...], uploaded to coef; dcoef stays zero.
int upload_coef_f64(Handle *h, const std::vector<double> &times) {
    std::vector<double> rec((size_t)h->ns_pad * h->sk, 0.0);
    const int L = h->L;
    for (int64_t s = 0; s < h->n_sam; ++s) {
        double *r = &rec[(size_t)s * h->sk];
        const double *k = &h->traces64[(size_t)s * L];
        r[0] = k[0];
        for (size_t i = 0; i < h->active.size(); ++i) r[1 + i] = k[h->active[i]];
        r[1 + h->active.size()] = times[(size_t)s];
    }
    HOE_TRY_HIP(h, hipMemcpy(h->coef.p, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice));
    return HOE_OK;
}

// The sample records coef ([k_1, k of the active terms, t, 0...] per sample, FP64) and
// their tau-derivatives dcoef, built on the device from traces_d: the traces resampled at
// s + tau/dwell with uniform Catmull-Rom weights (clamped ends; tau = 0 keeps them), and
// d/dtau of that resampling (all zero when dwell is unset). Queued on the handle's stream
// behind any work that still reads the previous records.
int rebuild_records(Handle *h) {
    const bool shift = h->tau != 0.0;
    int64_t m = 0;
    double w[4] = {0.0, 1.0, 0.0, 0.0}, dw[4] = {0.0, 0.0, 0.0, 0.0};
    if (h->dwell > 0.0) {
        const double u = h->tau / h->dwell;
        m = (int64_t)std::floor(u);
        catmull_rom(u - (double)m, w, dw);
    }
    if (m < INT32_MIN / 2 || m > INT32_MAX / 2) return set_err(h, HOE_EINVAL, "delay too large for the readout");
    HOE_TRY_HIP(h, hoe::launch_build_records(h, shift ? 1 : 0, (int)m, w, dw));
    return HOE_OK;
}

// Host traces -> traces_d, then the records; waits for the device.
int upload_coef(Handle *h) {
    HOE_TRY_HIP(h, hipMemcpyAsync(h->traces_d.p, h->traces.data(), sizeof(float) * h->traces.size(),
                                  hipMemcpyHostToDevice, h->stream));
    h->traces_stale = false;
    int rc = rebuild_records(h);
    if (rc) return rc;
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    return HOE_OK;
}

int check_open(Handle *h) {
    if (!h) return set_err(nullptr, HOE_EINVAL, "null handle");
    if (h->closed) return set_err(h, HOE_ECLOSED, "operator is closed");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return set_err(h, HOE_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return HOE_OK;
}

// Rebuild the FP32 backend's schedules after dt_on changed (the forward kernel instance,
// hence its occupancy and the stream-K slot count, depends on it); on failure dt_on goes
// back to prev_dt and the schedules are rebuilt for it.
int reschedule_direct(Handle *h, bool prev_dt) {
    hoe::drop_passes(h->sched_f[HOE_BACKEND_DIRECT]);
    hipError_t e = hoe::prepare_backend(h, HOE_BACKEND_DIRECT);
    if (e == hipSuccess) return HOE_OK;
    const std::string msg = std::string("schedule setup failed: ") + hipGetErrorString(e);
    h->dt_on = prev_dt;
    hoe::drop_passes(h->sched_f[HOE_BACKEND_DIRECT]);
    (void)hoe::prepare_backend(h, HOE_BACKEND_DIRECT);
    return set_err(h, e == hipErrorOutOfMemory ? HOE_ENOMEM : HOE_EHIP, msg);
}

// The FP32 forward's schedules exist (the direct backend is or was selected): they follow dt_on.
bool need_direct_resched(const Handle *h) {
    return h->backend == HOE_BACKEND_DIRECT || !h->sched_f[HOE_BACKEND_DIRECT].empty();
}

// The MFMA kernels' decision for the current records: anchored phases need the fieldmap
// bound (fixed at creation) and the device's Horner offset bound of the last record build.
bool anchored_now(const Handle *h) {
    if (!h->horner || !h->anchored || h->f64 || h->closed) return false;
    float b = std::numeric_limits<float>::infinity();
    if (hipStreamSynchronize(h->stream) != hipSuccess ||
        hipMemcpy(&b, (const unsigned char *)h->abound.p + 256, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    return (double)b <= hoe::kAnchorMaxOffsetCycles;
}

// Without a mask the masked voxel order is the grid order (mask_idx is the identity).
bool unmasked(const Handle *h) { return h->n_vox == h->n_grid; }

// The derivative model (kspha_dt) defines only the forward product.
int reject_dt(Handle *h) {
    return h->dt_on ? set_err(h, HOE_EINVAL,
                              "only the forward product is defined while kspha_dt is set "
                              "(array_operator.jl:153-172); clear it with hoe_set_kspha_dt(h, NULL)")
                    : HOE_OK;
}


// Upload normal weights D (already squared) into dvec [ns_pad][ncs], in the handle's
// precision (FP32, or FP64 for the FP64 backend).
template <typename T>
int set_weights_t(Handle *h, const T *D, int64_t n_d) {
    const int64_t ns = h->n_sam, nc = h->n_cha;
    if (D == nullptr && n_d == 0) {
        if (h->weights_unit) return HOE_OK;
    } else {
        if (D == nullptr) return set_err(h, HOE_EINVAL, "normal weights pointer is NULL");
        if (n_d != ns && n_d != ns * nc) return set_err(h, HOE_EDIM, "normal weights");
        for (int64_t i = 0; i < n_d; ++i) {
            if (!std::isfinite(D[i])) return set_err(h, HOE_EINVAL, "normal weights must be finite");
            if (D[i] < T(0)) return set_err(h, HOE_EINVAL, "normal weights must be nonnegative");
        }
    }
    auto fill = [&](auto zero) -> int {
        using W = decltype(zero);
        std::vector<W> w((size_t)h->ns_pad * h->ncs, W(0));
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t s = 0; s < ns; ++s) {
                W v = W(1);
                if (D) v = (W)((n_d == ns) ? D[s] : D[c * ns + s]);
                w[(size_t)s * h->ncs + c] = v;
            }
        // work queued with HOE_ASYNC may still read the current weights: the copy (on the
        // null stream, which a non-blocking stream does not wait for) goes in after it
        HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
        HOE_TRY_HIP(h, hipMemcpy(h->dvec.p, w.data(), w.size() * sizeof(W), hipMemcpyHostToDevice));
        return HOE_OK;
    };
    const int rc = h->f64 ? fill(0.0) : fill(0.0f);
    if (rc) return rc;
    h->weights_unit = (D == nullptr);
    return HOE_OK;
}

int set_weights(Handle *h, const float *D, int64_t n_d) { return set_weights_t(h, D, n_d); }

// Scale factors enter the FP32 kernels as float (rounded once, as before), the FP64 ones
// unrounded.
inline double prec_scale(const Handle *h, double v) { return h->f64 ? v : (double)(float)v; }

// Entry points the FP64 backend does not provide.
int reject_f64(Handle *h, const char *what) {
    return h->f64 ? set_err(h, HOE_EINVAL, std::string(what) + " is not provided on the FP64 backend") : HOE_OK;
}

double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
    return (double)ms;
}

void release(Handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    hoe::DeviceBuf *bufs[] = {&h->mask_idx, &h->basis,  &h->coef,   &h->dcoef,   &h->csm,    &h->xm,     &h->zbuf,
                              &h->hrec,     &h->vrun,   &h->beta,   &h->zfrag,   &h->ybuf,   &h->yfrag,  &h->part_f,
                              &h->part_a,   &h->part_g, &h->dvec,   &h->io_x,    &h->io_x2,  &h->io_y,   &h->io_y2,
                              &h->scal,     &h->cg_vec, &h->red,    &h->gk,      &h->cg_log, &h->gzfrag, &h->grfrag,
                              &h->gbfrag,   &h->gbscale, &h->part_gm, &h->ghbuf, &h->growmap, &h->dtrec,
                              &h->traces_d, &h->times_d,  &h->act_d,   &h->units,   &h->uvox,    &h->aanc,    &h->aud,
                              &h->abound};
    for (auto *b : bufs) dfree(*b);
    for (int b = 0; b < 4; ++b) {
        hoe::drop_passes(h->sched_f[b]);
        hoe::drop_passes(h->sched_a[b]);
    }
    hoe::Sched *scheds[] = {&h->sched_g, &h->sched_gm};
    for (auto *sc : scheds) {
        dfree(sc->seg);
        dfree(sc->chunk_off);
        dfree(sc->tile_off);
        dfree(sc->tile_list);
        sc->built = false;
    }
    h->gb_type = -1;
    for (auto &e : h->ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->stream = nullptr;
    h->closed = true;
}

}  // namespace

namespace {
template <typename R>
int apply_girf_entry(const R *g, int64_t n_point, int64_t n_other, const R *hw, int32_t n_freq, int32_t n_out,
                     double rbw, int32_t device, R *out, int flags) {
    g_err.clear();
    if (!g || !hw || !out) return set_err(nullptr, HOE_EINVAL, "gradient, response and output must not be NULL");
    if (n_point < 1 || n_other < 1) return set_err(nullptr, HOE_EINVAL, "the gradient waveforms are empty");
    if (n_freq < 1 || n_out < 1) return set_err(nullptr, HOE_EINVAL, "the response is empty");
    if (!std::isfinite(rbw)) return set_err(nullptr, HOE_EINVAL, "rbw must be finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(nullptr, HOE_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev) return set_err(nullptr, HOE_EINVAL, "invalid visible GPU ID");
    std::string err;
    const int rc = hoe::girf_run<R>(g, n_point, n_other, hw, n_freq, n_out, rbw, device, out, flags, err);
    if (rc) return set_err(nullptr, rc, err);
    return HOE_OK;
}
}  // namespace

extern "C" {

const char *hoe_version(void) { return "libhoe 0.1.0 (gfx950)"; }

const char *hoe_last_error(const hoe_handle *h) {
    if (h) return reinterpret_cast<const Handle *>(h)->err.c_str();
    return g_err.c_str();
}

int hoe_apply_girf(const float *g, int64_t n_point, int64_t n_other, const float *hw, int32_t n_freq, int32_t n_out,
                   double rbw, int32_t device, float *out, int flags) {
    return apply_girf_entry<float>(g, n_point, n_other, hw, n_freq, n_out, rbw, device, out, flags);
}

int hoe_apply_girf_f64(const double *g, int64_t n_point, int64_t n_other, const double *hw, int32_t n_freq,
                       int32_t n_out, double rbw, int32_t device, double *out, int flags) {
    return apply_girf_entry<double>(g, n_point, n_other, hw, n_freq, n_out, rbw, device, out, flags);
}

int64_t hoe_girf_fft_length(int64_t n) { return hoe::girf_fft_length(n); }

int hoe_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

namespace {

// hoe_create / hoe_create_f64: Desc is hoe_model_desc (FP32 data) or hoe_model_desc_f64.
template <typename Desc>
int create_impl(const Desc *d, hoe_handle **out) {
    constexpr bool F64 = std::is_same<Desc, hoe_model_desc_f64>::value;
    using T = typename std::conditional<F64, double, float>::type;   // element type of the inputs
    g_err.clear();
    if (!out) return set_err(nullptr, HOE_EINVAL, "out pointer is NULL");
    *out = nullptr;
    if (!d) return set_err(nullptr, HOE_EINVAL, "model description is NULL");
    if (d->n_x < 1 || d->n_y < 1 || d->n_z < 1) return set_err(nullptr, HOE_EINVAL, "grid dimensions must be positive");
    if (!d->x || !d->y || !d->z) return set_err(nullptr, HOE_EINVAL, "grid coordinates are NULL");
    if (d->n_term != 9 && d->n_term != 16) return set_err(nullptr, HOE_EINVAL, "kspha requires 9 or 16 terms");
    if (d->n_sam < 1 || d->n_cha < 1) return set_err(nullptr, HOE_EINVAL, "samples and coils must be positive");
    if (!d->kspha || !d->times) return set_err(nullptr, HOE_EINVAL, "kspha and times are required");
    if (!d->csm && d->n_cha != 1) return set_err(nullptr, HOE_EDIM, "csm and grid");
    int32_t backend = HOE_BACKEND_FP64;
    double dwell = 0.0;
    if constexpr (!F64) {
        backend = d->backend;
        dwell = d->dwell;
    }
    if (!F64 && (backend < HOE_BACKEND_DIRECT || backend > HOE_BACKEND_MIXED_BF16))
        return set_err(nullptr, HOE_EINVAL, "unknown explicit backend");
    if (!(dwell >= 0.0) || !std::isfinite(dwell)) return set_err(nullptr, HOE_EINVAL, "dwell must be >= 0");
    const int n_term = d->n_term;
    // order flags (prep_kspha.jl:8-33)
    bool t[4] = {true, true, true, n_term == 16};
    if (d->recon_terms) {
        const size_t expect = n_term == 9 ? 3 : 4;
        const size_t len = std::strlen(d->recon_terms);
        if (len != expect)
            return set_err(nullptr, HOE_EINVAL,
                           "recon_terms must contain " + std::to_string(expect) + " binary digits for nTerm=" +
                               std::to_string(n_term));
        for (size_t i = 0; i < len; ++i) {
            const char c = d->recon_terms[i];
            if (c != '0' && c != '1') return set_err(nullptr, HOE_EINVAL, "recon_terms must contain only '0' and '1'");
            t[i] = (c == '1');
        }
        if (n_term == 9) t[3] = false;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(nullptr, HOE_ENODEV, "no HIP device");
    if (d->device < 0 || d->device >= ndev) return set_err(nullptr, HOE_EINVAL, "invalid visible GPU ID");

    const int64_t n_grid = (int64_t)d->n_x * d->n_y * d->n_z;
    std::vector<int32_t> idx;
    idx.reserve((size_t)n_grid);
    for (int64_t i = 0; i < n_grid; ++i)
        if (!d->mask || d->mask[i]) idx.push_back((int32_t)i);
    if (idx.empty()) return set_err(nullptr, HOE_EINVAL, "mask must contain voxels");
    if (n_grid > INT32_MAX) return set_err(nullptr, HOE_EINVAL, "grid too large");

    Handle *h = new (std::nothrow) Handle();
    if (!h) return set_err(nullptr, HOE_ENOMEM, "host allocation failed");
    h->n_grid = n_grid;
    h->n_vox = (int64_t)idx.size();
    h->n_sam = d->n_sam;
    h->n_cha = d->n_cha;
    h->n_term_in = n_term;
    h->L = n_term + (d->concomitant ? 4 : 0);
    h->has_conc = d->concomitant != nullptr;
    h->nv_pad = round_up(h->n_vox, hoe::kRowPad);
    h->ns_pad = round_up(h->n_sam, hoe::kRowPad);
    // coil stride: one 32-coil MFMA group, or whole pairs of groups (the mixed kernels keep two
    // groups per wave when there are more than 32 coils)
    h->ncs = (int32_t)(h->n_cha <= 32 ? 32 : round_up(h->n_cha, 64));
    h->backend = backend;
    h->f64 = F64;
    h->cbytes = F64 ? 16 : 8;
    h->device = d->device;
    h->dwell = dwell;
    h->mask_idx_h = idx;

    // ---- prepared coefficient traces (term fastest), prep_kspha + prep_concomitant ----
    const int L = h->L;
    const int64_t ns = h->n_sam;
    for (int i = 0; i < 4; ++i) h->order[i] = t[i];
    std::vector<T> &traces = [&]() -> std::vector<T> & {
        if constexpr (F64) return h->traces64;
        else return h->traces;
    }();
    traces.assign((size_t)L * ns, T(0));
    prepare_traces_t<T>(h, traces, d->kspha, d->k_nominal);
    if (d->concomitant)
        for (int64_t s = 0; s < ns; ++s)
            for (int l = 0; l < 4; ++l) traces[(size_t)s * L + n_term + l] = d->concomitant[(size_t)s * 4 + l];
    if constexpr (!F64) h->times.assign(d->times, d->times + ns);
    const std::vector<double> times64(d->times, d->times + ns);

    // ---- basis columns in FP64; identically zero columns (z terms on a z = 0
    // plane, ...) contribute nothing and are dropped from the kernel records ----
    const int64_t nv = h->n_vox;
    std::vector<double> bfull((size_t)nv * L, 0.0);   // [v][l], column 0 = 1
    std::vector<char> nonzero(L, 0);
    for (int64_t v = 0; v < nv; ++v) {
        const int64_t g = idx[(size_t)v];
        const double x = d->x[g], y = d->y[g], z = d->z[g];
        double sh[15];
        solid_harmonics(x, y, z, sh);
        double *r = &bfull[(size_t)v * L];
        r[0] = 1.0;
        for (int l = 1; l < n_term; ++l) r[l] = sh[l - 1];
        if (d->concomitant) {
            r[n_term + 0] = z * z;
            r[n_term + 1] = x * x + y * y;
            r[n_term + 2] = x * z;
            r[n_term + 3] = y * z;
        }
        for (int l = 1; l < L; ++l)
            if (r[l] != 0.0) nonzero[l] = 1;
    }
    h->active.clear();
    for (int l = 1; l < L; ++l)
        if (nonzero[l]) h->active.push_back(l);
    // max |b| of each record entry (constant, active terms): the bound on psi that scales the
    // MFMA derivative forward's A operand
    h->rec_bmax.assign(h->active.size() + 1, 0.0);
    h->rec_bmax[0] = 1.0;
    for (int64_t v = 0; v < nv; ++v)
        for (size_t i = 0; i < h->active.size(); ++i)
            h->rec_bmax[1 + i] = std::max(h->rec_bmax[1 + i], std::fabs(bfull[(size_t)v * L + h->active[i]]));
    const int need = (int)h->active.size() + 1;     // active terms + fieldmap
    static const int kBuckets[] = {6, 10, 12, 16, 20};
    h->LK = 20;
    for (int b : kBuckets)
        if (need <= b) {
            h->LK = b;
            break;
        }
    h->sb = (int32_t)round_up(h->LK, 8);
    h->sk = (int32_t)round_up(h->LK + 1, 8);
    // voxel records [b_{a_1}..b_{a_m}, fieldmap, 0...]
    std::vector<double> brec((size_t)h->nv_pad * h->sb, 0.0);
    float csm_max = 0.0f;
    using C = typename std::conditional<F64, double2, float2>::type;
    std::vector<C> csmv((size_t)h->nv_pad * h->ncs, C{T(0), T(0)});
    const int na = (int)h->active.size();
    for (int64_t v = 0; v < nv; ++v) {
        const int64_t g = idx[(size_t)v];
        double *r = &brec[(size_t)v * h->sb];
        for (int i = 0; i < na; ++i) r[i] = bfull[(size_t)v * L + h->active[i]];
        r[na] = d->fieldmap ? (double)d->fieldmap[g] : 0.0;
        for (int64_t c = 0; c < h->n_cha; ++c) {
            C val{T(1), T(0)};
            if (d->csm) val = C{d->csm[2 * (c * n_grid + g)], d->csm[2 * (c * n_grid + g) + 1]};
            csmv[(size_t)v * h->ncs + c] = val;
            csm_max = std::max(csm_max, (float)std::sqrt(val.x * val.x + val.y * val.y));
        }
    }
    h->csm_max = csm_max;
    // power-of-two scale of each gradient basis column (constant, active terms):
    // max |b| * scale < 2^12, as the device's fragment scaling
    {
        auto pow2 = [](double m) {
            if (!(m > 0.0) || !std::isfinite(m)) return 1.0f;
            int e;
            (void)std::frexp(m, &e);
            return (float)std::ldexp(1.0, 12 - e);
        };
        h->gbscale_h.assign(32, 1.0f);
        h->gbscale_h[0] = pow2(1.0);
        for (int i = 0; i < na && 1 + i < 32; ++i) {
            double m = 0.0;
            for (int64_t v = 0; v < nv; ++v) m = std::max(m, std::fabs((double)(float)brec[(size_t)v * h->sb + i]));
            h->gbscale_h[1 + i] = pow2(m);
        }
    }

    // ---- runs of voxels sharing (y, z) and their x-polynomial tables (Horner forward) ----
    std::vector<double> hrec((size_t)h->nv_pad * 2, 0.0);
    std::vector<int32_t> vrun((size_t)h->nv_pad, 0);
    std::vector<double> beta;
    std::vector<hoe::HoeUnit> units;
    std::vector<hoe::HoeUnitVox> uvox;
    std::vector<hoe::HoeAdjAnchor> aanc;
    std::vector<float> aud;
    std::vector<double> abound;
    {
        int run = -1;
        T py = T(0), pz = T(0);
        for (int64_t v = 0; v < nv; ++v) {
            const int64_t g = idx[(size_t)v];
            const T yv = d->y[g], zv = d->z[g];
            if (run < 0 || yv != py || zv != pz) {
                ++run;
                py = yv;
                pz = zv;
                beta.resize((size_t)(run + 1) * h->LK * 4, 0.0);
                double *bt = &beta[(size_t)run * h->LK * 4];
                bt[0] = 1.0;                                   // k_1 * 1
                for (int i = 0; i < na; ++i) x_poly(h->active[i], n_term, (double)yv, (double)zv, bt + 4 * (1 + i));
            }
            vrun[(size_t)v] = run;
            hrec[2 * (size_t)v] = (double)d->x[g];
            hrec[2 * (size_t)v + 1] = d->fieldmap ? (double)d->fieldmap[g] : 0.0;
        }
        for (int64_t v = nv; v < h->nv_pad; ++v) vrun[(size_t)v] = run;
        h->n_runs = run + 1;
        h->t_index = na + 1;
        // consistency of the run polynomials with the basis records
        bool ok = true;
        for (int64_t v = 0; v < nv && ok; v += std::max<int64_t>(1, nv / 997)) {
            const double x = hrec[2 * (size_t)v];
            const double *bt = &beta[(size_t)vrun[(size_t)v] * h->LK * 4];
            for (int i = 0; i < na; ++i) {
                const double *p = bt + 4 * (1 + i);
                const double val = p[0] + x * (p[1] + x * (p[2] + x * p[3]));
                const double ref = bfull[(size_t)v * L + h->active[i]];
                if (std::fabs(val - ref) > 1e-12 * (1.0 + std::fabs(ref))) ok = false;
            }
        }
        h->horner = ok && (double)nv / (double)h->n_runs >= 16.0;
        // the MFMA adjoint evaluates a wave's 64 voxels from the Horner tables of at most
        // two runs; waves spanning more (short rows of a mask) take the generic phase
#ifndef HOE_ADJ_GENERIC_W
#define HOE_ADJ_GENERIC_W 1.6
#endif
#ifndef HOE_FWD_CROSS_W
#define HOE_FWD_CROSS_W 2.0
#endif
#ifndef HOE_HORNER_ADJ
#define HOE_HORNER_ADJ 1
#endif
        h->horner_adj = HOE_HORNER_ADJ && h->horner;
        // A workgroup of the MFMA adjoint lasts as long as its slowest wave; a tile holding a
        // wave on the generic phase (three or more runs) costs more than a Horner tile (its
        // workgroups took 1.3x the median on the 7 T spiral data). Its stream-K units weigh
        // 1.6, the best of 1.15-2.5 measured on those data.
        h->adj_tile_w.assign((size_t)(h->nv_pad / 256), 1.0);
        if (h->horner_adj)
            for (int64_t w = 0; w < h->nv_pad / 64; ++w)
                if (vrun[(size_t)(64 * w + 63)] - vrun[(size_t)(64 * w)] > 1)
                    h->adj_tile_w[(size_t)(w / 4)] = HOE_ADJ_GENERIC_W;
        // In the Horner forward a K unit whose 8 voxels cross a run boundary takes the generic
        // phase (L FP64 FMAs per phase and the full records instead of 4 FMAs); its weight in
        // the stream-K cut of the K range over the XCDs says so.
        h->fwd_k_w.assign((size_t)(h->nv_pad / 8), 1.0);
        if (h->horner)
            for (int64_t k = 0; k < h->nv_pad / 8; ++k)
                if (vrun[(size_t)(8 * k)] != vrun[(size_t)(8 * k + 7)]) h->fwd_k_w[(size_t)k] = HOE_FWD_CROSS_W;
        // K units of the MFMA Horner forward: the anchor is the middle of the real voxels'
        // x range (rounded to FP32) and the fieldmap of the first real voxel; padding
        // voxels get u = dfm = 0
        // group anchors of the MFMA Horner adjoint (waves spanning at most two runs)
        aanc.assign((size_t)(h->nv_pad / 64), hoe::HoeAdjAnchor{});
        aud.assign((size_t)h->nv_pad * 2, 0.0f);
        for (int64_t w = 0; w < h->nv_pad / 64; ++w) {
            const int r_lo = vrun[(size_t)(64 * w)];
            if (vrun[(size_t)(64 * w + 63)] - r_lo > 1) continue;
            double lo[4], hi[4];
            bool seen[4] = {false, false, false, false};
            for (int64_t v = 64 * w; v < std::min<int64_t>(64 * w + 64, nv); ++v) {
                const int c = 2 * (vrun[(size_t)v] - r_lo) + (int)((v % 64) / 32);
                const double x = hrec[2 * (size_t)v];
                if (!seen[c]) {
                    seen[c] = true;
                    lo[c] = hi[c] = x;
                    aanc[(size_t)w].fm[c] = (float)hrec[2 * (size_t)v + 1];
                }
                lo[c] = std::min(lo[c], x);
                hi[c] = std::max(hi[c], x);
            }
            for (int c = 0; c < 4; ++c)
                if (seen[c]) aanc[(size_t)w].xa[c] = (float)(0.5 * (lo[c] + hi[c]));
            for (int64_t v = 64 * w; v < std::min<int64_t>(64 * w + 64, nv); ++v) {
                const int c = 2 * (vrun[(size_t)v] - r_lo) + (int)((v % 64) / 32);
                aud[2 * (size_t)v] = (float)(hrec[2 * (size_t)v] - (double)aanc[(size_t)w].xa[c]);
                aud[2 * (size_t)v + 1] = (float)(hrec[2 * (size_t)v + 1] - (double)aanc[(size_t)w].fm[c]);
            }
        }
        units.assign((size_t)(h->nv_pad / 8), hoe::HoeUnit{});
        uvox.assign((size_t)(h->nv_pad / 8), hoe::HoeUnitVox{});
        for (int64_t k = 0; k < h->nv_pad / 8; ++k) {
            hoe::HoeUnit &a = units[(size_t)k];
            const int64_t v0 = 8 * k, v1 = std::min<int64_t>(8 * k + 8, nv);
            a.run = vrun[(size_t)v0] == vrun[(size_t)(8 * k + 7)] ? vrun[(size_t)v0] : -1;
            if (a.run < 0 || v0 >= v1) continue;
            double lo = hrec[2 * (size_t)v0], hi = lo;
            for (int64_t v = v0; v < v1; ++v) {
                lo = std::min(lo, hrec[2 * (size_t)v]);
                hi = std::max(hi, hrec[2 * (size_t)v]);
            }
            a.xa = (float)(0.5 * (lo + hi));
            a.fma = (float)hrec[2 * (size_t)v0 + 1];
            for (int64_t v = v0; v < v1; ++v) {
                uvox[(size_t)k].ud[v - v0][0] = (float)(hrec[2 * (size_t)v] - (double)a.xa);
                uvox[(size_t)k].ud[v - v0][1] = (float)(hrec[2 * (size_t)v + 1] - (double)a.fma);
            }
        }
        // Anchored phases add t * (fieldmap - anchor fieldmap) in FP32, whose rounding grows
        // with its size (about 2e-7 of it); the FP32 offset along x is bounded by the grid
        // (half a K unit or a wave half of voxels, times the phase change per voxel). Above
        // kAnchorMaxCycles of fieldmap offset (a rough map over a long readout) the MFMA
        // kernels keep the FP64 phase of every pair instead.
        {
            double tmax = 0.0, fmax_off = 0.0;
            for (int64_t s = 0; s < ns; ++s) tmax = std::max(tmax, std::fabs((double)d->times[s]));
            for (const hoe::HoeUnitVox &u : uvox)
                for (int j = 0; j < 8; ++j) fmax_off = std::max(fmax_off, std::fabs((double)u.ud[j][1]));
            for (int64_t v = 0; v < h->nv_pad; ++v) fmax_off = std::max(fmax_off, std::fabs((double)aud[2 * (size_t)v + 1]));
            h->anchor_fm_cycles = tmax * fmax_off;
            h->anchored = h->anchor_fm_cycles <= hoe::kAnchorMaxCycles;
        }
        // D_r: the largest change of record entry r's run polynomial between a voxel and its
        // anchor, over the forward's K units and the adjoint's groups. The device bounds the
        // Horner offsets of the current records by max_s sum_r |A_s[r]| D_r (entry 0 and t: 0).
        abound.assign(32, 0.0);
        if (h->horner) {
            auto poly = [&](int run, int r, double x) {
                const double *p = &beta[((size_t)run * h->LK + r) * 4];
                return p[0] + x * (p[1] + x * (p[2] + x * p[3]));
            };
            auto widen = [&](int run, double xa, double x) {
                for (int r = 1; r <= na; ++r)
                    abound[(size_t)r] = std::max(abound[(size_t)r], std::fabs(poly(run, r, x) - poly(run, r, xa)));
            };
            for (int64_t k = 0; k < h->nv_pad / 8; ++k) {
                const hoe::HoeUnit &a = units[(size_t)k];
                if (a.run < 0) continue;
                for (int64_t v = 8 * k; v < std::min<int64_t>(8 * k + 8, nv); ++v)
                    widen(a.run, (double)a.xa, hrec[2 * (size_t)v]);
            }
            for (int64_t w = 0; w < h->nv_pad / 64; ++w) {
                const int r_lo = vrun[(size_t)(64 * w)];
                if (vrun[(size_t)(64 * w + 63)] - r_lo > 1) continue;
                for (int64_t v = 64 * w; v < std::min<int64_t>(64 * w + 64, nv); ++v) {
                    const int c = 2 * (vrun[(size_t)v] - r_lo) + (int)((v % 64) / 32);
                    widen(vrun[(size_t)v], (double)aanc[(size_t)w].xa[c], hrec[2 * (size_t)v]);
                }
            }
        }
    }

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess && prop.multiProcessorCount > 0)
        h->cu_count = prop.multiProcessorCount;

    // ---- device allocation ----
    auto fail = [&](int code) {
        std::string msg = g_err;
        release(h);
        delete h;
        g_err = msg;
        return code;
    };
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return fail(set_err(nullptr, HOE_EHIP, "hipSetDevice failed"));
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(set_err(nullptr, HOE_EHIP, "hipStreamCreate failed"));
    for (auto &ev : h->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(set_err(nullptr, HOE_EHIP, "hipEventCreate failed"));
    const int n_gb = h->ncs / 32;
    const size_t cb = (size_t)h->cbytes;
    const int nbw = n_gb >= 2 ? 2 : 1;
    const int64_t n_gb_pad = ((n_gb + nbw - 1) / nbw) * nbw;
    struct {
        hoe::DeviceBuf *b;
        size_t bytes;
    } plan[] = {
        {&h->mask_idx, sizeof(int32_t) * (size_t)nv},
        {&h->basis, sizeof(double) * brec.size()},
        {&h->coef, sizeof(double) * (size_t)h->ns_pad * h->sk},
        {&h->dcoef, sizeof(double) * (size_t)h->ns_pad * h->sk},
        {&h->dtrec, sizeof(double) * (size_t)h->ns_pad * h->sk},
        {&h->ghbuf, sizeof(float) * (size_t)ns * L},
        {&h->growmap, sizeof(int32_t) * (size_t)L},
        {&h->traces_d, F64 ? 0 : sizeof(float) * (size_t)L * ns},
        {&h->times_d, F64 ? 0 : sizeof(float) * (size_t)ns},
        {&h->act_d, sizeof(int32_t) * std::max<size_t>(h->active.size(), 1)},
        {&h->csm, sizeof(C) * csmv.size()},
        {&h->hrec, sizeof(double) * hrec.size()},
        {&h->vrun, sizeof(int32_t) * vrun.size()},
        {&h->beta, sizeof(double) * std::max<size_t>(beta.size(), 1)},
        {&h->units, sizeof(hoe::HoeUnit) * units.size()},
        {&h->uvox, sizeof(hoe::HoeUnitVox) * uvox.size()},
        {&h->aanc, sizeof(hoe::HoeAdjAnchor) * aanc.size()},
        {&h->aud, sizeof(float) * aud.size()},
        {&h->abound, 512},
        {&h->xm, cb * (size_t)h->nv_pad},
        {&h->zbuf, cb * (size_t)h->nv_pad * h->ncs},
        {&h->zfrag, F64 ? 0 : (size_t)n_gb_pad * (size_t)(h->nv_pad / 8) * 64 * 64},
        {&h->ybuf, cb * (size_t)h->ns_pad * h->ncs},
        {&h->yfrag, F64 ? 0 : (size_t)n_gb_pad * (size_t)(h->ns_pad / 8) * 64 * 64},
        {&h->dvec, (cb / 2) * (size_t)h->ns_pad * h->ncs},
        {&h->io_x, cb * (size_t)n_grid},
        {&h->io_x2, cb * (size_t)n_grid},
        {&h->io_y, cb * (size_t)(ns * h->n_cha)},
        {&h->io_y2, cb * (size_t)(ns * h->n_cha)},
        {&h->scal, 4096},
        {&h->cg_vec, cb * (size_t)h->nv_pad * 4},
        {&h->red, sizeof(double) * 4096},
        {&h->gk, sizeof(double) * (size_t)h->LK * ns},
    };
    for (auto &p : plan) {
        e = dalloc(*p.b, p.bytes);
        if (e != hipSuccess)
            return fail(set_err(nullptr, e == hipErrorOutOfMemory ? HOE_ENOMEM : HOE_EHIP,
                                std::string("device allocation failed: ") + hipGetErrorString(e)));
    }
    e = hipMemcpy(h->mask_idx.p, idx.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->basis.p, brec.data(), sizeof(double) * brec.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->csm.p, csmv.data(), sizeof(C) * csmv.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->hrec.p, hrec.data(), sizeof(double) * hrec.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->vrun.p, vrun.data(), sizeof(int32_t) * vrun.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !beta.empty())
        e = hipMemcpy(h->beta.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(h->units.p, units.data(), sizeof(hoe::HoeUnit) * units.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(h->uvox.p, uvox.data(), sizeof(hoe::HoeUnitVox) * uvox.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(h->aanc.p, aanc.data(), sizeof(hoe::HoeAdjAnchor) * aanc.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->aud.p, aud.data(), sizeof(float) * aud.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // D_r, then the offset bound +inf (anchored phases off) until the first record build
        std::vector<unsigned char> ab(512, 0);
        std::memcpy(ab.data(), abound.data(), sizeof(double) * 32);
        const float inf = std::numeric_limits<float>::infinity();
        std::memcpy(ab.data() + 256, &inf, sizeof(float));
        e = hipMemcpy(h->abound.p, ab.data(), ab.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !F64)
        e = hipMemcpy(h->times_d.p, h->times.data(), sizeof(float) * (size_t)ns, hipMemcpyHostToDevice);
    if (e == hipSuccess && !h->active.empty())
        e = hipMemcpy(h->act_d.p, h->active.data(), sizeof(int32_t) * h->active.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // kernel row of each original term for the gradient (-1: identically zero basis)
        std::vector<int32_t> rowmap((size_t)L, -1);
        rowmap[0] = 0;
        for (int i = 0; i < na; ++i) rowmap[(size_t)h->active[i]] = 1 + i;
        e = hipMemcpy(h->growmap.p, rowmap.data(), sizeof(int32_t) * rowmap.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return fail(set_err(nullptr, HOE_EHIP, std::string("upload failed: ") + hipGetErrorString(e)));
    e = hoe::prepare_backend(h, h->backend);
    if (e == hipSuccess && !F64) e = hoe::prepare_grad(h);
    if (e != hipSuccess)
        return fail(set_err(nullptr, e == hipErrorOutOfMemory ? HOE_ENOMEM : HOE_EHIP,
                            std::string("schedule setup failed: ") + hipGetErrorString(e)));
    int rc = F64 ? upload_coef_f64(h, times64) : upload_coef(h);
    if (rc != HOE_OK) return fail(set_err(nullptr, rc, h->err));
    h->weights_unit = false;
    rc = set_weights(h, nullptr, 0);
    if (rc != HOE_OK) return fail(set_err(nullptr, rc, h->err));
    *out = reinterpret_cast<hoe_handle *>(h);
    return HOE_OK;
}

}  // namespace

extern "C" {

int hoe_create(const hoe_model_desc *d, hoe_handle **out) { return create_impl(d, out); }

int hoe_create_f64(const hoe_model_desc_f64 *d, hoe_handle **out) { return create_impl(d, out); }

int hoe_get_info(const hoe_handle *hh, hoe_info_t *info) {
    const Handle *h = reinterpret_cast<const Handle *>(hh);
    if (!h || !info) return set_err(nullptr, HOE_EINVAL, "null argument");
    info->n_grid = h->n_grid;
    info->n_vox = h->n_vox;
    info->n_sam = h->n_sam;
    info->n_cha = h->n_cha;
    info->n_term = h->L;
    info->backend = h->backend;
    info->device = h->device;
    info->closed = h->closed ? 1 : 0;
    info->horner = (h->horner ? 1 : 0) | (h->horner_adj ? 2 : 0) | (anchored_now(h) ? 4 : 0);
    info->n_runs = h->n_runs;
    return HOE_OK;
}

int hoe_set_backend(hoe_handle *hh, int32_t backend) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if (backend < HOE_BACKEND_DIRECT || backend > HOE_BACKEND_FP64)
        return set_err(h, HOE_EINVAL, "unknown explicit backend");
    if ((backend == HOE_BACKEND_FP64) != h->f64)
        return set_err(h, HOE_EINVAL,
                       "the FP64 backend takes ComplexF64 vectors: create it with hoe_create_f64; other "
                       "backends take ComplexF32 ones");
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    HOE_TRY_HIP(h, hoe::prepare_backend(h, backend));
    h->backend = backend;
    return HOE_OK;
}

// ---- vector staging helpers ----
namespace {

struct Staging {
    Handle *h;
    int flags;
    bool dev() const { return (flags & HOE_DEVICE_PTRS) != 0; }
};

int stage_in(Handle *h, int flags, const void *src, hoe::DeviceBuf &buf, size_t bytes, const void **dptr) {
    if (!src) return set_err(h, HOE_EINVAL, "vector pointer is NULL");
    if (flags & HOE_DEVICE_PTRS) {
        *dptr = src;
        return HOE_OK;
    }
    HOE_TRY_HIP(h, hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, h->stream));
    *dptr = buf.p;
    return HOE_OK;
}

int finish(Handle *h, int flags, void *dst_host, const void *dsrc, size_t bytes) {
    if (!(flags & HOE_DEVICE_PTRS) && dst_host)
        HOE_TRY_HIP(h, hipMemcpyAsync(dst_host, dsrc, bytes, hipMemcpyDeviceToHost, h->stream));
    // timing events only on synchronous calls: each event on the stream costs a gap of
    // several microseconds between the kernels it separates
    if (!(flags & HOE_ASYNC)) (void)hipEventRecord(h->ev[5], h->stream);
    if (!(flags & HOE_ASYNC) || !(flags & HOE_DEVICE_PTRS)) HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    HOE_TRY_HIP(h, hipGetLastError());
    return HOE_OK;
}

void record_times(Handle *h, bool fwd, bool adj) {
    h->last_ms[0] = fwd ? ev_ms(h->ev[0], h->ev[1]) : -1.0;
    h->last_ms[1] = adj ? ev_ms(h->ev[2], h->ev[3]) : -1.0;
    h->last_ms[2] = ev_ms(h->ev[4], h->ev[5]);
}

}  // namespace

int hoe_forward(hoe_handle *hh, const void *x, void *y, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if (!y) return set_err(h, HOE_EINVAL, "output pointer is NULL");
    if (!(flags & HOE_ASYNC)) (void)hipEventRecord(h->ev[4], h->stream);
    const void *dx;
    if ((rc = stage_in(h, flags, x, h->io_x, (size_t)h->cbytes * h->n_grid, &dx))) return rc;
    void *dy = (flags & HOE_DEVICE_PTRS) ? y : h->io_y.p;
    HOE_TRY_HIP(h, hoe::run_forward_user(h, dx, dy, prec_scale(h, 1.0 / std::sqrt((double)h->n_vox)),
                                         !(flags & HOE_ASYNC)));
    if ((rc = finish(h, flags, y, dy, (size_t)h->cbytes * h->n_sam * h->n_cha))) return rc;
    if (!(flags & HOE_ASYNC)) record_times(h, true, false);
    return HOE_OK;
}

int hoe_adjoint(hoe_handle *hh, const void *y, void *x, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if (!x) return set_err(h, HOE_EINVAL, "output pointer is NULL");
    if (!(flags & HOE_ASYNC)) (void)hipEventRecord(h->ev[4], h->stream);
    const void *dy;
    if ((rc = stage_in(h, flags, y, h->io_y, (size_t)h->cbytes * h->n_sam * h->n_cha, &dy))) return rc;
    void *dx = (flags & HOE_DEVICE_PTRS) ? x : h->io_x.p;
    HOE_TRY_HIP(h, hoe::launch_load_ybuf(h, dy, false));
    const double sc = prec_scale(h, 1.0 / std::sqrt((double)h->n_vox));
    if (unmasked(h)) {
        // the masked order is the grid order: the reduction writes the output directly
        HOE_TRY_HIP(h, hoe::run_adjoint_contract(h, sc, dx, !(flags & HOE_ASYNC), (int)h->n_vox));
    } else {
        HOE_TRY_HIP(h, hoe::run_adjoint_contract(h, sc, h->xm.p, !(flags & HOE_ASYNC)));
        HOE_TRY_HIP(h, hoe::launch_scatter(h, h->xm.p, dx, 1.0));
    }
    if ((rc = finish(h, flags, x, dx, (size_t)h->cbytes * h->n_grid))) return rc;
    if (!(flags & HOE_ASYNC)) record_times(h, false, true);
    return HOE_OK;
}

int hoe_set_weights(hoe_handle *hh, const float *D, int64_t n_d) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    return set_weights(h, D, n_d);
}

int hoe_set_weights_f64(hoe_handle *hh, const double *D, int64_t n_d) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    return set_weights_t(h, D, n_d);
}

int hoe_normal(hoe_handle *hh, const float *D, int64_t n_d, const void *x, void *out, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_dt(h))) return rc;
    if (!out) return set_err(h, HOE_EINVAL, "output pointer is NULL");
    if (!(D == nullptr && n_d == kCurrentWeights))
        if ((rc = set_weights(h, D, n_d))) return rc;
    if (!(flags & HOE_ASYNC)) (void)hipEventRecord(h->ev[4], h->stream);
    const void *dx;
    if ((rc = stage_in(h, flags, x, h->io_x, (size_t)h->cbytes * h->n_grid, &dx))) return rc;
    void *dout = (flags & HOE_DEVICE_PTRS) ? out : h->io_x2.p;
    const double sc = prec_scale(h, 1.0 / (double)h->n_vox);
    if (unmasked(h)) {
        // x is consumed by the gather before the reduction writes the output (in place is fine)
        HOE_TRY_HIP(h, hoe::run_normal_full(h, dx, dout, sc, !(flags & HOE_ASYNC), (int)h->n_vox));
    } else {
        void *tmp = h->cg_vec.p;   // masked output
        HOE_TRY_HIP(h, hoe::run_normal_full(h, dx, tmp, sc, !(flags & HOE_ASYNC)));
        HOE_TRY_HIP(h, hoe::launch_scatter(h, tmp, dout, 1.0));
    }
    if ((rc = finish(h, flags, out, dout, (size_t)h->cbytes * h->n_grid))) return rc;
    if (!(flags & HOE_ASYNC)) record_times(h, true, true);
    return HOE_OK;
}

}  // extern "C"

// ---- the normal operator without allocation: workspaces from a static device array ----

// Workspace of hoe_normal_noalloc on every GPU the library runs on: one file-scope array
// per device, held until the process ends. Its name is a host-side stand-in, not a device
// address: host code uses only the per-device address from hipGetSymbolAddress.
static constexpr size_t kScratchBytes = size_t(512) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (the device must be current).
int scratch_on(Handle *h, unsigned char **out) {
    if (h->device < 0 || h->device >= kMaxDevices) return set_err(h, HOE_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(h, HOE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[h->device] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[h->device]);
    return HOE_OK;
}

// The handle's workspaces that one normal-operator application writes, pointed into the
// scratch array for the duration of a call and restored afterwards (the kernels take the
// addresses when they are enqueued).
struct ScratchSwap {
    Handle *h;
    hoe::DeviceBuf *bufs[9];
    hoe::DeviceBuf saved[9];
    int n = 0;
    explicit ScratchSwap(Handle *hh) : h(hh) {}
    void put(hoe::DeviceBuf &b, unsigned char *p, size_t bytes) {
        bufs[n] = &b;
        saved[n] = b;
        b.p = p;
        b.bytes = bytes;
        ++n;
    }
    ~ScratchSwap() {
        for (int i = 0; i < n; ++i) *bufs[i] = saved[i];
    }
};

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of every workspace of a normal-operator application on the handle's backend,
// in the order they are carved out of the scratch array.
size_t max_segs(const std::vector<hoe::Sched> &passes) {
    size_t m = 0;
    for (const hoe::Sched &sc : passes) m = std::max(m, (size_t)sc.n_seg);
    return m;
}

void noalloc_plan(const Handle *h, size_t sizes[9]) {
    const bool mixed = h->backend != HOE_BACKEND_DIRECT;
    const size_t n_gb = (size_t)hoe::mixed_groups(h) * hoe::mixed_nb(h);
    sizes[0] = 4096;                                                              // scal: fragment scales
    sizes[1] = sizeof(float2) * (size_t)h->nv_pad;                                // xm
    sizes[2] = sizeof(float2) * (size_t)h->nv_pad;                                // cg_vec: masked output
    sizes[3] = sizeof(float2) * (size_t)h->ns_pad * h->ncs;                       // ybuf
    sizes[4] = sizeof(float2) * 256 * (size_t)h->ncs * max_segs(h->sched_f[h->backend]);   // part_f
    sizes[5] = sizeof(float2) * 256 * (size_t)hoe::adj_groups(h, h->backend) * max_segs(h->sched_a[h->backend]);   // part_a
    sizes[6] = mixed ? n_gb * (size_t)(h->nv_pad / 8) * 64 * 64 : 0;             // zfrag
    sizes[7] = mixed ? n_gb * (size_t)(h->ns_pad / 8) * 64 * 64 : 0;             // yfrag
    sizes[8] = mixed ? 0 : sizeof(float2) * (size_t)h->nv_pad * h->ncs;          // zbuf
}

}  // namespace

extern "C" {

int hoe_normal_noalloc(hoe_handle *hh, const void *x, void *out, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_dt(h))) return rc;
    if ((rc = reject_f64(h, "hoe_normal_noalloc"))) return rc;
    if (!x || !out) return set_err(h, HOE_EINVAL, "vector pointer is NULL");
    if (!(flags & HOE_DEVICE_PTRS))
        return set_err(h, HOE_EINVAL, "hoe_normal_noalloc takes device vectors (HOE_DEVICE_PTRS)");
    size_t sizes[9], need = 0;
    noalloc_plan(h, sizes);
    for (size_t s : sizes) need += align256(s);
    if (need > kScratchBytes)
        return set_err(h, HOE_EINVAL,
                       "hoe_normal_noalloc: this operator needs " + std::to_string(need >> 20) +
                           " MiB of workspace, more than the " + std::to_string(kScratchBytes >> 20) +
                           " MiB static array; use hoe_normal");
    unsigned char *base = nullptr;
    if ((rc = scratch_on(h, &base))) return rc;
    {
        ScratchSwap sw(h);
        hoe::DeviceBuf *targets[9] = {&h->scal,   &h->xm,     &h->cg_vec, &h->ybuf, &h->part_f,
                                      &h->part_a, &h->zfrag, &h->yfrag,  &h->zbuf};
        size_t off = 0;
        for (int i = 0; i < 9; ++i) {
            if (sizes[i] == 0) continue;
            sw.put(*targets[i], base + off, sizes[i]);
            off += align256(sizes[i]);
        }
        if (!(flags & HOE_ASYNC)) (void)hipEventRecord(h->ev[4], h->stream);
        const double sc = prec_scale(h, 1.0 / (double)h->n_vox);
        if (unmasked(h)) {
            HOE_TRY_HIP(h, hoe::run_normal_full(h, x, out, sc, !(flags & HOE_ASYNC), (int)h->n_vox));
        } else {
            void *tmp = h->cg_vec.p;
            HOE_TRY_HIP(h, hoe::run_normal_full(h, x, tmp, sc, !(flags & HOE_ASYNC)));
            HOE_TRY_HIP(h, hoe::launch_scatter(h, tmp, out, 1.0));
        }
    }
    if ((rc = finish(h, flags, out, out, sizeof(float2) * h->n_grid))) return rc;
    if (!(flags & HOE_ASYNC)) record_times(h, true, true);
    return HOE_OK;
}

int hoe_set_coefficients(hoe_handle *hh, const float *kspha, double tau, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_f64(h, "hoe_set_coefficients"))) return rc;
    if (!std::isfinite(tau)) return set_err(h, HOE_EINVAL, "tau must be finite");
    if (tau != 0.0 && !(h->dwell > 0.0)) return set_err(h, HOE_EINVAL, "a delay requires dwell > 0");
    if (kspha && (flags & HOE_DEVICE_PTRS)) {
        // device traces go straight to traces_d, behind the work queued on the stream
        HOE_TRY_HIP(h, hipMemcpyAsync(h->traces_d.p, kspha, sizeof(float) * h->traces.size(),
                                      hipMemcpyDeviceToDevice, h->stream));
        h->traces_stale = true;
    } else if (kspha) {
        HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
        std::memcpy(h->traces.data(), kspha, sizeof(float) * h->traces.size());
        HOE_TRY_HIP(h, hipMemcpyAsync(h->traces_d.p, h->traces.data(), sizeof(float) * h->traces.size(),
                                      hipMemcpyHostToDevice, h->stream));
        h->traces_stale = false;
    }
    const double tau_prev = h->tau;
    h->tau = tau;
    if ((rc = rebuild_records(h))) {
        h->tau = tau_prev;
        return rc;
    }
    if (!(flags & HOE_ASYNC) || !(flags & HOE_DEVICE_PTRS)) HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    return HOE_OK;
}

int hoe_set_kspha(hoe_handle *hh, const float *kspha, const float *k_nominal) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_f64(h, "hoe_set_kspha"))) return rc;
    if (!kspha) return set_err(h, HOE_EINVAL, "kspha is required");
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    if (h->traces_stale) {
        // the concomitant rows are kept from the traces last set from device memory
        HOE_TRY_HIP(h, hipMemcpy(h->traces.data(), h->traces_d.p, sizeof(float) * h->traces.size(),
                                 hipMemcpyDeviceToHost));
        h->traces_stale = false;
    }
    prepare_traces(h, kspha, k_nominal);
    h->tau = 0.0;
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    return upload_coef(h);
}

int hoe_set_kspha_dt(hoe_handle *hh, const float *kspha_dt) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_f64(h, "hoe_set_kspha_dt"))) return rc;
    if (!kspha_dt) {
        if (!h->dt_on) return HOE_OK;
        HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
        h->dt_on = false;
        return need_direct_resched(h) ? reschedule_direct(h, true) : HOE_OK;
    }
    if (h->has_conc)
        return set_err(h, HOE_EINVAL,
                       "kspha_dt cannot be combined with concomitant terms (the derivative model has none)");
    const int64_t ns = h->n_sam;
    const int n_term = h->n_term_in;
    for (int64_t i = 0; i < ns * n_term; ++i)
        if (!std::isfinite(kspha_dt[i])) return set_err(h, HOE_EINVAL, "kspha_dt must be finite");
    // records [kdt_1, kdt of the active terms, 0 (t slot), 0...]: terms whose basis is
    // identically zero on the masked voxels contribute nothing. psi_max bounds
    // |psi(s, v)| = |sum_i r_i(s) b_i(v)| over every pair.
    std::vector<double> rec((size_t)h->ns_pad * h->sk, 0.0);
    double psi_max = 0.0;
    for (int64_t s = 0; s < ns; ++s) {
        double *r = &rec[(size_t)s * h->sk];
        const float *k = kspha_dt + (size_t)s * n_term;
        r[0] = (double)k[0];
        for (size_t i = 0; i < h->active.size(); ++i) r[1 + i] = (double)k[h->active[i]];
        double b = 0.0;
        for (size_t i = 0; i < h->rec_bmax.size(); ++i) b += std::fabs(r[i]) * h->rec_bmax[i];
        psi_max = std::max(psi_max, b);
    }
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    HOE_TRY_HIP(h, hipMemcpy(h->dtrec.p, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice));
    // the MFMA forward's A operand is 2 pi sigma psi e^{i phi}: sigma, a power of two, keeps
    // it within [-1, 1] like the plain forward's e^{i phi}; the reduction divides it out
    const double m = 6.283185307179586 * psi_max;
    int ex = 0;
    if (m > 0.0) (void)std::frexp(m, &ex);   // m < 2^ex
    h->dt_sigma = std::ldexp(1.0, -ex);
    h->dt_w = 6.283185307179586 * h->dt_sigma;
    if (h->dt_on) return HOE_OK;
    h->dt_on = true;
    // the FP32 forward's schedule is sized by the occupancy of the kernel instance that will
    // run it: the derivative one from now on (the MFMA derivative forward keeps the plain
    // forward's schedule, its instance having the same occupancy)
    return need_direct_resched(h) ? reschedule_direct(h, false) : HOE_OK;
}

int hoe_grad(hoe_handle *hh, const void *x, const void *y, float *g_h, double *g_tau, double *loss, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_dt(h))) return rc;
    if ((rc = reject_f64(h, "hoe_grad"))) return rc;
    (void)hipEventRecord(h->ev[4], h->stream);   // the call always waits for its results
    const void *dx, *dy;
    if ((rc = stage_in(h, flags, x, h->io_x, sizeof(float2) * h->n_grid, &dx))) return rc;
    if ((rc = stage_in(h, flags, y, h->io_y, sizeof(float2) * h->n_sam * h->n_cha, &dy))) return rc;
    // residual r = E x - y into ybuf (coil fastest)
    HOE_TRY_HIP(h, hoe::run_forward_ybuf(h, dx, prec_scale(h, 1.0 / std::sqrt((double)h->n_vox)), false, dy, false));
    double *loss_dev = reinterpret_cast<double *>(h->scal.p) + 16;
    HOE_TRY_HIP(h, hoe::launch_norm2(h, h->ybuf.p, h->ns_pad * h->ncs, loss_dev));
    HOE_TRY_HIP(h, hoe::run_grad_contract(h, true));
    // dL/dh (through the delay interpolation) and dL/dtau on the device
    double *gtau_dev = reinterpret_cast<double *>(h->scal.p) + 17;
    const bool shift = h->tau != 0.0;
    int64_t m = 0;
    double w[4] = {0.0, 1.0, 0.0, 0.0}, dw[4];
    if (shift) {
        const double u = h->tau / h->dwell;
        m = (int64_t)std::floor(u);
        catmull_rom(u - (double)m, w, dw);
    }
    const bool dev_out = (flags & HOE_DEVICE_PTRS) && g_h;
    HOE_TRY_HIP(h, hoe::launch_grad_chain(h, gtau_dev, shift, (int)m, w, dev_out ? g_h : nullptr));
    (void)hipEventRecord(h->ev[5], h->stream);
    double host_sc[2] = {0.0, 0.0};
    if (g_h && !dev_out)
        HOE_TRY_HIP(h, hipMemcpyAsync(g_h, h->ghbuf.p, sizeof(float) * (size_t)h->n_sam * h->L,
                                      hipMemcpyDeviceToHost, h->stream));
    HOE_TRY_HIP(h, hipMemcpyAsync(host_sc, loss_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    h->last_ms[0] = ev_ms(h->ev[0], h->ev[1]);
    h->last_ms[1] = -1.0;
    h->last_ms[2] = ev_ms(h->ev[4], h->ev[5]);
    if (loss) *loss = host_sc[0];
    if (g_tau) *g_tau = h->dwell > 0.0 ? host_sc[1] : NAN;
    return HOE_OK;
}

int hoe_cg(hoe_handle *hh, const void *y, const float *D, int64_t n_d, int32_t iters, double lambda, double rel_tol,
           void *x_out, double *resid_log, int flags) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    if ((rc = reject_dt(h))) return rc;
    if (!x_out) return set_err(h, HOE_EINVAL, "output pointer is NULL");
    if (iters < 0) return set_err(h, HOE_EINVAL, "iterations must be nonnegative");
    if (!std::isfinite(lambda) || lambda < 0.0) return set_err(h, HOE_EINVAL, "lambda must be finite and >= 0");
    if (!(D == nullptr && n_d == kCurrentWeights))
        if ((rc = set_weights(h, D, n_d))) return rc;
    const size_t log_bytes = sizeof(double) * ((size_t)iters + 1);
    if (h->cg_log.bytes < log_bytes) {
        dfree(h->cg_log);
        HOE_TRY_HIP(h, dalloc(h->cg_log, log_bytes));
    }
    if (!(flags & HOE_ASYNC)) (void)hipEventRecord(h->ev[4], h->stream);
    const void *dy;
    if ((rc = stage_in(h, flags, y, h->io_y, (size_t)h->cbytes * h->n_sam * h->n_cha, &dy))) return rc;
    void *dx = (flags & HOE_DEVICE_PTRS) ? x_out : h->io_x.p;
    // rhs = E^H D y (masked) into xm
    HOE_TRY_HIP(h, hoe::launch_load_ybuf(h, dy, true));
    HOE_TRY_HIP(h, hoe::run_adjoint_contract(h, prec_scale(h, 1.0 / std::sqrt((double)h->n_vox)), h->xm.p, false));
    void *xs = static_cast<char *>(h->cg_vec.p) + (size_t)3 * h->nv_pad * h->cbytes;
    HOE_TRY_HIP(h, hoe::cg_run(h, h->xm.p, iters, lambda, rel_tol, xs, (double *)h->cg_log.p));
    HOE_TRY_HIP(h, hoe::launch_scatter(h, xs, dx, 1.0));
    if (resid_log)
        HOE_TRY_HIP(h, hipMemcpyAsync(resid_log, h->cg_log.p, log_bytes, hipMemcpyDeviceToHost, h->stream));
    if ((rc = finish(h, flags, x_out, dx, (size_t)h->cbytes * h->n_grid))) return rc;
    if (resid_log && (flags & HOE_ASYNC) && (flags & HOE_DEVICE_PTRS)) HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    h->last_ms[0] = h->last_ms[1] = -1.0;
    if (!(flags & HOE_ASYNC)) h->last_ms[2] = ev_ms(h->ev[4], h->ev[5]);
    return HOE_OK;
}

int hoe_close(hoe_handle *hh) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) return set_err(nullptr, HOE_EINVAL, "null handle");
    if (h->closed) return HOE_OK;
    release(h);
    return HOE_OK;
}

void hoe_destroy(hoe_handle *hh) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) return;
    if (!h->closed) release(h);
    delete h;
}

int hoe_sync(hoe_handle *hh) {
    Handle *h = reinterpret_cast<Handle *>(hh);
    int rc = check_open(h);
    if (rc) return rc;
    HOE_TRY_HIP(h, hipStreamSynchronize(h->stream));
    return HOE_OK;
}

double hoe_last_kernel_ms(const hoe_handle *hh, int which) {
    const Handle *h = reinterpret_cast<const Handle *>(hh);
    if (!h || which < 0 || which > 2) return -1.0;
    return h->last_ms[which];
}

}  // extern "C"

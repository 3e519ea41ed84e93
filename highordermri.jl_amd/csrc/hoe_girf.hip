// GIRF prediction on the device: the frequency-domain gradient response model
// of HighOrderMRI.jl's apply_girf (src/GIRF/apply_girf.jl:39-150).
//
// For every waveform (3 DCS channels x n_point samples) the reference pads to
// N = nextprod([2,3,5], max(n_point, n_freq)), takes the FFT, centres DC
// (fftshift), zeroes the outer bins when rbw < 1, multiplies by the transfer
// matrix Hw (n_freq x 3 x n_out, DC centred), undoes the shift, takes the
// inverse FFT and keeps the real part of the first n_point samples. When
// N > n_freq the response is first expanded in the time domain: ifft of the
// unshifted Hw, its first n_freq/2 samples kept at the start and samples
// n_freq/2+2.. (1-based) at the end of an N-long zero record, then fft
// (apply_girf.jl:72-90, including the sample it leaves out).
//
// Here everything stays in unshifted FFT order: the shift only decides which
// bins the rbw band limit clears (shifted index k = (m + N/2) mod N of bin m),
// and the response is unshifted once when it is built. The FFTs are hipFFT
// (C2C/Z2Z, batched over channels and waveforms); padding, response
// preparation, the 3 -> n_out transfer product with the band limit, and the
// real-part extraction are the kernels below.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "hoe_internal.h"

namespace hoe {
namespace {

template <typename R>
struct ctype;
template <>
struct ctype<float> {
    typedef float2 t;
    static constexpr hipfftType kind = HIPFFT_C2C;
};
template <>
struct ctype<double> {
    typedef double2 t;
    static constexpr hipfftType kind = HIPFFT_Z2Z;
};

hipfftResult fft_exec(hipfftHandle p, float2 *buf, int dir) {
    return hipfftExecC2C(p, reinterpret_cast<hipfftComplex *>(buf), reinterpret_cast<hipfftComplex *>(buf), dir);
}
hipfftResult fft_exec(hipfftHandle p, double2 *buf, int dir) {
    return hipfftExecZ2Z(p, reinterpret_cast<hipfftDoubleComplex *>(buf),
                         reinterpret_cast<hipfftDoubleComplex *>(buf), dir);
}

// buf[(w*3 + i)*N + m] = (g[(w*n_point + m)*3 + i], 0) for m < n_point, else 0
template <typename R, typename C>
__global__ void k_girf_pad(const R *__restrict__ g, int64_t n_point, int64_t N, int64_t total, C *__restrict__ buf) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t m = idx % N, wi = idx / N;
    const int64_t i = wi % 3, w = wi / 3;
    C v;
    v.x = m < n_point ? g[(w * n_point + m) * 3 + i] : (R)0;
    v.y = (R)0;
    buf[idx] = v;
}

// unshifted response column: dst[c*n + m] = src[c*n + (m + n/2) % n]   (ifftshift)
template <typename C>
__global__ void k_girf_unshift(const C *__restrict__ src, int64_t n, int64_t total, C *__restrict__ dst) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t m = idx % n, c = idx / n;
    dst[idx] = src[c * n + (m + n / 2) % n];
}

// time-domain expansion of each response column from n_freq to N samples;
// src holds the unnormalised inverse FFT (scaled here by 1/n_freq)
template <typename R, typename C>
__global__ void k_girf_expand(const C *__restrict__ src, int64_t n_freq, int64_t N, int64_t total,
                              C *__restrict__ dst) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t j = idx % N, c = idx / N;
    const int64_t half = n_freq / 2;
    const int64_t tail0 = N - (n_freq - half) + 1;   // first destination of the tail block
    int64_t s = -1;
    if (j < half) s = j;
    else if (j >= tail0) s = half + 1 + (j - tail0);
    C v;
    v.x = v.y = (R)0;
    if (s >= 0 && s < n_freq) {
        const C a = src[c * n_freq + s];
        const R sc = (R)1 / (R)n_freq;
        v.x = a.x * sc;
        v.y = a.y * sc;
    }
    dst[idx] = v;
}

// out[(w*n_out + o)*N + m] = keep(m) * sum_i hw[(o*3 + i)*N + m] * gf[(w*3 + i)*N + m]
template <typename R, typename C>
__global__ void k_girf_apply(const C *__restrict__ hw, const C *__restrict__ gf, int64_t N, int n_out, int64_t nz,
                             int64_t total, C *__restrict__ out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t m = idx % N, wo = idx / N;
    const int64_t o = wo % n_out, w = wo / n_out;
    const int64_t k = (m + N / 2) % N;   // position of bin m after fftshift
    C acc;
    acc.x = acc.y = (R)0;
    if (!(k < nz || k >= N - nz)) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const C h = hw[(o * 3 + i) * N + m];
            const C g = gf[(w * 3 + i) * N + m];
            acc.x += h.x * g.x - h.y * g.y;
            acc.y += h.x * g.y + h.y * g.x;
        }
    }
    out[idx] = acc;
}

// y[(w*n_point + t)*n_out + o] = Re(buf[(w*n_out + o)*N + t]) / N
template <typename R, typename C>
__global__ void k_girf_extract(const C *__restrict__ buf, int64_t n_point, int64_t N, int n_out, int64_t total,
                               R *__restrict__ y) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int64_t o = idx % n_out, wt = idx / n_out;
    const int64_t t = wt % n_point, w = wt / n_point;
    y[idx] = buf[(w * n_out + o) * N + t].x / (R)N;
}

unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// Plans are cached per (device, precision, length, batch) for the life of the
// process: planning dominates a call otherwise. The cache holds at most
// kPlanCache plans (the oldest is destroyed first); calls are serialised by
// g_plan_mutex, so one plan is never used by two streams at once.
struct PlanKey {
    int device, kind;
    int64_t n, batch;
    bool operator==(const PlanKey &o) const {
        return device == o.device && kind == o.kind && n == o.n && batch == o.batch;
    }
};
struct CachedPlan {
    PlanKey key;
    hipfftHandle h;
};
constexpr size_t kPlanCache = 16;
std::vector<CachedPlan> g_plans;
std::mutex g_plan_mutex;

struct Plan {
    hipfftHandle h = 0;
};

template <typename R>
hipfftResult make_plan(Plan &p, int64_t n, int64_t batch, hipStream_t s) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    const PlanKey key{dev, (int)ctype<R>::kind, n, batch};
    for (const CachedPlan &c : g_plans)
        if (c.key == key) {
            p.h = c.h;
            return hipfftSetStream(p.h, s);
        }
    int nn = (int)n;
    hipfftResult r = hipfftPlanMany(&p.h, 1, &nn, nullptr, 1, nn, nullptr, 1, nn, ctype<R>::kind, (int)batch);
    if (r != HIPFFT_SUCCESS) return r;
    if (g_plans.size() == kPlanCache) {
        (void)hipfftDestroy(g_plans.front().h);
        g_plans.erase(g_plans.begin());
    }
    g_plans.push_back({key, p.h});
    return hipfftSetStream(p.h, s);
}

struct DevMem {
    void *p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int64_t girf_fft_length(int64_t n) {
    for (int64_t m = n < 1 ? 1 : n;; ++m) {
        int64_t r = m;
        for (int64_t f : {2, 3, 5})
            while (r % f == 0) r /= f;
        if (r == 1) return m;
    }
}

template <typename R>
int girf_run(const R *g, int64_t n_point, int64_t n_other, const R *hw, int32_t n_freq, int32_t n_out, double rbw,
             int32_t device, R *out, int flags, std::string &err) {
    typedef typename ctype<R>::t C;
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    const int64_t N = girf_fft_length(n_point > n_freq ? n_point : n_freq);
    if (N > INT32_MAX) {
        err = "FFT length too large";
        return HOE_EINVAL;
    }
    int64_t nz = 0;
    if (rbw < 1.0) nz = (int64_t)std::nearbyint((double)N * (1.0 - rbw) / 2.0);
    if (nz < 0) nz = 0;
    int prev = 0;
    (void)hipGetDevice(&prev);
    struct Restore {
        int d;
        ~Restore() { (void)hipSetDevice(d); }
    } restore{prev};
    auto fail_hip = [&](hipError_t e, const char *what) {
        err = std::string("HIP error: ") + hipGetErrorString(e) + " at " + what;
        return e == hipErrorOutOfMemory ? HOE_ENOMEM : HOE_EHIP;
    };
    auto fail_fft = [&](hipfftResult r, const char *what) {
        err = std::string("hipFFT error ") + std::to_string((int)r) + " at " + what;
        return r == HIPFFT_ALLOC_FAILED ? HOE_ENOMEM : HOE_EHIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return fail_hip(e, "hipSetDevice");
    hipStream_t s = nullptr;
    if ((e = hipStreamCreate(&s)) != hipSuccess) return fail_hip(e, "hipStreamCreate");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sguard{s};
    const bool dev = (flags & HOE_DEVICE_PTRS) != 0;
    const size_t g_bytes = sizeof(R) * 3 * n_point * n_other;
    const size_t hw_bytes = sizeof(C) * (size_t)n_freq * 3 * n_out;
    const size_t y_bytes = sizeof(R) * (size_t)n_out * n_point * n_other;
    const size_t gf_elems = (size_t)3 * N * n_other, of_elems = (size_t)n_out * N * n_other;
    const size_t hwu_elems = (size_t)3 * n_out * N;
    const size_t tmp_elems = (size_t)3 * n_out * n_freq;
    DevMem m_g, m_hw, m_y, m_gf, m_of, m_hwu, m_tmp;
    if (!dev) {
        if ((e = hipMalloc(&m_g.p, g_bytes)) != hipSuccess) return fail_hip(e, "hipMalloc");
        if ((e = hipMalloc(&m_hw.p, hw_bytes)) != hipSuccess) return fail_hip(e, "hipMalloc");
        if ((e = hipMalloc(&m_y.p, y_bytes)) != hipSuccess) return fail_hip(e, "hipMalloc");
        if ((e = hipMemcpyAsync(m_g.p, g, g_bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
            return fail_hip(e, "copy of the gradients");
        if ((e = hipMemcpyAsync(m_hw.p, hw, hw_bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
            return fail_hip(e, "copy of the response");
    }
    const R *dg = dev ? g : static_cast<const R *>(m_g.p);
    const C *dhw = dev ? reinterpret_cast<const C *>(hw) : static_cast<const C *>(m_hw.p);
    R *dy = dev ? out : static_cast<R *>(m_y.p);
    if ((e = hipMalloc(&m_gf.p, sizeof(C) * gf_elems)) != hipSuccess) return fail_hip(e, "hipMalloc");
    if ((e = hipMalloc(&m_of.p, sizeof(C) * of_elems)) != hipSuccess) return fail_hip(e, "hipMalloc");
    if ((e = hipMalloc(&m_hwu.p, sizeof(C) * hwu_elems)) != hipSuccess) return fail_hip(e, "hipMalloc");
    C *gf = static_cast<C *>(m_gf.p), *of = static_cast<C *>(m_of.p), *hwu = static_cast<C *>(m_hwu.p);
    hipfftResult r;

    // ---- response in unshifted order, N samples per column ----
    if (N == n_freq) {
        k_girf_unshift<C><<<grid_of((int64_t)hwu_elems), 256, 0, s>>>(dhw, N, (int64_t)hwu_elems, hwu);
    } else {
        if ((e = hipMalloc(&m_tmp.p, sizeof(C) * tmp_elems)) != hipSuccess) return fail_hip(e, "hipMalloc");
        C *tmp = static_cast<C *>(m_tmp.p);
        k_girf_unshift<C><<<grid_of((int64_t)tmp_elems), 256, 0, s>>>(dhw, n_freq, (int64_t)tmp_elems, tmp);
        Plan pi;
        if ((r = make_plan<R>(pi, n_freq, 3 * n_out, s)) != HIPFFT_SUCCESS) return fail_fft(r, "plan (response)");
        if ((r = fft_exec(pi.h, tmp, HIPFFT_BACKWARD)) != HIPFFT_SUCCESS) return fail_fft(r, "inverse FFT (response)");
        k_girf_expand<R, C><<<grid_of((int64_t)hwu_elems), 256, 0, s>>>(tmp, n_freq, N, (int64_t)hwu_elems, hwu);
        Plan pf;
        if ((r = make_plan<R>(pf, N, 3 * n_out, s)) != HIPFFT_SUCCESS) return fail_fft(r, "plan (expanded response)");
        if ((r = fft_exec(pf.h, hwu, HIPFFT_FORWARD)) != HIPFFT_SUCCESS) return fail_fft(r, "FFT (expanded response)");
    }

    // ---- waveforms ----
    k_girf_pad<R, C><<<grid_of((int64_t)gf_elems), 256, 0, s>>>(dg, n_point, N, (int64_t)gf_elems, gf);
    Plan pg;
    if ((r = make_plan<R>(pg, N, 3 * n_other, s)) != HIPFFT_SUCCESS) return fail_fft(r, "plan (gradients)");
    if ((r = fft_exec(pg.h, gf, HIPFFT_FORWARD)) != HIPFFT_SUCCESS) return fail_fft(r, "FFT (gradients)");
    k_girf_apply<R, C><<<grid_of((int64_t)of_elems), 256, 0, s>>>(hwu, gf, N, n_out, nz, (int64_t)of_elems, of);
    Plan po;
    if ((r = make_plan<R>(po, N, (int64_t)n_out * n_other, s)) != HIPFFT_SUCCESS) return fail_fft(r, "plan (outputs)");
    if ((r = fft_exec(po.h, of, HIPFFT_BACKWARD)) != HIPFFT_SUCCESS) return fail_fft(r, "inverse FFT (outputs)");
    const int64_t ny = (int64_t)n_out * n_point * n_other;
    k_girf_extract<R, C><<<grid_of(ny), 256, 0, s>>>(of, n_point, N, n_out, ny, dy);
    if ((e = hipGetLastError()) != hipSuccess) return fail_hip(e, "kernel launch");
    if (!dev && (e = hipMemcpyAsync(out, dy, y_bytes, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return fail_hip(e, "copy of the prediction");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    return HOE_OK;
}

template int girf_run<float>(const float *, int64_t, int64_t, const float *, int32_t, int32_t, double, int32_t,
                             float *, int, std::string &);
template int girf_run<double>(const double *, int64_t, int64_t, const double *, int32_t, int32_t, double, int32_t,
                              double *, int, std::string &);

}  // namespace hoe

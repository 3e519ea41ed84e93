// gfx950 kernels of libhoe: the Explicit higher-order encoding operator.
//
// Phase model (HighOrderMRI.jl docs/src/theory/encoding-model.md:77-123,
// src/EncodingOperator/Explicit/array_operator.jl:269-271):
//   phi(s,v) = A_s[0] + sum_{i=1..L} A_s[i] * B_v[i-1]
// with the sample record A_s = [k_1(s), ..., k_L(s), t_s] and the voxel record
// B_v = [b_2(v), ..., b_L(v), fieldmap(v)] (b_1 = 1 is folded into A_s[0]).
// phi is accumulated in FP64 (every product of the FP32 inputs is exact in
// FP64), reduced modulo one cycle in FP64, and its sine and cosine are
// evaluated in FP32 by v_sin_f32 / v_cos_f32, which take revolutions.
//
// Work decomposition. One lane owns one "stationary" row (a sample for the
// forward, a voxel for the adjoint) and keeps its record in VGPRs; the other
// index is "streamed": its record has the same address in every lane, so it
// is fetched with scalar loads and broadcast through SGPRs. The streamed range
// is split across workgroups (split-K); partial sums go to a [split][row][coil]
// buffer and a second kernel reduces them in a fixed order (deterministic).
// Workgroup ids are remapped so that all workgroups of one split land on one
// XCD (blocks are dealt to the 8 XCDs round-robin), keeping that split's
// streamed data resident in that XCD's L2.
#include "hoe_internal.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace hoe {

// Resident waves per SIMD requested for the one-group MFMA kernels, forward
// and adjoint (register budget 512/N per lane).
#ifndef HOE_WAVES_FWD
#define HOE_WAVES_FWD 3
#endif
#ifndef HOE_WAVES_ADJ
#define HOE_WAVES_ADJ 3
#endif
// MFMA adjoint with 16 record terms: B fragments staged in LDS by the workgroup
// (0: each wave loads its own, as at the other record lengths).
#ifndef HOE_ADJ_LDS
#define HOE_ADJ_LDS 1
#endif
// Longest segment of the MFMA adjoint, in K units of 8 samples. Its FP32 accumulators
// lose accuracy with chain length (measured on a 3-D workload with 384 000 random
// samples: 2.8e-5 relative with whole K parts of 6 000 units, 1.05e-5 with 1 024,
// 7.2e-6 with 512), so longer runs are cut into segments with partial sums of their
// own. The forward's error on the same workload (whole K parts of 6 912 units) is
// 9.6e-6 for a random image, so its segments stay uncut.
#ifndef HOE_MFMA_ADJ_KMAX
#define HOE_MFMA_ADJ_KMAX 1024
#endif
// Largest partial-sum buffer of one contraction, in bytes. A schedule whose segments
// would need more runs as several passes over disjoint tile ranges (each scheduled over
// the whole device and reduced before the next starts), which share one buffer of the
// largest pass's size. A forward segment holds 256 x ncs complex values (64 KiB at 32
// FP32 coils); an adjoint segment holds 256 coil-combined values per coil group (2 KiB),
// so an MFMA adjoint schedule of a 96 x 96 x 48 grid at 384 000 samples, with
// 8 * ceil(nSam / 8 / 8 / 1024) segments per tile, holds about 170 MB. The environment
// variable HOE_PART_BUDGET_MIB (a positive integer, read when a backend's schedules are
// built) overrides it.
#ifndef HOE_PART_BUDGET
#define HOE_PART_BUDGET ((int64_t)512 << 20)
#endif
// K-part length (units) beyond which a schedule gives its chunks whole tiles first (see
// plan_schedule): 1 024 units of MFMA fragments are 4 MiB, one XCD's L2. On a 3-D
// workload (96 x 96 x 48, 384 000 samples) its 1 728 adjoint tiles split into 18 whole
// tiles per chunk, and the adjoint fetched 20 GB past L2 per launch; cut into 133-tile
// passes whose chunks start at scattered K offsets, 209 GB.
#ifndef HOE_DP_MIN_KX
#define HOE_DP_MIN_KX 1024
#endif

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f16x_t __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Stream-K work schedule (built on the host, see build_schedule): the
// flattened (row tile, K step) space of one contraction is cut into equal
// chunks, one per resident workgroup slot, so every workgroup does the same
// number of K steps and the whole grid runs in a single wave of residency.
// A chunk covers one or more segments (tile, k0, k1); every segment writes
// its own partial-sum block, reduced later in segment order (deterministic).
// The K range is first cut into 8 parts and chunk c of part x is launched as
// block 8*c + x: blocks are dealt round-robin to the 8 XCDs, so each XCD
// streams one eighth of the K data and keeps it in its own L2.
__device__ __forceinline__ int chunk_id(int cpx) { return (blockIdx.x & 7) * cpx + (blockIdx.x >> 3); }

// Fractional part of the phase in cycles (v_fract_f64: exact, in [0, 1)),
// then sin/cos(2*pi*r) from v_sin_f32 / v_cos_f32 (inputs in revolutions;
// measured max abs error 1.25e-7 on one period).
__device__ __forceinline__ void sincos_cycles(double ph, float &sn, float &cs) {
    const float r = (float)__builtin_amdgcn_fract(ph);
    sn = __builtin_amdgcn_sinf(r);
    cs = __builtin_amdgcn_cosf(r);
}

// The same from an FP32 phase in cycles of magnitude at most 64, taken without a reduction:
// v_sin_f32 / v_cos_f32 accept +-256 cycles and, measured on +-64, are as accurate there
// (1.2e-7) as on one cycle. Every FP32 phase is either reduced already (fract of an FP64
// phase) or an anchored one, whose size kAnchorMaxCycles and kAnchorMaxOffsetCycles bound
// (at most 41 cycles).
__device__ __forceinline__ void sincos_cycles(float ph, float &sn, float &cs) {
    sn = __builtin_amdgcn_sinf(ph);
    cs = __builtin_amdgcn_cosf(ph);
}

// exp(2 pi i ph) in FP64 (the FP64 backend): ph - rint(ph) in [-1/2, 1/2] is exact, then
// the device math library's sincospi of twice it.
__device__ __forceinline__ void sincos_cycles(double ph, double &sn, double &cs) {
    const double r = ph - rint(ph);
    sincospi(2.0 * r, &sn, &cs);
}

// Interleaved complex type of a real type (FP32 backends: float2; FP64 backend: double2).
template <typename R>
struct cplx;
template <>
struct cplx<float> {
    typedef float2 type;
};
template <>
struct cplx<double> {
    typedef double2 type;
};
template <typename R>
using cplx_t = typename cplx<R>::type;
template <typename R>
__device__ __forceinline__ cplx_t<R> mkc(R re, R im) {
    cplx_t<R> c;
    c.x = re;
    c.y = im;
    return c;
}
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Power-of-two scale that maps |values| <= maxabs below 2^12 (FP16 hi/lo
// split without overflow and without losing the low part to underflow).
__device__ __forceinline__ float pow2_scale(float maxabs) {
    if (!(maxabs > 0.0f) || !(maxabs < 3.0e38f)) return 1.0f;
    int e;
    (void)frexpf(maxabs, &e);      // maxabs < 2^e
    return ldexpf(1.0f, 12 - e);
}

// hi = T16(c), lo = T16(c - hi), packed (c, s) pairs: c == hi + lo to ~2^-21
// relative (FP16) or ~2^-16 (BF16) -- the compensated split of Markidis et
// al. (2018) and Ootomo & Yokota (2022) for matrix-core products. c - hi is
// exact in FP32. FP16 uses v_cvt_pkrtz_f16_f32 (round toward zero; the low
// part absorbs the truncation), BF16 uses v_cvt_pk_bf16_f32.
template <typename T16>
__device__ __forceinline__ void split2(float c, float s, unsigned &hi, unsigned &lo);

// c - (float)h.lo16 and c - (float)h.hi16 in one v_fma_mix_f32 each (the f16
// half is widened inside the FMA; exact because c - h is exact in FP32).
__device__ __forceinline__ float sub_f16lo(float c, unsigned h) {
    float r;
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(c));
    return r;
}
__device__ __forceinline__ float sub_f16hi(float c, unsigned h) {
    float r;
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(c));
    return r;
}

template <>
__device__ __forceinline__ void split2<_Float16>(float c, float s, unsigned &hi, unsigned &lo) {
    typedef __fp16 fp16x2_t __attribute__((ext_vector_type(2)));
    const fp16x2_t h = __builtin_amdgcn_cvt_pkrtz(c, s);
    const unsigned hb = __builtin_bit_cast(unsigned, h);
    const fp16x2_t l = __builtin_amdgcn_cvt_pkrtz(sub_f16lo(c, hb), sub_f16hi(s, hb));
    hi = hb;
    lo = __builtin_bit_cast(unsigned, l);
}

template <>
__device__ __forceinline__ void split2<__bf16>(float c, float s, unsigned &hi, unsigned &lo) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    const bf16x2_t h = {(__bf16)c, (__bf16)s};
    const unsigned hb = __builtin_bit_cast(unsigned, h);
    const float ch = __uint_as_float(hb << 16), sh = __uint_as_float(hb & 0xffff0000u);
    const bf16x2_t l = {(__bf16)(c - ch), (__bf16)(s - sh)};
    hi = hb;
    lo = __builtin_bit_cast(unsigned, l);
}

// Horner forward: voxels of one run share (y, z), so every record entry is a
// cubic in x and phi(s,v) = c0 + x (c1 + x (c2 + x c3)) + t_s * fieldmap_v with
// c_j(s) = sum_i A_s[i] * beta[run][i][j], set up once per run and lane.
// (S: stride between the record's entries, 1 in global memory, 64 in a wave's LDS copy)
template <int L, int S = 1>
__device__ __forceinline__ void run_setup(const double *__restrict__ arec, const double *__restrict__ beta_run,
                                          double c[4]) {
    c[0] = c[1] = c[2] = c[3] = 0.0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const double a = arec[i * S];
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = fma(a, beta_run[4 * i + j], c[j]);
    }
}

struct HornerArgs {
    const double2 *hrec;   // [nv_pad] (x, fieldmap)
    const int *vrun;       // [nv_pad] run of each voxel
    const double *beta;    // [n_runs][L][4]
    int t_index;           // position of t in the sample record
    const HoeUnit *units;     // [nv_pad / 8] K-unit anchors of the MFMA Horner forward
    const HoeUnitVox *uvox;   // [nv_pad / 8] voxel offsets from the anchors
    const HoeAdjAnchor *aanc; // [nv_pad / 64] group anchors of the MFMA Horner adjoint
    const float2 *aud;        // [nv_pad] voxel offsets from the adjoint group anchors
    int anchored;             // 0: the MFMA kernels take the FP64 generic phase on every step
    const double *dtrec;      // [ns_pad][sk] derivative records (MFMA derivative forward)
    double dtw;               // 2 pi sigma: psi's factor in the MFMA derivative forward's A operand
    const float *obound;      // bound of the current records' Horner offsets (cycles), set on the device
};

// Anchored phases for this launch: the fieldmap offsets allow them (host, at creation) and
// the Horner offsets of the current records stay within kAnchorMaxOffsetCycles (device, at
// each record build). One scalar load per wave.
__device__ __forceinline__ bool anchored_on(const HornerArgs &hx) {
    return hx.anchored && *hx.obound <= (float)kAnchorMaxOffsetCycles;
}

template <typename T16>
struct mfma_t;
template <>
struct mfma_t<_Float16> {
    typedef h8_t vec;
    static __device__ __forceinline__ f16x_t run(vec a, vec b, f16x_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct mfma_t<__bf16> {
    typedef b8_t vec;
    static __device__ __forceinline__ f16x_t run(vec a, vec b, f16x_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// ---------------------------------------------------------------------------
// direct FP32 kernels (VALU contraction)
// ---------------------------------------------------------------------------
//
// FWD:  part[seg][r][c] = sum_{v in seg} e(s,v) * z[v][c]          (s = 256*tile + r)
// !FWD: part[seg][b][r] = sum_{c in batch b} conj(csm[v][c]) sum_{s in seg} conj(e(s,v)) * y[s][c]
//                                                                  (v = 256*tile + r)
// Each lane keeps CB complex accumulators; the streamed record and the CB
// complex operands of the streamed row arrive through scalar loads. A K unit
// is 8 streamed rows; rows at or beyond n_stream (padding) are skipped.
//
// DT (forward only): the trajectory-derivative product B x of
// array_operator.jl:187-235, e(s,v) replaced by 2 pi i psi(s,v) e(s,v) with
// psi = sum_l kdt_l(s) b_l(v) from the derivative records dtrec (same layout
// as coef, t slot 0), evaluated like the phase.
template <bool FWD, bool HX, int L, int CB, bool DT = false, typename R = float>
__global__ __launch_bounds__(256) void k_direct(
    const double *__restrict__ coef, int sk, const double *__restrict__ basis, int sb,
    const cplx_t<R> *__restrict__ src, cplx_t<R> *__restrict__ part, int ncs, int n_stream,
    const int4 *__restrict__ segs, const int *__restrict__ chunk_off, int cpx, HornerArgs hx,
    const double *__restrict__ dtrec, const cplx_t<R> *__restrict__ csm) {
    static_assert(FWD || !DT, "the derivative product is a forward product");
    static_assert(!DT || std::is_same<R, float>::value, "the derivative product runs in FP32");
    const int chunk = chunk_id(cpx);
    const int cb0 = blockIdx.y * CB;
    const int sg0 = chunk_off[chunk], sg1 = chunk_off[chunk + 1];
    constexpr int NST = HX ? 1 : (FWD ? L + 1 : L);
    constexpr int NSD = DT ? (HX ? 1 : L + 1) : 1;
    for (int sgi = sg0; sgi < sg1; ++sgi) {
        const int4 d = segs[sgi];
        const int row = d.x * 256 + threadIdx.x;
        double st[NST], sd[NSD];
        if (HX) {
            st[0] = coef[(size_t)row * sk + hx.t_index];
        } else {
            const double *p = FWD ? coef + (size_t)row * sk : basis + (size_t)row * sb;
#pragma unroll
            for (int i = 0; i < NST; ++i) st[i] = p[i];
            if (DT) {
#pragma unroll
                for (int i = 0; i < NSD; ++i) sd[i] = dtrec[(size_t)row * sk + i];
            }
        }
        R ar[CB], ai[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) ar[c] = ai[c] = R(0);
        const int q0 = d.y * 8;
        const int q1 = min(d.z * 8, n_stream);
        double hc[4] = {0.0, 0.0, 0.0, 0.0}, hd[4] = {0.0, 0.0, 0.0, 0.0};
        int cur = -1;
        for (int q = q0; q < q1; ++q) {
            const double *sm = FWD ? basis + (size_t)q * sb : coef + (size_t)q * sk;
            double ph, psi = 0.0;
            if (HX) {
                const int r = hx.vrun[q];
                if (r != cur) {
                    run_setup<L>(coef + (size_t)row * sk, hx.beta + (size_t)r * L * 4, hc);
                    if (DT) run_setup<L>(dtrec + (size_t)row * sk, hx.beta + (size_t)r * L * 4, hd);
                    cur = r;
                }
                const double2 xd = hx.hrec[q];
                ph = fma(fma(fma(hc[3], xd.x, hc[2]), xd.x, hc[1]), xd.x, hc[0]);
                ph = fma(st[0], xd.y, ph);
                if (DT) psi = fma(fma(fma(hd[3], xd.x, hd[2]), xd.x, hd[1]), xd.x, hd[0]);
            } else if (FWD) {
                ph = st[0];
#pragma unroll
                for (int i = 1; i <= L; ++i) ph = fma(st[i], sm[i - 1], ph);
                if (DT) {
                    psi = sd[0];
#pragma unroll
                    for (int i = 1; i <= L; ++i) psi = fma(sd[i], sm[i - 1], psi);
                }
            } else {
                ph = sm[L] * st[L - 1];
#pragma unroll
                for (int i = 1; i < L; ++i) ph = fma(sm[i], st[i - 1], ph);
                ph += sm[0];
            }
            R sn, cs;
            sincos_cycles(ph, sn, cs);
            if (!FWD) sn = -sn;   // conj(e)
            if constexpr (DT) {
                // 2 pi i psi e = 2 pi psi (-sin + i cos)
                const float w = (float)(6.283185307179586 * psi);
                const float c2 = -sn * w;
                sn = cs * w;
                cs = c2;
            }
            const cplx_t<R> *w = src + (size_t)q * ncs + cb0;
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const cplx_t<R> u = w[c];
                ar[c] = fmaT(cs, u.x, ar[c]);
                ar[c] = fmaT(-sn, u.y, ar[c]);
                ai[c] = fmaT(cs, u.y, ai[c]);
                ai[c] = fmaT(sn, u.x, ai[c]);
            }
        }
        if constexpr (FWD) {
            cplx_t<R> *o = part + ((size_t)sgi * 256 + threadIdx.x) * ncs + cb0;
#pragma unroll
            for (int c = 0; c < CB; ++c) o[c] = mkc<R>(ar[c], ai[c]);
        } else {
            // coil combination of the batch, sum_c conj(csm[v][c]) * acc[c] (padding coils
            // have zero maps and zero data): one complex per voxel, segment and batch
            const cplx_t<R> *s = csm + (size_t)row * ncs + cb0;
            R xr = R(0), xi = R(0);
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const cplx_t<R> m = s[c];
                xr = fmaT(m.x, ar[c], fmaT(m.y, ai[c], xr));
                xi = fmaT(m.x, ai[c], fmaT(-m.y, ar[c], xi));
            }
            part[((size_t)sgi * gridDim.y + blockIdx.y) * 256 + threadIdx.x] = mkc<R>(xr, xi);
        }
    }
}

// ---------------------------------------------------------------------------
// cross-lane reduce-scatter of the MFMA adjoint's coil columns
// ---------------------------------------------------------------------------
//
// x[k], k = 0..15, in every lane of a 16-lane row; afterwards lane l holds the sum of
// x[l & 15] over the 16 lanes of its row. Each level pairs lane l with l ^ m (m = 15, 7,
// 2, 1: DPP row_mirror, row_half_mirror, quad_perm [2,3,0,1], quad_perm [1,0,3,2]);
// the lane whose bit b (b = 3, 2, 1, 0) is set keeps the upper half of its values and
// sends the lower half, its partner the reverse. The order of the additions is fixed.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int N, int CTRL>
__device__ __forceinline__ void rs_level(float (&x)[16], bool up) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const float keep = up ? x[k + N] : x[k];
        const float send = up ? x[k] : x[k + N];
        x[k] = keep + dpp_f<CTRL>(send);
    }
}
__device__ __forceinline__ float reduce_scatter16(float (&x)[16], int lane) {
    rs_level<8, 0x140>(x, (lane & 8) != 0);
    rs_level<4, 0x141>(x, (lane & 4) != 0);
    rs_level<2, 0x4e>(x, (lane & 2) != 0);
    rs_level<1, 0xb1>(x, (lane & 1) != 0);
    return x[0];
}

// ---------------------------------------------------------------------------
// mixed-precision phase-reuse kernels (MFMA contraction)
// ---------------------------------------------------------------------------
//
// A wave owns 64 stationary rows = two 32-row MFMA tiles. Per K step it
// evaluates exp(i*phi) for 8 streamed indices in all 64 lanes (lane = row),
// splits cos/sin into T16 hi+lo, and uses v_permlane32_swap to move the
// values into the v_mfma_f32_32x32x16 A-operand layout (lane l holds row
// l%32 and K entries 8*(l/32)..+7 = 4 complex (cos,sin) pairs). The B operand
// (coil data, precomputed per application in exactly the B-register layout,
// FP16/BF16 hi+lo, scaled by a power of two) is read with 16-byte loads.
// Three products per K step: Ahi*Bhi + Alo*Bhi + Ahi*Blo, FP32 accumulate.
// Output columns (PK = 0): tile 0 = real part, tile 1 = imaginary part of 32 coils.
// Up to 16 coils share one B tile (PK = 1): columns 0..15 the real part, 16..31 the
// imaginary part of coils 0..15, so each product yields both (6 MFMAs per K step
// instead of 12, two B fragments instead of four). Up to 8 coils (PK = 2): columns
// [re hi | im hi | re lo | im lo] of coils 0..7, so Ahi*B + Alo*B holds all three
// products (and the lo*lo term) in 4 MFMAs and one B fragment; column c + 16 is added
// to column c in the epilogue.
template <typename T16, bool FWD, bool HX, int L, int NB, int PK = 0, bool DT = false, bool PF = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NB == 1 ? (FWD ? HOE_WAVES_FWD : HOE_WAVES_ADJ) : 2))) void k_mixed(
    const double *__restrict__ coef, int sk, const double *__restrict__ basis, int sb,
    const uint4 *__restrict__ frag, int n_ks, float2 *__restrict__ part, int ncs,
    const int4 *__restrict__ segs, const int *__restrict__ chunk_off, int cpx, HornerArgs hx,
    const float2 *__restrict__ csm) {
    typedef typename mfma_t<T16>::vec vec_t;
    static_assert(PK == 0 || NB == 1, "packed coil layouts hold one group");
    static_assert(FWD || !DT, "the derivative product is a forward product");
    static_assert(!PF || (FWD && NB == 1 && PK == 0), "fragment prefetch: forward, one 32-coil group");
    constexpr int NQ = PK == 0 ? 4 : (PK == 1 ? 2 : 1);   // B fragments (uint4) per lane and K step
    constexpr int NA = PK == 0 ? 2 : 1;                   // accumulator tiles per row tile
    const int chunk = chunk_id(cpx);
    const int grp = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sg0 = chunk_off[chunk], sg1 = chunk_off[chunk + 1];
    constexpr int NST = FWD ? L + 1 : L;
    // Adjoint with one coil group and 16 or 20 record terms: B fragments staged in LDS, S K
    // steps per chunk. Staging removes the L1 misses of the fragment loads (the 4 waves
    // read the same fragments) but its barrier drains the Horner setup's prefetch each
    // chunk; measured, that takes 5 % off a 16-term 3-D adjoint and 1.2 % off a 20-term
    // 2-D one, and adds 4.5 % to a 10-term 2-D one, so the shorter records keep per-wave
    // loads. (At 20 terms the workgroup's 52.5 KiB of LDS still admits 3 per CU.)
    constexpr bool STAGE = !FWD && NB == 1 && HOE_ADJ_LDS && (L == 16 || L == 20);
    constexpr int S = 4;
    __shared__ uint4 fbuf[STAGE ? 2 : 1][STAGE ? S * 256 : 1];
    for (int sgi = sg0; sgi < sg1; ++sgi) {
        const int4 d = segs[sgi];
        const int row = d.x * 256 + wave * 64 + lane;
        f16x_t acc[2][NB][NA];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int i = 0; i < NA; ++i) acc[a][b][i] = f16x_t{};
        // K loop over this segment; phases_of(ks, ph, ps) fills 8 phases per lane:
        // with swap, those of the lane's own row for the 8 streamed rows of K
        // step ks (moved into the A layout by permlane32 swaps); without, the
        // A layout directly (rows lane%32 and 32 + lane%32, streamed rows
        // 4*(lane/32)..+3 each). DT: ps receives 2 pi sigma psi of the same pairs,
        // and the A operand is 2 pi sigma psi i e = 2 pi sigma psi (-sin + i cos)
        // (|2 pi sigma psi| <= 1 by the host's power-of-two sigma, undone in the
        // forward reduction).
        auto kloop = [&](auto phases_of, auto swap, auto ptype) {
            // one K step: phases of the 8 streamed rows (in cycles: FP64, or FP32 from the
            // anchored Horner forward), exp(i phi) split hi/lo into the A layout, 12 MFMAs
            // per coil group against the B fragments bfr
            auto step = [&](int ks, const uint4 (&bfr)[NB][NQ]) {
                decltype(ptype) ph[8];
                float ps[8];
                phases_of(ks, ph, ps);
                unsigned hp[8], lp[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    float sn, cs;
                    sincos_cycles(ph[j], sn, cs);
                    if constexpr (DT) {
                        const float c2 = -sn * ps[j];
                        sn = cs * ps[j];
                        cs = c2;
                    }
                    split2<T16>(cs, sn, hp[j], lp[j]);
                }
                if constexpr (decltype(swap)::value) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        auto a = __builtin_amdgcn_permlane32_swap(hp[j], hp[j + 4], false, false);
                        hp[j] = a[0];
                        hp[j + 4] = a[1];
                        auto b = __builtin_amdgcn_permlane32_swap(lp[j], lp[j + 4], false, false);
                        lp[j] = b[0];
                        lp[j + 4] = b[1];
                    }
                }
                vec_t ahi[2], alo[2];
                ahi[0] = __builtin_bit_cast(vec_t, make_uint4(hp[0], hp[1], hp[2], hp[3]));
                ahi[1] = __builtin_bit_cast(vec_t, make_uint4(hp[4], hp[5], hp[6], hp[7]));
                alo[0] = __builtin_bit_cast(vec_t, make_uint4(lp[0], lp[1], lp[2], lp[3]));
                alo[1] = __builtin_bit_cast(vec_t, make_uint4(lp[4], lp[5], lp[6], lp[7]));
                if constexpr (PK == 0) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const vec_t bhr = __builtin_bit_cast(vec_t, bfr[b][0]);
                        const vec_t bhi = __builtin_bit_cast(vec_t, bfr[b][1]);
                        const vec_t blr = __builtin_bit_cast(vec_t, bfr[b][2]);
                        const vec_t bli = __builtin_bit_cast(vec_t, bfr[b][3]);
#pragma unroll
                        for (int rt = 0; rt < 2; ++rt) {
                            acc[rt][b][0] = mfma_t<T16>::run(ahi[rt], bhr, acc[rt][b][0]);
                            acc[rt][b][NA - 1] = mfma_t<T16>::run(ahi[rt], bhi, acc[rt][b][NA - 1]);
                            acc[rt][b][0] = mfma_t<T16>::run(alo[rt], bhr, acc[rt][b][0]);
                            acc[rt][b][NA - 1] = mfma_t<T16>::run(alo[rt], bhi, acc[rt][b][NA - 1]);
                            acc[rt][b][0] = mfma_t<T16>::run(ahi[rt], blr, acc[rt][b][0]);
                            acc[rt][b][NA - 1] = mfma_t<T16>::run(ahi[rt], bli, acc[rt][b][NA - 1]);
                        }
                    }
                } else if constexpr (PK == 1) {
                    const vec_t bh = __builtin_bit_cast(vec_t, bfr[0][0]);
                    const vec_t bl = __builtin_bit_cast(vec_t, bfr[0][NQ - 1]);
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) {
                        acc[rt][0][0] = mfma_t<T16>::run(ahi[rt], bh, acc[rt][0][0]);
                        acc[rt][0][0] = mfma_t<T16>::run(alo[rt], bh, acc[rt][0][0]);
                        acc[rt][0][0] = mfma_t<T16>::run(ahi[rt], bl, acc[rt][0][0]);
                    }
                } else {
                    const vec_t bb = __builtin_bit_cast(vec_t, bfr[0][0]);
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) {
                        acc[rt][0][0] = mfma_t<T16>::run(ahi[rt], bb, acc[rt][0][0]);
                        acc[rt][0][0] = mfma_t<T16>::run(alo[rt], bb, acc[rt][0][0]);
                    }
                }
            };
            if constexpr (STAGE) {
                // B fragments staged in LDS by the workgroup, S steps per chunk, double
                // buffered: each wave copies its share of chunk c + 1 with direct-to-LDS
                // loads (no registers) while chunk c is consumed; one barrier per chunk.
                // The copy gathers the [lane][q] fragments into a [q][lane] LDS layout,
                // so the reads are conflict-free ds_read_b128.
                const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
                auto issue = [&](int c0, int bsel) {
                    for (int i = wv; i < S * NQ; i += 4) {
                        const int kk = c0 + i / NQ, q = i % NQ;
                        if (kk < d.z)
                            __builtin_amdgcn_global_load_lds(
                                (const __attribute__((address_space(1))) void *)(
                                    frag + (((size_t)grp * n_ks + kk) * 64 + lane) * NQ + q),
                                (__attribute__((address_space(3))) void *)&fbuf[bsel][(i / NQ) * 64 * NQ + q * 64],
                                16, 0, 0);
                    }
                };
                issue(d.y, 0);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                int bsel = 0;
                for (int c0 = d.y; c0 < d.z; c0 += S) {
                    if (c0 + S < d.z) issue(c0 + S, bsel ^ 1);
                    const int c1 = min(c0 + S, d.z);
                    for (int ks = c0; ks < c1; ++ks) {
                        uint4 bfr[NB][NQ];
                        const uint4 *lb = &fbuf[bsel][(ks - c0) * 64 * NQ + lane];
#pragma unroll
                        for (int q = 0; q < NQ; ++q) bfr[0][q] = lb[q * 64];
                        step(ks, bfr);
                    }
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                    bsel ^= 1;
                }
            } else if constexpr (PF) {
                // PF (a forward whose K part overflows an XCD's L2): the next step's B
                // fragments are loaded a step ahead. Where the part fits it costs 3 % instead,
                // and as a runtime branch in one kernel it raised every instance's registers.
                uint4 bfr[NB][NQ], nxt[NB][NQ];
                auto fetch = [&](int ks, uint4 (&o)[NB][NQ]) {
                    const uint4 *p = frag + (((size_t)grp * n_ks + min(ks, n_ks - 1)) * 64 + lane) * NQ;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) o[0][q] = p[q];
                };
                fetch(d.y, nxt);
                for (int ks = d.y; ks < d.z; ++ks) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) bfr[0][q] = nxt[0][q];
                    fetch(ks + 1, nxt);
                    step(ks, bfr);
                }
            } else {
                for (int ks = d.y; ks < d.z; ++ks) {
                    uint4 bfr[NB][NQ];
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const uint4 *p = frag + (((size_t)(grp * NB + b) * n_ks + ks) * 64 + lane) * NQ;
#pragma unroll
                        for (int q = 0; q < NQ; ++q) bfr[b][q] = p[q];
                    }
                    step(ks, bfr);
                }
            }
        };
        // generic phase: the lane's full record in registers, the streamed records
        // from scalar loads, L FP64 FMAs per phase (DT: and per psi)
        auto generic_path = [&]() {
            double st[NST], sd[DT ? L + 1 : 1];
            {
                const double *p = FWD ? coef + (size_t)row * sk : basis + (size_t)row * sb;
#pragma unroll
                for (int i = 0; i < NST; ++i) st[i] = p[i];
                if constexpr (DT) {
#pragma unroll
                    for (int i = 0; i <= L; ++i) sd[i] = hx.dtrec[(size_t)row * sk + i] * hx.dtw;
                }
            }
            kloop([&](int ks, double *ph, float *ps) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int q = ks * 8 + j;
                    const double *sm = FWD ? basis + (size_t)q * sb : coef + (size_t)q * sk;
                    double p;
                    if (FWD) {
                        p = st[0];
#pragma unroll
                        for (int i = 1; i <= L; ++i) p = fma(st[i], sm[i - 1], p);
                        if constexpr (DT) {
                            double q = sd[0];
#pragma unroll
                            for (int i = 1; i <= L; ++i) q = fma(sd[i], sm[i - 1], q);
                            ps[j] = (float)q;
                        }
                    } else {
                        p = sm[L] * st[L - 1];
#pragma unroll
                        for (int i = 1; i < L; ++i) p = fma(sm[i], st[i - 1], p);
                        p += sm[0];
                    }
                    ph[j] = p;
                }
            }, std::true_type{}, double{});
        };
        if constexpr (HX && FWD) {
            // Horner forward. Runs are contiguous, so the 8 rows of a K step share one run
            // when the first and the last do: then one run check per K step, and the step's
            // phases are anchored: P(x) = c0 + x (c1 + x (c2 + x c3)) of the run (c_j set up
            // per run in FP64) and t * fieldmap are evaluated in FP64 at the unit's anchor
            // (x_a, fm_a) and reduced modulo one cycle; the offsets of the 8 voxels,
            //   P(x_a + u) - P(x_a) = u (d1 + u (d2 + u c3)),  d1 = P'(x_a), d2 = P''(x_a) / 2,
            // and t * (fieldmap - fm_a), are added in FP32 (|u| <= half the unit's x range,
            // so the FP32 terms stay within a few cycles): 4 FP32 FMAs per phase instead of
            // 4 FP64 FMAs, a fract and a conversion.
            const double t = coef[(size_t)row * sk + hx.t_index];
            const float tf = (float)t;
            // the lane's record entries 0..L, copied once per segment into the wave's LDS
            // ([entry][lane]): the run setups and the run-crossing steps read it there. Read
            // from global memory at every run change, the records of all the XCD's tiles
            // (2 MB at 16 000 samples and L = 10, 3.9 MB at L = 20) competed with the K part's
            // fragments for its L2.
            __shared__ double rlds[HX && FWD ? 4 * (L + 1) * 64 : 1];
            double *myr = rlds + (threadIdx.x >> 6) * (L + 1) * 64 + lane;
            {
                const double *ap = coef + (size_t)row * sk;
#pragma unroll
                for (int i = 0; i <= L; ++i) myr[i * 64] = ap[i];
            }
            double hc[4] = {0.0, 0.0, 0.0, 0.0};
            float c1f = 0.0f, c2f = 0.0f, c2x2f = 0.0f, c3f = 0.0f, c3x3f = 0.0f;
            // DT: psi of the run, a cubic in x like P (scaled by 2 pi sigma), evaluated in FP32
            // as a Taylor expansion at the unit's anchor (no reduction: psi is no phase)
            float p0f = 0.0f, p1f = 0.0f, p2f = 0.0f, p2x2f = 0.0f, p3f = 0.0f, p3x3f = 0.0f;
            int cur = -1;
            // the anchor of the next K unit is fetched during this step, so that the step's
            // branch does not wait on a scalar load before the unit's voxel offsets are fetched
            HoeUnit an = hx.units[d.y];
            const bool anch = anchored_on(hx);
            kloop([&](int ks, float *ph, float *ps) {
                const int q0 = ks * 8;
                const HoeUnit a = an;
                const int r0 = a.run;
                if (r0 >= 0 && anch) {
                    const HoeUnitVox &uv = hx.uvox[ks];
                    if (r0 != cur) {
                        run_setup<L, 64>(myr, hx.beta + (size_t)r0 * L * 4, hc);
                        c1f = (float)hc[1];
                        c2f = (float)hc[2];
                        c2x2f = (float)(2.0 * hc[2]);
                        c3f = (float)hc[3];
                        c3x3f = (float)(3.0 * hc[3]);
                        if constexpr (DT) {
                            double hd[4];
                            run_setup<L>(hx.dtrec + (size_t)row * sk, hx.beta + (size_t)r0 * L * 4, hd);
                            p0f = (float)(hx.dtw * hd[0]);
                            p1f = (float)(hx.dtw * hd[1]);
                            p2f = (float)(hx.dtw * hd[2]);
                            p2x2f = (float)(2.0 * hx.dtw * hd[2]);
                            p3f = (float)(hx.dtw * hd[3]);
                            p3x3f = (float)(3.0 * hx.dtw * hd[3]);
                        }
                        cur = r0;
                    }
                    const double xa = (double)a.xa;
                    const double pa = fma(t, (double)a.fma, fma(fma(fma(hc[3], xa, hc[2]), xa, hc[1]), xa, hc[0]));
                    const float base = (float)__builtin_amdgcn_fract(pa);
                    const float d2 = fmaf(c3x3f, a.xa, c2f);
                    const float d1 = fmaf(fmaf(c3x3f, a.xa, c2x2f), a.xa, c1f);
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float u = uv.ud[j][0], dfm = uv.ud[j][1];
                        ph[j] = fmaf(tf, dfm, fmaf(fmaf(fmaf(c3f, u, d2), u, d1), u, base));
                    }
                    if constexpr (DT) {
                        const float qa = fmaf(fmaf(fmaf(p3f, a.xa, p2f), a.xa, p1f), a.xa, p0f);
                        const float q2 = fmaf(p3x3f, a.xa, p2f);
                        const float q1 = fmaf(fmaf(p3x3f, a.xa, p2x2f), a.xa, p1f);
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const float u = uv.ud[j][0];
                            ps[j] = fmaf(fmaf(fmaf(p3f, u, q2), u, q1), u, qa);
                        }
                    }
                } else {
                    // a run boundary inside the K step (or anchored phases off for this
                    // operator): generic phase from the full records
                    double a[L + 1];
#pragma unroll
                    for (int i = 0; i <= L; ++i) a[i] = myr[i * 64];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const double *sm = basis + (size_t)(q0 + j) * sb;
                        double p = a[0];
#pragma unroll
                        for (int i = 1; i <= L; ++i) p = fma(a[i], sm[i - 1], p);
                        ph[j] = (float)__builtin_amdgcn_fract(p);
                    }
                    if constexpr (DT) {
                        const double *dp = hx.dtrec + (size_t)row * sk;
                        double b[L + 1];
#pragma unroll
                        for (int i = 0; i <= L; ++i) b[i] = dp[i];
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const double *sm = basis + (size_t)(q0 + j) * sb;
                            double q = b[0];
#pragma unroll
                            for (int i = 1; i <= L; ++i) q = fma(b[i], sm[i - 1], q);
                            ps[j] = (float)(hx.dtw * q);
                        }
                    }
                }
                an = hx.units[min(ks + 1, n_ks - 1)];
            }, std::true_type{}, float{});
        } else if constexpr (HX && !FWD) {
            // A wave whose 64 voxels span three or more runs (short rows of a mask), or any
            // wave when the operator's fieldmap offsets are too large for anchored phases,
            // takes the generic phase; the decision is uniform over the wave.
            const int wv0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int base0 = d.x * 256 + wv0 * 64;
            if (hx.vrun[base0 + 63] - hx.vrun[base0] > 1 || !anchored_on(hx)) {
                generic_path();
            } else {
            // Horner adjoint. The wave's 64 voxels lie in at most two runs (r_lo, r_hi).
            // Its voxel rows fall into up to four groups c = 2 g + h: run r_lo + g, wave half
            // h (rows 0..31 or 32..63, the lane's two voxel rows). Each group has an anchor
            // (x_c, fm_c) and its voxels carry u = x - x_c, dfm = fieldmap - fm_c (FP32).
            // For each pair of K steps, FP64 matrix products give every coefficient the
            // pair needs: rows = its 16 streamed samples, columns = the Taylor coefficients
            // of the sample's run cubic at x_c, e_0 = P(x_c) + t_s fm_c, e_k = P^(k)(x_c) / k!
            // (k = 1..3), one product chain per group, from the run's x-polynomial table
            // re-centred at x_c (and fm_c in the t row). e_0 is reduced modulo one cycle in
            // FP64; the coefficients and the samples' t_s go to a wave-private LDS slot in
            // FP32, and each lane spends 4 FP32 FMAs per phase,
            //   phi = e_0 + u (e_1 + u (e_2 + u e_3)) + t_s dfm,
            // on its two voxel rows, already in the A-operand layout (|u| is at most half a
            // group's x range, so the FP32 terms stay within a few cycles).
            // The products are v_mfma_f64_4x4x4f64 (4 blocks of 4 samples x 4 coefficients):
            // 10-17 cycles each on gfx950 against 70-80 for v_mfma_f64_16x16x4f64, whose
            // 16 columns this product would fill only half. They run over k blocks of 4
            // record terms covering entries 0..L, so t (record entry L at most) is always
            // among them.
            constexpr int NM = L / 4 + 1;
            constexpr int SLOT = 272;             // floats: e [2 steps][4 groups][2 halves][4][4], t [16]
            __shared__ float hl[4][2 * SLOT];     // per wave, 2 slots
            __shared__ double hb[4][4][NM][16];   // per wave: B operands per group and k block (constant per segment)
            const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            float *cl = hl[wv];
            const int base = d.x * 256 + wv * 64;
            const int r_lo = hx.vrun[base], r_hi = hx.vrun[base + 63];
            const bool two = r_hi != r_lo;
            const HoeAdjAnchor &an = hx.aanc[base / 64];
            const int v0 = base + (lane & 31), v1 = v0 + 32;
            const float2 o0 = hx.aud[v0], o1 = hx.aud[v1];
            const int c0 = 2 * (hx.vrun[v0] - r_lo), c1 = 2 * (hx.vrun[v1] - r_lo) + 1;
            // a wave across two runs leaves one of its four groups empty (two when the run
            // boundary is the wave's middle): (r_lo, upper half) or (r_hi, lower half)
            const bool has1 = __any(c1 == 1), has2 = __any(c0 == 2);
            const int half = lane >> 5;
            const int kq = lane >> 4;
            // B entry (k, j) of group c and k block m: Taylor coefficient j at x_c of record
            // term 4 m + k's x-polynomial (run r_lo + c / 2), plus fm_c at (t, 0); the same for
            // the 4 blocks of a product
            for (int i = lane; i < 4 * NM * 16; i += 64) {
                const int c = i / (NM * 16), m = (i / 16) % NM, e = i & 15;
                const int term = 4 * m + (e >> 2), j = e & 3;
                const int rr = min(r_lo + (c >> 1), r_hi);
                const double x = (double)an.xa[c];
                double v = 0.0;
                if (term < L) {
                    const double *b = hx.beta + ((size_t)rr * L + term) * 4;
                    v = j == 0 ? fma(fma(fma(b[3], x, b[2]), x, b[1]), x, b[0])
                               : (j == 1 ? fma(fma(3.0 * b[3], x, 2.0 * b[2]), x, b[1])
                                         : (j == 2 ? fma(3.0 * b[3], x, b[2]) : b[3]));
                }
                if (term == hx.t_index && j == 0) v += (double)an.fm[c];
                hb[wv][c][m][e] = v;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // A operand of the pair starting at step kp: lane l holds record entry
            // l/16 + 4 m of row l%16; rows 0..7 = samples of kp, rows 8..15 = of kp + 1;
            // zero where the step is outside the segment (the address is then a valid row,
            // so the load is in bounds whatever the compiler does; entries up to 4 NM - 1
            // lie inside the sk >= L + 1 rounded up to 8 doubles of a record)
            auto load_a = [&](int kp, double *a) {
                const int r = lane & 15;
                const bool valid = kp < d.z && (r < 8 || kp + 1 < d.z);
                const double *ap = coef + ((size_t)(valid ? kp * 8 + r : d.y * 8)) * sk;
#pragma unroll
                for (int m = 0; m < NM; ++m) a[m] = valid ? ap[kq + 4 * m] : 0.0;
            };
            // setup products of one pair into an LDS slot. v_mfma_f64_4x4x4f64: A lane l =
            // row l%16, k l/16; B lane l = k l/16, column l%16 (block (l%16)/4, coefficient
            // l%4); D lane l = row R = 4 ((l%16)/4) + l/16, coefficient l%4. Row R is sample
            // R%4 of half (R/4)%2 of step R/8 of the pair; t of row R goes to slot[256 + R].
            const int hbe = (kq << 2) | (lane & 3);
            const int sidx = ((lane >> 3) & 1) * 128 + ((lane >> 2) & 1) * 16 + 4 * kq + (lane & 3);
            auto setup = [&](const double *a, float *slot) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (c == 2 && !two) break;
                    if ((c == 1 && !has1) || (c == 2 && !has2)) continue;
                    double acc = 0.0;
#pragma unroll
                    for (int m = 0; m < NM; ++m)
                        acc = __builtin_amdgcn_mfma_f64_4x4x4f64(a[m], hb[wv][c][m][hbe], acc, 0, 0, 0);
                    slot[sidx + c * 32] = (lane & 3) == 0 ? (float)__builtin_amdgcn_fract(acc) : (float)acc;
                }
                const int ti = hx.t_index;
                double tv = 0.0;
#pragma unroll
                for (int m = 0; m < NM; ++m) tv = (m == (ti >> 2)) ? a[m] : tv;
                if (kq == (ti & 3)) slot[256 + (lane & 15)] = (float)tv;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            };
            // software pipeline: the slot of pair p + 1 is filled during the second step of
            // pair p, and the sample records of pair p + 2 are fetched meanwhile
            double anx[NM];
            load_a(d.y, anx);
            if (d.y < d.z) setup(anx, cl);
            load_a(d.y + 2, anx);
            kloop([&](int ks, float *ph, float *) {
                const int rel = ks - d.y, k2 = rel & 1;
                const float *cur = cl + ((rel >> 1) & 1) * SLOT;
                if (k2 == 1 && ks + 1 < d.z) {
                    setup(anx, cl + (((rel >> 1) & 1) ^ 1) * SLOT);
                    load_a(ks + 3, anx);
                }
                const float *tp = cur + 256 + (k2 * 2 + half) * 4;
                const float4 *e0p = reinterpret_cast<const float4 *>(cur + ((k2 * 4 + c0) * 2 + half) * 16);
                const float4 *e1p = reinterpret_cast<const float4 *>(cur + ((k2 * 4 + c1) * 2 + half) * 16);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float t = tp[i];
                    const float4 e0 = e0p[i], e1 = e1p[i];
                    ph[i] = fmaf(t, o0.y, fmaf(fmaf(fmaf(e0.w, o0.x, e0.z), o0.x, e0.y), o0.x, e0.x));
                    ph[4 + i] = fmaf(t, o1.y, fmaf(fmaf(fmaf(e1.w, o1.x, e1.z), o1.x, e1.y), o1.x, e1.x));
                }
            }, std::false_type{}, float{});
            }
        } else {
            generic_path();
        }
        // D layout of v_mfma_f32_32x32x*: register r of lane l holds
        // row 8*(r/4) + 4*(l/32) + r%4, column l%32.
        const int col = lane & 31;
        const int hi32 = lane >> 5;
        if constexpr (!FWD) {
            // Adjoint: the coil combination sum_c conj(csm[v][c]) * acc[v][c] of the
            // workgroup's coils, then one complex per voxel into part[seg][grp][row].
            // Lane l of a half-wave ends up with row index k = l % 32 of its half:
            // tile rt = k / 16, register r = k % 16.
            float xr[2][16], xi[2][16];
            if constexpr (PK == 0) {
                // column = coil (grp * NB + b) * 32 + col; tile 0 real, tile 1 imaginary part
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int v = d.x * 256 + wave * 64 + rt * 32 + 8 * (r >> 2) + 4 * hi32 + (r & 3);
                        float sr = 0.0f, si = 0.0f;
#pragma unroll
                        for (int b = 0; b < NB; ++b) {
                            const float2 m = csm[(size_t)v * ncs + (grp * NB + b) * 32 + col];
                            const float ar = acc[rt][b][0][r], ai = acc[rt][b][NA - 1][r];
                            sr = fmaf(m.x, ar, fmaf(m.y, ai, sr));
                            si = fmaf(m.x, ai, fmaf(-m.y, ar, si));
                        }
                        xr[rt][r] = sr;
                        xi[rt][r] = si;
                    }
                // first level over the 32 coil lanes (l and l ^ 16) by row swaps: afterwards
                // row 0 of each half holds tile 0 and row 1 tile 1, each summed over the pair
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(xr[0][r]), __float_as_uint(xr[1][r]),
                                                              false, false);
                    auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(xi[0][r]), __float_as_uint(xi[1][r]),
                                                              false, false);
                    xr[0][r] = __uint_as_float(a[0]) + __uint_as_float(a[1]);
                    xi[0][r] = __uint_as_float(b[0]) + __uint_as_float(b[1]);
                }
            } else {
                // one FP32 component per column. A row swap of the two tiles' registers puts
                // tile 0 in row 0 and tile 1 in row 1 of each half: PK = 1 column c = (coil
                // c % 16, part c / 16), so the swap pairs each coil's real and imaginary
                // parts; PK = 2 columns c and c + 16 are the hi and lo halves of one
                // product, added by the swap, and lanes c % 16 = 0..7 hold the real parts of
                // coils 0..7, lanes 8..15 the imaginary parts (one row rotation by 8 pairs
                // them; the lanes 8..15 copies get zero weight).
                const bool row1 = (lane & 16) != 0;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(acc[0][0][0][r]),
                                                              __float_as_uint(acc[1][0][0][r]), false, false);
                    float re, im;
                    if constexpr (PK == 1) {
                        re = __uint_as_float(a[0]);
                        im = __uint_as_float(a[1]);
                    } else {
                        const float t = __uint_as_float(a[0]) + __uint_as_float(a[1]);
                        const float u = dpp_f<0x128>(t);   // row_ror:8
                        re = (lane & 8) ? u : t;
                        im = (lane & 8) ? t : u;
                    }
                    xr[0][r] = re;
                    xi[0][r] = im;
                }
                const int coil = PK == 1 ? (lane & 15) : (lane & 7);
                const bool live = PK == 1 || (lane & 8) == 0;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int v = d.x * 256 + wave * 64 + (row1 ? 32 : 0) + 8 * (r >> 2) + 4 * hi32 + (r & 3);
                    float2 m = csm[(size_t)v * ncs + coil];
                    if (!live) m = make_float2(0.0f, 0.0f);
                    const float ar = xr[0][r], ai = xi[0][r];
                    xr[0][r] = fmaf(m.x, ar, m.y * ai);
                    xi[0][r] = fmaf(m.x, ai, -m.y * ar);
                }
            }
            const float sr = reduce_scatter16(xr[0], lane);
            const float si = reduce_scatter16(xi[0], lane);
            const int k = lane & 31;
            const int rr = wave * 64 + (k >> 4) * 32 + 8 * ((k & 15) >> 2) + 4 * hi32 + (k & 3);
            part[((size_t)sgi * gridDim.y + grp) * 256 + rr] = make_float2(sr, si);
        } else if constexpr (PK == 0) {
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int coil = (grp * NB + b) * 32 + col;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int rr = wave * 64 + rt * 32 + 8 * (r >> 2) + 4 * hi32 + (r & 3);
                        // streaming stores: the partial sums are read once, by the reduction,
                        // and need not displace the K part's fragments from L2
                        float2 *dst = &part[((size_t)sgi * 256 + rr) * ncs + coil];
                        __builtin_nontemporal_store(acc[rt][b][0][r], &dst->x);
                        __builtin_nontemporal_store(acc[rt][b][NA - 1][r], &dst->y);
                    }
                }
        } else {
            // one FP32 component per lane: PK = 1 column c is (coil c % 16, part c / 16);
            // PK = 2 lanes with column c < 16 add column c + 16 (the lo half) and hold
            // (coil c % 8, part c / 8)
            float *pf = reinterpret_cast<float *>(part);
            const int coil = PK == 1 ? (col & 15) : (col & 7);
            const int cmp = PK == 1 ? (col >> 4) : ((col >> 3) & 1);
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    float v = acc[rt][0][0][r];
                    // lane ^ 16 within each 32-lane half: ds_swizzle bit mode, and 0x1f, xor 0x10
                    if constexpr (PK == 2)
                        v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401f));
                    const int rr = wave * 64 + rt * 32 + 8 * (r >> 2) + 4 * hi32 + (r & 3);
                    if (PK == 1 || col < 16)
                        __builtin_nontemporal_store(v, &pf[(((size_t)sgi * 256 + rr) * ncs + coil) * 2 + cmp]);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// gradient kernel: G[split][s][l] = sum_{v in split} b_l(v) * Im(e(s,v) * u(s,v)),
// u(s,v) = sum_c conj(r[s][c]) * z[v][c]   (FP32 products, FP64 accumulation)
// ---------------------------------------------------------------------------
template <int L, int CB>
__global__ __launch_bounds__(256) void k_grad(
    const double *__restrict__ coef, int sk, const double *__restrict__ basis, int sb,
    const float2 *__restrict__ z, const float2 *__restrict__ res, double *__restrict__ part, int ncs, int nv,
    const int4 *__restrict__ segs, const int *__restrict__ chunk_off, int cpx) {
    const int chunk = chunk_id(cpx);
    const int batch = blockIdx.y, n_batches = gridDim.y;
    const int cb0 = batch * CB;
    const int sg0 = chunk_off[chunk], sg1 = chunk_off[chunk + 1];
    for (int sgi = sg0; sgi < sg1; ++sgi) {
        const int4 d = segs[sgi];
        const int s = d.x * 256 + threadIdx.x;
        double st[L + 1];
#pragma unroll
        for (int i = 0; i <= L; ++i) st[i] = coef[(size_t)s * sk + i];
        float rr[CB], ri[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const float2 t = res[(size_t)s * ncs + cb0 + c];
            rr[c] = t.x;
            ri[c] = t.y;
        }
        double g[L];
#pragma unroll
        for (int l = 0; l < L; ++l) g[l] = 0.0;
        const int v0 = d.y * 8;
        const int v1 = min(d.z * 8, nv);
        for (int v = v0; v < v1; ++v) {
            const double *sm = basis + (size_t)v * sb;
            double ph = st[0];
#pragma unroll
            for (int i = 1; i <= L; ++i) ph = fma(st[i], sm[i - 1], ph);
            float sn, cs;
            sincos_cycles(ph, sn, cs);
            const float2 *w = z + (size_t)v * ncs + cb0;
            float ur = 0.0f, ui = 0.0f;
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const float2 u = w[c];
                ur = fmaf(rr[c], u.x, ur);
                ur = fmaf(ri[c], u.y, ur);
                ui = fmaf(rr[c], u.y, ui);
                ui = fmaf(-ri[c], u.x, ui);
            }
            const double m = (double)fmaf(cs, ui, sn * ur);
            g[0] += m;
#pragma unroll
            for (int l = 1; l < L; ++l) g[l] = fma(m, sm[l - 1], g[l]);
        }
        double *o = part + (((size_t)sgi * n_batches + batch) * 256 + threadIdx.x) * L;
#pragma unroll
        for (int l = 0; l < L; ++l) o[l] = g[l];
    }
}

// ---------------------------------------------------------------------------
// gradient on the matrix cores (mixed backends)
// ---------------------------------------------------------------------------
//
// G[s][l] = sum_v b_l(v) W[s][v],  W = Im(e(s,v) U[s][v]),  U = sum_c conj(r[s][c]) z[v][c],
// for a wave's 64 samples (two 32-sample tiles) and a segment of 32-voxel blocks:
//  1. U^T (32 voxels x 32 samples, per tile) on v_mfma_f32_32x32x16: A = Z
//     fragments (voxel rows, K = 2*ncs real: [zr | zi] for Re, [zi | -zr] for
//     Im), B = residual fragments (K = [rr | ri], sample columns), hi/lo split.
//     The residual fragments of the wave's samples sit in LDS for the segment;
//     the Z fragments of one voxel block are copied to LDS once for the 4 waves
//     (direct-to-LDS loads) and the next block's copy runs during stages 2 and 3
//     (measured 0.93 -> 0.89 ms at 200^2, 32 coils, 16 000 samples).
//  2. Phases with lane = sample (as in the forward: wave-uniform voxel
//     records), then v_permlane32_swap moves (cos, sin) of voxel pairs (j, j+4)
//     into the D layout of U^T: register r of lane l = voxel 8(r/4)+4(l/32)+r%4,
//     sample l%32 of the tile. W is formed there.
//  3. W (power-of-two scaled, hi/lo) is the A operand of the second product
//     with the basis fragments (columns = terms, each scaled by a power of
//     two); its K index 8(l/32)+i is the voxel of W's register 8h+i.
// Scales: Z by sZ (max|x| max|csm|), residual by sR (max|r|), W by sW from the
// bound 2 ncs max|z| max|r|; the reduction divides them out.
__device__ __forceinline__ void grad_scales(const unsigned *__restrict__ slots, float csm_max, int ncs, float &sz,
                                            float &sr, float &sw) {
    const float mz = __uint_as_float(slots[0]) * csm_max, mr = __uint_as_float(slots[1]);
    sz = pow2_scale(mz);
    sr = pow2_scale(mr);
    sw = pow2_scale(2.0f * (float)ncs * (mz * sz) * (mr * sr));
}

template <typename T16, bool HX, int L, int NKC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gmixed(
    const double *__restrict__ coef, int sk, const double *__restrict__ basis, int sb,
    const uint4 *__restrict__ zfr, const uint4 *__restrict__ rfr, const uint4 *__restrict__ bfr,
    float *__restrict__ part, const int4 *__restrict__ segs, const int *__restrict__ chunk_off, int cpx,
    HornerArgs hx, const unsigned *__restrict__ slots, float csm_max, int ncs) {
    constexpr int n_kc = NKC;   // 16-wide real K chunks: 2 ncs / 16
    typedef typename mfma_t<T16>::vec vec_t;
    extern __shared__ uint4 rlds[];   // [wave][kc][tile][hi, lo][64 lanes]
    const int chunk = chunk_id(cpx);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    float sz, sr, sw;
    grad_scales(slots, csm_max, ncs, sz, sr, sw);
    (void)sz;
    (void)sr;
    uint4 *my = rlds + (size_t)wave * n_kc * 4 * 64 + lane;
    // Z fragments of the current voxel block, shared by the 4 waves: [kc][q][lane]
    uint4 *zl = rlds + (size_t)4 * n_kc * 4 * 64;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    // direct-to-LDS copy of block kb's Z fragments, 4 of the 16 (kc, q) rows per wave
    auto zfetch = [&](int kb) {
#pragma unroll
        for (int i = wv; i < n_kc * 4; i += 4)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(zfr + ((((size_t)kb * n_kc + (i >> 2)) * 64 + lane) << 2) +
                                                                 (i & 3)),
                (__attribute__((address_space(3))) void *)&zl[i * 64], 16, 0, 0);
    };
    const int sg0 = chunk_off[chunk], sg1 = chunk_off[chunk + 1];
    for (int sgi = sg0; sgi < sg1; ++sgi) {
        const int4 d = segs[sgi];
        const int s0 = d.x * 256 + wave * 64;   // first sample of the wave
        const int row = s0 + lane;
        // residual fragments of the wave's two sample tiles -> LDS (each lane keeps its own)
#pragma unroll
        for (int kc = 0; kc < n_kc; ++kc)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const uint4 *src = rfr + ((((size_t)(s0 / 32 + t) * n_kc + kc) * 64 + lane) << 1);
                my[((kc * 2 + t) * 2 + 0) * 64] = src[0];
                my[((kc * 2 + t) * 2 + 1) * 64] = src[1];
            }
        f16x_t g[2] = {f16x_t{}, f16x_t{}};
        // Horner: t of the lane's sample, and its L coefficients for the run setups (in
        // registers: re-read at every run change, the records of all sample tiles competed
        // with the Z fragments for L2); generic: its full record
        double st[HX ? 1 : L + 1], rec[HX ? L : 1];
        double hc[4] = {0.0, 0.0, 0.0, 0.0};
        float c1f = 0.0f, c2f = 0.0f, c2x2f = 0.0f, c3f = 0.0f, c3x3f = 0.0f;
        int cur = -1;
        const bool anch = HX && anchored_on(hx);
        if (HX) {
            st[0] = coef[(size_t)row * sk + hx.t_index];
#pragma unroll
            for (int i = 0; i < L; ++i) rec[i] = coef[(size_t)row * sk + i];
        } else {
#pragma unroll
            for (int i = 0; i <= L; ++i) st[i] = coef[(size_t)row * sk + i];
        }
        zfetch(d.y);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int kb = d.y; kb < d.z; ++kb) {
            // basis fragments of both half blocks, in flight during stage 1
            uint4 bq[2][2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint4 *bp = bfr + ((((size_t)kb * 2 + h) * 64 + lane) << 1);
                bq[h][0] = bp[0];
                bq[h][1] = bp[1];
            }
            // ---- stage 1: U^T for both sample tiles ----
            f16x_t ure[2] = {f16x_t{}, f16x_t{}}, uim[2] = {f16x_t{}, f16x_t{}};
#pragma unroll
            for (int kc = 0; kc < n_kc; ++kc) {
                const uint4 *zp = zl + kc * 4 * 64 + lane;
                const vec_t zrh = __builtin_bit_cast(vec_t, zp[0]);
                const vec_t zrl = __builtin_bit_cast(vec_t, zp[64]);
                const vec_t zih = __builtin_bit_cast(vec_t, zp[128]);
                const vec_t zil = __builtin_bit_cast(vec_t, zp[192]);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const vec_t rh = __builtin_bit_cast(vec_t, my[((kc * 2 + t) * 2 + 0) * 64]);
                    const vec_t rl = __builtin_bit_cast(vec_t, my[((kc * 2 + t) * 2 + 1) * 64]);
                    ure[t] = mfma_t<T16>::run(zrh, rh, ure[t]);
                    uim[t] = mfma_t<T16>::run(zih, rh, uim[t]);
                    ure[t] = mfma_t<T16>::run(zrl, rh, ure[t]);
                    uim[t] = mfma_t<T16>::run(zil, rh, uim[t]);
                    ure[t] = mfma_t<T16>::run(zrh, rl, ure[t]);
                    uim[t] = mfma_t<T16>::run(zih, rl, uim[t]);
                }
            }
            // every wave has read block kb's Z: fetch block kb + 1 into the same buffer
            // while the phases of block kb run
            __syncthreads();
            if (kb + 1 < d.z) zfetch(kb + 1);
            // ---- stages 2 and 3, per half block of 16 voxels ----
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q0 = kb * 32 + h * 16;
                float cs[16], sn[16];
                if constexpr (HX) {
                    // the half block is K units 2 (q0 / 16) and 2 (q0 / 16) + 1; inside one run
                    // their phases are anchored as in the MFMA forward (FP64 cubic and
                    // fieldmap term at each unit's anchor, FP32 offsets of its 8 voxels)
                    const HoeUnit a0 = hx.units[q0 / 8], a1 = hx.units[q0 / 8 + 1];
                    const int r0 = a0.run;
                    if (r0 >= 0 && a1.run == r0 && anch) {
                        if (r0 != cur) {
                            run_setup<L>(rec, hx.beta + (size_t)r0 * L * 4, hc);
                            c1f = (float)hc[1];
                            c2f = (float)hc[2];
                            c2x2f = (float)(2.0 * hc[2]);
                            c3f = (float)hc[3];
                            c3x3f = (float)(3.0 * hc[3]);
                            cur = r0;
                        }
                        const float tf = (float)st[0];
#pragma unroll
                        for (int uh = 0; uh < 2; ++uh) {
                            const HoeUnit &a = uh ? a1 : a0;
                            const HoeUnitVox &uv = hx.uvox[q0 / 8 + uh];
                            const double xa = (double)a.xa;
                            const double pa =
                                fma(st[0], (double)a.fma, fma(fma(fma(hc[3], xa, hc[2]), xa, hc[1]), xa, hc[0]));
                            const float base = (float)__builtin_amdgcn_fract(pa);
                            const float d2 = fmaf(c3x3f, a.xa, c2f);
                            const float d1 = fmaf(fmaf(c3x3f, a.xa, c2x2f), a.xa, c1f);
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                const float u = uv.ud[j][0], dfm = uv.ud[j][1];
                                sincos_cycles(fmaf(tf, dfm, fmaf(fmaf(fmaf(c3f, u, d2), u, d1), u, base)),
                                              sn[8 * uh + j], cs[8 * uh + j]);
                            }
                        }
                    } else {
                        // a run boundary inside the half block (or anchored phases off):
                        // full records (generic phase)
                        double a[L + 1];
#pragma unroll
                        for (int i = 0; i < L; ++i) a[i] = rec[i];
                        a[L] = coef[(size_t)row * sk + L];
#pragma unroll
                        for (int j = 0; j < 16; ++j) {
                            const double *sm = basis + (size_t)(q0 + j) * sb;
                            double p = a[0];
#pragma unroll
                            for (int i = 1; i <= L; ++i) p = fma(a[i], sm[i - 1], p);
                            sincos_cycles(p, sn[j], cs[j]);
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const double *sm = basis + (size_t)(q0 + j) * sb;
                        double p = st[0];
#pragma unroll
                        for (int i = 1; i <= L; ++i) p = fma(st[i], sm[i - 1], p);
                        sincos_cycles(p, sn[j], cs[j]);
                    }
                }
                // (cos, sin) of voxel pairs (j, j+4) into the D layout of U^T
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    const int j = 8 * (a >> 2) + (a & 3);
                    auto c2 = __builtin_amdgcn_permlane32_swap(__float_as_uint(cs[j]), __float_as_uint(cs[j + 4]),
                                                               false, false);
                    cs[j] = __uint_as_float(c2[0]);
                    cs[j + 4] = __uint_as_float(c2[1]);
                    auto s2 = __builtin_amdgcn_permlane32_swap(__float_as_uint(sn[j]), __float_as_uint(sn[j + 4]),
                                                               false, false);
                    sn[j] = __uint_as_float(s2[0]);
                    sn[j + 4] = __uint_as_float(s2[1]);
                }
                const vec_t bh = __builtin_bit_cast(vec_t, bq[h][0]);
                const vec_t bl = __builtin_bit_cast(vec_t, bq[h][1]);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    unsigned wh[4], wl[4];
#pragma unroll
                    for (int i = 0; i < 8; i += 2) {
                        const int j0 = 8 * (i >> 2) + (i & 3) + 4 * t, j1 = j0 + 1;
                        const float w0 = fmaf(sn[j0], ure[t][8 * h + i], cs[j0] * uim[t][8 * h + i]) * sw;
                        const float w1 = fmaf(sn[j1], ure[t][8 * h + i + 1], cs[j1] * uim[t][8 * h + i + 1]) * sw;
                        split2<T16>(w0, w1, wh[i >> 1], wl[i >> 1]);
                    }
                    const vec_t ah = __builtin_bit_cast(vec_t, make_uint4(wh[0], wh[1], wh[2], wh[3]));
                    const vec_t al = __builtin_bit_cast(vec_t, make_uint4(wl[0], wl[1], wl[2], wl[3]));
                    g[t] = mfma_t<T16>::run(ah, bh, g[t]);
                    g[t] = mfma_t<T16>::run(al, bh, g[t]);
                    g[t] = mfma_t<T16>::run(ah, bl, g[t]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        // D layout: register r of lane l = sample 8(r/4)+4(l/32)+r%4 of the tile, term l%32
        const int col = lane & 31, hi32 = lane >> 5;
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int rr = wave * 64 + t * 32 + 8 * (r >> 2) + 4 * hi32 + (r & 3);
                __builtin_nontemporal_store(g[t][r], &part[((size_t)sgi * 256 + rr) * 32 + col]);
            }
    }
}

// Z fragments of the gradient (A operand, voxel rows):
// zfr[kb][kc][lane] = {Re hi, Re lo, Im hi, Im lo}, 8 x T16 each, lane l:
// voxel 32 kb + l%32, real K index k = 16 kc + 8 (l/32) + i: k < ncs ->
// (zr_k | zi_k), else (zi_{k-ncs} | -zr_{k-ncs}).
template <typename T16>
__global__ void k_build_gzfrag(const float2 *__restrict__ csm, const float2 *__restrict__ xm, uint4 *__restrict__ frag,
                               int n_kb, int n_kc, int ncs, const unsigned *__restrict__ slots, float csm_max) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_kb * n_kc * 64) return;
    const int lane = (int)(t & 63);
    const int kc = (int)((t >> 6) % n_kc);
    const int kb = (int)((t >> 6) / n_kc);
    float sz, sr, sw;
    grad_scales(slots, csm_max, ncs, sz, sr, sw);
    const int v = kb * 32 + (lane & 31);
    const float2 xv = xm[v];
    float re[8], im[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = 16 * kc + 8 * (lane >> 5) + i;
        const int c = k < ncs ? k : k - ncs;
        const float2 a = csm[(size_t)v * ncs + c];
        const float zr = (a.x * xv.x - a.y * xv.y) * sz;
        const float zi = (a.x * xv.y + a.y * xv.x) * sz;
        re[i] = k < ncs ? zr : zi;
        im[i] = k < ncs ? zi : -zr;
    }
    unsigned rh[4], rl[4], ih[4], il[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        split2<T16>(re[2 * i], re[2 * i + 1], rh[i], rl[i]);
        split2<T16>(im[2 * i], im[2 * i + 1], ih[i], il[i]);
    }
    uint4 *o = frag + (t << 2);
    o[0] = make_uint4(rh[0], rh[1], rh[2], rh[3]);
    o[1] = make_uint4(rl[0], rl[1], rl[2], rl[3]);
    o[2] = make_uint4(ih[0], ih[1], ih[2], ih[3]);
    o[3] = make_uint4(il[0], il[1], il[2], il[3]);
}

// Residual fragments (B operand, sample columns): rfr[st][kc][lane] = {hi, lo},
// lane l: sample 32 st + l%32, K index k = 16 kc + 8 (l/32) + i: rr_k (k < ncs)
// or ri_{k-ncs}.  r = ybuf (coil fastest).
template <typename T16>
__global__ void k_build_grfrag(const float2 *__restrict__ r, uint4 *__restrict__ frag, int n_st, int n_kc, int ncs,
                               const unsigned *__restrict__ slots, float csm_max) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_st * n_kc * 64) return;
    const int lane = (int)(t & 63);
    const int kc = (int)((t >> 6) % n_kc);
    const int st = (int)((t >> 6) / n_kc);
    float sz, sr, sw;
    grad_scales(slots, csm_max, ncs, sz, sr, sw);
    const int s = st * 32 + (lane & 31);
    float val[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = 16 * kc + 8 * (lane >> 5) + i;
        const float2 a = r[(size_t)s * ncs + (k < ncs ? k : k - ncs)];
        val[i] = (k < ncs ? a.x : a.y) * sr;
    }
    unsigned hh[4], ll[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) split2<T16>(val[2 * i], val[2 * i + 1], hh[i], ll[i]);
    uint4 *o = frag + (t << 1);
    o[0] = make_uint4(hh[0], hh[1], hh[2], hh[3]);
    o[1] = make_uint4(ll[0], ll[1], ll[2], ll[3]);
}

// Basis fragments (B operand of the second product): bfr[kb][h][lane] = {hi, lo},
// lane l: term n = l%32 (n = 0: the constant 1; 1 <= n < ncols: record entry
// n-1; else 0), K entry i: voxel 32 kb + 16 h + 8 (i/4) + 4 (l/32) + i%4, scaled by
// bscale[n] (a power of two).
template <typename T16>
__global__ void k_build_gbfrag(const double *__restrict__ basis, int sb, int ncols, const float *__restrict__ bscale,
                               uint4 *__restrict__ frag, int n_kb, int nv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_kb * 2 * 64) return;
    const int lane = (int)(t & 63);
    const int h = (int)((t >> 6) & 1);
    const int kb = (int)(t >> 7);
    const int n = lane & 31;
    float val[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int v = kb * 32 + 16 * h + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        double b = 0.0;
        if (v < nv && n < ncols) b = n == 0 ? 1.0 : basis[(size_t)v * sb + (n - 1)];
        val[i] = (float)b * bscale[n];
    }
    unsigned hh[4], ll[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) split2<T16>(val[2 * i], val[2 * i + 1], hh[i], ll[i]);
    uint4 *o = frag + (t << 1);
    o[0] = make_uint4(hh[0], hh[1], hh[2], hh[3]);
    o[1] = make_uint4(ll[0], ll[1], ll[2], ll[3]);
}

// gk[l][s] = scale / (sZ sR sW bscale[l]) * sum over the tile's segments of part[seg][s%256][l]  (FP64)
__global__ void k_reduce_gmixed(const float *__restrict__ part, const int *__restrict__ tile_off,
                                const int *__restrict__ tile_list, int L, int ns, double scale,
                                const float *__restrict__ bscale, const unsigned *__restrict__ slots, float csm_max,
                                int ncs, double *__restrict__ gk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ns * L) return;
    const int s = (int)(i / L), l = (int)(i % L);
    const int tile = s >> 8;
    float sz, sr, sw;
    grad_scales(slots, csm_max, ncs, sz, sr, sw);
    double acc = 0.0;
    for (int j = tile_off[tile]; j < tile_off[tile + 1]; ++j)
        acc += (double)part[((size_t)tile_list[j] * 256 + (s & 255)) * 32 + l];
    gk[(size_t)l * ns + s] = acc * (scale / ((double)sz * (double)sr * (double)sw * (double)bscale[l]));
}

// ---------------------------------------------------------------------------
// data movement and reductions
// ---------------------------------------------------------------------------

__global__ void k_zero_u32(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The slot is cleared before the launch and only grows during it, so a value
// read from it never exceeds the true maximum: when it already covers m the
// atomic is skipped (most waves, once a large value has landed), which keeps
// thousands of waves from serialising on one address.
__device__ __forceinline__ void wave_absmax_to(unsigned *dst, float m) {
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.0f &&
        __float_as_uint(m) > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(dst, __float_as_uint(m));
}

// xm[v] = x_full[mask_idx[v]] (zero for padded voxels); optional |.|max (FP32).
template <typename R>
__global__ void k_gather(const cplx_t<R> *__restrict__ x, const int *__restrict__ idx, cplx_t<R> *__restrict__ xm,
                         int nv, int nv_pad, unsigned *maxslot) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (v < nv_pad) {
        cplx_t<R> val = mkc<R>(R(0), R(0));
        if (v < nv) val = x[idx[v]];
        xm[v] = val;
        m = fmaxf(fabsf((float)val.x), fabsf((float)val.y));
    }
    if (maxslot) wave_absmax_to(maxslot, m);
}

// |.|max of a masked vector (float2) into a slot.
__global__ void k_absmax2(const float2 *__restrict__ a, int n, unsigned *maxslot) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (v < n) {
        const float2 t = a[v];
        m = fmaxf(fabsf(t.x), fabsf(t.y));
    }
    wave_absmax_to(maxslot, m);
}

template <typename R>
__global__ void k_scatter(const cplx_t<R> *__restrict__ xm, const int *__restrict__ idx, cplx_t<R> *__restrict__ x,
                          int nv, R scale) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) {
        const cplx_t<R> t = xm[v];
        x[idx[v]] = mkc<R>(t.x * scale, t.y * scale);
    }
}

template <typename R>
__global__ void k_fill0(cplx_t<R> *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = mkc<R>(R(0), R(0));
}

// z[v][c] = csm[v][c] * xm[v] (coil stride ncs; zero for padded coils/voxels).
template <typename R>
__global__ void k_build_z(const cplx_t<R> *__restrict__ csm, const cplx_t<R> *__restrict__ xm,
                          cplx_t<R> *__restrict__ z, int nv_pad, int ncs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nv_pad * ncs) return;
    const int v = (int)(i / ncs);
    const cplx_t<R> a = csm[i], b = xm[v];
    z[i] = mkc<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Packed B fragment (PK = 1, 2: see k_mixed) of one lane from its four (re tile,
// im tile) pairs: PK = 1 lane column c takes the re (c < 16) or im tile of coil c % 16,
// hi and lo in two uint4; PK = 2 column c takes tile (c / 8) % 2 of coil c % 8, hi
// (c < 16) or lo, in one uint4.
template <typename T16, int PK>
__device__ __forceinline__ void store_packed(uint4 *o, int col, const float (&re)[4][2], const float (&im)[4][2]) {
    unsigned h[4], l[4];
    const bool take_im = PK == 1 ? (col >= 16) : (((col >> 3) & 1) != 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float a = take_im ? im[j][0] : re[j][0], b = take_im ? im[j][1] : re[j][1];
        split2<T16>(a, b, h[j], l[j]);
    }
    if constexpr (PK == 1) {
        o[0] = make_uint4(h[0], h[1], h[2], h[3]);
        o[1] = make_uint4(l[0], l[1], l[2], l[3]);
    } else {
        o[0] = col < 16 ? make_uint4(h[0], h[1], h[2], h[3]) : make_uint4(l[0], l[1], l[2], l[3]);
    }
}

// MFMA B fragments of z = csm .* xm for the forward (K = voxels):
// frag[gb][ks][lane] = {hi_re, hi_im, lo_re, lo_im}, each 8 x T16:
//   lane l: coil 32*gb + l%32, voxels 8*ks + 4*(l/32) + i/2 (i = 0..7);
//   re tile: (Zr, -Zi) pairs, im tile: (Zi, Zr) pairs.
// PK = 1, 2: the packed layouts of k_mixed, 2 or 1 uint4 per lane (coils 0..15 / 0..7).
template <typename T16, int PK = 0>
__global__ void k_build_zfrag(const float2 *__restrict__ csm, const float2 *__restrict__ xm, uint4 *__restrict__ frag,
                              int n_ks, int ncs, int n_gb, const unsigned *__restrict__ maxslot, float csm_max) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_gb * n_ks * 64) return;
    const int lane = (int)(t & 63);
    const int64_t rest = t >> 6;
    const int ks = (int)(rest % n_ks);
    const int gb = (int)(rest / n_ks);
    const float scale = pow2_scale(__uint_as_float(*maxslot) * csm_max);
    if constexpr (PK != 0) {
        const int col = lane & 31, c = PK == 1 ? (col & 15) : (col & 7);
        float re[4][2], im[4][2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int v = ks * 8 + 4 * (lane >> 5) + j;
            const float2 a = csm[(size_t)v * ncs + c], b = xm[v];
            const float zr = (a.x * b.x - a.y * b.y) * scale;
            const float zi = (a.x * b.y + a.y * b.x) * scale;
            re[j][0] = zr;
            re[j][1] = -zi;
            im[j][0] = zi;
            im[j][1] = zr;
        }
        store_packed<T16, PK>(frag + t * (PK == 1 ? 2 : 1), col, re, im);
        return;
    }
    const int c = gb * 32 + (lane & 31);
    unsigned hr[4], hi[4], lr[4], li[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int v = ks * 8 + 4 * (lane >> 5) + j;
        const float2 a = csm[(size_t)v * ncs + c], b = xm[v];
        const float zr = (a.x * b.x - a.y * b.y) * scale;
        const float zi = (a.x * b.y + a.y * b.x) * scale;
        unsigned h0, l0, h1, l1;
        split2<T16>(zr, -zi, h0, l0);   // re tile
        split2<T16>(zi, zr, h1, l1);    // im tile
        hr[j] = h0; lr[j] = l0; hi[j] = h1; li[j] = l1;
    }
    uint4 *o = frag + (t << 2);
    o[0] = make_uint4(hr[0], hr[1], hr[2], hr[3]);
    o[1] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    o[2] = make_uint4(lr[0], lr[1], lr[2], lr[3]);
    o[3] = make_uint4(li[0], li[1], li[2], li[3]);
}

// MFMA B fragments of the weighted signal y[s][c] for the adjoint (K = samples):
//   re tile: (Yr, Yi) pairs, im tile: (Yi, -Yr) pairs (PK as in k_build_zfrag).
template <typename T16, int PK = 0>
__global__ void k_build_yfrag(const float2 *__restrict__ y, uint4 *__restrict__ frag, int n_ks, int ncs, int n_gb,
                              const unsigned *__restrict__ maxslot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_gb * n_ks * 64) return;
    const int lane = (int)(t & 63);
    const int64_t rest = t >> 6;
    const int ks = (int)(rest % n_ks);
    const int gb = (int)(rest / n_ks);
    const float scale = pow2_scale(__uint_as_float(*maxslot));
    if constexpr (PK != 0) {
        const int col = lane & 31, c = PK == 1 ? (col & 15) : (col & 7);
        float re[4][2], im[4][2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = ks * 8 + 4 * (lane >> 5) + j;
            const float2 a = y[(size_t)s * ncs + c];
            const float yr = a.x * scale, yi = a.y * scale;
            re[j][0] = yr;
            re[j][1] = yi;
            im[j][0] = yi;
            im[j][1] = -yr;
        }
        store_packed<T16, PK>(frag + t * (PK == 1 ? 2 : 1), col, re, im);
        return;
    }
    const int c = gb * 32 + (lane & 31);
    unsigned hr[4], hi[4], lr[4], li[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int s = ks * 8 + 4 * (lane >> 5) + j;
        const float2 a = y[(size_t)s * ncs + c];
        const float yr = a.x * scale, yi = a.y * scale;
        unsigned h0, l0, h1, l1;
        split2<T16>(yr, yi, h0, l0);
        split2<T16>(yi, -yr, h1, l1);
        hr[j] = h0; lr[j] = l0; hi[j] = h1; li[j] = l1;
    }
    uint4 *o = frag + (t << 2);
    o[0] = make_uint4(hr[0], hr[1], hr[2], hr[3]);
    o[1] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    o[2] = make_uint4(lr[0], lr[1], lr[2], lr[3]);
    o[3] = make_uint4(li[0], li[1], li[2], li[3]);
}

// Forward reduction over the segments of each sample tile (segment order).
// Tile of 32 samples x 32 coils per block.
// mode 0: user layout y_out[c*ns + s] = scale * sum
// mode 1: ybuf[s][c] = w[s][c] * (scale * sum - y_data[c*ns + s]) (w, y_data optional),
//         zero for padded rows/coils, optional |.|max into maxslot.
// The effective scale is scale / pow2_scale(max from scale_slot * scale_mul)
// when scale_slot is given (undoing the fragment scaling of the mixed path).
template <typename R>
__device__ __forceinline__ cplx_t<R> sum_segments(const cplx_t<R> *__restrict__ part, const int *__restrict__ list,
                                                  int n, int r, int c, int ncs, int i0 = 0, int step = 1) {
    cplx_t<R> a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = mkc<R>(R(0), R(0));
    int i = i0;
    for (; i + 3 * step < n; i += 4 * step) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const cplx_t<R> t = part[((size_t)list[i + q * step] * 256 + r) * ncs + c];
            a[q].x += t.x;
            a[q].y += t.y;
        }
    }
    for (; i < n; i += step) {
        const cplx_t<R> t = part[((size_t)list[i] * 256 + r) * ncs + c];
        a[0].x += t.x;
        a[0].y += t.y;
    }
    return mkc<R>((a[0].x + a[1].x) + (a[2].x + a[3].x), (a[0].y + a[1].y) + (a[2].y + a[3].y));
}

// Coil columns that one 32-lane row of a reduction covers: nc rounded up to a power of
// two when nc <= 16 (the 32 / that lanes of a coil then share its segment list, each
// summing every (32 / ncl)-th segment, and a fixed butterfly adds them), else 32.
__device__ __forceinline__ int reduce_cols(int nc) {
    return nc > 16 ? 32 : (nc <= 1 ? 1 : 1 << (32 - __clz(nc - 1)));
}

template <typename R>
__global__ __launch_bounds__(1024) void k_reduce_fwd(
    const cplx_t<R> *__restrict__ part, const int *__restrict__ tile_off, const int *__restrict__ tile_list, int ncs,
    int ns, int nc, R scale, const unsigned *__restrict__ scale_slot, float scale_mul, int mode,
    cplx_t<R> *__restrict__ out, const R *__restrict__ w, const cplx_t<R> *__restrict__ y_data, unsigned *maxslot,
    int s_base) {
    __shared__ cplx_t<R> tileb[32][33];
    __shared__ cplx_t<R> other[3][4][256];   // row sums of slices 1..H-1
    const int H = blockDim.x >> 8;              // segment slices of 256 threads (1, 2 or 4)
    const int s0 = s_base + blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int hf = threadIdx.x >> 8;
    const int tl = threadIdx.x & 255;
    const int tc = tl & 31, tr = tl >> 5;   // 8 row groups
    const int tile = s0 >> 8;
    const int *list = tile_list + tile_off[tile];
    const int nseg = tile_off[tile + 1] - tile_off[tile];
    R sc = scale;
    if constexpr (std::is_same<R, float>::value) {
        if (scale_slot) sc = scale / pow2_scale(__uint_as_float(*scale_slot) * scale_mul);
    }
    if (mode == 1 && y_data && hf == 0) {
        // stage the data tile transposed: y_data[c*ns + s]
        for (int k = tr; k < 32; k += 8) {
            const int c = c0 + k, s = s0 + tc;
            tileb[k][tc] = (c < nc && s < ns) ? y_data[(size_t)c * ns + s] : mkc<R>(R(0), R(0));
        }
    }
    // four rows per thread; slice hf of the block takes every H-th group of the thread's
    // segments, two per step in separate sums (8 loads in flight); lanes of coils beyond
    // nc read nothing (the contraction kernels need not write those columns). With nc <= 16
    // the lanes of one coil also split its segments (see reduce_cols); lane tc < ncl then
    // holds coil tc's total and the others, whose columns lie beyond nc, are discarded below.
    const int ncl = c0 == 0 ? reduce_cols(nc) : 32, nsub = 32 / ncl;
    const int S = H * nsub;   // segment stride of one thread
    const int cc = c0 + (tc & (ncl - 1));
    cplx_t<R> rows[4], rows2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) rows[q] = rows2[q] = mkc<R>(R(0), R(0));
    const int nsr = cc < nc ? nseg : 0;
    int j = tc / ncl + hf * nsub;
    for (; j + S < nsr; j += 2 * S) {
        const size_t b0 = (size_t)list[j] * 256, b1 = (size_t)list[j + S] * 256;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = (s0 + tr + 8 * q) & 255;
            const cplx_t<R> t0 = part[(b0 + r) * ncs + cc];
            const cplx_t<R> t1 = part[(b1 + r) * ncs + cc];
            rows[q].x += t0.x;
            rows[q].y += t0.y;
            rows2[q].x += t1.x;
            rows2[q].y += t1.y;
        }
    }
    if (j < nsr) {
        const size_t b0 = (size_t)list[j] * 256;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const cplx_t<R> t = part[(b0 + ((s0 + tr + 8 * q) & 255)) * ncs + cc];
            rows[q].x += t.x;
            rows[q].y += t.y;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        rows[q].x += rows2[q].x;
        rows[q].y += rows2[q].y;
        if (hf > 0) other[hf - 1][q][tl] = rows[q];
    }
    __syncthreads();   // the other slices' sums and the staged data tile
    float m = 0.0f;
    if (hf == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            for (int h = 1; h < H; ++h) {
                const cplx_t<R> o = other[h - 1][q][tl];
                rows[q].x += o.x;
                rows[q].y += o.y;
            }
            for (int k = ncl; k < 32; k <<= 1) {
                rows[q].x += __shfl_xor(rows[q].x, k, 32);
                rows[q].y += __shfl_xor(rows[q].y, k, 32);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = tr + 8 * q;
            const int s = s0 + k, c = c0 + tc;
            cplx_t<R> acc = rows[q];
            acc.x *= sc;
            acc.y *= sc;
            if (mode == 0) {
                tileb[k][tc] = acc;
            } else {
                const bool valid = (s < ns) && (c < nc);
                if (y_data) {
                    const cplx_t<R> d = tileb[tc][k];
                    acc.x -= d.x;
                    acc.y -= d.y;
                }
                if (w) {
                    const R ww = w[(size_t)s * ncs + c];
                    acc.x *= ww;
                    acc.y *= ww;
                }
                if (!valid) acc = mkc<R>(R(0), R(0));
                out[(size_t)s * ncs + c] = acc;
                m = fmaxf(m, fmaxf(fabsf((float)acc.x), fabsf((float)acc.y)));
            }
        }
    }
    if (mode == 0) {
        __syncthreads();
        if (hf == 0)
            for (int k = tr; k < 32; k += 8) {
                const int c = c0 + k, s = s0 + tc;
                if (c < nc && s < ns) out[(size_t)c * ns + s] = tileb[tc][k];
            }
    } else if (maxslot && hf == 0) {
        wave_absmax_to(maxslot, m);
    }
}

// ybuf[s][c] = w[s][c] * y_user[c*ns + s] (zero padded), optional |.|max.
template <typename R>
__global__ __launch_bounds__(256) void k_load_ybuf(const cplx_t<R> *__restrict__ y, const R *__restrict__ w,
                                                   cplx_t<R> *__restrict__ out, int ns, int nc, int ns_pad, int ncs,
                                                   unsigned *maxslot) {
    __shared__ cplx_t<R> tileb[32][33];
    const int s0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tc = threadIdx.x & 31, tr = threadIdx.x >> 5;
    for (int k = tr; k < 32; k += 8) {
        const int c = c0 + k, s = s0 + tc;
        tileb[k][tc] = (c < nc && s < ns) ? y[(size_t)c * ns + s] : mkc<R>(R(0), R(0));
    }
    __syncthreads();
    float m = 0.0f;
    for (int k = tr; k < 32; k += 8) {
        const int s = s0 + k, c = c0 + tc;
        cplx_t<R> v = tileb[tc][k];
        if (w) {
            const R ww = w[(size_t)s * ncs + c];
            v.x *= ww;
            v.y *= ww;
        }
        out[(size_t)s * ncs + c] = v;
        m = fmaxf(m, fmaxf(fabsf((float)v.x), fabsf((float)v.y)));
    }
    if (maxslot) wave_absmax_to(maxslot, m);
}

// Adjoint reduction: xm_out[v] = sc * sum_{seg} sum_{g < ngrp} part[seg][g][v % 256]
//                               (+ lambda * x_in[v]).
// The contraction kernels combine the coils of their group (conj(csm) weighting) in their
// epilogues, so one complex per voxel, segment and coil group remains. NL consecutive lanes
// share a voxel (NL = 1..16, a power of two chosen from the tile's segment count, so that
// small problems, whose tiles are cut into many segments, keep few loads per lane): lane j
// adds the entries j, j + NL, ... of the voxel's (segment, group) list, two in flight, and a
// fixed butterfly adds the lanes.
template <typename R>
__global__ __launch_bounds__(256) void k_reduce_adj(
    const cplx_t<R> *__restrict__ part, const int *__restrict__ tile_off, const int *__restrict__ tile_list,
    int ngrp, int nl, int nv, R scale, const unsigned *__restrict__ scale_slot, cplx_t<R> *__restrict__ xm_out,
    R lambda, const cplx_t<R> *__restrict__ x_in, int v_base, int v_end, int nv_write) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = v_base + (int)(gid / nl), j = (int)(gid % nl);
    const bool live = v < v_end;
    R sc = scale;
    if constexpr (std::is_same<R, float>::value) {
        if (scale_slot) sc = scale / pow2_scale(__uint_as_float(*scale_slot));
    }
    cplx_t<R> a = mkc<R>(R(0), R(0)), b = mkc<R>(R(0), R(0));
    if (live) {
        const int tile = v >> 8, r = v & 255;
        const int *list = tile_list + tile_off[tile];
        const int n = (tile_off[tile + 1] - tile_off[tile]) * ngrp;
        // entry i = (segment i / ngrp, group i % ngrp)
        auto at = [&](int i) { return part[((size_t)list[i / ngrp] * ngrp + i % ngrp) * 256 + r]; };
        int i = j;
        for (; i + nl < n; i += 2 * nl) {
            const cplx_t<R> t0 = at(i), t1 = at(i + nl);
            a.x += t0.x;
            a.y += t0.y;
            b.x += t1.x;
            b.y += t1.y;
        }
        if (i < n) {
            const cplx_t<R> t = at(i);
            a.x += t.x;
            a.y += t.y;
        }
    }
    R sx = a.x + b.x, sy = a.y + b.y;
    for (int o = 1; o < nl; o <<= 1) {
        sx += __shfl_xor(sx, o);
        sy += __shfl_xor(sy, o);
    }
    if (!live || j != 0) return;
    cplx_t<R> res = mkc<R>(sx * sc, sy * sc);
    if (v >= nv) res = mkc<R>(R(0), R(0));
    if (x_in && v < nv) {
        const cplx_t<R> t = x_in[v];
        res.x += lambda * t.x;
        res.y += lambda * t.y;
    }
    if (v < nv_write) xm_out[v] = res;
}

// gk[l][s] = scale * sum_segments sum_batches part[seg][batch][s%256][l]  (FP64)
__global__ void k_reduce_grad(const double *__restrict__ part, const int *__restrict__ tile_off,
                              const int *__restrict__ tile_list, int n_batches, int L, int ns, double scale,
                              double *__restrict__ gk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ns * L) return;
    const int s = (int)(i / L), l = (int)(i % L);
    const int tile = s >> 8;
    double acc = 0.0;
    for (int j = tile_off[tile]; j < tile_off[tile + 1]; ++j)
        for (int b = 0; b < n_batches; ++b)
            acc += part[(((size_t)tile_list[j] * n_batches + b) * 256 + (s & 255)) * L + l];
    gk[(size_t)l * ns + s] = acc * scale;
}

// ---- coefficient records from the traces ----

// coef[s][r] and dcoef[s][r] (r < sk) for every padded sample s from the traces [ns][L]
// (term fastest) resampled at s + tau/dwell with uniform Catmull-Rom weights w over taps
// clamp(s + m - 1 + t), t = 0..3 (shift == 0: the traces unchanged):
//   r = 0: term 1; r = 1..na: active term act[r-1]; r = na + 1: t_s (coef) or 0 (dcoef);
//   beyond, and for padded samples: 0.
// dcoef holds d k / d tau = sum_t dw[t] h(...) / dwell (all zero when dwell is unset).
__global__ void k_build_records(const float *__restrict__ traces, int ns, int ns_pad, int L,
                                const int *__restrict__ act, int na, const float *__restrict__ times, int sk,
                                int shift, int m, double4 w, double4 dw, double inv_dwell, double *__restrict__ coef,
                                double *__restrict__ dcoef) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ns_pad * sk) return;
    const int s = (int)(i / sk), r = (int)(i % sk);
    double c = 0.0, dc = 0.0;
    if (s < ns && r <= na + 1) {
        if (r == na + 1) {
            c = (double)times[s];
        } else {
            const int l = r == 0 ? 0 : act[r - 1];
            const double wt[4] = {w.x, w.y, w.z, w.w}, dwt[4] = {dw.x, dw.y, dw.z, dw.w};
            double a = 0.0, da = 0.0;
            for (int t = 0; t < 4; ++t) {
                const int j = min(max(s + m - 1 + t, 0), ns - 1);
                const double v = (double)traces[(size_t)j * L + l];
                a = fma(wt[t], v, a);
                da = fma(dwt[t], v, da);
            }
            c = shift ? a : (double)traces[(size_t)s * L + l];
            dc = da * inv_dwell;
        }
    }
    coef[i] = c;
    dcoef[i] = dc;
}

// Bound of the anchored phases' Horner offsets for the records just built:
// max_s sum_{r <= nr} |coef[s][r]| * D[r] into *out (float bits; cleared before the
// launch, non-negative, so the unsigned order is the float order). 256 threads per block.
__global__ __launch_bounds__(256) void k_anchor_bound(const double *__restrict__ coef, int sk, int ns,
                                                      const double *__restrict__ D, int nr, unsigned *out) {
    __shared__ float red[256];
    float m = 0.0f;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        double b = 0.0;
        for (int r = 1; r < nr; ++r) b = fma(fabs(coef[(size_t)s * sk + r]), D[r], b);
        m = fmaxf(m, (float)b);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

// ---- gradient chain rule through the delay interpolation ----

// gh[j][l] = sum_t w[t] sum_{s : clamp(s+m-1+t) = j} gk[row(l)][s]   (shift != 0)
//          = gk[row(l)][j]                                            (shift == 0)
// (the adjoint of the Catmull-Rom resampling k(s) = sum_t w[t] h(clamp(s+m-1+t)));
// rows of identically zero basis terms give 0. FP64 sums, FP32 output.
__global__ void k_grad_chain(const double *__restrict__ gk, int ns, const int *__restrict__ rowmap, int L,
                             int shift, int m, double w0, double w1, double w2, double w3, float *__restrict__ gh) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ns * L) return;
    const int j = (int)(i / L), l = (int)(i % L);
    const int row = rowmap[l];
    if (row < 0) {
        gh[i] = 0.0f;
        return;
    }
    const double *g = gk + (size_t)row * ns;
    if (!shift) {
        gh[i] = (float)g[j];
        return;
    }
    const double w[4] = {w0, w1, w2, w3};
    double acc = 0.0;
    for (int t = 0; t < 4; ++t) {
        // s with clamp(s + m - 1 + t) == j
        int lo = j - m + 1 - t, hi = lo;
        if (j == 0) lo = INT_MIN / 2;
        if (j == ns - 1) hi = INT_MAX / 2;
        lo = max(lo, 0);
        hi = min(hi, ns - 1);
        double a = 0.0;
        for (int s = lo; s <= hi; ++s) a += g[s];
        acc += w[t] * a;
    }
    gh[i] = (float)acc;
}

// block partial sums of sum_{s, r < nrow} gk[r][s] * dk[s][r]  (dL/dtau)
__global__ __launch_bounds__(256) void k_gtau_partial(const double *__restrict__ gk, const double *__restrict__ dk,
                                                      int ns, int nrow, int sk, double *__restrict__ red) {
    __shared__ double sh[4];
    double acc = 0.0;
    const int64_t n = (int64_t)ns * nrow;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / ns), s = (int)(i % ns);
        acc += gk[i] * dk[(size_t)s * sk + r];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---- FP64 reductions for CG and losses ----

// block partial sums of Re<a,b> (b == nullptr: |a|^2)
template <typename R>
__global__ __launch_bounds__(256) void k_dot_partial(const cplx_t<R> *__restrict__ a, const cplx_t<R> *__restrict__ b,
                                                     int64_t n, double *__restrict__ red) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const cplx_t<R> x = a[i];
        const cplx_t<R> y = b ? b[i] : x;
        acc += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// CG scalar logic, one block. op: 0 = store sum into out[0];
// 1 = pAp: cg[2] = sum, cg[3] = alpha = rr/pAp (0 if done or pAp <= 0);
// 2 = rr_new: cg[6] = sum; beta = rr_new/rr; log[it] = sqrt(rr_new/rr0); done test; rr = rr_new;
// 3 = init: rr = rr0 = sum; log[0] = 1.
__global__ __launch_bounds__(256) void k_cg_scalar(const double *__restrict__ red, int nred, double *cg, int op,
                                                   double *log, int it, double rel_tol) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nred; i += 256) acc += red[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double sum = sh[0];
    if (op == 0) {
        cg[0] = sum;
    } else if (op == 3) {
        cg[0] = sum;   // rr
        cg[1] = sum;   // rr0
        cg[5] = (sum > 0.0) ? 0.0 : 1.0;
        if (log) log[0] = (sum > 0.0) ? 1.0 : 0.0;
    } else if (op == 1) {
        cg[2] = sum;
        cg[3] = (cg[5] != 0.0 || !(sum > 0.0)) ? 0.0 : cg[0] / sum;
        if (!(sum > 0.0)) cg[5] = 1.0;
    } else if (op == 2) {
        const bool done = cg[5] != 0.0;
        if (!done) {
            cg[6] = sum;
            cg[4] = (cg[0] > 0.0) ? sum / cg[0] : 0.0;
            const double rel = (cg[1] > 0.0) ? sqrt(sum / cg[1]) : 0.0;
            if (log) log[it] = rel;
            if (rel <= rel_tol) cg[5] = 1.0;
            cg[0] = sum;
        } else {
            cg[4] = 0.0;
            if (log) log[it] = 0.0;
        }
    }
}

// x += alpha p; r -= alpha q; partial |r|^2 (skipped when done)
template <typename R>
__global__ __launch_bounds__(256) void k_cg_update(cplx_t<R> *__restrict__ x, cplx_t<R> *__restrict__ r,
                                                   const cplx_t<R> *__restrict__ p, const cplx_t<R> *__restrict__ q,
                                                   int64_t n, const double *__restrict__ cg, double *__restrict__ red) {
    __shared__ double sh[4];
    const bool done = cg[5] != 0.0;
    const R alpha = (R)cg[3];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        cplx_t<R> rv = r[i];
        if (!done) {
            const cplx_t<R> pv = p[i], qv = q[i];
            cplx_t<R> xv = x[i];
            xv.x = fmaT(alpha, pv.x, xv.x);
            xv.y = fmaT(alpha, pv.y, xv.y);
            rv.x = fmaT(-alpha, qv.x, rv.x);
            rv.y = fmaT(-alpha, qv.y, rv.y);
            x[i] = xv;
            r[i] = rv;
        }
        acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// p = r + beta p (skipped when done)
template <typename R>
__global__ void k_cg_pupdate(cplx_t<R> *__restrict__ p, const cplx_t<R> *__restrict__ r, int64_t n,
                             const double *__restrict__ cg) {
    if (cg[5] != 0.0) return;
    const R beta = (R)cg[4];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const cplx_t<R> rv = r[i], pv = p[i];
        p[i] = mkc<R>(fmaT(beta, pv.x, rv.x), fmaT(beta, pv.y, rv.y));
    }
}

template <typename R>
__global__ void k_copy2(cplx_t<R> *__restrict__ d, const cplx_t<R> *__restrict__ s, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------

static inline unsigned *slot(Handle *h, int i) { return reinterpret_cast<unsigned *>(h->scal.p) + 64 + i; }

static inline double *cgs(Handle *h) { return reinterpret_cast<double *>(h->scal.p); }

static inline int blocks_for(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

static void zero_slots(Handle *h) { k_zero_u32<<<1, 64, 0, h->stream>>>(slot(h, 0), 64); }

// f(R{}) with R the real type of the handle's vectors: double for the FP64 backend.
template <class F>
static void by_prec(const Handle *h, F &&f) {
    if (h->f64) f(double{});
    else f(float{});
}

hipError_t launch_gather(Handle *h, const void *x_full, void *xm) {
    by_prec(h, [&](auto r) {
        using R = decltype(r);
        k_gather<R><<<blocks_for(h->nv_pad), 256, 0, h->stream>>>(
            (const cplx_t<R> *)x_full, (const int *)h->mask_idx.p, (cplx_t<R> *)xm, (int)h->n_vox, (int)h->nv_pad,
            nullptr);
    });
    return hipGetLastError();
}

// Gather that also leaves max|xm| in slot 0 (all slots cleared first): the
// fragment scale of the mixed forward, without a separate pass over xm.
static hipError_t launch_gather_max(Handle *h, const void *x_full, void *xm) {
    if (h->f64) return launch_gather(h, x_full, xm);
    zero_slots(h);
    k_gather<float><<<blocks_for(h->nv_pad), 256, 0, h->stream>>>(
        (const float2 *)x_full, (const int *)h->mask_idx.p, (float2 *)xm, (int)h->n_vox, (int)h->nv_pad, slot(h, 0));
    return hipGetLastError();
}

hipError_t launch_scatter(Handle *h, const void *xm, void *x_full, double scale) {
    by_prec(h, [&](auto r) {
        using R = decltype(r);
        // without a mask every grid entry is written below
        if (h->n_vox != h->n_grid)
            k_fill0<R><<<blocks_for(h->n_grid), 256, 0, h->stream>>>((cplx_t<R> *)x_full, h->n_grid);
        k_scatter<R><<<blocks_for(h->n_vox), 256, 0, h->stream>>>(
            (const cplx_t<R> *)xm, (const int *)h->mask_idx.p, (cplx_t<R> *)x_full, (int)h->n_vox, (R)scale);
    });
    return hipGetLastError();
}

// ---- kernel instances over (L, coil batch): visitors hand the instance to a launcher ----

template <bool FWD, bool HX, int L, bool DT, typename R, class F>
static void visit_direct_cb(int cb, F &&f) {
    if (cb == 8) f(k_direct<FWD, HX, L, 8, DT, R>);
    else if (cb == 16) f(k_direct<FWD, HX, L, 16, DT, R>);
    else if constexpr (std::is_same<R, float>::value) f(k_direct<FWD, HX, L, 32, DT, R>);
}
// FP64 instances batch at most 16 coils (direct_cb), FP32 ones up to 32.
template <bool FWD, bool HX, bool DT = false, typename R = float, class F>
static void visit_direct(int L, int cb, F &&f) {
    switch (L) {
        case 6: visit_direct_cb<FWD, HX, 6, DT, R>(cb, f); break;
        case 10: visit_direct_cb<FWD, HX, 10, DT, R>(cb, f); break;
        case 12: visit_direct_cb<FWD, HX, 12, DT, R>(cb, f); break;
        case 16: visit_direct_cb<FWD, HX, 16, DT, R>(cb, f); break;
        default: visit_direct_cb<FWD, HX, 20, DT, R>(cb, f); break;
    }
}
// shape: coil groups per wave (1, 2) for 32-coil groups, or 3 / 4 for the packed layouts
// of up to 16 / 8 coils (see k_mixed)
template <typename T16, bool FWD, bool HX, int L, bool DT, class F>
static void visit_mixed_nb(int shape, F &&f, bool pf) {
    if constexpr (FWD && HX && !DT) {
        if (pf && shape == 1) {
            f(k_mixed<T16, FWD, HX, L, 1, 0, DT, true>);
            return;
        }
    }
    switch (shape) {
        case 1: f(k_mixed<T16, FWD, HX, L, 1, 0, DT>); break;
        case 3: f(k_mixed<T16, FWD, HX, L, 1, 1, DT>); break;
        case 4: f(k_mixed<T16, FWD, HX, L, 1, 2, DT>); break;
        default: f(k_mixed<T16, FWD, HX, L, 2, 0, DT>); break;
    }
}
template <typename T16, bool FWD, bool HX, bool DT = false, class F>
static void visit_mixed(int L, int shape, F &&f, bool pf = false) {
    switch (L) {
        case 6: visit_mixed_nb<T16, FWD, HX, 6, DT>(shape, f, pf); break;
        case 10: visit_mixed_nb<T16, FWD, HX, 10, DT>(shape, f, pf); break;
        case 12: visit_mixed_nb<T16, FWD, HX, 12, DT>(shape, f, pf); break;
        case 16: visit_mixed_nb<T16, FWD, HX, 16, DT>(shape, f, pf); break;
        default: visit_mixed_nb<T16, FWD, HX, 20, DT>(shape, f, pf); break;
    }
}
template <int L, class F>
static void visit_grad_cb(int cb, F &&f) {
    if (cb == 8) f(k_grad<L, 8>);
    else if (cb == 16) f(k_grad<L, 16>);
    else f(k_grad<L, 32>);
}
template <class F>
static void visit_grad(int L, int cb, F &&f) {
    switch (L) {
        case 6: visit_grad_cb<6>(cb, f); break;
        case 10: visit_grad_cb<10>(cb, f); break;
        case 12: visit_grad_cb<12>(cb, f); break;
        case 16: visit_grad_cb<16>(cb, f); break;
        default: visit_grad_cb<20>(cb, f); break;
    }
}
template <typename T16, bool HX, int L, class F>
static void visit_gmixed_nkc(int nkc, F &&f) {
    (void)nkc;   // 32 coils: 4 chunks (the residual fragments of 4 waves fill 64 KB of LDS)
    f(k_gmixed<T16, HX, L, 4>);
}
template <typename T16, bool HX, class F>
static void visit_gmixed_l(int L, int nkc, F &&f) {
    switch (L) {
        case 6: visit_gmixed_nkc<T16, HX, 6>(nkc, f); break;
        case 10: visit_gmixed_nkc<T16, HX, 10>(nkc, f); break;
        case 12: visit_gmixed_nkc<T16, HX, 12>(nkc, f); break;
        case 16: visit_gmixed_nkc<T16, HX, 16>(nkc, f); break;
        default: visit_gmixed_nkc<T16, HX, 20>(nkc, f); break;
    }
}
// The MFMA gradient kernel of the handle's mixed backend (Horner phases when the runs allow).
template <class F>
static void visit_gmixed(Handle *h, F &&f) {
    const int nkc = h->ncs / 8;
    if (h->backend == HOE_BACKEND_MIXED_FP16) {
        if (h->horner) visit_gmixed_l<_Float16, true>(h->LK, nkc, f);
        else visit_gmixed_l<_Float16, false>(h->LK, nkc, f);
    } else {
        if (h->horner) visit_gmixed_l<__bf16, true>(h->LK, nkc, f);
        else visit_gmixed_l<__bf16, false>(h->LK, nkc, f);
    }
}
// Visit the contraction kernel of a backend and direction (the forward uses
// the Horner variant when the handle's runs are long enough).
template <bool FWD, class F>
static void visit_contract_direct(Handle *h, F &&f) {
    if (h->f64) {
        if (FWD && h->horner) visit_direct<FWD, true, false, double>(h->LK, direct_cb(h, FWD), f);
        else visit_direct<FWD, false, false, double>(h->LK, direct_cb(h, FWD), f);
        return;
    }
    if constexpr (FWD) {
        if (h->dt_on) {
            if (h->horner) visit_direct<true, true, true>(h->LK, direct_cb(h, true), f);
            else visit_direct<true, false, true>(h->LK, direct_cb(h, true), f);
        } else if (h->horner) {
            visit_direct<true, true>(h->LK, direct_cb(h, true), f);
        } else {
            visit_direct<true, false>(h->LK, direct_cb(h, true), f);
        }
    } else {
        visit_direct<false, false>(h->LK, direct_cb(h, false), f);
    }
}
// Packed MFMA coil layout of the mixed backends: 2 for up to 8 coils, 1 for up to 16,
// 0 for groups of 32 coils.
#ifndef HOE_PACK8
#define HOE_PACK8 1
#endif
static inline int mixed_pack(const Handle *h) {
    return (HOE_PACK8 && h->n_cha <= 8) ? 2 : (h->n_cha <= 16 ? 1 : 0);
}
int mixed_nb(const Handle *h);
static inline int mixed_shape(const Handle *h) {
    const int pk = mixed_pack(h);
    return pk ? 2 + pk : mixed_nb(h);
}
// dt: the derivative forward (the schedules are built for the plain forward, whose
// instance has the same occupancy, so that setting kspha_dt keeps them)
template <bool FWD, class F>
static void visit_contract_mixed(Handle *h, int backend, F &&f, bool dt = false) {
    if constexpr (FWD) {
        const int sh = mixed_shape(h);
        // K parts (one per XCD) longer than HOE_DP_MIN_KX units overflow an XCD's L2
        const bool pf = (h->nv_pad / 8) / 8 > HOE_DP_MIN_KX;
        if (backend == HOE_BACKEND_MIXED_FP16) {
            if (dt) {
                if (h->horner) visit_mixed<_Float16, true, true, true>(h->LK, sh, f);
                else visit_mixed<_Float16, true, false, true>(h->LK, sh, f);
            } else if (h->horner) {
                visit_mixed<_Float16, true, true>(h->LK, sh, f, pf);
            } else {
                visit_mixed<_Float16, true, false>(h->LK, sh, f);
            }
        } else {
            if (dt) {
                if (h->horner) visit_mixed<__bf16, true, true, true>(h->LK, sh, f);
                else visit_mixed<__bf16, true, false, true>(h->LK, sh, f);
            } else if (h->horner) {
                visit_mixed<__bf16, true, true>(h->LK, sh, f, pf);
            } else {
                visit_mixed<__bf16, true, false>(h->LK, sh, f);
            }
        }
    } else {
        if (backend == HOE_BACKEND_MIXED_FP16) {
            if (h->horner_adj) visit_mixed<_Float16, false, true>(h->LK, mixed_shape(h), f);
            else visit_mixed<_Float16, false, false>(h->LK, mixed_shape(h), f);
        } else {
            if (h->horner_adj) visit_mixed<__bf16, false, true>(h->LK, mixed_shape(h), f);
            else visit_mixed<__bf16, false, false>(h->LK, mixed_shape(h), f);
        }
    }
}
template <bool FWD, class F>
static void visit_contract(Handle *h, int backend, F &&f) {
    if (backend == HOE_BACKEND_DIRECT || backend == HOE_BACKEND_FP64) visit_contract_direct<FWD>(h, f);
    else visit_contract_mixed<FWD>(h, backend, f);
}

// ---- coil batching ----
// Coils per batch of the FP32/FP64 direct kernels. The FP32 Horner forward (cheap phases,
// the streamed coil row read through scalar loads) takes batches of 16 above 16 coils: a
// 32-coil row (64 SGPRs) leaves no registers to keep the next row's loads in flight, and
// the forward waited on them (200^2, 32 coils, 16 000 samples: 3.96 ms with 32-coil
// batches, 3.17 ms with two 16-coil batches). The adjoint and the gradient, whose phases
// cost more, are faster with one phase per 32 coils (3.11 vs 3.20 ms).
int direct_cb(const Handle *h, bool fwd) {
    if (h->f64) return h->n_cha <= 8 ? 8 : 16;   // 4 FP64 registers per complex accumulator
    if (h->n_cha <= 8) return 8;
    if (h->n_cha <= 16) return 16;
    return (fwd && h->horner && !h->dt_on) ? 16 : 32;
}
int direct_batches(const Handle *h, bool fwd) {
    const int cb = direct_cb(h, fwd);
    return (int)((h->n_cha + cb - 1) / cb);
}
int mixed_nb(const Handle *h) { return h->ncs / 32 >= 2 ? 2 : 1; }
int mixed_groups(const Handle *h) { return (h->ncs / 32 + mixed_nb(h) - 1) / mixed_nb(h); }
// fragment groups of 32 coils, rounded up to whole per-wave batches
static inline int frag_groups(const Handle *h) { return mixed_groups(h) * mixed_nb(h); }

// ---- stream-K schedules ----

static hipError_t upload_ints(DeviceBuf &b, const std::vector<int> &v) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = std::max<size_t>(v.size(), 1) * sizeof(int);
    hipError_t e = hipMalloc(&b.p, b.bytes);
    if (e != hipSuccess) return e;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice);
}

// Host side of a schedule: the arrays build_schedule uploads.
struct SchedPlan {
    int cpx = 0, n_chunks = 0, n_seg = 0;
    std::vector<int> seg, chunk_off, tile_off, tile_list;
};

// Cut tiles [t_lo, t_hi) x n_ks units into 8 K-parts (one per XCD) and each part into
// cpx = slots/8 chunks of equal work; record the segments. When a K part is longer
// than HOE_DP_MIN_KX units and the tiles carry no weights, each chunk first takes
// floor(nt / cpx) whole tiles (the part's full K range each): all chunks of an XCD
// then stream the same K rows at the same time and share them in its L2 (a shorter
// part fits in the L2 whatever the order). The tiles left over, and all tiles
// otherwise, are cut stream-K: into cpx equal runs of consecutive (tile, k) units,
// appended to the chunks.
// kmax > 0 cuts a chunk's run of one tile into near-equal segments of at most kmax
// units. k_w (optional) weights the units of each K index; the K parts and the
// stream-K runs then hold equal weighted work instead of equal unit counts. tile_off
// spans all `tiles` (tiles outside the range have no segments).
static void plan_schedule(SchedPlan &sc, int tiles, int n_ks, int slots, int kmax, const std::vector<double> *tile_w,
                          const std::vector<double> *k_w, int t_lo, int t_hi) {
    const int X = 8;
    sc.cpx = std::max(1, slots / X);
    const bool tweighted = tile_w && (int)tile_w->size() == tiles;
    const bool dp = !tweighted && n_ks / X > HOE_DP_MIN_KX;
    const int m_dp = dp ? (t_hi - t_lo) / sc.cpx : 0;   // whole tiles per chunk
    const int t_sk = t_lo + m_dp * sc.cpx;                      // first stream-K tile
    const int nt = t_hi - t_sk;
    sc.n_chunks = sc.cpx * X;
    std::vector<int> &seg = sc.seg, &chunk_off = sc.chunk_off;
    seg.clear();
    chunk_off.assign((size_t)sc.n_chunks + 1, 0);
    const bool kw = k_w && (int)k_w->size() == n_ks;
    std::vector<double> kcum;   // prefix sums of the K weights
    std::vector<int> kb(X + 1);
    if (kw) {
        kcum.assign((size_t)n_ks + 1, 0.0);
        for (int k = 0; k < n_ks; ++k) kcum[(size_t)k + 1] = kcum[(size_t)k] + (*k_w)[(size_t)k];
        for (int x = 0; x <= X; ++x) {
            const double target = kcum[(size_t)n_ks] * x / X;
            const int k = (int)(std::lower_bound(kcum.begin(), kcum.end(), target) - kcum.begin());
            kb[x] = x == 0 ? 0 : (x == X ? n_ks : std::max(kb[x - 1], std::min(k, n_ks)));
        }
    } else {
        for (int x = 0; x <= X; ++x) kb[x] = (int)((int64_t)x * n_ks / X);
    }
    for (int x = 0; x < X; ++x) {
        const int K0 = kb[x], K1 = kb[x + 1];
        const int Kx = K1 - K0;
        const int64_t units = (int64_t)nt * Kx;
        // first stream-K unit of each chunk (ub[cpx] = units), counted from tile t_sk
        std::vector<int64_t> ub((size_t)sc.cpx + 1, units);
        if (tweighted) {
            std::vector<double> cum((size_t)nt + 1, 0.0);
            for (int t = 0; t < nt; ++t) cum[(size_t)t + 1] = cum[(size_t)t] + (*tile_w)[(size_t)(t_sk + t)] * Kx;
            int t = 0;
            for (int c = 0; c < sc.cpx; ++c) {
                const double target = cum[(size_t)nt] * c / sc.cpx;
                while (t + 1 < nt && cum[(size_t)t + 1] <= target) ++t;
                const int64_t k = std::min<int64_t>(
                    Kx, (int64_t)std::ceil((target - cum[(size_t)t]) / (*tile_w)[(size_t)(t_sk + t)]));
                ub[(size_t)c] = std::max<int64_t>(c ? ub[(size_t)c - 1] : 0, (int64_t)t * Kx + std::max<int64_t>(0, k));
            }
        } else if (kw && Kx > 0 && nt > 0) {
            const double wt = kcum[(size_t)K1] - kcum[(size_t)K0];   // weighted work of one tile in the part
            for (int c = 0; c < sc.cpx; ++c) {
                const double target = wt * nt * c / sc.cpx;
                const int t = std::min(nt - 1, (int)(target / wt));
                const double rem = kcum[(size_t)K0] + (target - wt * t);
                const int k = (int)(std::lower_bound(kcum.begin() + K0, kcum.begin() + K1 + 1, rem) - kcum.begin());
                ub[(size_t)c] = std::max<int64_t>(c ? ub[(size_t)c - 1] : 0, (int64_t)t * Kx + (k - K0));
            }
        } else {
            const int64_t U = (units + sc.cpx - 1) / sc.cpx;
            for (int c = 0; c < sc.cpx; ++c) ub[(size_t)c] = std::min(units, (int64_t)c * U);
        }
        // segments of tile t over K0 + [kk, kk + len), cut at kmax
        auto run = [&](int t, int kk, int len) {
            const int pieces = kmax > 0 ? (len + kmax - 1) / kmax : 1;
            for (int p = 0; p < pieces; ++p) {
                seg.push_back(t);
                seg.push_back(K0 + kk + (int)((int64_t)len * p / pieces));
                seg.push_back(K0 + kk + (int)((int64_t)len * (p + 1) / pieces));
                seg.push_back(0);
            }
        };
        for (int c = 0; c < sc.cpx; ++c) {
            chunk_off[x * sc.cpx + c] = (int)(seg.size() / 4);
            if (Kx > 0)
                for (int j = 0; j < m_dp; ++j) run(t_lo + c * m_dp + j, 0, Kx);
            int64_t u = ub[(size_t)c];
            const int64_t u1 = ub[(size_t)c + 1];
            while (u < u1) {
                const int t = (int)(u / Kx), kk = (int)(u % Kx);
                const int len = (int)std::min<int64_t>(u1 - u, Kx - kk);
                run(t_sk + t, kk, len);
                u += len;
            }
        }
    }
    sc.n_seg = (int)(seg.size() / 4);
    chunk_off[sc.n_chunks] = sc.n_seg;
    std::vector<int> &tile_off = sc.tile_off, &tile_list = sc.tile_list;
    tile_off.assign((size_t)tiles + 1, 0);
    tile_list.assign((size_t)sc.n_seg, 0);
    for (int i = 0; i < sc.n_seg; ++i) tile_off[seg[4 * i] + 1]++;
    for (int t = 0; t < tiles; ++t) tile_off[t + 1] += tile_off[t];
    std::vector<int> fill(tile_off.begin(), tile_off.end() - 1);
    for (int i = 0; i < sc.n_seg; ++i) tile_list[fill[seg[4 * i]]++] = i;
}

static hipError_t upload_plan(Sched &sc, const SchedPlan &pl, int tiles, int n_ks, int t_lo, int t_hi) {
    sc.tiles = tiles;
    sc.n_ks = n_ks;
    sc.t_lo = t_lo;
    sc.t_hi = t_hi;
    sc.cpx = pl.cpx;
    sc.n_chunks = pl.n_chunks;
    sc.n_seg = pl.n_seg;
    sc.max_tile_segs = 0;
    for (int t = 0; t + 1 < (int)pl.tile_off.size(); ++t)
        sc.max_tile_segs = std::max(sc.max_tile_segs, pl.tile_off[(size_t)t + 1] - pl.tile_off[(size_t)t]);
    hipError_t e;
    if ((e = upload_ints(sc.seg, pl.seg)) != hipSuccess) return e;
    if ((e = upload_ints(sc.chunk_off, pl.chunk_off)) != hipSuccess) return e;
    if ((e = upload_ints(sc.tile_off, pl.tile_off)) != hipSuccess) return e;
    if ((e = upload_ints(sc.tile_list, pl.tile_list)) != hipSuccess) return e;
    sc.built = true;
    return hipSuccess;
}

// One schedule over all tiles.
hipError_t build_schedule(Sched &sc, int tiles, int n_ks, int slots, int kmax,
                          const std::vector<double> *tile_w, const std::vector<double> *k_w) {
    SchedPlan pl;
    plan_schedule(pl, tiles, n_ks, slots, kmax, tile_w, k_w, 0, tiles);
    return upload_plan(sc, pl, tiles, n_ks, 0, tiles);
}

void drop_passes(std::vector<Sched> &passes) {
    for (Sched &sc : passes)
        for (DeviceBuf *b : {&sc.seg, &sc.chunk_off, &sc.tile_off, &sc.tile_list}) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->bytes = 0;
        }
    passes.clear();
}

// Schedules of one contraction as passes over contiguous tile ranges, as few as keep
// each pass's partial sums (n_seg * seg_bytes) within `budget` (one tile per pass at
// most). Every pass is scheduled over `slots` workgroups, so each still fills the
// device; the passes run one after another and reuse one partial-sum buffer. When
// plan_schedule gives chunks whole tiles, the passes are made of whole blocks of cpx
// tiles (the tiles left over join the last pass), so that every pass keeps that
// order; otherwise, or when a one-block pass is still too large, the ranges hold
// near-equal (weighted) work. tb receives the P + 1 tile bounds.
static void plan_passes(std::vector<SchedPlan> &plans, std::vector<int> &tb, int tiles, int n_ks, int slots,
                        int kmax, const std::vector<double> *tile_w, const std::vector<double> *k_w,
                        int64_t seg_bytes, int64_t budget) {
    const bool tw = tile_w && (int)tile_w->size() == tiles;
    auto try_bounds = [&]() {
        const int P = (int)tb.size() - 1;
        plans.assign((size_t)P, SchedPlan());
        int64_t worst = 0;
        for (int p = 0; p < P; ++p) {
            plan_schedule(plans[(size_t)p], tiles, n_ks, slots, kmax, tile_w, k_w, tb[(size_t)p], tb[(size_t)p + 1]);
            worst = std::max(worst, (int64_t)plans[(size_t)p].n_seg * seg_bytes);
        }
        return worst;
    };
    const int cpx = std::max(1, slots / 8);
    const int nblk = (tw || n_ks / 8 <= HOE_DP_MIN_KX) ? 0 : tiles / cpx;
    int last_bp = 0;
    for (int P = 1; P <= nblk; ++P) {
        const int bp = (nblk + P - 1) / P;   // most blocks in one pass
        if (bp == last_bp) continue;
        last_bp = bp;
        tb.assign((size_t)P + 1, tiles);
        for (int p = 0; p < P; ++p) tb[(size_t)p] = cpx * (int)((int64_t)nblk * p / P);
        if (try_bounds() <= budget) return;
    }
    std::vector<double> cum((size_t)tiles + 1, 0.0);
    for (int t = 0; t < tiles; ++t) cum[(size_t)t + 1] = cum[(size_t)t] + (tw ? (*tile_w)[(size_t)t] : 1.0);
    int P = std::min(tiles, nblk + 1);
    for (;;) {
        tb.assign((size_t)P + 1, tiles);
        tb[0] = 0;
        for (int p = 1; p < P; ++p) {
            const double target = cum[(size_t)tiles] * p / P;
            const int t = (int)(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
            tb[(size_t)p] = std::min(tiles - (P - p), std::max(tb[(size_t)p - 1] + 1, t));
        }
        const int64_t worst = try_bounds();
        if (worst <= budget || P >= tiles) return;
        // at least 25 % more passes per step, so that planning ends after O(log tiles) steps
        P = (int)std::min<int64_t>(tiles, std::max<int64_t>(P + std::max(1, P / 4), (P * worst + budget - 1) / budget));
    }
}

static hipError_t build_passes(std::vector<Sched> &out, int tiles, int n_ks, int slots, int kmax,
                               const std::vector<double> *tile_w, const std::vector<double> *k_w,
                               int64_t seg_bytes, int64_t budget) {
    drop_passes(out);
    std::vector<SchedPlan> plans;
    std::vector<int> tb;
    plan_passes(plans, tb, tiles, n_ks, slots, kmax, tile_w, k_w, seg_bytes, budget);
    out.resize(plans.size());
    for (size_t p = 0; p < plans.size(); ++p) {
        hipError_t e = upload_plan(out[p], plans[p], tiles, n_ks, tb[p], tb[p + 1]);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static int64_t part_budget() {
    const char *v = std::getenv("HOE_PART_BUDGET_MIB");
    if (v && *v) {
        char *end = nullptr;
        const long long mib = std::strtoll(v, &end, 10);
        if (end && *end == '\0' && mib > 0 && mib < ((long long)1 << 40)) return (int64_t)mib << 20;
    }
    return HOE_PART_BUDGET;
}

static int max_segments(const std::vector<Sched> &passes) {
    int m = 0;
    for (const Sched &sc : passes) m = std::max(m, sc.n_seg);
    return m;
}

static hipError_t grow(DeviceBuf &b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return hipSuccess;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess) b.bytes = bytes;
    return e;
}

template <class K>
static int blocks_per_cu(K kern) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

hipError_t prepare_backend(Handle *h, int backend) {
    std::vector<Sched> &sf = h->sched_f[backend], &sa = h->sched_a[backend];
    if (!sf.empty() && !sa.empty()) return hipSuccess;
    const bool is_mixed = backend == HOE_BACKEND_MIXED_FP16 || backend == HOE_BACKEND_MIXED_BF16;
    const int groups_f = is_mixed ? mixed_groups(h) : direct_batches(h, true);
    const int groups_a = is_mixed ? mixed_groups(h) : direct_batches(h, false);
    int occ_f = 1, occ_a = 1;
    visit_contract<true>(h, backend, [&](auto k) { occ_f = blocks_per_cu(k); });
    visit_contract<false>(h, backend, [&](auto k) { occ_a = blocks_per_cu(k); });
    const int slots_f = std::max(8, h->cu_count * occ_f / groups_f);
    const int slots_a = std::max(8, h->cu_count * occ_a / groups_a);
    hipError_t e;
    const int kmax_a = is_mixed ? HOE_MFMA_ADJ_KMAX : 0;
    // forward segments hold every coil; adjoint segments one coil-combined value per group
    const int64_t seg_bytes = (int64_t)256 * h->ncs * h->cbytes, budget = part_budget();
    const int64_t seg_bytes_a = (int64_t)256 * groups_a * h->cbytes;
    // the Horner forward (MFMA and FP32) weighs its run-crossing K units
    const bool fweighted = h->horner && !(h->dt_on && !is_mixed) &&
                           std::any_of(h->fwd_k_w.begin(), h->fwd_k_w.end(), [](double w) { return w != 1.0; });
    if ((e = build_passes(sf, (int)(h->ns_pad / 256), (int)(h->nv_pad / 8), slots_f, 0, nullptr,
                          fweighted ? &h->fwd_k_w : nullptr, seg_bytes, budget)) != hipSuccess)
        return e;
    const bool weighted = is_mixed &&
                          std::any_of(h->adj_tile_w.begin(), h->adj_tile_w.end(), [](double w) { return w != 1.0; });
    if ((e = build_passes(sa, (int)(h->nv_pad / 256), (int)(h->ns_pad / 8), slots_a, kmax_a,
                          weighted ? &h->adj_tile_w : nullptr, nullptr, seg_bytes_a, budget)) != hipSuccess)
        return e;
    h->occ_f[backend] = occ_f;
    h->occ_a[backend] = occ_a;
    if ((e = grow(h->part_f, (size_t)max_segments(sf) * seg_bytes)) != hipSuccess) return e;
    return grow(h->part_a, (size_t)max_segments(sa) * seg_bytes_a);
}

int adj_groups(const Handle *h, int backend) {
    const bool is_mixed = backend == HOE_BACKEND_MIXED_FP16 || backend == HOE_BACKEND_MIXED_BF16;
    return is_mixed ? mixed_groups(h) : direct_batches(h, false);
}

hipError_t prepare_grad(Handle *h) {
    Sched &sg = h->sched_g;
    if (sg.built) return hipSuccess;
    int occ = 1;
    visit_grad(h->LK, direct_cb(h, false), [&](auto k) { occ = blocks_per_cu(k); });
    const int nb = direct_batches(h, false);
    const int slots = std::max(8, h->cu_count * occ / nb);
    hipError_t e = build_schedule(sg, (int)(h->ns_pad / 256), (int)(h->nv_pad / 8), slots, 0);
    if (e != hipSuccess) return e;
    return grow(h->part_g, (size_t)sg.n_seg * nb * 256 * h->LK * sizeof(double));
}

// residual fragments (waves x chunks x (tile, hi/lo) x lanes x uint4) and one voxel block's
// Z fragments (chunks x 4 x lanes x uint4)
static constexpr int kGradLds = 4 * 4 * 4 * 64 * 16 + 4 * 4 * 64 * 16;

bool grad_on_mfma(const Handle *h) { return h->backend != HOE_BACKEND_DIRECT && h->ncs == 32; }

static hipError_t prepare_grad_mfma(Handle *h) {
    hipError_t e;
    Sched &sg = h->sched_gm;
    if (!sg.built) {
        int occ = 1;
        hipError_t ea = hipSuccess;
        visit_gmixed(h, [&](auto k) {
            ea = hipFuncSetAttribute(reinterpret_cast<const void *>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     kGradLds);
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 256, kGradLds) == hipSuccess && nb > 0) occ = nb;
        });
        if (ea != hipSuccess) return ea;
        const int slots = std::max(8, h->cu_count * occ);
        if ((e = build_schedule(sg, (int)(h->ns_pad / 256), (int)(h->nv_pad / 32), slots, 0)) != hipSuccess) return e;
        if ((e = grow(h->part_gm, (size_t)sg.n_seg * 256 * 32 * sizeof(float))) != hipSuccess) return e;
        const size_t nkc = (size_t)h->ncs / 8;
        if ((e = grow(h->gzfrag, (size_t)(h->nv_pad / 32) * nkc * 64 * 64)) != hipSuccess) return e;
        if ((e = grow(h->grfrag, (size_t)(h->ns_pad / 32) * nkc * 64 * 32)) != hipSuccess) return e;
        if ((e = grow(h->gbfrag, (size_t)(h->nv_pad / 32) * 2 * 64 * 32)) != hipSuccess) return e;
        if ((e = grow(h->gbscale, 32 * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMemcpy(h->gbscale.p, h->gbscale_h.data(), 32 * sizeof(float), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return e;
    }
    if (h->gb_type != h->backend) {
        const int n_kb = (int)(h->nv_pad / 32);
        const int64_t nt = (int64_t)n_kb * 2 * 64;
        if (h->backend == HOE_BACKEND_MIXED_FP16)
            k_build_gbfrag<_Float16><<<blocks_for(nt), 256, 0, h->stream>>>(
                (const double *)h->basis.p, h->sb, h->t_index, (const float *)h->gbscale.p, (uint4 *)h->gbfrag.p,
                n_kb, (int)h->n_vox);
        else
            k_build_gbfrag<__bf16><<<blocks_for(nt), 256, 0, h->stream>>>(
                (const double *)h->basis.p, h->sb, h->t_index, (const float *)h->gbscale.p, (uint4 *)h->gbfrag.p,
                n_kb, (int)h->n_vox);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        h->gb_type = h->backend;
    }
    return hipSuccess;
}

// ---- contraction launches ----

// Launch of a k_direct instance; the complex type of its operand and partial sums follows
// from the instance (float2 or double2).
template <typename C>
static void launch_direct(void (*k)(const double *, int, const double *, int, const C *, C *, int, int, const int4 *,
                                    const int *, int, HornerArgs, const double *, const C *),
                          dim3 grid, hipStream_t st, const double *coef, int sk, const double *basis, int sb,
                          const void *operand, void *part, int ncs, int n_stream, const int4 *segs, const int *coff,
                          int cpx, HornerArgs hx, const double *dtrec, const void *csm) {
    k<<<grid, 256, 0, st>>>(coef, sk, basis, sb, (const C *)operand, (C *)part, ncs, n_stream, segs, coff, cpx, hx,
                            dtrec, (const C *)csm);
}

template <bool FWD>
static void launch_contract(Handle *h, const Sched &sc, const void *operand, void *part) {
    const int backend = h->backend;
    const double *coef = (const double *)h->coef.p, *basis = (const double *)h->basis.p;
    const int4 *segs = (const int4 *)sc.seg.p;
    const int *coff = (const int *)sc.chunk_off.p;
    HornerArgs hx{(const double2 *)h->hrec.p, (const int *)h->vrun.p, (const double *)h->beta.p, h->t_index,
                  (const HoeUnit *)h->units.p, (const HoeUnitVox *)h->uvox.p, (const HoeAdjAnchor *)h->aanc.p,
                  (const float2 *)h->aud.p, h->anchored ? 1 : 0, (const double *)h->dtrec.p, h->dt_w,
                  (const float *)((const unsigned char *)h->abound.p + 256)};
    if (backend == HOE_BACKEND_DIRECT || backend == HOE_BACKEND_FP64) {
        dim3 grid(sc.n_chunks, direct_batches(h, FWD));
        const int n_stream = (int)(FWD ? h->n_vox : h->n_sam);
        visit_contract_direct<FWD>(h, [&](auto k) {
            launch_direct(k, grid, h->stream, coef, h->sk, basis, h->sb, operand, part, h->ncs, n_stream, segs, coff,
                          sc.cpx, hx, (const double *)h->dtrec.p, h->csm.p);
        });
    } else {
        dim3 grid(sc.n_chunks, mixed_groups(h));
        visit_contract_mixed<FWD>(h, backend, [&](auto k) {
            k<<<grid, 256, 0, h->stream>>>(coef, h->sk, basis, h->sb, (const uint4 *)operand, sc.n_ks,
                                           (float2 *)part, h->ncs, segs, coff, sc.cpx, hx, (const float2 *)h->csm.p);
        }, FWD && h->dt_on);
    }
}

// The FP32 direct kernels serve the direct backend, the derivative product and (in FP64)
// the FP64 backend: no MFMA fragments.
static inline bool direct_path(int backend) {
    return backend == HOE_BACKEND_DIRECT || backend == HOE_BACKEND_FP64;
}

// Contraction passes of one direction, each followed by its reduction red(pass); the
// events (when timed) bracket the contractions, with the reductions of all passes but
// the last in between.
template <bool FWD, class Red>
static hipError_t contract_passes(Handle *h, const void *operand, void *part, bool time_it, int ev0, Red red) {
    const std::vector<Sched> &passes = FWD ? h->sched_f[h->backend] : h->sched_a[h->backend];
    if (time_it) (void)hipEventRecord(h->ev[ev0], h->stream);
    for (size_t p = 0; p < passes.size(); ++p) {
        launch_contract<FWD>(h, passes[p], operand, part);
        if (time_it && p + 1 == passes.size()) (void)hipEventRecord(h->ev[ev0 + 1], h->stream);
        const hipError_t e = red(passes[p]);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

// Builds the forward operand from masked xm (device) and runs the forward
// contraction into part_f and its reduction red(pass), pass by pass. have_max: slot 0
// already holds max|xm| (gathered).
template <class Red>
static hipError_t forward_from_xm(Handle *h, const void *xmv, bool time_it, bool have_max, Red red) {
    if (direct_path(h->backend)) {
        const int64_t nz = h->nv_pad * h->ncs;
        by_prec(h, [&](auto r) {
            using R = decltype(r);
            k_build_z<R><<<blocks_for(nz), 256, 0, h->stream>>>((const cplx_t<R> *)h->csm.p, (const cplx_t<R> *)xmv,
                                                                (cplx_t<R> *)h->zbuf.p, (int)h->nv_pad, h->ncs);
        });
        return contract_passes<true>(h, h->zbuf.p, h->part_f.p, time_it, 0, red);
    } else {
        const float2 *xm = (const float2 *)xmv;
        const int n_ks = (int)(h->nv_pad / 8);
        const int n_gb = frag_groups(h);
        if (!have_max) {
            zero_slots(h);
            k_absmax2<<<blocks_for(h->nv_pad), 256, 0, h->stream>>>(xm, (int)h->nv_pad, slot(h, 0));
        }
        const int64_t nt = (int64_t)n_gb * n_ks * 64;
        auto build = [&](auto k) {
            k<<<blocks_for(nt), 256, 0, h->stream>>>((const float2 *)h->csm.p, xm, (uint4 *)h->zfrag.p, n_ks, h->ncs,
                                                      n_gb, slot(h, 0), h->csm_max);
        };
        const int pk = mixed_pack(h);
        if (h->backend == HOE_BACKEND_MIXED_FP16)
            build(pk == 2 ? k_build_zfrag<_Float16, 2> : (pk == 1 ? k_build_zfrag<_Float16, 1> : k_build_zfrag<_Float16, 0>));
        else
            build(pk == 2 ? k_build_zfrag<__bf16, 2> : (pk == 1 ? k_build_zfrag<__bf16, 1> : k_build_zfrag<__bf16, 0>));
        return contract_passes<true>(h, h->zfrag.p, h->part_f.p, time_it, 0, red);
    }
}

hipError_t run_normal_full(Handle *h, const void *x_full, void *xm_out, double scale, bool time_it, int nv_write) {
    hipError_t e = launch_gather_max(h, x_full, h->xm.p);
    if (e != hipSuccess) return e;
    return run_normal_masked(h, h->xm.p, xm_out, scale, 0.0, time_it, true, nv_write);
}

static inline bool mixed(Handle *h) { return !direct_path(h->backend); }

// Segment slices of the forward reduction's blocks (1, 2 or 4 x 256 threads): about 3
// segments per thread, counting the lanes that already split a coil's segments when the
// operator has 16 coils or fewer.
static int fwd_reduce_slices(const Handle *h, const Sched &sc) {
    const int nsub = 32 / (h->n_cha > 16 ? 32 : (h->n_cha <= 1 ? 1 : 1 << (32 - __builtin_clz((unsigned)h->n_cha - 1))));
    int H = 1;
    while (H < 4 && H * nsub * 3 < sc.max_tile_segs) H <<= 1;
    return H;
}

// Reduction of one forward pass (its 32-row blocks of samples) into the user signal
// layout (complex [c][s]) with scale.
static hipError_t launch_reduce_fwd_user(Handle *h, const Sched &sc, void *y_out, double scale) {
    const bool unscale = !direct_path(h->backend);
    // the MFMA derivative forward's A operand carries the power of two dt_sigma
    if (unscale && h->dt_on) scale /= h->dt_sigma;
    dim3 grid((unsigned)((sc.t_hi - sc.t_lo) * 8), (unsigned)(h->ncs / 32));
    const int thr = 256 * fwd_reduce_slices(h, sc);
    by_prec(h, [&](auto r) {
        using R = decltype(r);
        k_reduce_fwd<R><<<grid, thr, 0, h->stream>>>(
            (const cplx_t<R> *)h->part_f.p, (const int *)sc.tile_off.p, (const int *)sc.tile_list.p, h->ncs,
            (int)h->n_sam, (int)h->n_cha, (R)scale, unscale ? slot(h, 0) : nullptr, h->csm_max, 0,
            (cplx_t<R> *)y_out, nullptr, nullptr, nullptr, sc.t_lo * 256);
    });
    return hipGetLastError();
}

// Reduction of one forward pass into ybuf (coil fastest) with scale, optional weights
// dvec and optional subtraction of data (user layout).
static hipError_t launch_reduce_fwd_ybuf(Handle *h, const Sched &sc, double scale, bool weights,
                                         const void *y_data) {
    dim3 grid((unsigned)((sc.t_hi - sc.t_lo) * 8), (unsigned)(h->ncs / 32));
    const int thr = 256 * fwd_reduce_slices(h, sc);
    by_prec(h, [&](auto r) {
        using R = decltype(r);
        k_reduce_fwd<R><<<grid, thr, 0, h->stream>>>(
            (const cplx_t<R> *)h->part_f.p, (const int *)sc.tile_off.p, (const int *)sc.tile_list.p, h->ncs,
            (int)h->n_sam, (int)h->n_cha, (R)scale, mixed(h) ? slot(h, 0) : nullptr, h->csm_max, 1,
            (cplx_t<R> *)h->ybuf.p, weights ? (const R *)h->dvec.p : nullptr, (const cplx_t<R> *)y_data,
            h->f64 ? nullptr : slot(h, 1), sc.t_lo * 256);
    });
    return hipGetLastError();
}

hipError_t run_forward_user(Handle *h, const void *x_full, void *y_out, double scale, bool time_it) {
    hipError_t e = launch_gather_max(h, x_full, h->xm.p);
    if (e != hipSuccess) return e;
    return forward_from_xm(h, h->xm.p, time_it, true,
                           [&](const Sched &sc) { return launch_reduce_fwd_user(h, sc, y_out, scale); });
}

hipError_t run_forward_ybuf(Handle *h, const void *x_full, double scale, bool weights, const void *y_data,
                            bool time_it) {
    hipError_t e = launch_gather_max(h, x_full, h->xm.p);
    if (e != hipSuccess) return e;
    return forward_from_xm(h, h->xm.p, time_it, true, [&](const Sched &sc) {
        return launch_reduce_fwd_ybuf(h, sc, scale, weights, y_data);
    });
}

hipError_t launch_load_ybuf(Handle *h, const void *y_user, bool weights) {
    dim3 grid((unsigned)(h->ns_pad / 32), (unsigned)(h->ncs / 32));
    if (!h->f64) zero_slots(h);
    by_prec(h, [&](auto r) {
        using R = decltype(r);
        k_load_ybuf<R><<<grid, 256, 0, h->stream>>>((const cplx_t<R> *)y_user, weights ? (const R *)h->dvec.p : nullptr,
                                                     (cplx_t<R> *)h->ybuf.p, (int)h->n_sam, (int)h->n_cha,
                                                     (int)h->ns_pad, h->ncs, h->f64 ? nullptr : slot(h, 1));
    });
    return hipGetLastError();
}

static hipError_t adjoint_reduce(Handle *h, const Sched &sc, double scale, void *xm_out, double lambda,
                                 const void *x_in, int nv_write);

// Adjoint contraction from ybuf (slot 1 must hold |ybuf|max for the mixed path) and its
// reduction into xm_out (+ lambda x_in), pass by pass.
static hipError_t adjoint_passes(Handle *h, bool time_it, double scale, void *xm_out, double lambda,
                                 const void *x_in, int nv_write) {
    if (nv_write < 0) nv_write = (int)h->nv_pad;
    auto red = [&](const Sched &sc) { return adjoint_reduce(h, sc, scale, xm_out, lambda, x_in, nv_write); };
    if (direct_path(h->backend)) {
        return contract_passes<false>(h, h->ybuf.p, h->part_a.p, time_it, 2, red);
    } else {
        const int n_ks = (int)(h->ns_pad / 8);
        const int n_gb = frag_groups(h);
        const int64_t nt = (int64_t)n_gb * n_ks * 64;
        auto build = [&](auto k) {
            k<<<blocks_for(nt), 256, 0, h->stream>>>((const float2 *)h->ybuf.p, (uint4 *)h->yfrag.p, n_ks, h->ncs,
                                                      n_gb, slot(h, 1));
        };
        const int pk = mixed_pack(h);
        if (h->backend == HOE_BACKEND_MIXED_FP16)
            build(pk == 2 ? k_build_yfrag<_Float16, 2> : (pk == 1 ? k_build_yfrag<_Float16, 1> : k_build_yfrag<_Float16, 0>));
        else
            build(pk == 2 ? k_build_yfrag<__bf16, 2> : (pk == 1 ? k_build_yfrag<__bf16, 1> : k_build_yfrag<__bf16, 0>));
        return contract_passes<false>(h, h->yfrag.p, h->part_a.p, time_it, 2, red);
    }
}

// Reduction of one adjoint pass (the voxels of its tiles); entries at nv_write and beyond
// are not written (a caller's grid-sized output has no padding).
static hipError_t adjoint_reduce(Handle *h, const Sched &sc, double scale, void *xm_out, double lambda,
                                 const void *x_in, int nv_write) {
    // lanes per voxel: about 4 entries per lane, at most 16 lanes
    const int ngrp = adj_groups(h, h->backend);
    int nl = 1;
    while (nl < 16 && nl * 4 < sc.max_tile_segs * ngrp) nl <<= 1;
    const int64_t threads = (int64_t)(sc.t_hi - sc.t_lo) * 256 * nl;
    by_prec(h, [&](auto r) {
        using R = decltype(r);
        k_reduce_adj<R><<<blocks_for(threads), 256, 0, h->stream>>>(
            (const cplx_t<R> *)h->part_a.p, (const int *)sc.tile_off.p, (const int *)sc.tile_list.p, ngrp, nl,
            (int)h->n_vox, (R)scale, mixed(h) ? slot(h, 1) : nullptr, (cplx_t<R> *)xm_out, (R)lambda,
            (const cplx_t<R> *)x_in, sc.t_lo * 256, sc.t_hi * 256, nv_write);
    });
    return hipGetLastError();
}

hipError_t run_adjoint_contract(Handle *h, double scale, void *xm_out, bool time_it, int nv_write) {
    return adjoint_passes(h, time_it, scale, xm_out, 0.0, nullptr, nv_write);
}

hipError_t run_normal_masked(Handle *h, const void *xm_in, void *xm_out, double scale, double lambda,
                             bool time_it, bool have_max, int nv_write) {
    // slot 0 / slot 1 hold the fragment scales of the mixed path
    if (!have_max && !h->f64) zero_slots(h);
    hipError_t e = forward_from_xm(h, xm_in, time_it, have_max, [&](const Sched &sc) {
        return launch_reduce_fwd_ybuf(h, sc, 1.0, true, nullptr);
    });
    if (e != hipSuccess) return e;
    // FP32 backends: lambda enters the reduction as float, as it always has
    const double lam = h->f64 ? lambda : (double)(float)lambda;
    return adjoint_passes(h, time_it, scale, xm_out, lam, lambda != 0.0 ? xm_in : nullptr, nv_write);
}

// Mixed backends with 32 coils: Z and residual fragments (slot 0 = max|xm|,
// slot 1 = max|r| from the forward and its reduction), k_gmixed, reduction.
static hipError_t run_grad_mfma(Handle *h, bool time_it) {
    hipError_t e = prepare_grad_mfma(h);
    if (e != hipSuccess) return e;
    const Sched &sc = h->sched_gm;
    const int nkc = h->ncs / 8;
    const int n_kb = (int)(h->nv_pad / 32), n_st = (int)(h->ns_pad / 32);
    const int64_t nz = (int64_t)n_kb * nkc * 64, nr = (int64_t)n_st * nkc * 64;
    const unsigned *slots = slot(h, 0);
    if (h->backend == HOE_BACKEND_MIXED_FP16) {
        k_build_gzfrag<_Float16><<<blocks_for(nz), 256, 0, h->stream>>>(
            (const float2 *)h->csm.p, (const float2 *)h->xm.p, (uint4 *)h->gzfrag.p, n_kb, nkc, h->ncs, slots,
            h->csm_max);
        k_build_grfrag<_Float16><<<blocks_for(nr), 256, 0, h->stream>>>(
            (const float2 *)h->ybuf.p, (uint4 *)h->grfrag.p, n_st, nkc, h->ncs, slots, h->csm_max);
    } else {
        k_build_gzfrag<__bf16><<<blocks_for(nz), 256, 0, h->stream>>>(
            (const float2 *)h->csm.p, (const float2 *)h->xm.p, (uint4 *)h->gzfrag.p, n_kb, nkc, h->ncs, slots,
            h->csm_max);
        k_build_grfrag<__bf16><<<blocks_for(nr), 256, 0, h->stream>>>(
            (const float2 *)h->ybuf.p, (uint4 *)h->grfrag.p, n_st, nkc, h->ncs, slots, h->csm_max);
    }
    HornerArgs hx{(const double2 *)h->hrec.p, (const int *)h->vrun.p, (const double *)h->beta.p, h->t_index,
                  (const HoeUnit *)h->units.p, (const HoeUnitVox *)h->uvox.p, (const HoeAdjAnchor *)h->aanc.p,
                  (const float2 *)h->aud.p, h->anchored ? 1 : 0, (const double *)h->dtrec.p, h->dt_w,
                  (const float *)((const unsigned char *)h->abound.p + 256)};
    if (time_it) (void)hipEventRecord(h->ev[0], h->stream);
    visit_gmixed(h, [&](auto k) {
        k<<<sc.n_chunks, 256, kGradLds, h->stream>>>(
            (const double *)h->coef.p, h->sk, (const double *)h->basis.p, h->sb, (const uint4 *)h->gzfrag.p,
            (const uint4 *)h->grfrag.p, (const uint4 *)h->gbfrag.p, (float *)h->part_gm.p, (const int4 *)sc.seg.p,
            (const int *)sc.chunk_off.p, sc.cpx, hx, slots, h->csm_max, h->ncs);
    });
    if (time_it) (void)hipEventRecord(h->ev[1], h->stream);
    const double scale = -4.0 * M_PI / std::sqrt((double)h->n_vox);
    k_reduce_gmixed<<<blocks_for(h->n_sam * h->LK), 256, 0, h->stream>>>(
        (const float *)h->part_gm.p, (const int *)sc.tile_off.p, (const int *)sc.tile_list.p, h->LK, (int)h->n_sam,
        scale, (const float *)h->gbscale.p, slots, h->csm_max, h->ncs, (double *)h->gk.p);
    return hipGetLastError();
}

hipError_t run_grad_contract(Handle *h, bool time_it) {
    if (grad_on_mfma(h)) return run_grad_mfma(h, time_it);
    // direct kernel: zbuf = csm .* xm, ybuf = the residual (coil fastest)
    hipError_t e = launch_build_zbuf(h, h->xm.p);
    if (e != hipSuccess) return e;
    const Sched &sc = h->sched_g;
    const int nb = direct_batches(h, false);
    dim3 grid(sc.n_chunks, nb);
    const double *coef = (const double *)h->coef.p, *basis = (const double *)h->basis.p;
    const float2 *z = (const float2 *)h->zbuf.p, *res = (const float2 *)h->ybuf.p;
    double *part = (double *)h->part_g.p;
    if (time_it) (void)hipEventRecord(h->ev[0], h->stream);
    visit_grad(h->LK, direct_cb(h, false), [&](auto k) {
        k<<<grid, 256, 0, h->stream>>>(coef, h->sk, basis, h->sb, z, res, part, h->ncs, (int)h->n_vox,
                                       (const int4 *)sc.seg.p, (const int *)sc.chunk_off.p, sc.cpx);
    });
    if (time_it) (void)hipEventRecord(h->ev[1], h->stream);
    const double scale = -4.0 * M_PI / std::sqrt((double)h->n_vox);
    k_reduce_grad<<<blocks_for(h->n_sam * h->LK), 256, 0, h->stream>>>(
        part, (const int *)sc.tile_off.p, (const int *)sc.tile_list.p, nb, h->LK, (int)h->n_sam, scale,
        (double *)h->gk.p);
    return hipGetLastError();
}

hipError_t launch_build_zbuf(Handle *h, const void *xm) {
    const int64_t nz = h->nv_pad * h->ncs;
    k_build_z<float><<<blocks_for(nz), 256, 0, h->stream>>>((const float2 *)h->csm.p, (const float2 *)xm,
                                                             (float2 *)h->zbuf.p, (int)h->nv_pad, h->ncs);
    return hipGetLastError();
}

hipError_t launch_build_records(Handle *h, int shift, int m, const double w[4],
                                const double dw[4]) {
    const int64_t n = h->ns_pad * h->sk;
    const double inv = h->dwell > 0.0 ? 1.0 / h->dwell : 0.0;
    k_build_records<<<blocks_for(n), 256, 0, h->stream>>>(
        (const float *)h->traces_d.p, (int)h->n_sam, (int)h->ns_pad, h->L, (const int *)h->act_d.p, (int)h->active.size(),
        (const float *)h->times_d.p, h->sk, shift, m, make_double4(w[0], w[1], w[2], w[3]),
        make_double4(dw[0], dw[1], dw[2], dw[3]), inv, (double *)h->coef.p, (double *)h->dcoef.p);
    // the Horner offset bound of these records (entries 1..na: the active terms)
    unsigned *ob = reinterpret_cast<unsigned *>((unsigned char *)h->abound.p + 256);
    k_zero_u32<<<1, 64, 0, h->stream>>>(ob, 1);
    k_anchor_bound<<<std::max(1, std::min(256, blocks_for(h->n_sam))), 256, 0, h->stream>>>(
        (const double *)h->coef.p, h->sk, (int)h->n_sam, (const double *)h->abound.p, (int)h->active.size() + 1, ob);
    return hipGetLastError();
}

hipError_t launch_grad_chain(Handle *h, double *gtau_dev, bool shift, int m, const double w[4], float *gh_out) {
    const int ns = (int)h->n_sam;
    k_grad_chain<<<blocks_for((int64_t)ns * h->L), 256, 0, h->stream>>>(
        (const double *)h->gk.p, ns, (const int *)h->growmap.p, h->L, shift ? 1 : 0, m, w[0], w[1], w[2], w[3],
        gh_out ? gh_out : (float *)h->ghbuf.p);
    const int nb = 512;
    k_gtau_partial<<<nb, 256, 0, h->stream>>>((const double *)h->gk.p, (const double *)h->dcoef.p, ns,
                                              (int)h->active.size() + 1, h->sk, (double *)h->red.p);
    k_cg_scalar<<<1, 256, 0, h->stream>>>((const double *)h->red.p, nb, gtau_dev, 0, nullptr, 0, 0.0);
    return hipGetLastError();
}

hipError_t launch_norm2(Handle *h, const void *v, int64_t n, double *out_dev) {
    const int nb = 512;
    k_dot_partial<float><<<nb, 256, 0, h->stream>>>((const float2 *)v, nullptr, n, (double *)h->red.p);
    k_cg_scalar<<<1, 256, 0, h->stream>>>((const double *)h->red.p, nb, out_dev, 0, nullptr, 0, 0.0);
    return hipGetLastError();
}

hipError_t cg_run(Handle *h, const void *rhs_m, int iters, double lambda, double rel_tol, void *x_m,
                  double *log_dev) {
    hipError_t err = hipSuccess;
    by_prec(h, [&](auto rr) {
        using R = decltype(rr);
        using C = cplx_t<R>;
        const int64_t n = h->nv_pad;
        C *x = (C *)x_m;
        C *r = (C *)h->cg_vec.p;
        C *p = r + n;
        C *q = p + n;
        double *cg = cgs(h);
        double *red = (double *)h->red.p;
        const int nb = 512;
        const double nscale = 1.0 / (double)h->n_vox;
        k_fill0<R><<<blocks_for(n), 256, 0, h->stream>>>(x, n);
        k_copy2<R><<<blocks_for(n), 256, 0, h->stream>>>(r, (const C *)rhs_m, n);
        k_copy2<R><<<blocks_for(n), 256, 0, h->stream>>>(p, (const C *)rhs_m, n);
        k_dot_partial<R><<<nb, 256, 0, h->stream>>>(r, nullptr, n, red);
        k_cg_scalar<<<1, 256, 0, h->stream>>>(red, nb, cg, 3, log_dev, 0, rel_tol);
        for (int it = 1; it <= iters && err == hipSuccess; ++it) {
            err = run_normal_masked(h, p, q, h->f64 ? nscale : (double)(float)nscale, lambda, false);
            if (err != hipSuccess) break;
            k_dot_partial<R><<<nb, 256, 0, h->stream>>>(p, q, n, red);
            k_cg_scalar<<<1, 256, 0, h->stream>>>(red, nb, cg, 1, log_dev, it, rel_tol);
            k_cg_update<R><<<nb, 256, 0, h->stream>>>(x, r, p, q, n, cg, red);
            k_cg_scalar<<<1, 256, 0, h->stream>>>(red, nb, cg, 2, log_dev, it, rel_tol);
            k_cg_pupdate<R><<<blocks_for(n), 256, 0, h->stream>>>(p, r, n, cg);
        }
    });
    if (err != hipSuccess) return err;
    return hipGetLastError();
}

}  // namespace hoe

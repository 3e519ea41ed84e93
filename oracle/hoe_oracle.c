/*
 * hoe_oracle.c -- CPU restatements of HighOrderMRI.jl's Explicit operator.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (libhoe, the hoe Python
 * package) may link, load or call this file; it is the checker that tests,
 * the smoke test and the CPU-baseline leg of the benchmark use.
 *
 * Two families:
 *  1. FP64 oracle (oracle_*): brute-force sums in double precision, basis and
 *     phase evaluated in FP64 from the FP32 inputs, as the reference's own
 *     independent oracle does (test/test_explicit_gpu.jl:8-35).
 *  2. cpu_ref_*: a restatement of the reference CPU Array path
 *     explicit_array_forward / explicit_array_adjoint
 *     (src/EncodingOperator/Explicit/array_operator.jl:241-348) with its
 *     precision choices: FP32 basis (basisfunction.jl:47-78 in Float32
 *     arithmetic), FP32 phase ("times[p] .* fieldmap' .+ (bf*kspha[:,p])'"),
 *     exp(2*1im*pi*phi) in Complex{Float64} (the literal 2*1im*pi promotes),
 *     FP64 accumulation of the mixed-type product, per-block storage to
 *     ComplexF32, then ./ sqrt(nVox), blocks of nSam/nBlock samples.
 *
 * Complex arrays are interleaved (re, im). Signals are [c][s] (samples
 * fastest), images are [v] over masked voxels, csm is [c][v].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI 6.283185307179586

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---------------- basis ---------------- */

/* FP64 solid harmonics 1..16 from FP32 coordinates, the polynomials of
 * test/test_explicit_gpu.jl:8-13. out: [n][16]. */
void oracle_basis_f64(const float *x, const float *y, const float *z, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double X = x[i], Y = y[i], Z = z[i];
        const double r2 = X * X + Y * Y + Z * Z;
        double *o = out + i * 16;
        o[0] = 1.0;
        o[1] = X;
        o[2] = Y;
        o[3] = Z;
        o[4] = X * Y;
        o[5] = Z * Y;
        o[6] = 3 * Z * Z - r2;
        o[7] = X * Z;
        o[8] = X * X - Y * Y;
        o[9] = 3 * Y * X * X - Y * Y * Y;
        o[10] = X * Y * Z;
        o[11] = (5 * Z * Z - r2) * Y;
        o[12] = 5 * Z * Z * Z - 3 * Z * r2;
        o[13] = (5 * Z * Z - r2) * X;
        o[14] = (X * X - Y * Y) * Z;
        o[15] = X * X * X - 3 * X * Y * Y;
    }
}

/* As oracle_basis_f64 from FP64 coordinates (used to check the polynomials). */
void oracle_basis_f64d(const double *x, const double *y, const double *z, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double X = x[i], Y = y[i], Z = z[i];
        const double r2 = X * X + Y * Y + Z * Z;
        double *o = out + i * 16;
        o[0] = 1.0;
        o[1] = X;
        o[2] = Y;
        o[3] = Z;
        o[4] = X * Y;
        o[5] = Z * Y;
        o[6] = 3 * Z * Z - r2;
        o[7] = X * Z;
        o[8] = X * X - Y * Y;
        o[9] = 3 * Y * X * X - Y * Y * Y;
        o[10] = X * Y * Z;
        o[11] = (5 * Z * Z - r2) * Y;
        o[12] = 5 * Z * Z * Z - 3 * Z * r2;
        o[13] = (5 * Z * Z - r2) * X;
        o[14] = (X * X - Y * Y) * Z;
        o[15] = X * X * X - 3 * X * Y * Y;
    }
}

/* FP32 basis with Julia's Float32 operation order (basisfunction.jl:47-78):
 * literal powers are repeated products, broadcasts associate left. */
void cpu_ref_basis_f32(const float *x, const float *y, const float *z, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) {
        const float X = x[i], Y = y[i], Z = z[i];
        const float x2 = X * X, y2 = Y * Y, z2 = Z * Z;
        const float r2 = (x2 + y2) + z2;
        float *o = out + i * 16;
        o[0] = 1.0f;
        o[1] = X;
        o[2] = Y;
        o[3] = Z;
        o[4] = X * Y;
        o[5] = Z * Y;
        o[6] = 3.0f * z2 - r2;
        o[7] = X * Z;
        o[8] = x2 - y2;
        o[9] = (3.0f * Y) * x2 - (Y * Y * Y);
        o[10] = (X * Z) * Y;
        o[11] = (5.0f * z2 - r2) * Y;
        o[12] = 5.0f * (Z * Z * Z) - (3.0f * Z) * r2;
        o[13] = (5.0f * z2 - r2) * X;
        o[14] = x2 * Z - y2 * Z;
        o[15] = (X * X * X) - (3.0f * X) * y2;
    }
}

/* ---------------- FP64 oracle ---------------- */

/* phi(s,v) = sum_l coef[s][l] * basis[v][l] + times[s]*fieldmap[v] (cycles)
 * basis: [nv][L] (column 0 = 1), coef: [ns][L]. */
static inline double phase64(const double *b, const double *k, int L, double t, double df) {
    double ph = t * df;
    for (int l = 0; l < L; ++l) ph += k[l] * b[l];
    return ph;
}

/* y[c][s] = 1/sqrt(nv) sum_v exp(2 pi i phi) csm[c][v] x[v] */
void oracle_forward_f64(int64_t nv, int64_t ns, int64_t nc, int L, const double *basis, const double *coef,
                        const double *times, const double *fieldmap, const double *csm, const double *x,
                        double *y) {
    const double scale = 1.0 / sqrt((double)nv);
    double *z = (double *)malloc(sizeof(double) * 2 * nv * nc);
    for (int64_t c = 0; c < nc; ++c)
        for (int64_t v = 0; v < nv; ++v) {
            const double ar = csm[2 * (c * nv + v)], ai = csm[2 * (c * nv + v) + 1];
            const double br = x[2 * v], bi = x[2 * v + 1];
            z[2 * (v * nc + c)] = ar * br - ai * bi;
            z[2 * (v * nc + c) + 1] = ar * bi + ai * br;
        }
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * 2 * nc);
#pragma omp for schedule(static)
        for (int64_t s = 0; s < ns; ++s) {
            memset(acc, 0, sizeof(double) * 2 * nc);
            for (int64_t v = 0; v < nv; ++v) {
                const double ph = phase64(basis + v * L, coef + s * L, L, times[s], fieldmap[v]);
                const double a = TWO_PI * ph;
                const double er = cos(a), ei = sin(a);
                const double *zv = z + 2 * v * nc;
                for (int64_t c = 0; c < nc; ++c) {
                    acc[2 * c] += er * zv[2 * c] - ei * zv[2 * c + 1];
                    acc[2 * c + 1] += er * zv[2 * c + 1] + ei * zv[2 * c];
                }
            }
            for (int64_t c = 0; c < nc; ++c) {
                y[2 * (c * ns + s)] = acc[2 * c] * scale;
                y[2 * (c * ns + s) + 1] = acc[2 * c + 1] * scale;
            }
        }
        free(acc);
    }
    free(z);
}

/* Derivative product (explicit_array_derivative, array_operator.jl:187-235):
 * y[c][s] = 1/sqrt(nv) sum_v exp(2 pi i phi) * 2 pi i (sum_l kdt[s][l] b[v][l]) * csm[c][v] x[v]
 * (the t * fieldmap term of phi is not differentiated). */
void oracle_forward_dt_f64(int64_t nv, int64_t ns, int64_t nc, int L, const double *basis, const double *coef,
                           const double *kdt, const double *times, const double *fieldmap, const double *csm,
                           const double *x, double *y) {
    const double scale = 1.0 / sqrt((double)nv);
    double *z = (double *)malloc(sizeof(double) * 2 * nv * nc);
    for (int64_t c = 0; c < nc; ++c)
        for (int64_t v = 0; v < nv; ++v) {
            const double ar = csm[2 * (c * nv + v)], ai = csm[2 * (c * nv + v) + 1];
            const double br = x[2 * v], bi = x[2 * v + 1];
            z[2 * (v * nc + c)] = ar * br - ai * bi;
            z[2 * (v * nc + c) + 1] = ar * bi + ai * br;
        }
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * 2 * nc);
#pragma omp for schedule(static)
        for (int64_t s = 0; s < ns; ++s) {
            memset(acc, 0, sizeof(double) * 2 * nc);
            for (int64_t v = 0; v < nv; ++v) {
                const double ph = phase64(basis + v * L, coef + s * L, L, times[s], fieldmap[v]);
                const double psi = phase64(basis + v * L, kdt + s * L, L, 0.0, 0.0);
                const double a = TWO_PI * ph;
                /* 2 pi i psi (cos a + i sin a) = 2 pi psi (-sin a + i cos a) */
                const double er = -TWO_PI * psi * sin(a), ei = TWO_PI * psi * cos(a);
                const double *zv = z + 2 * v * nc;
                for (int64_t c = 0; c < nc; ++c) {
                    acc[2 * c] += er * zv[2 * c] - ei * zv[2 * c + 1];
                    acc[2 * c + 1] += er * zv[2 * c + 1] + ei * zv[2 * c];
                }
            }
            for (int64_t c = 0; c < nc; ++c) {
                y[2 * (c * ns + s)] = acc[2 * c] * scale;
                y[2 * (c * ns + s) + 1] = acc[2 * c + 1] * scale;
            }
        }
        free(acc);
    }
    free(z);
}

/* x[v] = 1/sqrt(nv) sum_c conj(csm[c][v]) sum_s conj(exp(2 pi i phi)) y[c][s] */
void oracle_adjoint_f64(int64_t nv, int64_t ns, int64_t nc, int L, const double *basis, const double *coef,
                        const double *times, const double *fieldmap, const double *csm, const double *y,
                        double *x) {
    const double scale = 1.0 / sqrt((double)nv);
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * 2 * nc);
#pragma omp for schedule(static)
        for (int64_t v = 0; v < nv; ++v) {
            memset(acc, 0, sizeof(double) * 2 * nc);
            for (int64_t s = 0; s < ns; ++s) {
                const double ph = phase64(basis + v * L, coef + s * L, L, times[s], fieldmap[v]);
                const double a = TWO_PI * ph;
                const double er = cos(a), ei = -sin(a);
                for (int64_t c = 0; c < nc; ++c) {
                    const double yr = y[2 * (c * ns + s)], yi = y[2 * (c * ns + s) + 1];
                    acc[2 * c] += er * yr - ei * yi;
                    acc[2 * c + 1] += er * yi + ei * yr;
                }
            }
            double xr = 0.0, xi = 0.0;
            for (int64_t c = 0; c < nc; ++c) {
                const double sr = csm[2 * (c * nv + v)], si = -csm[2 * (c * nv + v) + 1];
                xr += sr * acc[2 * c] - si * acc[2 * c + 1];
                xi += sr * acc[2 * c + 1] + si * acc[2 * c];
            }
            x[2 * v] = xr * scale;
            x[2 * v + 1] = xi * scale;
        }
        free(acc);
    }
}

/* g[l][s] = dL/dk_l(s) of L = ||E x - y||^2 given the residual r = E x - y ([c][s]):
 *   -(4 pi / sqrt(nv)) Im sum_v b_l(v) exp(2 pi i phi) sum_c conj(r[c][s]) csm[c][v] x[v] */
void oracle_grad_f64(int64_t nv, int64_t ns, int64_t nc, int L, const double *basis, const double *coef,
                     const double *times, const double *fieldmap, const double *csm, const double *x,
                     const double *r, double *g) {
    const double scale = -4.0 * M_PI / sqrt((double)nv);
    double *z = (double *)malloc(sizeof(double) * 2 * nv * nc);
    for (int64_t c = 0; c < nc; ++c)
        for (int64_t v = 0; v < nv; ++v) {
            const double ar = csm[2 * (c * nv + v)], ai = csm[2 * (c * nv + v) + 1];
            const double br = x[2 * v], bi = x[2 * v + 1];
            z[2 * (v * nc + c)] = ar * br - ai * bi;
            z[2 * (v * nc + c) + 1] = ar * bi + ai * br;
        }
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * L);
#pragma omp for schedule(static)
        for (int64_t s = 0; s < ns; ++s) {
            memset(acc, 0, sizeof(double) * L);
            for (int64_t v = 0; v < nv; ++v) {
                const double ph = phase64(basis + v * L, coef + s * L, L, times[s], fieldmap[v]);
                const double a = TWO_PI * ph;
                const double er = cos(a), ei = sin(a);
                double ur = 0.0, ui = 0.0;
                for (int64_t c = 0; c < nc; ++c) {
                    const double rr = r[2 * (c * ns + s)], ri = -r[2 * (c * ns + s) + 1];
                    const double zr = z[2 * (v * nc + c)], zi = z[2 * (v * nc + c) + 1];
                    ur += rr * zr - ri * zi;
                    ui += rr * zi + ri * zr;
                }
                const double m = er * ui + ei * ur;
                for (int l = 0; l < L; ++l) acc[l] += basis[v * L + l] * m;
            }
            for (int l = 0; l < L; ++l) g[l * ns + s] = acc[l] * scale;
        }
        free(acc);
    }
    free(z);
}

/* ---------------- restatement of the reference CPU Array path ---------------- */

/* Sample blocks of explicit_array_forward: nBlock capped at ns, n = ns / nBlock,
 * blocks [n*(i-1), n*i) and a final remainder block (array_operator.jl:129-135). */
static int64_t block_bounds(int64_t ns, int64_t nblock, int64_t *starts) {
    if (nblock > ns) nblock = ns;
    const int64_t n = ns / nblock;
    int64_t cnt = 0;
    for (int64_t i = 0; i < nblock; ++i) starts[cnt++] = n * i;
    if (ns % nblock != 0) starts[cnt++] = n * nblock;
    starts[cnt] = ns;
    return cnt;
}

/* phi in Float32: times[s]*fieldmap[v] + sum_l bf[v][l]*kspha[s][l] (the sgemm sum) */
static inline float phase32(const float *bf, const float *k, int L, float t, float df) {
    float acc = 0.0f;
    for (int l = 0; l < L; ++l) acc = fmaf(bf[l], k[l], acc);
    return t * df + acc;
}

/* y[c][s] (ComplexF32). basis: [nv][L] FP32 (column 0 = 1); kspha: [ns][L] FP32;
 * csm [c][v] and x [v] ComplexF32. */
void cpu_ref_forward(int64_t nv, int64_t ns, int64_t nc, int L, const float *bf, const float *kspha,
                     const float *times, const float *fieldmap, const float *csm, const float *x, float *y,
                     int64_t nblock) {
    /* x .* csm in ComplexF32 */
    float *z = (float *)malloc(sizeof(float) * 2 * nv * nc);
    for (int64_t c = 0; c < nc; ++c)
        for (int64_t v = 0; v < nv; ++v) {
            const float ar = x[2 * v], ai = x[2 * v + 1];
            const float br = csm[2 * (c * nv + v)], bi = csm[2 * (c * nv + v) + 1];
            z[2 * (v * nc + c)] = ar * br - ai * bi;
            z[2 * (v * nc + c) + 1] = ar * bi + ai * br;
        }
    const double sq = sqrt((double)nv);
    (void)nblock; /* every row of a block is an independent dot product */
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * 2 * nc);
#pragma omp for schedule(static)
        for (int64_t s = 0; s < ns; ++s) {
            memset(acc, 0, sizeof(double) * 2 * nc);
            for (int64_t v = 0; v < nv; ++v) {
                const float ph = phase32(bf + v * L, kspha + s * L, L, times[s], fieldmap[v]);
                const double a = TWO_PI * (double)ph;
                const double er = cos(a), ei = sin(a);
                const float *zv = z + 2 * v * nc;
                for (int64_t c = 0; c < nc; ++c) {
                    const double zr = zv[2 * c], zi = zv[2 * c + 1];
                    acc[2 * c] += er * zr - ei * zi;
                    acc[2 * c + 1] += er * zi + ei * zr;
                }
            }
            for (int64_t c = 0; c < nc; ++c) {
                const float r32 = (float)acc[2 * c], i32 = (float)acc[2 * c + 1];
                y[2 * (c * ns + s)] = (float)((double)r32 / sq);
                y[2 * (c * ns + s) + 1] = (float)((double)i32 / sq);
            }
        }
        free(acc);
    }
    free(z);
}

/* x[v] (ComplexF32) = sum_c conj(csm) * (sum_blocks sum_{s in block} conj(e) y) / sqrt(nv) */
void cpu_ref_adjoint(int64_t nv, int64_t ns, int64_t nc, int L, const float *bf, const float *kspha,
                     const float *times, const float *fieldmap, const float *csm, const float *y, float *x,
                     int64_t nblock) {
    int64_t *starts = (int64_t *)malloc(sizeof(int64_t) * (nblock + 3));
    const int64_t nb = block_bounds(ns, nblock, starts);
    const double sq = sqrt((double)nv);
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * 2 * nc);
        double *blk = (double *)malloc(sizeof(double) * 2 * nc);
#pragma omp for schedule(static)
        for (int64_t v = 0; v < nv; ++v) {
            memset(acc, 0, sizeof(double) * 2 * nc);
            for (int64_t b = 0; b < nb; ++b) {
                memset(blk, 0, sizeof(double) * 2 * nc);
                for (int64_t s = starts[b]; s < starts[b + 1]; ++s) {
                    const float ph = phase32(bf + v * L, kspha + s * L, L, times[s], fieldmap[v]);
                    const double a = TWO_PI * (double)ph;
                    const double er = cos(a), ei = -sin(a);
                    for (int64_t c = 0; c < nc; ++c) {
                        const double yr = y[2 * (c * ns + s)], yi = y[2 * (c * ns + s) + 1];
                        blk[2 * c] += er * yr - ei * yi;
                        blk[2 * c + 1] += er * yi + ei * yr;
                    }
                }
                for (int64_t c = 0; c < 2 * nc; ++c) acc[c] += blk[c];
            }
            double xr = 0.0, xi = 0.0;
            for (int64_t c = 0; c < nc; ++c) {
                const double ar = acc[2 * c] / sq, ai = acc[2 * c + 1] / sq;
                const double sr = csm[2 * (c * nv + v)], si = -(double)csm[2 * (c * nv + v) + 1];
                xr += ar * sr - ai * si;
                xi += ar * si + ai * sr;
            }
            x[2 * v] = (float)xr;
            x[2 * v + 1] = (float)xi;
        }
        free(acc);
        free(blk);
    }
    free(starts);
}

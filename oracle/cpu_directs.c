/*
 * cpu_directs.c -- restatement of the reference CPU directs (the CPU fallback
 * of the near-field path), used as the CPU baseline and as a parity target.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's cpu_baseline
 * leg, never by the product library.
 *
 * It is compiled with the reference's CPU flags (common.mk.in:97-109:
 * -Ofast -mprefer-vector-width=512, here -march=x86-64-v4 so the same object
 * runs on any AVX-512 host) and OpenMP, so 1.0f/sqrtf becomes rsqrt14 plus a
 * Newton step exactly as in the reference build, and the source loop is
 * auto-vectorised 16-wide like the production fallback (or_execute_slab_cpu).
 * or_execute_slab_avx512 runs the same cell-pair loop on an explicit AVX-512
 * kernel in the structure of AVX512Direct::compute (avx512direct.cc:36-153).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { OR_PLUMMER = 0, OR_SINGLE_SPLINE = 1, OR_CUBIC_SPLINE = 2, OR_CUBIC_PLUMMER = 3 };

/* grid::WrapSlab, grid.cpp:70-74 */
static inline int wrapi(int i, int cpd)
{
    while (i >= cpd) i -= cpd;
    while (i < 0) i += cpd;
    return i;
}

/* FP32 kernels of the CPU fallback, DirectCPUKernels.cc:53-80 (single
 * spline: branchless FMIN), :108-129 (Plummer).  No density.  Compiled with
 * the reference's -Ofast, 1.0f/sqrtf becomes rsqrt + one Newton step. */
static inline void or_kernel_cpu_f32(int kind, float sx, float sy, float sz,
                                     float kx, float ky, float kz, float eps,
                                     float *ax, float *ay, float *az)
{
    float dx = sx - kx, dy = sy - ky, dz = sz - kz;
    float f;
    if (kind == OR_SINGLE_SPLINE) {
        float dr2 = (dx * dx + dy * dy + dz * dz) * eps + 1e-32f;
        f = 1.0f / sqrtf(dr2);
        f = fminf(f * f * f, (-15.0f * f + 6.0f) * dr2 + 10.0f);
    } else {
        float r = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz + eps);
        f = r * r * r;
    }
    *ax -= f * dx;
    *ay -= f * dy;
    *az -= f * dz;
}

void or_pair_cpu_f32(int kind, const float *src3, const float *sink3, float eps, float *a3)
{
    a3[0] = a3[1] = a3[2] = 0.0f;
    or_kernel_cpu_f32(kind, src3[0], src3[1], src3[2], sink3[0], sink3[1], sink3[2], eps,
                      &a3[0], &a3[1], &a3[2]);
}

/* ------------------------------------------------------------------ */
/* FP32 CPU fallback restatement                                       */
/* ------------------------------------------------------------------ */

/* Direct::Execute (direct.cpp:49-59): acc per sink over one source cell, with
 * the sink shifted by delta, then SA[q] += acc. */
static void or_direct_execute(int kind, const float *sx, const float *sy, const float *sz,
                              int64_t nsrc, const float *kx, const float *ky, const float *kz,
                              int64_t nsink, float dlx, float dly, float dlz, float eps,
                              float *SA /* 3 per sink */)
{
    for (int64_t q = 0; q < nsink; q++) {
        float ax = 0, ay = 0, az = 0;
        float px = dlx + kx[q], py = dly + ky[q], pz = dlz + kz[q];
        for (int64_t p = 0; p < nsrc; p++)
            or_kernel_cpu_f32(kind, sx[p], sy[p], sz[p], px, py, pz, eps, &ax, &ay, &az);
        SA[3 * q + 0] += ax;
        SA[3 * q + 1] += ay;
        SA[3 * q + 2] += az;
    }
}

/* NearFieldDriver::ExecuteSlabCPU (extras_directdriver.cpp:59-122) for slab x,
 * rows [y0, y1): zero the acc, then for each sink cell and each of the
 * (2R+1)^3 neighbours add the Direct::Execute result with
 * delta = CellCenter(sink) - CellCenter(source) (grid.cpp:98-107) formed in
 * double and cast to float.  acc4 is float4 per particle of slab x (w left 0:
 * the CPU path computes no FOF density, extras_directdriver.cpp:100-104).
 * Returns the number of pair interactions done. */
uint64_t or_execute_slab_cpu(int cpd, const float *const *pos, const int64_t *nslab,
                             const int64_t *const *start, const int32_t *const *count, int R,
                             int kind, float eps, int x, int y0, int y1, float *acc4)
{
    const double invcpd = 1.0 / (double)cpd;
    const int cpdhalf = (cpd - 1) / 2;
    uint64_t di = 0;
    const float *K = pos[x];
    int64_t n = nslab[x];
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : di)
    for (int y = y0; y < y1; y++) {
        float *tmp3 = NULL;
        int64_t tmpcap = 0;
        for (int z = 0; z < cpd; z++) {
            int c = y * cpd + z;
            int64_t st = start[x][c];
            int32_t ns = count[x][c];
            if (ns == 0) continue;
            if (tmpcap < ns) {
                free(tmp3);
                tmpcap = ns;
                tmp3 = (float *)malloc(sizeof(float) * 3 * (size_t)ns);
            }
            memset(tmp3, 0, sizeof(float) * 3 * (size_t)ns);
            double cx = (x - cpdhalf) * invcpd, cy = (y - cpdhalf) * invcpd, cz = (z - cpdhalf) * invcpd;
            for (int i = x - R; i <= x + R; i++) {
                int xs = wrapi(i, cpd);
                const float *Q = pos[xs];
                int64_t m = nslab[xs];
                for (int j = y - R; j <= y + R; j++)
                    for (int k = z - R; k <= z + R; k++) {
                        int cs = wrapi(j, cpd) * cpd + wrapi(k, cpd);
                        int64_t ss = start[xs][cs];
                        int32_t nsrc = count[xs][cs];
                        if (nsrc <= 0) continue;
                        float dlx = (float)(cx - (i - cpdhalf) * invcpd);
                        float dly = (float)(cy - (j - cpdhalf) * invcpd);
                        float dlz = (float)(cz - (k - cpdhalf) * invcpd);
                        or_direct_execute(kind, Q + ss, Q + m + ss, Q + 2 * m + ss, nsrc, K + st,
                                          K + n + st, K + 2 * n + st, ns, dlx, dly, dlz, eps, tmp3);
                        di += (uint64_t)ns * (uint64_t)nsrc;
                    }
            }
            for (int32_t q = 0; q < ns; q++) {
                acc4[4 * (st + q) + 0] = tmp3[3 * q + 0];
                acc4[4 * (st + q) + 1] = tmp3[3 * q + 1];
                acc4[4 * (st + q) + 2] = tmp3[3 * q + 2];
                acc4[4 * (st + q) + 3] = 0.0f;
            }
        }
        free(tmp3);
    }
    return di;
}


/* ------------------------------------------------------------------ */
/* Explicit AVX-512 form                                               */
/* ------------------------------------------------------------------ */

#include <immintrin.h>

/* 1/sqrt(x) to ~FP32 accuracy: the 14-bit hardware estimate refined by one
 * Newton-Raphson step, y1 = y0 (3 - x y0^2) / 2 (relative error ~1e-7; the
 * estimate alone errs by up to 2^-14). */
static inline __m512 rsqrt_nr(__m512 x)
{
    __m512 y = _mm512_rsqrt14_ps(x);
    __m512 xyy = _mm512_mul_ps(_mm512_mul_ps(x, y), y);
    return _mm512_mul_ps(_mm512_mul_ps(_mm512_set1_ps(0.5f), y), _mm512_sub_ps(_mm512_set1_ps(3.0f), xyy));
}

/* One sink against nsrc sources, 16 sources per vector, the sink broadcast;
 * the structure of AVX512Direct::compute (avx512direct.cc:36-153): full
 * vectors, then one masked vector for the remainder, then horizontal sums.
 * Single spline: f = min(q^-3/2, (-15 f + 6) q + 10) on q = r^2 eps + 1e-32
 * (avx512direct.cc:73-107); Plummer: (r^2 + eps)^-3/2.  Unlike the
 * reference's form it refines rsqrt14 by one Newton step, so that its
 * accuracy is that of the -Ofast 1.0f/sqrtf of the production fallback. */
static inline void avx512_sink(int kind, const float *sx, const float *sy, const float *sz, int64_t nsrc,
                               float px, float py, float pz, float eps, float *a3)
{
    const __m512 kx = _mm512_set1_ps(px), ky = _mm512_set1_ps(py), kz = _mm512_set1_ps(pz);
    const __m512 ve = _mm512_set1_ps(eps), small = _mm512_set1_ps(1e-32f);
    const __m512 m15 = _mm512_set1_ps(-15.0f), six = _mm512_set1_ps(6.0f), ten = _mm512_set1_ps(10.0f);
    __m512 ax = _mm512_setzero_ps(), ay = _mm512_setzero_ps(), az = _mm512_setzero_ps();
    for (int64_t j = 0; j < nsrc; j += 16) {
        __mmask16 m = nsrc - j >= 16 ? (__mmask16)0xFFFF : (__mmask16)((1u << (nsrc - j)) - 1u);
        __m512 dx = _mm512_sub_ps(_mm512_maskz_loadu_ps(m, sx + j), kx);
        __m512 dy = _mm512_sub_ps(_mm512_maskz_loadu_ps(m, sy + j), ky);
        __m512 dz = _mm512_sub_ps(_mm512_maskz_loadu_ps(m, sz + j), kz);
        __m512 r2 = _mm512_fmadd_ps(dz, dz, _mm512_fmadd_ps(dy, dy, _mm512_mul_ps(dx, dx)));
        __m512 f;
        if (kind == OR_SINGLE_SPLINE) {
            __m512 q = _mm512_fmadd_ps(r2, ve, small);
            __m512 y = rsqrt_nr(q);
            __m512 f3 = _mm512_mul_ps(y, _mm512_mul_ps(y, y));
            __m512 soft = _mm512_fmadd_ps(_mm512_fmadd_ps(m15, y, six), q, ten);
            f = _mm512_min_ps(f3, soft);
        } else {
            __m512 y = rsqrt_nr(_mm512_add_ps(r2, ve));
            f = _mm512_mul_ps(y, _mm512_mul_ps(y, y));
        }
        /* masked lanes keep their accumulators (their loads were zeros) */
        ax = _mm512_mask3_fnmadd_ps(f, dx, ax, m);
        ay = _mm512_mask3_fnmadd_ps(f, dy, ay, m);
        az = _mm512_mask3_fnmadd_ps(f, dz, az, m);
    }
    a3[0] += _mm512_reduce_add_ps(ax);
    a3[1] += _mm512_reduce_add_ps(ay);
    a3[2] += _mm512_reduce_add_ps(az);
}

/* or_execute_slab_cpu's cell-pair loop (ExecuteSlabCPU, extras_directdriver.cpp:
 * 59-122) with the explicit AVX-512 kernel above: per sink cell, per
 * neighbour cell, each sink (shifted by delta) against the cell's sources.
 * Same arguments and result. */
uint64_t or_execute_slab_avx512(int cpd, const float *const *pos, const int64_t *nslab,
                                const int64_t *const *start, const int32_t *const *count, int R,
                                int kind, float eps, int x, int y0, int y1, float *acc4)
{
    const double invcpd = 1.0 / (double)cpd;
    const int cpdhalf = (cpd - 1) / 2;
    uint64_t di = 0;
    const float *K = pos[x];
    int64_t n = nslab[x];
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : di)
    for (int y = y0; y < y1; y++) {
        for (int z = 0; z < cpd; z++) {
            int c = y * cpd + z;
            int64_t st = start[x][c];
            int32_t ns = count[x][c];
            if (ns == 0) continue;
            for (int32_t q = 0; q < ns; q++)
                acc4[4 * (st + q) + 0] = acc4[4 * (st + q) + 1] = acc4[4 * (st + q) + 2] =
                    acc4[4 * (st + q) + 3] = 0.0f;
            double cx = (x - cpdhalf) * invcpd, cy = (y - cpdhalf) * invcpd, cz = (z - cpdhalf) * invcpd;
            for (int i = x - R; i <= x + R; i++) {
                int xs = wrapi(i, cpd);
                const float *Q = pos[xs];
                int64_t m = nslab[xs];
                for (int j = y - R; j <= y + R; j++)
                    for (int k = z - R; k <= z + R; k++) {
                        int cs = wrapi(j, cpd) * cpd + wrapi(k, cpd);
                        int64_t ss = start[xs][cs];
                        int32_t nsrc = count[xs][cs];
                        if (nsrc <= 0) continue;
                        float dlx = (float)(cx - (i - cpdhalf) * invcpd);
                        float dly = (float)(cy - (j - cpdhalf) * invcpd);
                        float dlz = (float)(cz - (k - cpdhalf) * invcpd);
                        for (int32_t q = 0; q < ns; q++) {
                            float a3[3] = {0.0f, 0.0f, 0.0f};
                            avx512_sink(kind, Q + ss, Q + m + ss, Q + 2 * m + ss, nsrc, dlx + K[st + q],
                                        dly + K[n + st + q], dlz + K[2 * n + st + q], eps, a3);
                            acc4[4 * (st + q) + 0] += a3[0];
                            acc4[4 * (st + q) + 1] += a3[1];
                            acc4[4 * (st + q) + 2] += a3[2];
                        }
                        di += (uint64_t)ns * (uint64_t)nsrc;
                    }
            }
        }
    }
    return di;
}


#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads the OpenMP loops of this library use. */
int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

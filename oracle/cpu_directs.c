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
 * auto-vectorised 16-wide like the production fallback.
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

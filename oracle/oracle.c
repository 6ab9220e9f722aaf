/*
 * oracle.c -- CPU restatement of Abacus's near-field direct force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and never by the product library.  Nothing here is shipped or
 * measured as the GPU path.
 *
 * Contents (each function cites the reference file:line it restates):
 *   or_kernel_gpu_f32   the FP32 pair kernels of DirectGPUKernels.cuh, with
 *                       the FOF density term (direct_density<0>)
 *   or_kernel_f64       the same forms evaluated in FP64 (no guards)
 *   or_direct_f64       FP64 brute-force direct sum over the (2R+1)^3 cell
 *                       neighbourhood on the exact FP32 cell-relative inputs
 *   or_floor_f32        the same sum of FP32 pair terms formed as the batched
 *                       kernel forms them (the error floor of FP32 terms)
 *   or_calc_di          brute-force interaction count (DirectBench.cpp CalcDI)
 *   or_gather_*         restatement of the pencil gathers (PencilPlan.cc)
 *   or_coadd_f32        restatement of the k-ordered set/add co-add
 *   or_pencil_execute   restatement of SetInteractionCollection::CPUExecute
 *                       (pencil-structured executor), any softening + density
 *
 * The CPU-fallback kernels and ExecuteSlabCPU live in cpu_directs.c, which is
 * built with the reference's fast-math CPU flags; this file is built without
 * fast-math so its FP32 arithmetic is IEEE-exact and its FP64 sums unfused.
 *
 * Conventions: a "box" is cpd^3 cells; slab x holds cells (x, y, z) ordered
 * y-major then z (cellparticles.cpp:30-36); particles of a slab are stored
 * PosXYZ-style: X[n] | Y[n] | Z[n] (cellparticles.cpp:218-228), with cell
 * (y,z) occupying [start[y*cpd+z], start+count).  Positions are FP32 relative
 * to the cell centre (grid.cpp:174-178).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { OR_PLUMMER = 0, OR_SINGLE_SPLINE = 1, OR_CUBIC_SPLINE = 2, OR_CUBIC_PLUMMER = 3 };

/* ------------------------------------------------------------------ */
/* Pair kernels                                                        */
/* ------------------------------------------------------------------ */

/* FP32 kernels with FOF density, as DirectGPUKernels.cuh:68-108 (single
 * spline, template R=0), :144-174 (Plummer), :20-66 (cubic spline),
 * :110-143 (cubic Plummer).  d = source - sink; a -= f*d; if r2<b2: w += b2-r2.
 * The reference GPU rsqrt is approximate; here 1/sqrtf is used (the
 * restatement is exact-rounded, the GPU result is compared by tolerance). */
static inline void or_kernel_gpu_f32(int kind, float sx, float sy, float sz,
                                     float kx, float ky, float kz, float eps, float b2,
                                     float *ax, float *ay, float *az, float *aw)
{
    float dx = sx - kx, dy = sy - ky, dz = sz - kz;
    float r2 = dx * dx + dy * dy + dz * dz;
    float f;
    if (r2 < b2) *aw = *aw + b2 - r2;
    if (kind == OR_SINGLE_SPLINE) {
        float q = r2 * eps + 1e-32f;
        f = 1.0f / sqrtf(q);
        if (q > 1.0f) f = f * f * f;
        else f = (-15.0f * f + 6.0f) * q + 10.0f;
    } else if (kind == OR_PLUMMER) {
        float r = 1.0f / sqrtf(r2 + eps);
        f = r * r * r;
    } else if (kind == OR_CUBIC_SPLINE) {
        float rinv = 1.0f / sqrtf(r2);
        float u = eps / rinv; /* eps = 1/eps_int here */
        if (u >= 2.0f) f = rinv * rinv * rinv;
        else if (u >= 1.0f)
            f = rinv * rinv * rinv * ((float)(-1.0 / 15.0) + (float)(8.0 / 3.0) * u * u * u -
                                      3.0f * u * u * u * u + (float)(6.0 / 5.0) * u * u * u * u * u -
                                      (float)(1.0 / 6.0) * u * u * u * u * u * u);
        else
            f = eps * eps * eps * ((float)(4.0 / 3.0) - (float)(6.0 / 5.0) * u * u + 0.5f * u * u * u);
    } else { /* cubic Plummer, DirectGPUKernels.cuh:110-143, eps = eps^3 */
        float r = r2 + 1e-16f;
        r *= r / sqrtf(r);
        r += eps;
        f = 1.0f / r;
    }
    *ax -= f * dx;
    *ay -= f * dy;
    *az -= f * dz;
}

/* FP64 form of the same kernels, used by the brute-force oracle: the pair
 * term g(r), with a -= g d.  The mathematics follows DirectGPUKernels.cuh
 * without the 1e-32 guard: the single spline is F(u)=u^-3 for u>1 and
 * 10-15u+6u^2 for u<=1 with u=r/eps_int (q=u^2); the self pair (r=0)
 * contributes nothing to a (g = 0) and b2 to w. */
static inline double or_g_f64(int kind, double r2, double eps)
{
    if (r2 == 0.0) return 0.0;
    if (kind == OR_SINGLE_SPLINE) {
        double q = r2 * eps, u = sqrt(q);
        return (q > 1.0) ? 1.0 / (q * u) : (10.0 - 15.0 * u + 6.0 * q);
    }
    if (kind == OR_PLUMMER) {
        double t = r2 + eps;
        return 1.0 / (t * sqrt(t));
    }
    if (kind == OR_CUBIC_SPLINE) {
        double r = sqrt(r2), rinv = 1.0 / r, u = eps * r, u3 = u * u * u;
        if (u >= 2.0) return rinv * rinv * rinv;
        if (u >= 1.0)
            return rinv * rinv * rinv * (-1.0 / 15.0 + 8.0 / 3.0 * u3 - 3.0 * u3 * u +
                                         6.0 / 5.0 * u3 * u * u - 1.0 / 6.0 * u3 * u3);
        return eps * eps * eps * (4.0 / 3.0 - 6.0 / 5.0 * u * u + 0.5 * u3);
    }
    double r = r2 + 1e-16;
    return 1.0 / (r * sqrt(r) + eps);
}

static inline void or_kernel_f64(int kind, double dx, double dy, double dz, double eps,
                                 double b2, double *a)
{
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < b2) a[3] += b2 - r2;
    double g = or_g_f64(kind, r2, eps);
    a[0] -= g * dx;
    a[1] -= g * dy;
    a[2] -= g * dz;
}

/* Exported scalar access to the kernels, for known-answer tests. */
void or_pair_f64(int kind, double dx, double dy, double dz, double eps, double b2, double *a4)
{
    a4[0] = a4[1] = a4[2] = a4[3] = 0.0;
    or_kernel_f64(kind, dx, dy, dz, eps, b2, a4);
}

void or_pair_gpu_f32(int kind, const float *src3, const float *sink3, float eps, float b2, float *a4)
{
    a4[0] = a4[1] = a4[2] = a4[3] = 0.0f;
    or_kernel_gpu_f32(kind, src3[0], src3[1], src3[2], sink3[0], sink3[1], sink3[2], eps, b2,
                      &a4[0], &a4[1], &a4[2], &a4[3]);
}

/* ------------------------------------------------------------------ */
/* Box access helpers                                                  */
/* ------------------------------------------------------------------ */

/* grid::WrapSlab, grid.cpp:70-74 */
static inline int wrapi(int i, int cpd)
{
    while (i >= cpd) i -= cpd;
    while (i < 0) i += cpd;
    return i;
}


/* ------------------------------------------------------------------ */
/* FP64 brute-force oracle                                             */
/* ------------------------------------------------------------------ */

/* For each requested sink (slab x, global index i within the slab, cell
 * (y,z)), sum over the (2R+1)^3 neighbouring cells with the exact offset
 * (source cell centre - sink cell centre) = (di,dj,dk)/cpd in FP64, on the
 * FP32 cell-relative positions promoted to double.  The neighbourhood is
 * that of ExecuteSlabCPU (extras_directdriver.cpp:84-99).  Output a4 is the
 * kernel sum in the reference's units (no eps_int^-3 prefactor; density
 * includes the self term b2, removed later by kick.cpp:105).  If soft_norm is
 * not NULL it receives |sum of the pair terms with r below the softening
 * length|, used to flag softening-dominated sinks.  If abs_norm is not NULL it
 * receives sum_j |g_j| |d_j|, the sum of the pair terms' magnitudes: its ratio
 * to |a| is the condition number of the sink's force sum. */
void or_direct_f64(int cpd, const float *const *pos, const int64_t *nslab,
                   const int64_t *const *start, const int32_t *const *count, int R, int kind,
                   double eps, double b2, int64_t nsinks, const int32_t *sink_slab,
                   const int64_t *sink_index, const int32_t *sink_cell, double *out,
                   double *soft_norm, double *abs_norm)
{
    const double invcpd = 1.0 / (double)cpd;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsinks; s++) {
        int x = sink_slab[s];
        int64_t i = sink_index[s];
        int y = sink_cell[s] / cpd, z = sink_cell[s] % cpd;
        const float *P = pos[x];
        int64_t n = nslab[x];
        double kx = P[i], ky = P[n + i], kz = P[2 * n + i];
        double a[4] = {0, 0, 0, 0};
        double as[4] = {0, 0, 0, 0}; /* contribution of softened pairs (r < eps_int) */
        double mag = 0.0;            /* sum of |pair term| */
        /* r^2 below which a pair is softened: single spline r < e (eps = 1/e^2),
         * Plummer r < e (eps = e^2), cubic spline u = r/e < 2 (eps = 1/e),
         * cubic Plummer r < e (eps = e^3). */
        const double soft_r2 = kind == OR_SINGLE_SPLINE ? 1.0 / eps
                             : kind == OR_PLUMMER       ? eps
                             : kind == OR_CUBIC_SPLINE  ? 4.0 / (eps * eps)
                                                        : cbrt(eps * eps);
        for (int di = -R; di <= R; di++) {
            int xs = wrapi(x + di, cpd);
            const float *Q = pos[xs];
            int64_t m = nslab[xs];
            double ox = di * invcpd - kx;
            for (int dj = -R; dj <= R; dj++) {
                int ys = wrapi(y + dj, cpd);
                double oy = dj * invcpd - ky;
                for (int dk = -R; dk <= R; dk++) {
                    int zs = wrapi(z + dk, cpd);
                    double oz = dk * invcpd - kz;
                    int c = ys * cpd + zs;
                    int64_t st = start[xs][c];
                    int32_t cn = count[xs][c];
                    for (int32_t p = 0; p < cn; p++) {
                        double dx = (double)Q[st + p] + ox;
                        double dy = (double)Q[m + st + p] + oy;
                        double dz = (double)Q[2 * m + st + p] + oz;
                        double r2 = dx * dx + dy * dy + dz * dz;
                        /* the pair term of or_kernel_f64, evaluated once */
                        if (r2 < b2) a[3] += b2 - r2;
                        double g = or_g_f64(kind, r2, eps);
                        a[0] -= g * dx;
                        a[1] -= g * dy;
                        a[2] -= g * dz;
                        if (soft_norm && r2 < soft_r2) {
                            as[0] -= g * dx;
                            as[1] -= g * dy;
                            as[2] -= g * dz;
                        }
                        if (abs_norm) mag += fabs(g) * sqrt(r2);
                    }
                }
            }
        }
        out[4 * s + 0] = a[0];
        out[4 * s + 1] = a[1];
        out[4 * s + 2] = a[2];
        out[4 * s + 3] = a[3];
        if (soft_norm) soft_norm[s] = sqrt(as[0] * as[0] + as[1] * as[1] + as[2] * as[2]);
        if (abs_norm) abs_norm[s] = mag;
    }
}

/* The FP32 pair-term floor of the batched kernel's arithmetic: for each
 * requested sink, every pair term evaluated in FP32 exactly as the batched
 * path forms it, the terms summed in FP64.  Compared with or_direct_f64 it
 * isolates the error that FP32 pair terms alone carry, whatever the
 * summation: the force sum of a sink can be no more accurate than this.
 * The frame is the pencil's (extras_PencilOnPencil.cc:80, PencilPlan.cc:35-68,
 * 226-238, with the batched path's FP64 offsets rounded once): the source x
 * moved by its slab offset di/cpd, the sink y by -dj/cpd and the sink z by
 * -dk/cpd, each rounded once to FP32; then d = s - k in FP32,
 * r2 = fma(dz,dz, fma(dy,dy, fma(dx,dx, tau))), f = 1/sqrt(r2) rounded once
 * (the hardware v_rsq_f32 is within an ulp of it), g = f*f*f in box units,
 * or min(f^3, (c1 f + c2) r2 + c3) for softened pairs.  Single spline only;
 * out is scaled to the reference's units (times e^3) like the co-add. */
void or_floor_f32(int cpd, const float *const *pos, const int64_t *nslab,
                  const int64_t *const *start, const int32_t *const *count, int R,
                  double eps, int64_t nsinks, const int32_t *sink_slab,
                  const int64_t *sink_index, const int32_t *sink_cell, double *out)
{
    const double invcpd = 1.0 / (double)cpd;
    const double e = eps; /* eps = 1/e_int^2 */
    const float c1 = (float)(-15.0 * e * e), c2 = (float)(6.0 * e * e * sqrt(e)),
                c3 = (float)(10.0 * e * sqrt(e));
    const double out_scale = 1.0 / (e * sqrt(e));
    const float soft_r2 = (float)(1.0 / e), tau = 1e-24f;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsinks; s++) {
        int x = sink_slab[s];
        int64_t i = sink_index[s];
        int y = sink_cell[s] / cpd, z = sink_cell[s] % cpd;
        const float *P = pos[x];
        int64_t n = nslab[x];
        const float kx = P[i], ky = P[n + i], kz = P[2 * n + i];
        double a[3] = {0, 0, 0};
        for (int di = -R; di <= R; di++) {
            int xs = wrapi(x + di, cpd);
            const float *Q = pos[xs];
            int64_t m = nslab[xs];
            for (int dj = -R; dj <= R; dj++) {
                int ys = wrapi(y + dj, cpd);
                const float kyy = (float)((double)ky - dj * invcpd);
                for (int dk = -R; dk <= R; dk++) {
                    int zs = wrapi(z + dk, cpd);
                    const float kzz = (float)((double)kz - dk * invcpd);
                    int c = ys * cpd + zs;
                    int64_t st = start[xs][c];
                    int32_t cn = count[xs][c];
                    for (int32_t p = 0; p < cn; p++) {
                        const float sx = (float)((double)Q[st + p] + di * invcpd);
                        const float dx = sx - kx, dy = Q[m + st + p] - kyy, dz = Q[2 * m + st + p] - kzz;
                        float r2 = fmaf(dx, dx, tau);
                        r2 = fmaf(dy, dy, r2);
                        r2 = fmaf(dz, dz, r2);
                        const float f = (float)(1.0 / sqrt((double)r2));
                        float g = f * f;
                        g = g * f;
                        if (r2 <= soft_r2) {
                            const float poly = fmaf(fmaf(c1, f, c2), r2, c3);
                            g = poly < g ? poly : g;
                        }
                        if (dx == 0.0f && dy == 0.0f && dz == 0.0f) continue; /* the self pair: no force */
                        a[0] -= (double)g * dx;
                        a[1] -= (double)g * dy;
                        a[2] -= (double)g * dz;
                    }
                }
            }
        }
        out[3 * s + 0] = a[0] * out_scale;
        out[3 * s + 1] = a[1] * out_scale;
        out[3 * s + 2] = a[2] * out_scale;
    }
}

/* Brute-force count of unpadded pair interactions for slab x, including
 * self pairs: sum over sink cells of n_sink * n_source over the
 * neighbourhood.  Restates CalcDI (DirectBench.cpp:111-125). */
uint64_t or_calc_di(int cpd, const int32_t *const *count, int R, int x)
{
    uint64_t di = 0;
    for (int y = 0; y < cpd; y++)
        for (int z = 0; z < cpd; z++) {
            uint64_t ns = (uint64_t)count[x][y * cpd + z];
            for (int a = -R; a <= R; a++)
                for (int b = -R; b <= R; b++)
                    for (int c = -R; c <= R; c++)
                        di += ns * (uint64_t)count[wrapi(x + a, cpd)][wrapi(y + b, cpd) * cpd + wrapi(z + c, cpd)];
        }
    return di;
}

/* ------------------------------------------------------------------ */
/* Pencil gathers, co-add and the pencil-structured executor           */
/* ------------------------------------------------------------------ */

/* Plans as the planner lays them out: per pencil W (start, N) pairs, and
 * for sink pencils the wrapcell (SetInteractionCollection.hh:76-132). */

/* SinkPencilPlan::copy_into_pinned_memory<XYZ> (PencilPlan.cc:22-90) for
 * one slab: Z gets + (c-R)*cellsize per cell, with cellsize the FLOAT
 * invcpd; X and Y are copied in two chunks split at wrapcell.  Output is
 * float4 (x,y,z,0) at SinkSetStart. */
void or_gather_sinks(int nsets, int W, int R, const uint32_t *plan_start, const int32_t *plan_n,
                     const int32_t *wrapcell, const int32_t *set_start, const float *slab,
                     int64_t nslab, float cellsize, float *out4)
{
    for (int s = 0; s < nsets; s++) {
        const uint32_t *st = plan_start + (size_t)s * W;
        const int32_t *nn = plan_n + (size_t)s * W;
        int w = wrapcell[s];
        int64_t o = set_start[s];
        int nw = 0;
        for (int c = 0; c < W; c++) {
            float dz = (float)(c - R) * cellsize;
            for (int i = 0; i < nn[c]; i++) {
                float v = slab[2 * nslab + st[c] + i];
                out4[4 * (o + nw + i) + 2] = (dz != 0) ? v + dz : v;
            }
            nw += nn[c];
        }
        int N1 = (w > 0) ? (int)(st[w - 1] + nn[w - 1] - st[0]) : 0;
        int N2 = (int)(st[W - 1] + nn[W - 1] - st[w]);
        for (int i = 0; i < N1; i++) {
            out4[4 * (o + i) + 0] = slab[st[0] + i];
            out4[4 * (o + i) + 1] = slab[nslab + st[0] + i];
        }
        for (int i = 0; i < N2; i++) {
            out4[4 * (o + N1 + i) + 0] = slab[st[w] + i];
            out4[4 * (o + N1 + i) + 1] = slab[nslab + st[w] + i];
        }
        for (int i = 0; i < nw; i++) out4[4 * (o + i) + 3] = 0.0f;
    }
}

/* SourcePencilPlan::copy_into_pinned_memory<XYZ> (PencilPlan.cc:214-255):
 * cells from W slabs; X gets + (c-R)*cellsize; Y, Z raw. */
void or_gather_sources(int nsets, int W, int R, const uint32_t *plan_start, const int32_t *plan_n,
                       const int32_t *set_start, const float *const *slabs, const int64_t *nslab,
                       float cellsize, float *out4)
{
    for (int s = 0; s < nsets; s++) {
        const uint32_t *st = plan_start + (size_t)s * W;
        const int32_t *nn = plan_n + (size_t)s * W;
        int64_t o = set_start[s];
        int nw = 0;
        for (int c = 0; c < W; c++) {
            float dx = (float)(c - R) * cellsize;
            const float *P = slabs[c];
            int64_t m = nslab[c];
            for (int i = 0; i < nn[c]; i++) {
                float v = P[st[c] + i];
                out4[4 * (o + nw + i) + 0] = (dx != 0) ? v + dx : v;
                out4[4 * (o + nw + i) + 1] = P[m + st[c] + i];
                out4[4 * (o + nw + i) + 2] = P[2 * m + st[c] + i];
                out4[4 * (o + nw + i) + 3] = 0.0f;
            }
            nw += nn[c];
        }
    }
}

/* SinkPencilPlan::copy_from_pinned_memory (PencilPlan.cc:140-207), driven in
 * k order per skewer (PencilTask.cu:237-247): k==0 sets the whole pencil;
 * otherwise add the cells before the wrap (or before the last cell); past
 * the wrap add, else set the last cell.  pinacc is float4 per padded sink
 * slot; acc is float4 per particle of the slab. */
void or_coadd_f32(int nsets, int Nk, int W, const uint32_t *plan_start, const int32_t *plan_n,
                  const int32_t *wrapcell, const int32_t *set_start, const int32_t *set_count,
                  const float *pinacc, float *acc)
{
    int k = 0;
    for (int s = 0; s < nsets; s++) {
        const uint32_t *st = plan_start + (size_t)s * W;
        const int32_t *nn = plan_n + (size_t)s * W;
        int w = wrapcell[s], total = set_count[s];
        const float *p = pinacc + 4 * (int64_t)set_start[s];
        float *a = acc + 4 * (int64_t)st[0];
        if (k == 0) {
            memcpy(a, p, sizeof(float) * 4 * (size_t)total);
        } else {
            int last_N = 0, contig;
            if (w > 0) contig = (int)(st[w - 1] + nn[w - 1] - st[0]);
            else {
                last_N = nn[W - 1];
                contig = total - last_N;
            }
            for (int t = 0; t < 4 * contig; t++) a[t] += p[t];
            p += 4 * contig;
            if (w > 0) {
                a = acc + 4 * (int64_t)st[w];
                int rest = total - contig;
                for (int t = 0; t < 4 * rest; t++) a[t] += p[t];
            } else {
                a = acc + 4 * (int64_t)st[W - 1];
                memcpy(a, p, sizeof(float) * 4 * (size_t)last_N);
            }
        }
        if (++k == Nk) k = 0;
    }
}

/* SetInteractionCollection::CPUExecute (extras_PencilOnPencil.cc:38-130)
 * restated for any softening and with the FOF density: per padded sink id,
 * loop over the W source pencils of its parent pencil, dry = sy - ky + yoffset,
 * with the per-64 BLOCK_ACCUM tile sums of PencilKernel.cu:23-55 for full
 * tiles (w accumulated directly).  Writes float4 for ids < SinkSetIdMax. */
void or_pencil_execute(int kind, int W, int nblocks, int block, const int32_t *parent,
                       const uint32_t *idmax, const int32_t *src_start, const int32_t *src_count,
                       const int32_t *ilist, const float *yoff, const float *sinks4,
                       const float *sources4, float eps, float b2, float *out4)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < nblocks; b++) {
        int s = parent[b];
        for (int t = 0; t < block; t++) {
            int64_t id = (int64_t)b * block + t;
            float kx = 0, ky = 0, kz = 0;
            if (id < (int64_t)idmax[s]) {
                kx = sinks4[4 * id];
                ky = sinks4[4 * id + 1];
                kz = sinks4[4 * id + 2];
            }
            float ax = 0, ay = 0, az = 0, aw = 0;
            for (int c = s * W; c < s * W + W; c++) {
                int src = ilist[c];
                int st = src_start[src], cnt = src_count[src];
                float yo = yoff[c];
                int nB = cnt / block;
                for (int bb = 0; bb < nB; bb++) {
                    float tx = 0, ty = 0, tz = 0;
                    for (int i = 0; i < block; i++) {
                        const float *q = sources4 + 4 * (int64_t)(st + bb * block + i);
                        or_kernel_gpu_f32(kind, q[0], q[1] + yo, q[2], kx, ky, kz, eps, b2, &tx, &ty, &tz, &aw);
                    }
                    ax += tx;
                    ay += ty;
                    az += tz;
                }
                for (int i = nB * block; i < cnt; i++) {
                    const float *q = sources4 + 4 * (int64_t)(st + i);
                    or_kernel_gpu_f32(kind, q[0], q[1] + yo, q[2], kx, ky, kz, eps, b2, &ax, &ay, &az, &aw);
                }
            }
            if (id < (int64_t)idmax[s]) {
                out4[4 * id] = ax;
                out4[4 * id + 1] = ay;
                out4[4 * id + 2] = az;
                out4[4 * id + 3] = aw;
            }
        }
    }
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

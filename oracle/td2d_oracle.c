/*
 * td2d_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A from-scratch CPU restatement of TorchWI's time-domain 2D acoustic
 * propagator and its adjoint (imaging-condition) gradient.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library; the product path (torchwi_amd) never does.
 *
 * Two arithmetic modes:
 *   - "f32": mirrors the reference CPU path's mixed precision exactly
 *     (float storage; Laplacian and Keys intermediates in double, rounded
 *     back to float at the points the reference rounds; float-only
 *     arithmetic where the reference uses float).  The goal is bit-exact
 *     agreement with the reference's own compiled CPU path.
 *   - "f64": the same recursion entirely in double; used to state the
 *     rounding drift of every FP32 result.
 *
 * Reference files followed (paths relative to the reference repo):
 *   torchwi/propagator/td2d/td2d_cpu.cpp        stencils (:6-98), Keys (:110-146)
 *   torchwi/propagator/td2d/propagator_cpu.cpp  forward loop (:5-79),
 *                                               backward loop (:82-152)
 *   torchwi/propagator/td2d/propagator_forward_cpu.cpp  forward-only (:5-71)
 *   torchwi/propagator/td2d/td2d_cuda.cu        EXA semantics (:182-205),
 *   torchwi/propagator/td2d/propagator_exa_cuda.cu      EXA loops (:5-159)
 *
 * Build with -ffp-contract=off (no fused multiply-add), as the reference
 * is built for baseline x86-64 where gcc has no FMA to contract into.
 *
 * Grid convention (the reference's): padded grid dimx x dimy stored
 * x-fastest, row stride dimx; halo ne = order/2; interior cells are
 * ix in [ne, dimx-ne), iy in [ne, dimy-ne).  Rows iy < ne (top) are never
 * written and stay zero (free surface).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* Central-difference second-derivative weights, as the reference writes
 * them (double literals): w[0] is the centre weight, w[k] the weight of
 * the +-k neighbours.  td2d_cpu.cpp:4-5 (order 2), :31-33 (order 4),
 * :62-66 (order 8).  These are the textbook values -2,1; -5/2,4/3,-1/12;
 * -205/72,8/5,-1/5,8/315,-1/560. */
static const double LAPW2[2] = {-2.0, 1.0};
static const double LAPW4[3] = {-2.5, 1.3333333333333333, -0.08333333333333333};
static const double LAPW8[5] = {-2.8472222222222223, 1.6, -0.2,
                                0.025396825396825397, -0.0017857142857142857};

/* Keys two-angle paraxial ABC constants, cos(pi/6)*cos(pi/12) and
 * cos(pi/6)+cos(pi/12) (td2d_cpu.cpp:107-108). */
#define KEYS_CM 0.8365163037378079
#define KEYS_CP 1.831951230073507

static const double *lap_weights(int order)
{
    if (order == 2) return LAPW2;
    if (order == 4) return LAPW4;
    if (order == 8) return LAPW8;
    return NULL;
}

/* ------------------------------------------------------------------ */
/* FP32-mirror mode                                                    */
/* ------------------------------------------------------------------ */

/* Interior leapfrog update of one cell (td2d_cpu.cpp:23-26, :52-55,
 * :90-95).  Neighbour sums are float, added in the reference's order
 * (order 2: -1, -dimx, +1, +dimx; orders 4/8: -k*dimx, -k, +k, +k*dimx);
 * the weighted sum is double, outermost ring first; the result is
 * rounded to float before the float time update.  The time update is
 * ((2u-u_prev) + m*lap) for order 2 and ((lap*m + 2u) - u_prev) for 4/8,
 * with m = (v*v)*dtoh2 in float. */
static inline float interior_f32(int order, const float *u1, const float *u2,
                                 const float *v, size_t i, size_t sy, float dtoh2)
{
    float m = (v[i] * v[i]) * dtoh2;
    if (order == 2) {
        float s = ((u2[i - 1] + u2[i - sy]) + u2[i + 1]) + u2[i + sy];
        float lap = (float)(LAPW2[1] * (double)s + (LAPW2[0] * 2.0) * (double)u2[i]);
        return (2.f * u2[i] - u1[i]) + m * lap;
    }
    const double *W = lap_weights(order);
    int ne = order / 2;
    size_t ks = (size_t)ne * sy;
    float s = ((u2[i - ks] + u2[i - ne]) + u2[i + ne]) + u2[i + ks];
    double acc = W[ne] * (double)s;
    for (int k = ne - 1; k >= 1; --k) {
        ks = (size_t)k * sy;
        s = ((u2[i - ks] + u2[i - k]) + u2[i + k]) + u2[i + ks];
        acc = acc + W[k] * (double)s;
    }
    acc = acc + (W[0] * 2.0) * (double)u2[i];
    float lap = (float)acc;
    return (lap * m + 2.f * u2[i]) - u1[i];
}

/* Keys update of one boundary cell, with `in` the offset of the first
 * inward neighbour (-1 right edge, +1 left edge, -dimx bottom edge)
 * (td2d_cpu.cpp:122-124, :132-134, :142-144).  uxx is a double expression
 * rounded to float; uxt is float; the final combination is double. */
static inline float keys_f32(const float *um, const float *uo, const float *v,
                             size_t i, ptrdiff_t in, float h2, float dt2, float hdt)
{
    float uxx = (float)(((double)uo[i] - 2.0 * (double)uo[i + in] + (double)uo[i + 2 * in]) / (double)h2);
    float uxt = ((uo[i] - um[i]) - (uo[i + in] - um[i + in])) / hdt;
    float a = ((-v[i]) * v[i]) * dt2;
    double r = (double)a / KEYS_CM * ((double)uxx + KEYS_CP / (double)v[i] * (double)uxt)
               + 2.0 * (double)uo[i] - (double)um[i];
    return (float)r;
}

typedef struct {
    int order, ne, dimx, dimy, nx, ny;
    float h, dt, dt2, h2, dtoh2, hdt, cons;
} geom32;

static int make_geom32(geom32 *g, int order, int dimx, int dimy, float h, float dt)
{
    if (!lap_weights(order)) return -1;
    g->order = order;
    g->ne = order / 2;
    g->dimx = dimx;
    g->dimy = dimy;
    g->nx = dimx - order;
    g->ny = dimy - order;
    g->h = h;
    g->dt = dt;
    /* propagator_cpu.cpp:36-41 */
    g->dt2 = dt * dt;
    g->h2 = h * h;
    g->dtoh2 = g->dt2 / g->h2;
    g->hdt = h * dt;
    g->cons = (float)(-2.0 / (double)g->dt2);
    return 0;
}

/* One full wavefield update u3 <- F(u1, u2): interior, then Keys on the
 * right/left columns (rows ne..dimy-ne) and the bottom rows (all x).
 * The source / residual injection happens between the two in the
 * reference; it touches interior cells only, so it commutes with Keys. */
static void step_interior_f32(const geom32 *g, const float *u1, const float *u2,
                              float *u3, const float *v)
{
    const size_t sy = (size_t)g->dimx;
    #pragma omp parallel for schedule(static)
    for (int iy = g->ne; iy < g->dimy - g->ne; ++iy) {
        for (int ix = g->ne; ix < g->dimx - g->ne; ++ix) {
            size_t i = (size_t)iy * sy + ix;
            u3[i] = interior_f32(g->order, u1, u2, v, i, sy, g->dtoh2);
        }
    }
}

static void step_keys_f32(const geom32 *g, const float *u1, const float *u2,
                          float *u3, const float *v)
{
    const size_t sy = (size_t)g->dimx;
    const int ne = g->ne;
    #pragma omp parallel for schedule(static)
    for (int iy = ne; iy < g->dimy - ne; ++iy) {
        for (int ix = g->dimx - ne; ix < g->dimx; ++ix) {
            size_t i = (size_t)iy * sy + ix;
            u3[i] = keys_f32(u1, u2, v, i, -1, g->h2, g->dt2, g->hdt);
        }
        for (int ix = 0; ix < ne; ++ix) {
            size_t i = (size_t)iy * sy + ix;
            u3[i] = keys_f32(u1, u2, v, i, +1, g->h2, g->dt2, g->hdt);
        }
    }
    #pragma omp parallel for schedule(static)
    for (int iy = g->dimy - ne; iy < g->dimy; ++iy) {
        for (int ix = 0; ix < g->dimx; ++ix) {
            size_t i = (size_t)iy * sy + ix;
            u3[i] = keys_f32(u1, u2, v, i, -(ptrdiff_t)sy, g->h2, g->dt2, g->hdt);
        }
    }
}

/* Source / receiver geometry exactly as the reference truncates it
 * (propagator_cpu.cpp:44-48): float division, then C truncation. */
static void geometry32(const geom32 *g, float sx, float sy, float ry,
                       long *isrc, long *ir0)
{
    int isx = (int)(sx / g->h);
    int isy = (int)(sy / g->h);
    int iry = (int)(ry / g->h);
    if (isrc) *isrc = (long)(isy + g->ne) * g->dimx + (isx + g->ne);
    if (ir0) *ir0 = (long)(iry + g->ne) * g->dimx + g->ne;
}

/* Forward modelling (propagator_cpu.cpp:5-79 with virt != NULL,
 * propagator_forward_cpu.cpp:5-71 with virt == NULL).  frd is (nt, nx),
 * virt is (nt, dimx*dimy).  Returns 0, or -1 for an unsupported order. */
int orc_forward_f32(float *frd, float *virt, float *u1, float *u2, float *u3,
                    const float *v, const float *w, int order, int dimx, int dimy,
                    int nt, float h, float dt, float sx, float sy, float ry)
{
    geom32 g;
    if (make_geom32(&g, order, dimx, dimy, h, dt)) return -1;
    long isrc, ir0;
    geometry32(&g, sx, sy, ry, &isrc, &ir0);
    const size_t n = (size_t)dimx * dimy;
    memset(u1, 0, n * sizeof(float));
    memset(u2, 0, n * sizeof(float));
    memset(u3, 0, n * sizeof(float));
    for (int it = 0; it < nt; ++it) {
        step_interior_f32(&g, u1, u2, u3, v);
        u3[isrc] += ((w[it] * v[isrc]) * v[isrc]) * g.dt2;
        step_keys_f32(&g, u1, u2, u3, v);
        memcpy(frd + (size_t)it * g.nx, u3 + ir0, (size_t)g.nx * sizeof(float));
        if (virt) {
            float *vt = virt + (size_t)it * n;
            #pragma omp parallel for schedule(static)
            for (long i = 0; i < (long)n; ++i) {
                float v3 = (v[i] * v[i]) * v[i];
                vt[i] = (g.cons / v3) * ((u1[i] - 2.f * u2[i]) + u3[i]);
            }
        }
        float *t = u1; u1 = u2; u2 = u3; u3 = t;
    }
    return 0;
}

/* Adjoint propagation with imaging condition (propagator_cpu.cpp:82-152).
 * grad (dimx*dimy) is accumulated into (the reference's wrapper zeroes it
 * first, wrapper_cpu.cpp:101).  resid is (nt, nx). */
int orc_backward_f32(float *grad, const float *virt, float *u1, float *u2, float *u3,
                     const float *v, const float *resid, int order, int dimx, int dimy,
                     int nt, float h, float dt, float ry)
{
    geom32 g;
    if (make_geom32(&g, order, dimx, dimy, h, dt)) return -1;
    long ir0;
    geometry32(&g, 0.f, 0.f, ry, NULL, &ir0);
    const size_t n = (size_t)dimx * dimy;
    memset(u1, 0, n * sizeof(float));
    memset(u2, 0, n * sizeof(float));
    memset(u3, 0, n * sizeof(float));
    for (int it = nt - 1; it >= 0; --it) {
        step_interior_f32(&g, u1, u2, u3, v);
        const float *r = resid + (size_t)it * g.nx;
        for (int ix = 0; ix < g.nx; ++ix) {
            size_t j = (size_t)ir0 + ix;
            u3[j] += ((r[ix] * v[j]) * v[j]) * g.dt2;
        }
        step_keys_f32(&g, u1, u2, u3, v);
        const float *vt = virt + (size_t)it * n;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i)
            grad[i] = grad[i] + (-vt[i]) * u3[i];
        float *t = u1; u1 = u2; u2 = u3; u3 = t;
    }
    return 0;
}

/* EXA (excitation amplitude) forward: per cell keep the virtual source of
 * largest magnitude and the step it occurred at (td2d_cuda.cu:182-195,
 * propagator_exa_cuda.cu:5-84).  The virtual source here is the EXA
 * kernel's own expression cons/(v*v*v) * ((u3 - 2u2) + u1).  Ties keep the
 * earliest step (strict >).  The reference exists only for CUDA, so this
 * function has no runnable counterpart; it is pinned (tests/test_oracle.py)
 * against the reference's full-storage virtual sources: on every cell whose
 * largest source is clearly separated, the same step and the same signed
 * value to rounding (the two kernels associate the sum differently). */
int orc_exa_forward_f32(float *frd, float *exa, short *iexa, float *u1, float *u2,
                        float *u3, const float *v, const float *w, int order, int dimx,
                        int dimy, int nt, float h, float dt, float sx, float sy, float ry)
{
    geom32 g;
    if (make_geom32(&g, order, dimx, dimy, h, dt)) return -1;
    long isrc, ir0;
    geometry32(&g, sx, sy, ry, &isrc, &ir0);
    const size_t n = (size_t)dimx * dimy;
    memset(u1, 0, n * sizeof(float));
    memset(u2, 0, n * sizeof(float));
    memset(u3, 0, n * sizeof(float));
    memset(exa, 0, n * sizeof(float));
    memset(iexa, 0, n * sizeof(short));
    for (int it = 0; it < nt; ++it) {
        step_interior_f32(&g, u1, u2, u3, v);
        u3[isrc] += ((w[it] * v[isrc]) * v[isrc]) * g.dt2;
        step_keys_f32(&g, u1, u2, u3, v);
        memcpy(frd + (size_t)it * g.nx, u3 + ir0, (size_t)g.nx * sizeof(float));
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            float vv = v[i];
            float vs = (g.cons / ((vv * vv) * vv)) * ((u3[i] - 2.f * u2[i]) + u1[i]);
            if (fabsf(vs) > fabsf(exa[i])) {
                exa[i] = vs;
                iexa[i] = (short)it;
            }
        }
        float *t = u1; u1 = u2; u2 = u3; u3 = t;
    }
    return 0;
}

/* EXA gradient: grad = -exa * lambda at the step recorded in iexa
 * (assignment, not accumulation) (td2d_cuda.cu:197-205,
 * propagator_exa_cuda.cu:87-159).  grad must be zeroed by the caller.
 * Pinned bit for bit against the reference's full-storage backward run with
 * a virtual source that is exa at step iexa and zero at every other step. */
int orc_exa_backward_f32(float *grad, const float *exa, const short *iexa, float *u1,
                         float *u2, float *u3, const float *v, const float *resid,
                         int order, int dimx, int dimy, int nt, float h, float dt,
                         float ry)
{
    geom32 g;
    if (make_geom32(&g, order, dimx, dimy, h, dt)) return -1;
    long ir0;
    geometry32(&g, 0.f, 0.f, ry, NULL, &ir0);
    const size_t n = (size_t)dimx * dimy;
    memset(u1, 0, n * sizeof(float));
    memset(u2, 0, n * sizeof(float));
    memset(u3, 0, n * sizeof(float));
    for (int it = nt - 1; it >= 0; --it) {
        step_interior_f32(&g, u1, u2, u3, v);
        const float *r = resid + (size_t)it * g.nx;
        for (int ix = 0; ix < g.nx; ++ix) {
            size_t j = (size_t)ir0 + ix;
            u3[j] += ((r[ix] * v[j]) * v[j]) * g.dt2;
        }
        step_keys_f32(&g, u1, u2, u3, v);
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i)
            if (iexa[i] == it) grad[i] = (-exa[i]) * u3[i];
        float *t = u1; u1 = u2; u2 = u3; u3 = t;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 mode: the same recursion with every quantity in double.         */
/* Geometry is truncated exactly as in FP32 mode so both runs use the   */
/* same cells.                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    int order, ne, dimx, dimy, nx, ny;
    double dt2, h2, dtoh2, hdt, cons;
} geom64;

static int make_geom64(geom64 *g, int order, int dimx, int dimy, double h, double dt)
{
    if (!lap_weights(order)) return -1;
    g->order = order;
    g->ne = order / 2;
    g->dimx = dimx;
    g->dimy = dimy;
    g->nx = dimx - order;
    g->ny = dimy - order;
    g->dt2 = dt * dt;
    g->h2 = h * h;
    g->dtoh2 = g->dt2 / g->h2;
    g->hdt = h * dt;
    g->cons = -2.0 / g->dt2;
    return 0;
}

static void step_f64(const geom64 *g, const double *u1, const double *u2, double *u3,
                     const double *v)
{
    const size_t sy = (size_t)g->dimx;
    const int ne = g->ne;
    const double *W = lap_weights(g->order);
    #pragma omp parallel for schedule(static)
    for (int iy = ne; iy < g->dimy - ne; ++iy) {
        for (int ix = ne; ix < g->dimx - ne; ++ix) {
            size_t i = (size_t)iy * sy + ix;
            double acc = 2.0 * W[0] * u2[i];
            for (int k = 1; k <= ne; ++k) {
                size_t ks = (size_t)k * sy;
                acc += W[k] * (u2[i - ks] + u2[i - k] + u2[i + k] + u2[i + ks]);
            }
            u3[i] = acc * (v[i] * v[i] * g->dtoh2) + 2.0 * u2[i] - u1[i];
        }
    }
}

static inline double keys_f64(const double *um, const double *uo, const double *v,
                              size_t i, ptrdiff_t in, const geom64 *g)
{
    double uxx = (uo[i] - 2.0 * uo[i + in] + uo[i + 2 * in]) / g->h2;
    double uxt = ((uo[i] - um[i]) - (uo[i + in] - um[i + in])) / g->hdt;
    return -v[i] * v[i] * g->dt2 / KEYS_CM * (uxx + KEYS_CP / v[i] * uxt) + 2.0 * uo[i] - um[i];
}

static void keys_f64_all(const geom64 *g, const double *u1, const double *u2, double *u3,
                         const double *v)
{
    const size_t sy = (size_t)g->dimx;
    const int ne = g->ne;
    #pragma omp parallel for schedule(static)
    for (int iy = ne; iy < g->dimy - ne; ++iy) {
        for (int ix = g->dimx - ne; ix < g->dimx; ++ix)
            u3[(size_t)iy * sy + ix] = keys_f64(u1, u2, v, (size_t)iy * sy + ix, -1, g);
        for (int ix = 0; ix < ne; ++ix)
            u3[(size_t)iy * sy + ix] = keys_f64(u1, u2, v, (size_t)iy * sy + ix, +1, g);
    }
    #pragma omp parallel for schedule(static)
    for (int iy = g->dimy - ne; iy < g->dimy; ++iy)
        for (int ix = 0; ix < g->dimx; ++ix)
            u3[(size_t)iy * sy + ix] =
                keys_f64(u1, u2, v, (size_t)iy * sy + ix, -(ptrdiff_t)sy, g);
}

/* FP64 forward.  Indices come from the float truncation rule. */
int orc_forward_f64(double *frd, double *virt, double *u1, double *u2, double *u3,
                    const double *v, const double *w, int order, int dimx, int dimy,
                    int nt, float h, float dt, float sx, float sy, float ry)
{
    geom64 g;
    if (make_geom64(&g, order, dimx, dimy, (double)h, (double)dt)) return -1;
    geom32 g32;
    make_geom32(&g32, order, dimx, dimy, h, dt);
    long isrc, ir0;
    geometry32(&g32, sx, sy, ry, &isrc, &ir0);
    const size_t n = (size_t)dimx * dimy;
    memset(u1, 0, n * sizeof(double));
    memset(u2, 0, n * sizeof(double));
    memset(u3, 0, n * sizeof(double));
    for (int it = 0; it < nt; ++it) {
        step_f64(&g, u1, u2, u3, v);
        u3[isrc] += w[it] * v[isrc] * v[isrc] * g.dt2;
        keys_f64_all(&g, u1, u2, u3, v);
        memcpy(frd + (size_t)it * g.nx, u3 + ir0, (size_t)g.nx * sizeof(double));
        if (virt) {
            double *vt = virt + (size_t)it * n;
            #pragma omp parallel for schedule(static)
            for (long i = 0; i < (long)n; ++i)
                vt[i] = g.cons / (v[i] * v[i] * v[i]) * (u1[i] - 2.0 * u2[i] + u3[i]);
        }
        double *t = u1; u1 = u2; u2 = u3; u3 = t;
    }
    return 0;
}

int orc_backward_f64(double *grad, const double *virt, double *u1, double *u2,
                     double *u3, const double *v, const double *resid, int order,
                     int dimx, int dimy, int nt, float h, float dt, float ry)
{
    geom64 g;
    if (make_geom64(&g, order, dimx, dimy, (double)h, (double)dt)) return -1;
    geom32 g32;
    make_geom32(&g32, order, dimx, dimy, h, dt);
    long ir0;
    geometry32(&g32, 0.f, 0.f, ry, NULL, &ir0);
    const size_t n = (size_t)dimx * dimy;
    memset(u1, 0, n * sizeof(double));
    memset(u2, 0, n * sizeof(double));
    memset(u3, 0, n * sizeof(double));
    for (int it = nt - 1; it >= 0; --it) {
        step_f64(&g, u1, u2, u3, v);
        const double *r = resid + (size_t)it * g.nx;
        for (int ix = 0; ix < g.nx; ++ix) {
            size_t j = (size_t)ir0 + ix;
            u3[j] += r[ix] * v[j] * v[j] * g.dt2;
        }
        keys_f64_all(&g, u1, u2, u3, v);
        const double *vt = virt + (size_t)it * n;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i)
            grad[i] += -vt[i] * u3[i];
        double *t = u1; u1 = u2; u2 = u3; u3 = t;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Single-step entry points, used by tests that pin one update against  */
/* brute force from arbitrary (u1, u2) states.                          */
/* ------------------------------------------------------------------ */

/* u3 <- interior + Keys update from (u1, u2); top rows of u3 untouched. */
int orc_step_f32(float *u3, const float *u1, const float *u2, const float *v,
                 int order, int dimx, int dimy, float h, float dt)
{
    geom32 g;
    if (make_geom32(&g, order, dimx, dimy, h, dt)) return -1;
    step_interior_f32(&g, u1, u2, u3, v);
    step_keys_f32(&g, u1, u2, u3, v);
    return 0;
}

int orc_step_f64(double *u3, const double *u1, const double *u2, const double *v,
                 int order, int dimx, int dimy, double h, double dt)
{
    geom64 g;
    if (make_geom64(&g, order, dimx, dimy, h, dt)) return -1;
    step_f64(&g, u1, u2, u3, v);
    keys_f64_all(&g, u1, u2, u3, v);
    return 0;
}

/*
 * cda_oracle.c -- CPU oracle for PySIT's time-domain constant-density acoustic
 * (CDA) leapfrog + PML solver and its adjoint-state imaging condition.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, the smoke() entry and the cpu_baseline
 * leg of bench.py load it as the CPU result the HIP library is compared
 * against.  The product package (pysit_amd) never links, loads or calls it.
 *
 * What it restates (reference paths relative to the PySIT source tree):
 *   orc_step2d_{d,f}   pysit/solvers/constant_density_acoustic/time/scalar/
 *                      constant_density_acoustic_time_scalar_2D.h:43-180 with
 *                      the tap tables of pysit/solvers/fd_tools/fd_manual.hpp:6-69
 *   orc_step3d_{d,f}   .../constant_density_acoustic_time_scalar_3D.h:46-236
 *   orc_forward        pysit/modeling/temporal_modeling.py:60-169 (record before
 *                      the step, rhs_k, step, dWaveOp of
 *                      pysit/solvers/constant_density_acoustic/
 *                      constant_density_acoustic_base.py:24-40)
 *   orc_migrate        temporal_modeling.py:171-214 + adjoint_model :228-387
 *                      (time-reversed loop, ic += q_k*dWaveOp_k, ic *= -dt*p);
 *                      the forward dWaveOp is recomputed from state checkpoints
 *                      every `ckpt_interval` steps, which gives the same bits
 *                      as storing every step (the loop is deterministic).
 *   orc_fold           pysit/solvers/model_parameter.py:537-593 (add_padding)
 *
 * The node loops of the steps, dWaveOp and the imaging condition run over
 * OpenMP threads when built with -fopenmp: every node is computed by one
 * thread with the same operations, so the results do not depend on the
 * thread count.
 *
 * Floating point follows the reference's source literally: the tap weights are
 * double literals, the taps are summed left to right in double and then
 * rounded to T, and every update expression keeps the reference's evaluation
 * order and its float/double promotions.  Built without fast-math and without
 * FP contraction, the FP64 and FP32 steps reproduce the reference headers bit
 * for bit (checked against oracle/_ref and tests/golden).
 *
 * How each function is pinned:
 *   steps        bit-exact vs the reference headers compiled into oracle/_ref
 *                (all orders, FP64 and FP32, 2D and 3D, asymmetric PML), and
 *                vs the committed golden vectors tests/golden/step_*.npz;
 *                plus closed forms (polynomial exactness of the taps).
 *   forward      loop order vs the reference step driven by a Python
 *                restatement of temporal_modeling.py on tiny grids.
 *   migrate      dot-product test <F m1, d> = <m1, F* d> and central finite
 *                differences of the misfit with the PML switched off (there
 *                PySIT's time-reversed recursion is the exact discrete
 *                adjoint); checkpoint interval invariance (bit-exact).
 *   fold         exact transpose of edge padding: <pad x, y> = <x, fold y>.
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* Central-difference weights for accuracy 2r (r = 1..4).  W1[r][j-1] is the
 * first-derivative weight at offset +j (offset -j carries the opposite sign);
 * W2[r][0] is the second-derivative centre weight and W2[r][j] the weight at
 * offsets +-j.  These are the textbook (Fornberg) weights; they are written as
 * the same double-literal quotients that the reference evaluates so that the
 * rounded values agree. */
static const double W1[5][4] = {
    {0.0, 0.0, 0.0, 0.0},
    {1. / 2., 0.0, 0.0, 0.0},
    {2. / 3., -1. / 12., 0.0, 0.0},
    {3. / 4., -3. / 20., 1. / 60., 0.0},
    {4. / 5., -1. / 5., 4. / 105., -1. / 280.}};
static const double W2[5][5] = {
    {0.0, 0.0, 0.0, 0.0, 0.0},
    {-2.0, 1.0, 0.0, 0.0, 0.0},
    {-5. / 2., 4. / 3., -1. / 12., 0.0, 0.0},
    {-49. / 18., 3. / 2., -3. / 20., 1. / 90., 0.0},
    {-205. / 72., 8. / 5., -1. / 5., 8. / 315., -1. / 560.}};

/* ------------------------------------------------------------------------ */
/* Generic pieces, instantiated for T = double and T = float.               */
/* A tap outside [0, n) reads 0: that is what the shifted stencils of the   */
/* reference do next to the grid edge (fd_manual.hpp:29-37, 61-69).          */
/* ------------------------------------------------------------------------ */
#define ORC_DEFINE_TAPS(T, SUF)                                                   \
    static T orc_d1_##SUF(const T *v, long idx, long s, int p, int n, int r,      \
                          T delta)                                                \
    {                                                                             \
        double acc = 0.0;                                                         \
        for (int j = r; j >= 1; --j) {                                            \
            double x = (p - j >= 0) ? (double)v[idx - j * s] : 0.0;               \
            acc += (-W1[r][j - 1]) * x;                                           \
        }                                                                         \
        acc += 0.0;                                                               \
        for (int j = 1; j <= r; ++j) {                                            \
            double x = (p + j < n) ? (double)v[idx + j * s] : 0.0;                \
            acc += W1[r][j - 1] * x;                                              \
        }                                                                         \
        return (T)(acc / (double)delta);                                          \
    }                                                                             \
    static T orc_d2_##SUF(const T *v, long idx, long s, int p, int n, int r,      \
                          T delta2)                                               \
    {                                                                             \
        double acc = 0.0;                                                         \
        for (int j = r; j >= 1; --j) {                                            \
            double x = (p - j >= 0) ? (double)v[idx - j * s] : 0.0;               \
            acc += W2[r][j] * x;                                                  \
        }                                                                         \
        acc += W2[r][0] * (double)v[idx];                                         \
        for (int j = 1; j <= r; ++j) {                                            \
            double x = (p + j < n) ? (double)v[idx + j * s] : 0.0;                \
            acc += W2[r][j] * x;                                                  \
        }                                                                         \
        return (T)(acc / (double)delta2);                                         \
    }                                                                             \
    /* PML profile lookup with the kernel's indexing: left array by node, right \
     * array offset by one (the node next to the outer edge reads r[n-1]).      \
     * Callers never ask for the edge node itself. */                            \
    static T orc_sigma_##SUF(const T *l, int nl, const T *rr, int nr, int i,      \
                             int n)                                               \
    {                                                                             \
        if (nl > 0 && i < nl) return l[i];                                        \
        if (nr > 0 && i >= n - nr) return rr[nr - ((n - 1) - i)];                 \
        return (T)0;                                                              \
    }

ORC_DEFINE_TAPS(double, d)
ORC_DEFINE_TAPS(float, f)

/* ------------------------------------------------------------------------ */
/* 2D step: one leapfrog update of u, Phix, Phiz on the padded nx x nz grid, */
/* idx = i*nz + k (z contiguous).                                            */
/* ------------------------------------------------------------------------ */
#define ORC_DEFINE_STEP2D(T, SUF)                                                 \
    ORC_EXPORT int orc_step2d_##SUF(                                              \
        int order, const T *km1_u, const T *k_phix, const T *k_phiz,              \
        const T *k_u, const T *C, const T *rhs, const T *xl, int nxl,             \
        const T *xr, int nxr, const T *zl, int nzl, const T *zr, int nzr,         \
        double dt, double dx, double dz, int nx, int nz, T *kp1_phix,             \
        T *kp1_phiz, T *kp1_u)                                                    \
    {                                                                             \
        if (order != 2 && order != 4 && order != 6 && order != 8) return -1;      \
        const int r = order / 2;                                                  \
        const T hx = (T)dx, hz = (T)dz;                                           \
        const T hx2 = hx * hx, hz2 = hz * hz;                                     \
        const T dt2 = (T)(dt * dt);                                               \
        _Pragma("omp parallel for schedule(static)")                              \
        for (int i = 0; i < nx; ++i) {                                            \
            for (int k = 0; k < nz; ++k) {                                        \
                const long idx = (long)i * nz + k;                                \
                kp1_phix[idx] = (T)0;                                             \
                kp1_phiz[idx] = (T)0;                                             \
                kp1_u[idx] = (T)0;                                                \
                if (i == 0 || i == nx - 1 || k == 0 || k == nz - 1) continue;     \
                /* x axis first, then z: lap accumulates in that order */         \
                T lap = (T)0;                                                     \
                const T dux = orc_d1_##SUF(k_u, idx, nz, i, nx, r, hx);           \
                const T dpx = orc_d1_##SUF(k_phix, idx, nz, i, nx, r, hx);        \
                lap += orc_d2_##SUF(k_u, idx, nz, i, nx, r, hx2);                 \
                const T duz = orc_d1_##SUF(k_u, idx, 1, k, nz, r, hz);            \
                const T dpz = orc_d1_##SUF(k_phiz, idx, 1, k, nz, r, hz);         \
                lap += orc_d2_##SUF(k_u, idx, 1, k, nz, r, hz2);                  \
                const T sx = orc_sigma_##SUF(xl, nxl, xr, nxr, i, nx);            \
                const T sz = orc_sigma_##SUF(zl, nzl, zr, nzr, k, nz);            \
                const T u = k_u[idx], um1 = km1_u[idx];                           \
                const T c2 = C[idx] * C[idx];                                     \
                const T src = ((rhs[idx] + lap) + dpx) + dpz;                     \
                if (sz != (T)0 || sx != (T)0) {                                   \
                    const T px = k_phix[idx], pz = k_phiz[idx];                   \
                    kp1_phix[idx] = (T)((px - dt * sx * px) + dt * (sz - sx) * dux); \
                    kp1_phiz[idx] = (T)((pz - dt * sz * pz) + dt * (sx - sz) * duz); \
                    const T fac1 = (T)(2.0 * dt2 / (2.0 + dt * (sx + sz)));      \
                    const T fac2 = (T)(c2 * src - (um1 - 2.0 * u) / dt2 +         \
                                       (sx + sz) * um1 / (2.0 * dt) -             \
                                       (sx * sz) * u);                            \
                    kp1_u[idx] = fac1 * fac2;                                     \
                } else {                                                          \
                    kp1_phix[idx] = k_phix[idx];                                  \
                    kp1_phiz[idx] = k_phiz[idx];                                  \
                    kp1_u[idx] = (T)(dt2 * c2 * src - (um1 - 2.0 * u));           \
                }                                                                 \
            }                                                                     \
        }                                                                         \
        return 0;                                                                 \
    }

ORC_DEFINE_STEP2D(double, d)
ORC_DEFINE_STEP2D(float, f)

/* ------------------------------------------------------------------------ */
/* 3D step: u, Phix, Phiy, Phiz, psi on nx x ny x nz, idx = (i*ny + j)*nz + k */
/* ------------------------------------------------------------------------ */
#define ORC_DEFINE_STEP3D(T, SUF)                                                 \
    ORC_EXPORT int orc_step3d_##SUF(                                              \
        int order, const T *km1_u, const T *k_phix, const T *k_phiy,              \
        const T *k_phiz, const T *k_psi, const T *k_u, const T *C, const T *rhs,  \
        const T *xl, int nxl, const T *xr, int nxr, const T *yl, int nyl,         \
        const T *yr, int nyr, const T *zl, int nzl, const T *zr, int nzr,         \
        double dt, double dx, double dy, double dz, int nx, int ny, int nz,       \
        T *kp1_phix, T *kp1_phiy, T *kp1_phiz, T *kp1_psi, T *kp1_u)              \
    {                                                                             \
        if (order != 2 && order != 4 && order != 6 && order != 8) return -1;      \
        const int r = order / 2;                                                  \
        const T hx = (T)dx, hy = (T)dy, hz = (T)dz;                               \
        const T hx2 = hx * hx, hy2 = hy * hy, hz2 = hz * hz;                      \
        const T dt2 = (T)(dt * dt);                                               \
        const long sxs = (long)ny * nz, sys = nz;                                 \
        _Pragma("omp parallel for schedule(static)")                              \
        for (int i = 0; i < nx; ++i) {                                            \
            for (int j = 0; j < ny; ++j) {                                        \
                for (int k = 0; k < nz; ++k) {                                    \
                    const long idx = (long)i * sxs + (long)j * sys + k;           \
                    kp1_u[idx] = (T)0;                                            \
                    kp1_phix[idx] = (T)0;                                         \
                    kp1_phiy[idx] = (T)0;                                         \
                    kp1_phiz[idx] = (T)0;                                         \
                    kp1_psi[idx] = (T)0;                                          \
                    if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 ||         \
                        k == 0 || k == nz - 1)                                    \
                        continue;                                                 \
                    T lap = (T)0;                                                 \
                    const T dux = orc_d1_##SUF(k_u, idx, sxs, i, nx, r, hx);      \
                    const T dpx = orc_d1_##SUF(k_phix, idx, sxs, i, nx, r, hx);   \
                    const T dsx = orc_d1_##SUF(k_psi, idx, sxs, i, nx, r, hx);    \
                    lap += orc_d2_##SUF(k_u, idx, sxs, i, nx, r, hx2);            \
                    const T duy = orc_d1_##SUF(k_u, idx, sys, j, ny, r, hy);      \
                    const T dpy = orc_d1_##SUF(k_phiy, idx, sys, j, ny, r, hy);   \
                    const T dsy = orc_d1_##SUF(k_psi, idx, sys, j, ny, r, hy);    \
                    lap += orc_d2_##SUF(k_u, idx, sys, j, ny, r, hy2);            \
                    const T duz = orc_d1_##SUF(k_u, idx, 1, k, nz, r, hz);        \
                    const T dpz = orc_d1_##SUF(k_phiz, idx, 1, k, nz, r, hz);     \
                    const T dsz = orc_d1_##SUF(k_psi, idx, 1, k, nz, r, hz);      \
                    lap += orc_d2_##SUF(k_u, idx, 1, k, nz, r, hz2);              \
                    const T sx = orc_sigma_##SUF(xl, nxl, xr, nxr, i, nx);        \
                    const T sy = orc_sigma_##SUF(yl, nyl, yr, nyr, j, ny);        \
                    const T sz = orc_sigma_##SUF(zl, nzl, zr, nzr, k, nz);        \
                    const T u = k_u[idx], um1 = km1_u[idx], ps = k_psi[idx];      \
                    const T c2 = C[idx] * C[idx];                                 \
                    const T src = (((rhs[idx] + lap) + dpx) + dpy) + dpz;         \
                    kp1_psi[idx] = (T)(ps + dt * u);                              \
                    if (sz != (T)0 || sy != (T)0 || sx != (T)0) {                 \
                        const T px = k_phix[idx], py = k_phiy[idx],               \
                                pz = k_phiz[idx];                                 \
                        kp1_phix[idx] = (T)((px - dt * sx * px) +                 \
                                            dt * (sy + sz - sx) * dux +           \
                                            dt * (sy * sz * dsx));                \
                        kp1_phiy[idx] = (T)((py - dt * sy * py) +                 \
                                            dt * (sz + sx - sy) * duy +           \
                                            dt * (sz * sx * dsy));                \
                        kp1_phiz[idx] = (T)((pz - dt * sz * pz) +                 \
                                            dt * (sx + sy - sz) * duz +           \
                                            dt * (sx * sy * dsz));                \
                        const T fac1 =                                            \
                            (T)(2.0 * dt2 / (2.0 + dt * (sx + sy + sz)));         \
                        const T fac2 =                                            \
                            (T)(c2 * (src - (sx * sy * sz) * ps) -                \
                                (um1 - 2.0 * u) / dt2 +                           \
                                (sx + sy + sz) * um1 / (2.0 * dt) -               \
                                (sx * sy + sy * sz + sz * sx) * u);               \
                        kp1_u[idx] = fac1 * fac2;                                 \
                    } else {                                                      \
                        kp1_phix[idx] = k_phix[idx];                              \
                        kp1_phiy[idx] = k_phiy[idx];                              \
                        kp1_phiz[idx] = k_phiz[idx];                              \
                        kp1_u[idx] = (T)(dt2 * c2 * src - (um1 - 2.0 * u));       \
                    }                                                             \
                }                                                                 \
            }                                                                     \
        }                                                                         \
        return 0;                                                                 \
    }

ORC_DEFINE_STEP3D(double, d)
ORC_DEFINE_STEP3D(float, f)

/* ------------------------------------------------------------------------ */
/* Shot loops (FP64, the reference's parity precision).                     */
/* ------------------------------------------------------------------------ */

/* Geometry and model of one propagation, shared by the shot loops.  Arrays
 * are padded and z-contiguous; ny = 1 and the y arrays are unused in 2D. */
typedef struct {
    int dim, order, nx, ny, nz;
    double dx, dy, dz, dt;
    const double *C;                           /* padded wave speed        */
    const double *xl, *xr, *yl, *yr, *zl, *zr; /* kernel PML arrays        */
    int nxl, nxr, nyl, nyr, nzl, nzr;
    const double *sx, *sy, *sz; /* mesh sigma per axis (len nx, ny, nz), the
                                   values build_sigma spreads over the grid */
} orc_grid;

/* The wave state of one time level: u plus the auxiliary fields
 * (2D: Phix, Phiz; 3D: Phix, Phiy, Phiz, psi). */
typedef struct {
    double *u, *px, *py, *pz, *psi;
} orc_level;

static long orc_npts(const orc_grid *g) { return (long)g->nx * g->ny * g->nz; }

static int orc_level_alloc(orc_level *L, long n, int dim)
{
    memset(L, 0, sizeof(*L));
    L->u = (double *)calloc((size_t)n, sizeof(double));
    L->px = (double *)calloc((size_t)n, sizeof(double));
    L->pz = (double *)calloc((size_t)n, sizeof(double));
    if (dim == 3) {
        L->py = (double *)calloc((size_t)n, sizeof(double));
        L->psi = (double *)calloc((size_t)n, sizeof(double));
        if (!L->py || !L->psi) return -1;
    }
    return (L->u && L->px && L->pz) ? 0 : -1;
}

static void orc_level_free(orc_level *L)
{
    free(L->u); free(L->px); free(L->py); free(L->pz); free(L->psi);
    memset(L, 0, sizeof(*L));
}

static void orc_level_copy(orc_level *dst, const orc_level *src, long n, int dim)
{
    memcpy(dst->u, src->u, (size_t)n * sizeof(double));
    memcpy(dst->px, src->px, (size_t)n * sizeof(double));
    memcpy(dst->pz, src->pz, (size_t)n * sizeof(double));
    if (dim == 3) {
        memcpy(dst->py, src->py, (size_t)n * sizeof(double));
        memcpy(dst->psi, src->psi, (size_t)n * sizeof(double));
    }
}

static void orc_level_zero(orc_level *L, long n, int dim)
{
    memset(L->u, 0, (size_t)n * sizeof(double));
    memset(L->px, 0, (size_t)n * sizeof(double));
    memset(L->pz, 0, (size_t)n * sizeof(double));
    if (dim == 3) {
        memset(L->py, 0, (size_t)n * sizeof(double));
        memset(L->psi, 0, (size_t)n * sizeof(double));
    }
}

/* kp1 <- step(km1.u, k) with right-hand side rhs */
static int orc_step(const orc_grid *g, const orc_level *km1, const orc_level *k,
                    const double *rhs, orc_level *kp1)
{
    if (g->dim == 2)
        return orc_step2d_d(g->order, km1->u, k->px, k->pz, k->u, g->C, rhs, g->xl,
                            g->nxl, g->xr, g->nxr, g->zl, g->nzl, g->zr, g->nzr,
                            g->dt, g->dx, g->dz, g->nx, g->nz, kp1->px, kp1->pz,
                            kp1->u);
    return orc_step3d_d(g->order, km1->u, k->px, k->py, k->pz, k->psi, k->u, g->C,
                        rhs, g->xl, g->nxl, g->xr, g->nxr, g->yl, g->nyl, g->yr,
                        g->nyr, g->zl, g->nzl, g->zr, g->nzr, g->dt, g->dx, g->dy,
                        g->dz, g->nx, g->ny, g->nz, kp1->px, kp1->py, kp1->pz,
                        kp1->psi, kp1->u);
}

/* dWaveOp_k from u_{k+1}, u_k, u_{k-1} (and psi_k in 3D) with the mesh sigma,
 * in numpy's evaluation order:
 *   ((u+ - 2u) + u-)/dt**2 + [ (S1*(u+ - u-))/(2dt) + S2*u (+ S3*psi) ]
 * with S1 = sum of sigmas, S2 = sum of pairwise products, S3 = product
 * (constant_density_acoustic_base.py:24-40; the sums are formed as in
 * constant_density_acoustic_time_scalar_{2D,3D}.py _rebuild_operators). */
static void orc_dwaveop(const orc_grid *g, const orc_level *km1, const orc_level *k,
                        const orc_level *kp1, double *out)
{
    const double dt2 = g->dt * g->dt, tdt = 2.0 * g->dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < g->nx; ++i)
        for (int j = 0; j < g->ny; ++j)
            for (int q = 0; q < g->nz; ++q) {
                const long idx = ((long)i * g->ny + j) * g->nz + q;
                const double up = kp1->u[idx], u = k->u[idx], um = km1->u[idx];
                const double tt = ((up - 2.0 * u) + um) / dt2;
                double pml;
                if (g->dim == 2) {
                    const double a = g->sx[i], c = g->sz[q];
                    pml = (a + c) * (up - um) / tdt + (a * c) * u;
                } else {
                    const double a = g->sx[i], b = g->sy[j], c = g->sz[q];
                    pml = ((a + b) + c) * (up - um) / tdt +
                          ((a * b + a * c) + b * c) * u + ((a * b) * c) * k->psi[idx];
                }
                out[idx] = tt + pml;
            }
}

/* Sparse point operators.  Injection: rhs[idx[e]] += w[e] * amp[chan[e]] in
 * entry order (the order scipy's sparse products accumulate in).  Sampling:
 * out[chan] = sum over the channel's entries, in entry order, of w*u[idx]. */
static void orc_inject(double *rhs, int n, const long *idx, const int *chan,
                       const double *w, const double *amp)
{
    for (int e = 0; e < n; ++e) rhs[idx[e]] += w[e] * amp[chan[e]];
}

static void orc_clear(double *rhs, int n, const long *idx)
{
    for (int e = 0; e < n; ++e) rhs[idx[e]] = 0.0;
}

static void orc_sample(const double *u, int n, const long *idx, const int *chan,
                       const double *w, int nchan, double *out)
{
    for (int c = 0; c < nchan; ++c) out[c] = 0.0;
    for (int e = 0; e < n; ++e) out[chan[e]] += w[e] * u[idx[e]];
}

static void orc_grid_fill(orc_grid *g, int dim, int order, int nx, int ny, int nz,
                          double dx, double dy, double dz, double dt,
                          const double *C, const double *xl, int nxl,
                          const double *xr, int nxr, const double *yl, int nyl,
                          const double *yr, int nyr, const double *zl, int nzl,
                          const double *zr, int nzr, const double *sx,
                          const double *sy, const double *sz)
{
    g->dim = dim; g->order = order;
    g->nx = nx; g->ny = (dim == 3) ? ny : 1; g->nz = nz;
    g->dx = dx; g->dy = dy; g->dz = dz; g->dt = dt;
    g->C = C;
    g->xl = xl; g->nxl = nxl; g->xr = xr; g->nxr = nxr;
    g->yl = yl; g->nyl = nyl; g->yr = yr; g->nyr = nyr;
    g->zl = zl; g->nzl = nzl; g->zr = zr; g->nzr = nzr;
    g->sx = sx; g->sy = sy; g->sz = sz;
}

/* Forward model of one shot (temporal_modeling.py:60-169).
 *   src_*: injection entries, amp[k*n_src_chan + c] = intensity*w(k*dt)
 *   rec_*: sampling entries (weights already carry prod(deltas))
 *   simdata[k*n_rec_chan + r] (or NULL); dwaveop[(k/p)*N] for k % p == 0
 *   (or NULL); wavefield[k*N] = padded u_k (or NULL). */
ORC_EXPORT int orc_forward(int dim, int order, int nx, int ny, int nz, double dx,
                           double dy, double dz, double dt, int nsteps,
                           const double *C, const double *xl, int nxl,
                           const double *xr, int nxr, const double *yl, int nyl,
                           const double *yr, int nyr, const double *zl, int nzl,
                           const double *zr, int nzr, const double *sx,
                           const double *sy, const double *sz, int n_src,
                           const long *src_idx, const int *src_chan,
                           const double *src_w, int n_src_chan, const double *amp,
                           int n_rec, const long *rec_idx, const int *rec_chan,
                           const double *rec_w, int n_rec_chan, double *simdata,
                           int imaging_period, double *dwaveop, double *wavefield)
{
    orc_grid g;
    orc_grid_fill(&g, dim, order, nx, ny, nz, dx, dy, dz, dt, C, xl, nxl, xr, nxr,
                  yl, nyl, yr, nyr, zl, nzl, zr, nzr, sx, sy, sz);
    const long N = orc_npts(&g);
    if (imaging_period < 1) return -2;
    orc_level lv[3];
    double *rhs = (double *)calloc((size_t)N, sizeof(double));
    int bad = (rhs == NULL);
    for (int t = 0; t < 3; ++t) bad |= orc_level_alloc(&lv[t], N, dim);
    int rc = bad ? -3 : 0;
    /* us[0] = kp1, us[1] = k, us[2] = km1; advance() rotates them */
    orc_level *kp1 = &lv[0], *k = &lv[1], *km1 = &lv[2];
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        if (wavefield) memcpy(wavefield + (size_t)s * N, k->u, (size_t)N * sizeof(double));
        if (simdata)
            orc_sample(k->u, n_rec, rec_idx, rec_chan, rec_w, n_rec_chan,
                       simdata + (size_t)s * n_rec_chan);
        orc_inject(rhs, n_src, src_idx, src_chan, src_w, amp + (size_t)s * n_src_chan);
        rc = orc_step(&g, km1, k, rhs, kp1);
        orc_clear(rhs, n_src, src_idx);
        if (dwaveop && s % imaging_period == 0)
            orc_dwaveop(&g, km1, k, kp1, dwaveop + (size_t)(s / imaging_period) * N);
        if (s == nsteps - 1) break;
        orc_level *old = km1;
        km1 = k; k = kp1; kp1 = old;
        orc_level_zero(kp1, N, dim);
    }
    for (int t = 0; t < 3; ++t) orc_level_free(&lv[t]);
    free(rhs);
    return rc;
}

/* migrate_shot (temporal_modeling.py:171-214, adjoint_model :228-387).
 * operand[k*n_adj_chan + r] is the data-space operand (the residual for a
 * gradient); adj_* are the adjoint injection entries (weights without
 * prod(deltas)).  The forward dWaveOp is regenerated segment by segment from
 * state checkpoints taken every ckpt_interval steps (<= 0: store every step).
 * ic_padded (N doubles) receives -dt * p * sum_k q_k * dWaveOp_k, i.e. the
 * imaging condition before the PML fold. */
ORC_EXPORT int orc_migrate(int dim, int order, int nx, int ny, int nz, double dx,
                           double dy, double dz, double dt, int nsteps,
                           const double *C, const double *xl, int nxl,
                           const double *xr, int nxr, const double *yl, int nyl,
                           const double *yr, int nyr, const double *zl, int nzl,
                           const double *zr, int nzr, const double *sx,
                           const double *sy, const double *sz, int n_src,
                           const long *src_idx, const int *src_chan,
                           const double *src_w, int n_src_chan, const double *amp,
                           int n_adj, const long *adj_idx, const int *adj_chan,
                           const double *adj_w, int n_adj_chan,
                           const double *operand, int imaging_period,
                           int ckpt_interval, double *ic_padded)
{
    orc_grid g;
    orc_grid_fill(&g, dim, order, nx, ny, nz, dx, dy, dz, dt, C, xl, nxl, xr, nxr,
                  yl, nyl, yr, nyr, zl, nzl, zr, nzr, sx, sy, sz);
    const long N = orc_npts(&g);
    const int p = imaging_period;
    if (p < 1 || nsteps < 1) return -2;
    const int K = (ckpt_interval <= 0 || ckpt_interval > nsteps) ? nsteps : ckpt_interval;
    const int nseg = (nsteps + K - 1) / K;

    /* checkpoint s holds (u_{sK-1}, u_{sK} with its aux fields): the state the
     * forward loop has when it is about to take step sK. */
    orc_level *ck_km1 = (orc_level *)calloc((size_t)nseg, sizeof(orc_level));
    orc_level *ck_k = (orc_level *)calloc((size_t)nseg, sizeof(orc_level));
    double *seg_dw = (double *)calloc((size_t)K * N, sizeof(double));
    double *rhs = (double *)calloc((size_t)N, sizeof(double));
    orc_level lv[3], q[3];
    int bad = (!ck_km1 || !ck_k || !seg_dw || !rhs);
    for (int s = 0; s < nseg && !bad; ++s) {
        bad |= orc_level_alloc(&ck_km1[s], N, dim);
        bad |= orc_level_alloc(&ck_k[s], N, dim);
    }
    for (int t = 0; t < 3 && !bad; ++t) {
        bad |= orc_level_alloc(&lv[t], N, dim);
        bad |= orc_level_alloc(&q[t], N, dim);
    }
    int rc = bad ? -3 : 0;

    /* pass 1: forward, keeping checkpoints */
    if (rc == 0) {
        orc_level *kp1 = &lv[0], *k = &lv[1], *km1 = &lv[2];
        for (int s = 0; s < nsteps && rc == 0; ++s) {
            if (s % K == 0) {
                orc_level_copy(&ck_km1[s / K], km1, N, dim);
                orc_level_copy(&ck_k[s / K], k, N, dim);
            }
            orc_inject(rhs, n_src, src_idx, src_chan, src_w, amp + (size_t)s * n_src_chan);
            rc = orc_step(&g, km1, k, rhs, kp1);
            orc_clear(rhs, n_src, src_idx);
            orc_level *old = km1;
            km1 = k; k = kp1; kp1 = old;
            orc_level_zero(kp1, N, dim);
        }
    }

    /* pass 2: segments from last to first; regenerate dWaveOp, then run the
     * time-reversed adjoint recursion across the segment */
    for (int i = 0; i < N; ++i) ic_padded[i] = 0.0;
    orc_level *qkp1 = &q[0], *qk = &q[1], *qkm1 = &q[2];
    for (int sg = nseg - 1; sg >= 0 && rc == 0; --sg) {
        const int k0 = sg * K, k1 = (k0 + K < nsteps) ? k0 + K : nsteps;
        orc_level *kp1 = &lv[0], *k = &lv[1], *km1 = &lv[2];
        orc_level_copy(km1, &ck_km1[sg], N, dim);
        orc_level_copy(k, &ck_k[sg], N, dim);
        orc_level_zero(kp1, N, dim);
        for (int s = k0; s < k1 && rc == 0; ++s) {
            orc_inject(rhs, n_src, src_idx, src_chan, src_w, amp + (size_t)s * n_src_chan);
            rc = orc_step(&g, km1, k, rhs, kp1);
            orc_clear(rhs, n_src, src_idx);
            if (s % p == 0) orc_dwaveop(&g, km1, k, kp1, seg_dw + (size_t)(s - k0) * N);
            orc_level *old = km1;
            km1 = k; k = kp1; kp1 = old;
            orc_level_zero(kp1, N, dim);
        }
        /* adjoint over [k0, k1): us = [q_kp1, q_k, q_km1]; the solver state
         * runs "forward" in reversed time, its kp1 is q_{k-1} */
        for (int s = k1 - 1; s >= k0 && rc == 0; --s) {
            if (s % p == 0) {
                const double *dw = seg_dw + (size_t)(s - k0) * N;
#pragma omp parallel for schedule(static)
                for (long i = 0; i < N; ++i) ic_padded[i] += qk->u[i] * dw[i];
            }
            orc_inject(rhs, n_adj, adj_idx, adj_chan, adj_w,
                       operand + (size_t)s * n_adj_chan);
            rc = orc_step(&g, qkm1, qk, rhs, qkp1);
            orc_clear(rhs, n_adj, adj_idx);
            if (s == 0) break;
            orc_level *old = qkm1;
            qkm1 = qk; qk = qkp1; qkp1 = old;
            orc_level_zero(qkp1, N, dim);
        }
    }
    /* ic *= (-1*dt); ic *= imaging_period */
    for (long i = 0; i < N; ++i) {
        ic_padded[i] *= (-1.0 * dt);
        ic_padded[i] *= (double)p;
    }

    for (int s = 0; s < nseg; ++s) {
        if (ck_km1) orc_level_free(&ck_km1[s]);
        if (ck_k) orc_level_free(&ck_k[s]);
    }
    for (int t = 0; t < 3; ++t) { orc_level_free(&lv[t]); orc_level_free(&q[t]); }
    free(ck_km1); free(ck_k); free(seg_dw); free(rhs);
    return rc;
}

/* add_padding (model_parameter.py:537-593): fold the PML slabs onto the first
 * and last physical planes, x first, then y, then z, each pass on the data
 * the previous pass left; then crop.  `a` (padded) is modified in place like
 * the reference's reshape view; out receives the unpadded block. */
ORC_EXPORT void orc_fold(int dim, int nx, int ny, int nz, int lx, int rx, int ly,
                         int ry, int lz, int rz, double *a, double *out)
{
    if (dim == 2) { ny = 1; ly = ry = 0; }
    const int ox = nx - lx - rx, oy = ny - ly - ry, oz = nz - lz - rz;
#define A3(i, j, k) a[((long)(i) * ny + (j)) * nz + (k)]
    /* x: row lx += sum of rows [0, lx); row lx+ox-1 += sum of rows past it */
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            double s = 0.0;
            for (int i = 0; i < lx; ++i) s += A3(i, j, k);
            A3(lx, j, k) = A3(lx, j, k) + s;
            s = 0.0;
            for (int i = lx + ox; i < nx; ++i) s += A3(i, j, k);
            A3(lx + ox - 1, j, k) = A3(lx + ox - 1, j, k) + s;
        }
    if (dim == 3)
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) {
                double s = 0.0;
                for (int j = 0; j < ly; ++j) s += A3(i, j, k);
                A3(i, ly, k) = A3(i, ly, k) + s;
                s = 0.0;
                for (int j = ly + oy; j < ny; ++j) s += A3(i, j, k);
                A3(i, ly + oy - 1, k) = A3(i, ly + oy - 1, k) + s;
            }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            double s = 0.0;
            for (int k = 0; k < lz; ++k) s += A3(i, j, k);
            A3(i, j, lz) = A3(i, j, lz) + s;
            s = 0.0;
            for (int k = lz + oz; k < nz; ++k) s += A3(i, j, k);
            A3(i, j, lz + oz - 1) = A3(i, j, lz + oz - 1) + s;
        }
    for (int i = 0; i < ox; ++i)
        for (int j = 0; j < oy; ++j)
            for (int k = 0; k < oz; ++k)
                out[((long)i * oy + j) * oz + k] = A3(i + lx, j + ly, k + lz);
#undef A3
}

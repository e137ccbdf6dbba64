// cda2d.hip -- 2D constant-density acoustic leapfrog + PML step for gfx950.
//
// Semantics: one time step of pysit/solvers/constant_density_acoustic/time/
// scalar/constant_density_acoustic_time_scalar_2D.h:43-180 (u, Phix, Phiz on
// the padded grid, taps beyond the grid read 0, edge nodes held at 0, PML
// nodes take the damped update), with the reference's index quirk on the
// right PML side (kernel profile one node stronger than the mesh profile).
//
// Mapping to CDNA: one wavefront owns a tile of 64 consecutive z (the
// contiguous axis, so every row access is one coalesced 256 B transaction)
// by `tile_x` rows along x.  The wave walks down its rows keeping a register
// queue of the 2r+1 rows of u (and Phix) it needs for the x derivatives; the
// centre row plus an r-wide halo on each side goes through a per-wave LDS
// row for the z derivatives.  Tiles farther than r nodes from any PML node
// take a path that never touches Phi.  Source injection (per-tile entry
// lists), receiver sampling (extra waves that gather u_k before it is
// overwritten), the dWaveOp store (forward) and the FP64 imaging condition
// (adjoint) are fused into the same launch.  u_{k+1} overwrites u_{k-1} in
// place; Phi is double-buffered.  No MFMA: there is no dense contraction,
// the step is bound by HBM bandwidth.
#include "common.h"

// No implicit FP contraction: every fused multiply-add below is written out,
// so all kernel modes (propagate, store dWaveOp, adjoint) evaluate the step
// with the same roundings and a checkpoint recompute reproduces the stored
// history bit for bit.
#pragma clang fp contract(off)

namespace pa {

template <typename T>
struct Step2DParams {
    AxisCoef<T> cx, cz;
    TimeCoef<T> tc;
    const T *rhs;          // dense rhs (step entry point only), may be null
    int single;            // 1: one shot given by value, tiles by arithmetic
    pysit_amd_shot shot0;  // used when single
};

template <typename T, int R, int MODE>
__device__ __forceinline__ void tile2d(const pysit_amd_batch &B, const Step2DParams<T> &P,
                                       const pysit_amd_shot &S, int tile, int flags, int tloc,
                                       int slot, int seg_base, int lane, T *su, T *sp)
{
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    const int nx = S.nx, nz = S.nz;
    const int tz = tloc % S.ntz, tx = tloc / S.ntz;
    const int z0 = tz * 64, z = z0 + lane;
    const int x0 = tx * B.tile_x;
    const int x1 = min(x0 + B.tile_x, nx);
    const bool zin = z < nz;
    const bool full = flags & 1;
    // which buffer holds time level k: parity of the steps this propagation took
    const int taken = (MODE == 2) ? (S.nt - 1 - k) : k;
    const int cur = taken & 1;
    const long long off = S.off;
    const T *__restrict__ u = (const T *)B.u[cur] + off;
    T *__restrict__ uo = (T *)B.u[cur ^ 1] + off;  // u_{k-1} in, u_{k+1} out
    const T *__restrict__ px = (const T *)B.phix[cur] + off;
    T *__restrict__ pxo = (T *)B.phix[cur ^ 1] + off;
    const T *__restrict__ pz = (const T *)B.phiz[cur] + off;
    T *__restrict__ pzo = (T *)B.phiz[cur ^ 1] + off;
    const T *__restrict__ model = (const T *)S.model;
    const int mcol = clampi(z - S.mz_off, S.mz_lo, S.mz_hi);
    const T *xl = (const T *)S.xl, *xr = (const T *)S.xr;
    const T *zl = (const T *)S.zl, *zr = (const T *)S.zr;

    // z profile of this lane (kernel and mesh indexing)
    T szk = T(0), szm = T(0);
    if (full && zin) {
        szk = (z <= nz - 2) ? sigma_kernel<T>(zl, S.nzl, zr, S.nzr, z, nz) : T(0);
        szm = sigma_mesh<T>(zl, S.nzl, zr, S.nzr, z, nz);
    }
    const bool img = (MODE == 2) && (k % B.imaging_period == 0) && B.dw;
    const bool store = (MODE == 1) && (k % B.imaging_period == 0) && B.dw;
    const long long soff = (long long)(slot - seg_base) * B.total_n + off;
    const T *dwi = (img || MODE == 3) ? (const T *)B.dw + soff : nullptr;
    T *dwo = store ? (T *)B.dw + soff : nullptr;
    T *fst = B.fstore ? (T *)B.fstore + soff : nullptr;
    double *ic = (MODE == 2) ? B.ic + off : nullptr;
    const T *m1 = (MODE == 3) ? (const T *)B.born_m1 : nullptr;

    T q[2 * R + 1], pq[2 * R + 1];
#pragma unroll
    for (int j = 0; j < 2 * R; ++j) {
        const int xx = x0 - R + j;
        const bool in = zin && xx >= 0 && xx < nx;
        q[j] = in ? u[(long long)xx * nz + z] : T(0);
        pq[j] = (full && in) ? px[(long long)xx * nz + z] : T(0);
    }
    q[2 * R] = T(0);
    pq[2 * R] = T(0);

    int e = 0, e_end = 0;
    const int4 *inj = nullptr;
    if (B.inj_ptr) {
        e = B.inj_ptr[tile];
        e_end = B.inj_ptr[tile + 1];
        inj = reinterpret_cast<const int4 *>(B.inj_pos);
    }
    const double *amp_k = S.amp ? S.amp + (long long)k * S.namp : nullptr;

    // halo lane roles for the LDS row: lanes [0, R) fetch z0-R+lane, lanes
    // [R, 2R) fetch z0+64+(lane-R)
    const int hz = (lane < R) ? (z0 - R + lane) : (z0 + 64 + lane - R);
    const int hpos = (lane < R) ? lane : (64 + lane);
    const bool hlane = lane < 2 * R;
    const bool hin = hlane && hz >= 0 && hz < nz;

    for (int x = x0; x < x1; ++x) {
        const int xa = x + R;
        const bool ain = zin && xa < nx;
        q[2 * R] = ain ? u[(long long)xa * nz + z] : T(0);
        if (full) pq[2 * R] = ain ? px[(long long)xa * nz + z] : T(0);
        const long long row = (long long)x * nz;

        const T uc = q[R];
        su[R + lane] = uc;
        if (hlane) su[hpos] = hin ? u[row + hz] : T(0);
        T pzc = T(0);
        if (full) {
            pzc = zin ? pz[row + z] : T(0);
            sp[R + lane] = pzc;
            if (hlane) sp[hpos] = hin ? pz[row + hz] : T(0);
        }
        // source / adjoint-source term of this row
        double facc = 0.0;
        while (e < e_end && inj[e].x == x) {
            const int4 en = inj[e];
            if (en.z == z) facc += B.inj_w[e] * amp_k[en.w];
            ++e;
        }
        T f = (T)facc;
        if (P.rhs && zin) f += P.rhs[row + z];
        const long long mrow = (long long)clampi(x - S.mx_off, S.mx_lo, S.mx_hi) * S.mnz;
        // Born source: rhs_k = m1_pad * (-dWaveOp0_k)
        if (MODE == 3 && zin) f += m1[mrow + mcol] * (-dwi[row + z]);

        const T um = zin ? uo[row + z] : T(0);
        const T cv = model[mrow + mcol];
        wave_sync();

        // second derivatives: symmetric pairs, far taps first
        T lx = T(0), lz = T(0);
#pragma unroll
        for (int j = R; j >= 1; --j) {
            lx = fma(P.cx.d2[j], q[R + j] + q[R - j], lx);
            lz = fma(P.cz.d2[j], su[R + lane + j] + su[R + lane - j], lz);
        }
        lx = fma(P.cx.d2[0], uc, lx);
        lz = fma(P.cz.d2[0], uc, lz);
        const T lap = lx + lz;
        const T c2 = cv * cv;
        const bool edge = (x == 0) | (x == nx - 1) | (z == 0) | (z >= nz - 1);

        T up, pxn = T(0), pzn = T(0), dw;
        if (!full) {
            const T src = f + lap;
            up = fma(P.tc.dt2, c2 * src, T(2) * uc - um);
            dw = c2 * src;
        } else {
            T dux = T(0), duz = T(0), dpx = T(0), dpz = T(0);
#pragma unroll
            for (int j = R; j >= 1; --j) {
                dux = fma(P.cx.d1[j - 1], q[R + j] - q[R - j], dux);
                duz = fma(P.cz.d1[j - 1], su[R + lane + j] - su[R + lane - j], duz);
                dpx = fma(P.cx.d1[j - 1], pq[R + j] - pq[R - j], dpx);
                dpz = fma(P.cz.d1[j - 1], sp[R + lane + j] - sp[R + lane - j], dpz);
            }
            const T src = ((f + lap) + dpx) + dpz;
            const T sx = sigma_kernel<T>(xl, S.nxl, xr, S.nxr, min(x, nx - 2), nx);
            const T sxm = sigma_mesh<T>(xl, S.nxl, xr, S.nxr, x, nx);
            const T dt = P.tc.dt;
            if (sx != T(0) || szk != T(0)) {
                const T pxc = pq[R];
                pxn = (pxc - dt * sx * pxc) + dt * (szk - sx) * dux;
                pzn = (pzc - dt * szk * pzc) + dt * (sx - szk) * duz;
                const T ss = sx + szk;
                const T fac1 = P.tc.two_dt2 / (T(2) + dt * ss);
                const T fac2 = c2 * src - (um - T(2) * uc) * P.tc.inv_dt2 +
                               ss * um * P.tc.inv_2dt - (sx * szk) * uc;
                up = fac1 * fac2;
            } else {
                pxn = pq[R];
                pzn = pzc;
                up = fma(P.tc.dt2, c2 * src, T(2) * uc - um);
            }
            // dWaveOp with the mesh profile: the step identity gives
            // (u+ - 2u + u-)/dt^2 = c^2 src - sum(sk)(u+ - u-)/(2dt) - prod(sk) u
            dw = c2 * src + ((sxm + szm) - (sx + szk)) * (up - um) * P.tc.inv_2dt +
                 (sxm * szm - sx * szk) * uc;
        }
        if (edge) {
            up = T(0);
            pxn = T(0);
            pzn = T(0);
            dw = T(0);
        }
        if (zin) {
            if (MODE == 2 && img) ic[row + z] += (double)uc * (double)dwi[row + z];
            uo[row + z] = up;
            if (full) {
                pxo[row + z] = pxn;
                pzo[row + z] = pzn;
            }
            if (MODE == 1 && store) dwo[row + z] = dw;
            if (fst) fst[row + z] = uc;
        }
        wave_sync();
#pragma unroll
        for (int j = 0; j < 2 * R; ++j) {
            q[j] = q[j + 1];
            pq[j] = pq[j + 1];
        }
    }
}

template <typename T>
__device__ __forceinline__ void sample2d(const pysit_amd_batch &B, int slot, int rl)
{
    const int4 rd = reinterpret_cast<const int4 *>(B.recs)[rl];  // {shot, r, e0, e1}
    const pysit_amd_shot &S = B.shots[rd.x];
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    const T *u = (const T *)B.u[k & 1] + S.off;
    double acc = 0.0;
    for (int e = rd.z; e < rd.w; ++e) acc += B.rec_w[e] * (double)u[B.rec_idx[e]];
    S.trace[(long long)k * S.nrec + rd.y] = acc;
}

template <typename T, int R, int MODE>
__global__ __launch_bounds__(256) void k_step2d(const pysit_amd_batch B, const Step2DParams<T> P,
                                                int slot, int seg_base, int n_tile_waves,
                                                int n_rec_lanes)
{
    __shared__ T lds_u[4][64 + 2 * R];
    __shared__ T lds_p[4][64 + 2 * R];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gw = xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    if (gw < n_tile_waves) {
        if (P.single) {
            tile2d<T, R, MODE>(B, P, P.shot0, gw, 1, gw, slot, seg_base, lane, lds_u[w], lds_p[w]);
        } else {
            const int4 td = reinterpret_cast<const int4 *>(B.tiles)[gw];
            tile2d<T, R, MODE>(B, P, B.shots[td.x], gw, td.z, td.y, slot, seg_base, lane,
                               lds_u[w], lds_p[w]);
        }
        return;
    }
    if (MODE != 2) {
        const int rl = (gw - n_tile_waves) * 64 + lane;
        if (rl < n_rec_lanes) sample2d<T>(B, slot, rl);
    }
}

template <typename T, int R, int MODE>
static int launch2d(const pysit_amd_batch &B, const Step2DParams<T> &P, int slot, int seg_base,
                    int n_tiles, int n_recs, hipStream_t st)
{
    const int rec_waves = (MODE != 2) ? (n_recs + 63) / 64 : 0;
    const int waves = n_tiles + rec_waves;
    if (waves == 0) return PYSIT_AMD_OK;
    const int blocks = (waves + 3) / 4;
    hipLaunchKernelGGL((k_step2d<T, R, MODE>), dim3(blocks), dim3(256), 0, st, B, P, slot, seg_base,
                       n_tiles, (MODE != 2) ? n_recs : 0);
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

template <typename T, int MODE>
static int launch2d_r(int r, const pysit_amd_batch &B, const Step2DParams<T> &P, int slot,
                      int seg_base, int n_tiles, int n_recs, hipStream_t st)
{
    switch (r) {
    case 1: return launch2d<T, 1, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    case 2: return launch2d<T, 2, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    case 3: return launch2d<T, 3, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    case 4: return launch2d<T, 4, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    }
    set_error("order must be 2, 4, 6 or 8");
    return PYSIT_AMD_EINVAL;
}

template <typename T>
int run2d(const pysit_amd_batch &B, int mode, int slot_begin, int slot_end, int direction,
          int seg_base, const int *active_tiles, const int *active_recs, hipStream_t st)
{
    Step2DParams<T> P{};
    const int r = B.order / 2;
    P.cx = axis_coef<T>(r, B.dx);
    P.cz = axis_coef<T>(r, B.dz);
    P.tc = time_coef<T>(B.dt);
    P.rhs = nullptr;
    P.single = 0;
    const int n = slot_end - slot_begin;
    for (int i = 0; i < n; ++i) {
        const int j = direction > 0 ? slot_begin + i : slot_end - 1 - i;
        const int nt = active_tiles ? active_tiles[j] : B.n_tiles;
        const int nr = active_recs ? active_recs[j] : B.n_rec;
        int rc;
        if (mode == 0) rc = launch2d_r<T, 0>(r, B, P, j, seg_base, nt, nr, st);
        else if (mode == 1) rc = launch2d_r<T, 1>(r, B, P, j, seg_base, nt, nr, st);
        else if (mode == 2) rc = launch2d_r<T, 2>(r, B, P, j, seg_base, nt, 0, st);
        else rc = launch2d_r<T, 3>(r, B, P, j, seg_base, nt, nr, st);
        if (rc) return rc;
    }
    return PYSIT_AMD_OK;
}

template int run2d<float>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                          const int *, hipStream_t);
template int run2d<double>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                           const int *, hipStream_t);

// One step with the reference contract (dense rhs, caller-owned outputs).
template <typename T>
int step2d_entry(int order, const T *km1_u, const T *k_phix, const T *k_phiz, const T *k_u,
                 const T *C, const T *rhs, const T *xl, int nxl, const T *xr, int nxr,
                 const T *zl, int nzl, const T *zr, int nzr, double dt, double dx, double dz,
                 int nx, int nz, T *kp1_phix, T *kp1_phiz, T *kp1_u, hipStream_t st)
{
    PA_CHECK(order == 2 || order == 4 || order == 6 || order == 8, "order must be 2, 4, 6 or 8");
    PA_CHECK(nx >= 1 && nz >= 1, "grid must be non-empty");
    PA_CHECK(km1_u && k_phix && k_phiz && k_u && C && kp1_phix && kp1_phiz && kp1_u,
             "null array argument");
    PA_CHECK((nxl == 0 || xl) && (nxr == 0 || xr) && (nzl == 0 || zl) && (nzr == 0 || zr),
             "null PML array with nonzero length");
    PA_CHECK(nxl >= 0 && nxr >= 0 && nzl >= 0 && nzr >= 0 && nxl <= nx && nxr <= nx &&
                 nzl <= nz && nzr <= nz,
             "PML length out of range");
    PA_CHECK(kp1_u != km1_u && kp1_u != k_u && kp1_phix != k_phix && kp1_phiz != k_phiz,
             "outputs must not alias inputs");
    const long long N = (long long)nx * nz;
    // the kernel updates u in place over u_{k-1}: seed the output with it
    PA_HIP(hipMemcpyAsync(kp1_u, km1_u, N * sizeof(T), hipMemcpyDeviceToDevice, st));

    pysit_amd_batch B{};
    B.dtype = sizeof(T) == 8;
    B.dim = 2;
    B.order = order;
    B.tile_x = 16;
    B.dt = dt;
    B.dx = dx;
    B.dz = dz;
    B.u[0] = (void *)k_u;
    B.u[1] = (void *)kp1_u;
    B.phix[0] = (void *)k_phix;
    B.phix[1] = (void *)kp1_phix;
    B.phiz[0] = (void *)k_phiz;
    B.phiz[1] = (void *)kp1_phiz;
    B.total_n = N;
    B.imaging_period = 1;

    Step2DParams<T> P{};
    const int r = order / 2;
    P.cx = axis_coef<T>(r, dx);
    P.cz = axis_coef<T>(r, dz);
    P.tc = time_coef<T>(dt);
    P.rhs = rhs;
    P.single = 1;
    pysit_amd_shot &S = P.shot0;
    S.nx = nx;
    S.ny = 1;
    S.nz = nz;
    S.nt = 1;
    S.kshift = 0;
    S.ntz = (nz + 63) / 64;
    S.ntx = (nx + B.tile_x - 1) / B.tile_x;
    S.nty = 1;
    S.off = 0;
    S.model = C;
    S.mny = 1;
    S.mnz = nz;
    S.mx_off = 0;
    S.mx_lo = 0;
    S.mx_hi = nx - 1;
    S.mz_off = 0;
    S.mz_lo = 0;
    S.mz_hi = nz - 1;
    S.xl = xl;
    S.xr = xr;
    S.zl = zl;
    S.zr = zr;
    S.nxl = nxl;
    S.nxr = nxr;
    S.nzl = nzl;
    S.nzr = nzr;
    S.amp = nullptr;
    S.namp = 0;
    S.nrec = 0;
    S.trace = nullptr;
    return launch2d_r<T, 0>(r, B, P, 0, 0, S.ntz * S.ntx, 0, st);
}

template int step2d_entry<float>(int, const float *, const float *, const float *, const float *,
                                 const float *, const float *, const float *, int, const float *,
                                 int, const float *, int, const float *, int, double, double,
                                 double, int, int, float *, float *, float *, hipStream_t);
template int step2d_entry<double>(int, const double *, const double *, const double *,
                                  const double *, const double *, const double *, const double *,
                                  int, const double *, int, const double *, int, const double *,
                                  int, double, double, double, int, int, double *, double *,
                                  double *, hipStream_t);

}  // namespace pa

// cda3d.hip -- 3D constant-density acoustic leapfrog + PML step for gfx950.
//
// Semantics: one time step of pysit/solvers/constant_density_acoustic/time/
// scalar/constant_density_acoustic_time_scalar_3D.h:46-236 (u, Phix, Phiy,
// Phiz, psi; taps beyond the grid read 0; edge nodes held at 0; psi += dt*u
// on every updated node; right-side PML index quirk as in 2D).
//
// Mapping to CDNA: a 512-thread workgroup owns a 64 (z, contiguous) x 8 (y)
// column of the grid and walks `tile_x` planes along x.  Each thread keeps
// register queues along x (u, and in PML tiles Phix and psi); the current
// y-z plane of u (and of psi / Phiy / Phiz in PML tiles) is staged in LDS
// with an r-node halo for the y and z derivatives.  Tiles farther than r
// nodes from every PML node never touch Phi or psi: nothing there can read
// them (Phi is zero off the PML, and psi only enters through sigma-weighted
// terms and PML-node derivatives), so they are neither read nor written.
// The step entry point with the reference contract marks every tile as PML
// so that it returns all five fields exactly as the reference defines them.
#include "common.h"

// No implicit FP contraction (see cda2d.hip): all modes round identically.
#pragma clang fp contract(off)

namespace pa {

constexpr int TY3 = 8;  // y rows per workgroup (one wavefront each)
constexpr int TX3 = 8;  // x planes per tile

int tile_rows_3d(int axis) { return axis == 0 ? TX3 : TY3; }

template <typename T>
struct Step3DParams {
    AxisCoef<T> cx, cy, cz;
    TimeCoef<T> tc;
    const T *rhs;
    int single;
    pysit_amd_shot shot0;
};

template <typename T, int R>
struct Planes3D {
    T u[TY3 + 2 * R][64 + 2 * R];
    T ps[TY3 + 2 * R][64 + 2 * R];
    T py[TY3 + 2 * R][64];
    T pz[TY3][64 + 2 * R];
};

template <typename T, int R, int MODE>
__device__ __forceinline__ void tile3d(const pysit_amd_batch &B, const Step3DParams<T> &P,
                                       const pysit_amd_shot &S, int tile, int flags, int tloc,
                                       int slot, int seg_base, Planes3D<T, R> &L, T *lds_f)
{
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    constexpr int LZ = 64 + 2 * R;
    const int tid = threadIdx.x;
    const int lane = tid & 63, ty = tid >> 6;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int tz = tloc % S.ntz;
    const int tyi = (tloc / S.ntz) % S.nty;
    const int tx = tloc / (S.ntz * S.nty);
    const int z0 = tz * 64, y0 = tyi * TY3;
    const int z = z0 + lane, y = y0 + ty;
    const int x0 = tx * TX3, x1 = min(x0 + TX3, nx);
    const bool valid = (z < nz) && (y < ny);
    const bool full = flags & 1;
    const int taken = (MODE == 2) ? (S.nt - 1 - k) : k;
    const int cur = taken & 1;
    const long long off = S.off;
    const long long sx_ = (long long)ny * nz;  // x stride
    const T *__restrict__ u = (const T *)B.u[cur] + off;
    T *__restrict__ uo = (T *)B.u[cur ^ 1] + off;
    const T *__restrict__ px = (const T *)B.phix[cur] + off;
    T *__restrict__ pxo = (T *)B.phix[cur ^ 1] + off;
    const T *__restrict__ py = (const T *)B.phiy[cur] + off;
    T *__restrict__ pyo = (T *)B.phiy[cur ^ 1] + off;
    const T *__restrict__ pz = (const T *)B.phiz[cur] + off;
    T *__restrict__ pzo = (T *)B.phiz[cur ^ 1] + off;
    const T *__restrict__ ps = (const T *)B.psi[cur] + off;
    T *__restrict__ pso = (T *)B.psi[cur ^ 1] + off;
    const T *__restrict__ model = (const T *)S.model;
    const long long myz = (long long)clampi(y - S.my_off, S.my_lo, S.my_hi) * S.mnz +
                          clampi(z - S.mz_off, S.mz_lo, S.mz_hi);
    const T *xl = (const T *)S.xl, *xr = (const T *)S.xr;
    const T *yl = (const T *)S.yl, *yr = (const T *)S.yr;
    const T *zl = (const T *)S.zl, *zr = (const T *)S.zr;

    T syk = T(0), sym = T(0), szk = T(0), szm = T(0);
    if (full && valid) {
        syk = (y <= ny - 2) ? sigma_kernel<T>(yl, S.nyl, yr, S.nyr, y, ny) : T(0);
        sym = sigma_mesh<T>(yl, S.nyl, yr, S.nyr, y, ny);
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

    const long long yz = (long long)y * nz + z;
    T q[2 * R + 1], qx[2 * R + 1], qs[2 * R + 1];
#pragma unroll
    for (int j = 0; j < 2 * R; ++j) {
        const int xx = x0 - R + j;
        const bool in = valid && xx >= 0 && xx < nx;
        q[j] = in ? u[xx * sx_ + yz] : T(0);
        qx[j] = (full && in) ? px[xx * sx_ + yz] : T(0);
        qs[j] = (full && in) ? ps[xx * sx_ + yz] : T(0);
    }
    q[2 * R] = qx[2 * R] = qs[2 * R] = T(0);

    // sources of this tile, gathered into LDS (x-plane, y row, z lane): one
    // thread per destination node, entries summed in the reference's order
    bool has_inj = false;
    if (B.inj_ptr) {
        const int n0 = B.inj_ptr[tile], n1 = B.inj_ptr[tile + 1];
        if (n1 > n0) {
            const double *amp_k = S.amp + (long long)k * S.namp;
            for (int i = tid; i < TX3 * TY3 * 64; i += TY3 * 64) lds_f[i] = T(0);
            __syncthreads();
            for (int n = n0 + tid; n < n1; n += TY3 * 64) {
                double acc = 0.0;
                for (int e = B.inj_nptr[n]; e < B.inj_nptr[n + 1]; ++e)
                    acc += B.inj_w[e] * amp_k[B.inj_chan[e]];
                lds_f[B.inj_node[n]] = (T)acc;
            }
            __syncthreads();
            has_inj = true;
        }
    }

    // halo elements of an LY x LZ plane outside the TY3 x 64 centre
    constexpr int HU = 2 * R * LZ + TY3 * 2 * R;
    auto halo_uz = [&](int h, int &ly, int &lz) {
        if (h < 2 * R * LZ) {
            const int rr = h / LZ;
            ly = rr < R ? rr : rr + TY3;
            lz = h % LZ;
        } else {
            const int hh = h - 2 * R * LZ;
            ly = R + hh / (2 * R);
            const int c = hh % (2 * R);
            lz = c < R ? c : c + 64;
        }
    };

    for (int x = x0; x < x1; ++x) {
        const int xa = x + R;
        const bool ain = valid && xa < nx;
        q[2 * R] = ain ? u[xa * sx_ + yz] : T(0);
        if (full) {
            qx[2 * R] = ain ? px[xa * sx_ + yz] : T(0);
            qs[2 * R] = ain ? ps[xa * sx_ + yz] : T(0);
        }
        const long long pl = (long long)x * sx_;
        const T uc = q[R];
        L.u[R + ty][R + lane] = uc;
        T pyc = T(0), pzc = T(0);
        if (full) {
            L.ps[R + ty][R + lane] = qs[R];
            pyc = valid ? py[pl + yz] : T(0);
            pzc = valid ? pz[pl + yz] : T(0);
            L.py[R + ty][lane] = pyc;
            L.pz[ty][R + lane] = pzc;
        }
        for (int h = tid; h < HU; h += TY3 * 64) {
            int ly, lz;
            halo_uz(h, ly, lz);
            const int gy = y0 - R + ly, gz = z0 - R + lz;
            const bool in = gy >= 0 && gy < ny && gz >= 0 && gz < nz;
            const long long g = pl + (long long)gy * nz + gz;
            L.u[ly][lz] = in ? u[g] : T(0);
            if (full) L.ps[ly][lz] = in ? ps[g] : T(0);
        }
        if (full) {
            for (int h = tid; h < 2 * R * 64; h += TY3 * 64) {
                const int rr = h / 64, c = h % 64;
                const int ly = rr < R ? rr : rr + TY3;
                const int gy = y0 - R + ly, gz = z0 + c;
                const bool in = gy >= 0 && gy < ny && gz < nz;
                L.py[ly][c] = in ? py[pl + (long long)gy * nz + gz] : T(0);
            }
            for (int h = tid; h < TY3 * 2 * R; h += TY3 * 64) {
                const int rr = h / (2 * R), c = h % (2 * R);
                const int lz = c < R ? c : c + 64;
                const int gy = y0 + rr, gz = z0 - R + lz;
                const bool in = gy < ny && gz >= 0 && gz < nz;
                L.pz[rr][lz] = in ? pz[pl + (long long)gy * nz + gz] : T(0);
            }
        }
        T f = has_inj ? lds_f[((x - x0) * TY3 + ty) * 64 + lane] : T(0);
        if (P.rhs && valid) f += P.rhs[pl + yz];
        const long long mi = (long long)clampi(x - S.mx_off, S.mx_lo, S.mx_hi) * S.mny * S.mnz + myz;
        if (MODE == 3 && valid) f += m1[mi] * (-dwi[pl + yz]);
        const T um = valid ? uo[pl + yz] : T(0);
        const T cv = model[mi];
        __syncthreads();

        if (valid) {
            T lx = T(0), lyy = T(0), lz = T(0);
#pragma unroll
            for (int j = R; j >= 1; --j) {
                lx = fma(P.cx.d2[j], q[R + j] + q[R - j], lx);
                lyy = fma(P.cy.d2[j], L.u[R + ty + j][R + lane] + L.u[R + ty - j][R + lane], lyy);
                lz = fma(P.cz.d2[j], L.u[R + ty][R + lane + j] + L.u[R + ty][R + lane - j], lz);
            }
            lx = fma(P.cx.d2[0], uc, lx);
            lyy = fma(P.cy.d2[0], uc, lyy);
            lz = fma(P.cz.d2[0], uc, lz);
            const T lap = (lx + lyy) + lz;
            const T c2 = cv * cv;
            const bool edge = (x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) |
                              (z == nz - 1);
            T up, dw, pxn = T(0), pyn = T(0), pzn = T(0), psn = T(0);
            if (!full) {
                const T src = f + lap;
                up = fma(P.tc.dt2, c2 * src, T(2) * uc - um);
                dw = c2 * src;
            } else {
                T dux = T(0), duy = T(0), duz = T(0), dpx = T(0), dpy = T(0), dpz = T(0);
                T dsx = T(0), dsy = T(0), dsz = T(0);
#pragma unroll
                for (int j = R; j >= 1; --j) {
                    const T ax = P.cx.d1[j - 1], ay = P.cy.d1[j - 1], az = P.cz.d1[j - 1];
                    dux = fma(ax, q[R + j] - q[R - j], dux);
                    dpx = fma(ax, qx[R + j] - qx[R - j], dpx);
                    dsx = fma(ax, qs[R + j] - qs[R - j], dsx);
                    duy = fma(ay, L.u[R + ty + j][R + lane] - L.u[R + ty - j][R + lane], duy);
                    dpy = fma(ay, L.py[R + ty + j][lane] - L.py[R + ty - j][lane], dpy);
                    dsy = fma(ay, L.ps[R + ty + j][R + lane] - L.ps[R + ty - j][R + lane], dsy);
                    duz = fma(az, L.u[R + ty][R + lane + j] - L.u[R + ty][R + lane - j], duz);
                    dpz = fma(az, L.pz[ty][R + lane + j] - L.pz[ty][R + lane - j], dpz);
                    dsz = fma(az, L.ps[R + ty][R + lane + j] - L.ps[R + ty][R + lane - j], dsz);
                }
                const T src = (((f + lap) + dpx) + dpy) + dpz;
                const T sxk = sigma_kernel<T>(xl, S.nxl, xr, S.nxr, min(x, nx - 2), nx);
                const T sxm = sigma_mesh<T>(xl, S.nxl, xr, S.nxr, x, nx);
                const T dt = P.tc.dt;
                const T psc = qs[R];
                psn = psc + dt * uc;
                const T s1k = (sxk + syk) + szk;
                const T s2k = (sxk * syk + syk * szk) + szk * sxk;
                const T s3k = (sxk * syk) * szk;
                if (sxk != T(0) || syk != T(0) || szk != T(0)) {
                    const T pxc = qx[R];
                    pxn = (pxc - dt * sxk * pxc) + dt * ((syk + szk) - sxk) * dux + dt * (syk * szk * dsx);
                    pyn = (pyc - dt * syk * pyc) + dt * ((szk + sxk) - syk) * duy + dt * (szk * sxk * dsy);
                    pzn = (pzc - dt * szk * pzc) + dt * ((sxk + syk) - szk) * duz + dt * (sxk * syk * dsz);
                    const T fac1 = P.tc.two_dt2 / (T(2) + dt * s1k);
                    const T fac2 = c2 * (src - s3k * psc) - (um - T(2) * uc) * P.tc.inv_dt2 +
                                   s1k * um * P.tc.inv_2dt - s2k * uc;
                    up = fac1 * fac2;
                } else {
                    pxn = qx[R];
                    pyn = pyc;
                    pzn = pzc;
                    up = fma(P.tc.dt2, c2 * src, T(2) * uc - um);
                }
                const T s1m = (sxm + sym) + szm;
                const T s2m = (sxm * sym + sxm * szm) + sym * szm;
                const T s3m = (sxm * sym) * szm;
                dw = c2 * (src - s3k * psc) + (s1m - s1k) * (up - um) * P.tc.inv_2dt +
                     (s2m - s2k) * uc + s3m * psc;
            }
            if (edge) {
                up = dw = pxn = pyn = pzn = psn = T(0);
            }
            const long long g = pl + yz;
            if (MODE == 2 && img) ic[g] += (double)uc * (double)dwi[g];
            uo[g] = up;
            if (full) {
                pxo[g] = pxn;
                pyo[g] = pyn;
                pzo[g] = pzn;
                pso[g] = psn;
            }
            if (MODE == 1 && store) dwo[g] = dw;
            if (fst) fst[g] = uc;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 2 * R; ++j) {
            q[j] = q[j + 1];
            qx[j] = qx[j + 1];
            qs[j] = qs[j + 1];
        }
    }
}

// Receiver sampling, 16 lanes per receiver with a fixed reduction order
// (as in 2D): entries g, g+16, ... per lane, then a butterfly.
template <typename T>
__device__ __forceinline__ void sample_group3(const pysit_amd_batch &B, int slot, int rec, int g)
{
    const bool act = rec < B.n_rec;
    int4 rd = make_int4(0, 0, 0, 0);
    if (act) rd = reinterpret_cast<const int4 *>(B.recs)[rec];
    const pysit_amd_shot *S = act ? &B.shots[rd.x] : nullptr;
    const int k = act ? slot - S->kshift : -1;
    const bool live = act && k >= 0 && k < S->nt;
    double acc = 0.0;
    if (live) {
        const T *u = (const T *)B.u[k & 1] + S->off;
        for (int e = rd.z + g; e < rd.w; e += 16) acc += B.rec_w[e] * (double)u[B.rec_idx[e]];
    }
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 16);
    if (live && g == 0) S->trace[(long long)k * S->nrec + rd.y] = acc;
}

template <typename T, int R, int MODE>
__global__ __launch_bounds__(512) void k_step3d(const pysit_amd_batch B, const Step3DParams<T> P,
                                                int slot, int seg_base, int n_tile_blocks,
                                                int n_rec_lanes)
{
    __shared__ Planes3D<T, R> L;
    __shared__ T lds_f[TX3 * TY3 * 64];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    if (blk < n_tile_blocks) {
        if (P.single) {
            tile3d<T, R, MODE>(B, P, P.shot0, blk, 1, blk, slot, seg_base, L, lds_f);
        } else {
            const int4 td = reinterpret_cast<const int4 *>(B.tiles)[blk];
            tile3d<T, R, MODE>(B, P, B.shots[td.x], blk, td.z, td.y, slot, seg_base, L, lds_f);
        }
        return;
    }
    if (MODE != 2) {
        const int rec = (blk - n_tile_blocks) * (TY3 * 4) + (threadIdx.x >> 4);
        sample_group3<T>(B, slot, rec < n_rec_lanes ? rec : B.n_rec, threadIdx.x & 15);
    }
}

template <typename T, int R, int MODE>
static int launch3d(const pysit_amd_batch &B, const Step3DParams<T> &P, int slot, int seg_base,
                    int n_tiles, int n_recs, hipStream_t st)
{
    const int rec_blocks = (MODE != 2) ? (n_recs + TY3 * 4 - 1) / (TY3 * 4) : 0;
    const int blocks = n_tiles + rec_blocks;
    if (blocks == 0) return PYSIT_AMD_OK;
    hipLaunchKernelGGL((k_step3d<T, R, MODE>), dim3(blocks), dim3(TY3 * 64), 0, st, B, P, slot,
                       seg_base, n_tiles, (MODE != 2) ? n_recs : 0);
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

template <typename T, int MODE>
static int launch3d_r(int r, const pysit_amd_batch &B, const Step3DParams<T> &P, int slot,
                      int seg_base, int n_tiles, int n_recs, hipStream_t st)
{
    switch (r) {
    case 1: return launch3d<T, 1, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    case 2: return launch3d<T, 2, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    case 3: return launch3d<T, 3, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    case 4: return launch3d<T, 4, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st);
    }
    set_error("order must be 2, 4, 6 or 8");
    return PYSIT_AMD_EINVAL;
}

template <typename T>
int run3d(const pysit_amd_batch &B, int mode, int slot_begin, int slot_end, int direction,
          int seg_base, const int *active_tiles, const int *active_recs, hipStream_t st)
{
    PA_CHECK(B.tile_x == TX3 && B.tile_y == TY3, "3D tiles are 8 x 8 (tile_x, tile_y)");
    Step3DParams<T> P{};
    const int r = B.order / 2;
    P.cx = axis_coef<T>(r, B.dx);
    P.cy = axis_coef<T>(r, B.dy);
    P.cz = axis_coef<T>(r, B.dz);
    P.tc = time_coef<T>(B.dt);
    P.rhs = nullptr;
    P.single = 0;
    for (int i = 0; i < slot_end - slot_begin; ++i) {
        const int j = direction > 0 ? slot_begin + i : slot_end - 1 - i;
        const int nt = active_tiles ? active_tiles[j] : B.n_tiles;
        const int nr = active_recs ? active_recs[j] : B.n_rec;
        int rc;
        if (mode == 0) rc = launch3d_r<T, 0>(r, B, P, j, seg_base, nt, nr, st);
        else if (mode == 1) rc = launch3d_r<T, 1>(r, B, P, j, seg_base, nt, nr, st);
        else if (mode == 2) rc = launch3d_r<T, 2>(r, B, P, j, seg_base, nt, 0, st);
        else rc = launch3d_r<T, 3>(r, B, P, j, seg_base, nt, nr, st);
        if (rc) return rc;
    }
    return PYSIT_AMD_OK;
}

template int run3d<float>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                          const int *, hipStream_t);
template int run3d<double>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                           const int *, hipStream_t);

template <typename T>
int step3d_entry(int order, const T *km1_u, const T *k_phix, const T *k_phiy, const T *k_phiz,
                 const T *k_psi, const T *k_u, const T *C, const T *rhs, const T *xl, int nxl,
                 const T *xr, int nxr, const T *yl, int nyl, const T *yr, int nyr, const T *zl,
                 int nzl, const T *zr, int nzr, double dt, double dx, double dy, double dz, int nx,
                 int ny, int nz, T *kp1_phix, T *kp1_phiy, T *kp1_phiz, T *kp1_psi, T *kp1_u,
                 hipStream_t st)
{
    PA_CHECK(order == 2 || order == 4 || order == 6 || order == 8, "order must be 2, 4, 6 or 8");
    PA_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid must be non-empty");
    PA_CHECK(km1_u && k_phix && k_phiy && k_phiz && k_psi && k_u && C && kp1_phix && kp1_phiy &&
                 kp1_phiz && kp1_psi && kp1_u,
             "null array argument");
    PA_CHECK((nxl == 0 || xl) && (nxr == 0 || xr) && (nyl == 0 || yl) && (nyr == 0 || yr) &&
                 (nzl == 0 || zl) && (nzr == 0 || zr),
             "null PML array with nonzero length");
    PA_CHECK(nxl >= 0 && nxr >= 0 && nyl >= 0 && nyr >= 0 && nzl >= 0 && nzr >= 0 &&
                 nxl <= nx && nxr <= nx && nyl <= ny && nyr <= ny && nzl <= nz && nzr <= nz,
             "PML length out of range");
    PA_CHECK(kp1_u != km1_u && kp1_u != k_u && kp1_phix != k_phix && kp1_phiy != k_phiy &&
                 kp1_phiz != k_phiz && kp1_psi != k_psi,
             "outputs must not alias inputs");
    const long long N = (long long)nx * ny * nz;
    PA_HIP(hipMemcpyAsync(kp1_u, km1_u, N * sizeof(T), hipMemcpyDeviceToDevice, st));
    pysit_amd_batch B{};
    B.dtype = sizeof(T) == 8;
    B.dim = 3;
    B.order = order;
    B.tile_x = TX3;
    B.tile_y = TY3;
    B.dt = dt;
    B.dx = dx;
    B.dy = dy;
    B.dz = dz;
    B.u[0] = (void *)k_u;
    B.u[1] = (void *)kp1_u;
    B.phix[0] = (void *)k_phix;
    B.phix[1] = (void *)kp1_phix;
    B.phiy[0] = (void *)k_phiy;
    B.phiy[1] = (void *)kp1_phiy;
    B.phiz[0] = (void *)k_phiz;
    B.phiz[1] = (void *)kp1_phiz;
    B.psi[0] = (void *)k_psi;
    B.psi[1] = (void *)kp1_psi;
    B.total_n = N;
    B.imaging_period = 1;
    Step3DParams<T> P{};
    const int r = order / 2;
    P.cx = axis_coef<T>(r, dx);
    P.cy = axis_coef<T>(r, dy);
    P.cz = axis_coef<T>(r, dz);
    P.tc = time_coef<T>(dt);
    P.rhs = rhs;
    P.single = 1;
    pysit_amd_shot &S = P.shot0;
    S.nx = nx;
    S.ny = ny;
    S.nz = nz;
    S.nt = 1;
    S.kshift = 0;
    S.ntz = (nz + 63) / 64;
    S.nty = (ny + TY3 - 1) / TY3;
    S.ntx = (nx + TX3 - 1) / TX3;
    S.off = 0;
    S.model = C;
    S.mny = ny;
    S.mnz = nz;
    S.mx_off = 0;
    S.mx_lo = 0;
    S.mx_hi = nx - 1;
    S.my_off = 0;
    S.my_lo = 0;
    S.my_hi = ny - 1;
    S.mz_off = 0;
    S.mz_lo = 0;
    S.mz_hi = nz - 1;
    S.xl = xl;
    S.xr = xr;
    S.yl = yl;
    S.yr = yr;
    S.zl = zl;
    S.zr = zr;
    S.nxl = nxl;
    S.nxr = nxr;
    S.nyl = nyl;
    S.nyr = nyr;
    S.nzl = nzl;
    S.nzr = nzr;
    return launch3d_r<T, 0>(r, B, P, 0, 0, S.ntz * S.nty * S.ntx, 0, st);
}

template int step3d_entry<float>(int, const float *, const float *, const float *, const float *,
                                 const float *, const float *, const float *, const float *,
                                 const float *, int, const float *, int, const float *, int,
                                 const float *, int, const float *, int, const float *, int,
                                 double, double, double, double, int, int, int, float *, float *,
                                 float *, float *, float *, hipStream_t);
template int step3d_entry<double>(int, const double *, const double *, const double *,
                                  const double *, const double *, const double *, const double *,
                                  const double *, const double *, int, const double *, int,
                                  const double *, int, const double *, int, const double *, int,
                                  const double *, int, double, double, double, double, int, int,
                                  int, double *, double *, double *, double *, double *,
                                  hipStream_t);

}  // namespace pa

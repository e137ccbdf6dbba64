// cda2d.hip -- 2D constant-density acoustic leapfrog + PML step for gfx950.
//
// Semantics: one time step of pysit/solvers/constant_density_acoustic/time/
// scalar/constant_density_acoustic_time_scalar_2D.h:43-180 (u, Phix, Phiz on
// the padded grid, taps beyond the grid read 0, edge nodes held at 0, PML
// nodes take the damped update), with the reference's index quirk on the
// right PML side (kernel profile one node stronger than the mesh profile).
//
// Mapping to CDNA: one wavefront owns a tile of 64 consecutive z (the
// contiguous axis: every row access is one coalesced 256 B transaction) by
// TX rows along x (16 in FP32, 8 in FP64).  A tile is processed in register
// blocks of NB rows: all loads of a block (the NB + 2r rows of u its x
// derivatives need, u_{k-1}, the imaging accumulator) are issued up front so
// each wave keeps tens of loads in flight, then the NB rows are computed
// fully unrolled.  z derivatives take their neighbours from other lanes with
// ds_bpermute (no LDS storage, no fences, so the compiler is free to hoist
// loads); lanes 0..r-1 and 64-r..63 also carry the r-node halo beyond each
// side of the tile.  Tiles farther than r nodes from every PML node take a
// path without Phi; in PML tiles Phi is read and written only where it can
// be nonzero (PML nodes), except for the reference-contract step entry point,
// which treats Phi as arbitrary input.
//
// Fused per launch: source / adjoint-source injection (per-tile entry
// lists, gathered into LDS before the rows are computed), receiver sampling
// (extra waves gathering u_k before it is overwritten), the dWaveOp store
// (forward), the FP64 imaging condition (adjoint), the Born source, and the
// field store.  u_{k+1} overwrites u_{k-1} in place; Phi is double-buffered.
// No MFMA: there is no dense contraction; the step is bound by HBM bandwidth.
#include "common.h"

// No implicit FP contraction: every fused multiply-add below is written out,
// so all kernel modes (propagate, store dWaveOp, adjoint, Born) evaluate the
// step with the same roundings and a checkpoint recompute reproduces the
// stored history bit for bit.
#pragma clang fp contract(off)

namespace pa {

template <typename T> struct Tile2D;
template <> struct Tile2D<float> { static constexpr int TX = 16; };
template <> struct Tile2D<double> { static constexpr int TX = 8; };

// rows per register block: fewer where more state is live per row
template <typename T, int MODE, bool FULL>
constexpr int block_rows()
{
    return Tile2D<T>::TX / ((FULL ? 2 : 1) * (MODE == 2 ? 2 : 1));
}

template <typename T>
struct Step2DParams {
    AxisCoef<T> cx, cz;
    TimeCoef<T> tc;
    const T *rhs;          // dense rhs (step entry point only), may be null
    int single;            // 1: one shot given by value, tiles by arithmetic
    int general_phi;       // 1: Phi may be nonzero anywhere (step entry point)
    pysit_amd_shot shot0;  // used when single
};

// value of lane `src_byte/4` of v (32- or 64-bit)
__device__ __forceinline__ float lane_read(float v, int src_byte)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_byte, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int src_byte)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src_byte, (int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_ds_bpermute(src_byte, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Per-tile constants shared by the row blocks.  Every field access goes
// through a raw buffer resource based at this shot's region: masked lanes use
// an out-of-range offset, so loads and stores need no branches and the whole
// register block is one straight-line region the scheduler can reorder.
template <typename T, int R>
struct Ctx2D {
    Rsrc u, uo, px, pxo, pz, pzo, model, m1, dwi, dwo, fst, ic, rhs;
    bool has_fst, has_dwo, has_rhs, has_inj;
    const T *finj;  // LDS row sources of this wave
    const T *xl, *xr;
    int nxl, nxr, nzl, nzr;
    int nx, nz, z, zc, hz, hzc, lane, mcol, mx_off, mx_lo, mx_hi, mnz;
    bool zin, hin, general_phi, img;
    T szk, szm;  // lane z profile (kernel, mesh)
    int srcp[R + 1], srcm[R + 1];
    bool inp[R + 1], inm[R + 1];
};

template <typename T, int R, int NB, int MODE, bool FULL>
__device__ __forceinline__ void rows2d(const Ctx2D<T, R> &c, const Step2DParams<T> &P, int xb, int jb)
{
    const int nz = c.nz, nx = c.nx, z = c.z;
    // ---- loads of the block -------------------------------------------------
    T uc[NB + 2 * R];
#pragma unroll
    for (int j = 0; j < NB + 2 * R; ++j) {
        const int xx = xb - R + j;
        uc[j] = buf_ld<T>(c.u, boff<T>(c.zin && xx >= 0 && xx < nx, (long long)xx * nz + z));
    }
    T uh[NB], um[NB], cv[NB], fv[NB], dv[NB], mv[NB];
    double icv[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int x = xb + j;
        const long long row = (long long)x * nz;
        const bool st = c.zin && x < nx;
        uh[j] = buf_ld<T>(c.u, boff<T>(c.hin && x < nx, row + c.hz));
        um[j] = buf_ld<T>(c.uo, boff<T>(st, row + z));
        const long long mrow = (long long)clampi(x - c.mx_off, c.mx_lo, c.mx_hi) * c.mnz;
        cv[j] = buf_ld<T>(c.model, boff<T>(true, mrow + c.mcol));
        fv[j] = c.has_rhs ? buf_ld<T>(c.rhs, boff<T>(st, row + z)) : T(0);
        if (c.has_inj) fv[j] += c.finj[(jb + j) * 64 + c.lane];
        dv[j] = (MODE == 3 || (MODE == 2 && c.img)) ? buf_ld<T>(c.dwi, boff<T>(st, row + z)) : T(0);
        mv[j] = (MODE == 3) ? buf_ld<T>(c.m1, boff<T>(true, mrow + c.mcol)) : T(0);
        icv[j] = (MODE == 2 && c.img) ? buf_ld<double>(c.ic, boff<double>(st, row + z)) : 0.0;
    }
    // PML tiles: Phix along x for the block, Phiz centre and halo per row,
    // read only where Phi can be nonzero (PML nodes)
    T pc[FULL ? NB + 2 * R : 1], pv[FULL ? NB : 1], ph[FULL ? NB : 1];
    if (FULL) {
        const bool zp = c.general_phi || c.z < c.nzl || c.z >= nz - c.nzr;
        const bool hzp = c.general_phi || c.hz < c.nzl || c.hz >= nz - c.nzr;
#pragma unroll
        for (int j = 0; j < NB + 2 * R; ++j) {
            const int xx = xb - R + j;
            const bool xp = xx < c.nxl || xx >= nx - c.nxr;
            pc[j] = buf_ld<T>(c.px, boff<T>((xp || zp) && c.zin && xx >= 0 && xx < nx, (long long)xx * nz + z));
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int x = xb + j;
            const long long row = (long long)x * nz;
            const bool xp = x < c.nxl || x >= nx - c.nxr;
            pv[j] = buf_ld<T>(c.pz, boff<T>((xp || zp) && c.zin && x < nx, row + z));
            ph[j] = buf_ld<T>(c.pz, boff<T>((xp || hzp) && c.hin && x < nx, row + c.hz));
        }
    }
    // ---- rows ------------------------------------------------------------------
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int x = xb + j;
        const bool st = c.zin && x < nx;
        const long long row = (long long)x * nz;
        const T v = uc[R + j];
        T f = fv[j];
        if (MODE == 3) f += mv[j] * (-dv[j]);  // Born source: m1_pad * (-dWaveOp0_k)

        // z neighbours from other lanes (halo lanes supply the tile edges)
        T zp[R + 1], zm[R + 1];
#pragma unroll
        for (int i = 1; i <= R; ++i) {
            const T a = lane_read(v, c.srcp[i]), b = lane_read(uh[j], c.srcp[i]);
            const T d = lane_read(v, c.srcm[i]), e = lane_read(uh[j], c.srcm[i]);
            zp[i] = c.inp[i] ? a : b;
            zm[i] = c.inm[i] ? d : e;
        }
        T lx = T(0), lz = T(0);
#pragma unroll
        for (int i = R; i >= 1; --i) {
            lx = fma(P.cx.d2[i], uc[R + j + i] + uc[R + j - i], lx);
            lz = fma(P.cz.d2[i], zp[i] + zm[i], lz);
        }
        lx = fma(P.cx.d2[0], v, lx);
        lz = fma(P.cz.d2[0], v, lz);
        const T lap = lx + lz;
        const T c2 = cv[j] * cv[j];
        const T u_m = um[j];
        const bool edge = (x == 0) | (x == nx - 1) | (z == 0) | (z >= nz - 1);

        T up, dw, pxn = T(0), pzn = T(0);
        bool phi_store = false;
        if (!FULL) {
            const T src = f + lap;
            up = fma(P.tc.dt2, c2 * src, T(2) * v - u_m);
            dw = c2 * src;
        } else {
            T pzp[R + 1], pzm[R + 1];
#pragma unroll
            for (int i = 1; i <= R; ++i) {
                const T a = lane_read(pv[j], c.srcp[i]), b = lane_read(ph[j], c.srcp[i]);
                const T d = lane_read(pv[j], c.srcm[i]), e = lane_read(ph[j], c.srcm[i]);
                pzp[i] = c.inp[i] ? a : b;
                pzm[i] = c.inm[i] ? d : e;
            }
            T dux = T(0), duz = T(0), dpx = T(0), dpz = T(0);
#pragma unroll
            for (int i = R; i >= 1; --i) {
                dux = fma(P.cx.d1[i - 1], uc[R + j + i] - uc[R + j - i], dux);
                duz = fma(P.cz.d1[i - 1], zp[i] - zm[i], duz);
                dpx = fma(P.cx.d1[i - 1], pc[R + j + i] - pc[R + j - i], dpx);
                dpz = fma(P.cz.d1[i - 1], pzp[i] - pzm[i], dpz);
            }
            const T src = ((f + lap) + dpx) + dpz;
            const T sx = sigma_kernel<T>(c.xl, c.nxl, c.xr, c.nxr, min(x, nx - 2), nx);
            const T sxm = sigma_mesh<T>(c.xl, c.nxl, c.xr, c.nxr, min(x, nx - 1), nx);
            const T szk = c.szk, szm = c.szm;
            const T dt = P.tc.dt;
            const T pxc = pc[R + j], pzc = pv[j];
            const bool pml = sx != T(0) || szk != T(0);
            // damped update (PML node) and plain update, selected per lane
            const T pxd = (pxc - dt * sx * pxc) + dt * (szk - sx) * dux;
            const T pzd = (pzc - dt * szk * pzc) + dt * (sx - szk) * duz;
            const T ss = sx + szk;
            const T fac1 = P.tc.two_dt2 / (T(2) + dt * ss);
            const T fac2 = c2 * src - (u_m - T(2) * v) * P.tc.inv_dt2 + ss * u_m * P.tc.inv_2dt -
                           (sx * szk) * v;
            const T upd = fac1 * fac2;
            const T upi = fma(P.tc.dt2, c2 * src, T(2) * v - u_m);
            pxn = pml ? pxd : pxc;
            pzn = pml ? pzd : pzc;
            up = pml ? upd : upi;
            phi_store = pml || c.general_phi;
            // dWaveOp with the mesh profile: the step identity gives
            // (u+ - 2u + u-)/dt^2 = c^2 src - sum(sk)(u+ - u-)/(2dt) - prod(sk) u
            dw = c2 * src + ((sxm + szm) - (sx + szk)) * (up - u_m) * P.tc.inv_2dt +
                 (sxm * szm - sx * szk) * v;
        }
        if (edge) {
            up = T(0);
            pxn = T(0);
            pzn = T(0);
            dw = T(0);
            phi_store = FULL;
        }
        const int o = boff<T>(st, row + z);
        if (MODE == 2 && c.img) buf_st(c.ic, boff<double>(st, row + z), icv[j] + (double)v * (double)dv[j]);
        buf_st(c.uo, o, up);
        if (FULL) {
            const int op = (st && phi_store) ? o : kOOB;
            buf_st(c.pxo, op, pxn);
            buf_st(c.pzo, op, pzn);
        }
        if (MODE == 1 && c.has_dwo) buf_st(c.dwo, o, dw);
        if (c.has_fst) buf_st(c.fst, o, v);
    }
}

template <typename T, int R, int MODE>
__device__ __forceinline__ void tile2d(const pysit_amd_batch &B, const Step2DParams<T> &P,
                                       const pysit_amd_shot &S, int tile, int flags, int tloc,
                                       int slot, int seg_base, int lane, T *lds_f)
{
    constexpr int TX = Tile2D<T>::TX;
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    Ctx2D<T, R> c;
    c.nx = S.nx;
    c.nz = S.nz;
    const int tz = tloc % S.ntz, tx = tloc / S.ntz;
    const int z0 = tz * 64;
    c.lane = lane;
    c.z = z0 + lane;
    c.zin = c.z < c.nz;
    const int x0 = tx * TX;
    const bool full = flags & 1;
    // which buffer holds time level k: parity of the steps this propagation took
    const int taken = (MODE == 2) ? (S.nt - 1 - k) : k;
    const int cur = taken & 1;
    const long long off = S.off;
    const long long nbytes = (long long)c.nx * c.nz * (long long)sizeof(T);
    c.u = make_rsrc((const T *)B.u[cur] + off, nbytes);
    c.uo = make_rsrc((T *)B.u[cur ^ 1] + off, nbytes);
    c.px = make_rsrc((const T *)B.phix[cur] + off, nbytes);
    c.pxo = make_rsrc((T *)B.phix[cur ^ 1] + off, nbytes);
    c.pz = make_rsrc((const T *)B.phiz[cur] + off, nbytes);
    c.pzo = make_rsrc((T *)B.phiz[cur ^ 1] + off, nbytes);
    c.model = make_rsrc(S.model, (long long)(S.mx_hi + 1) * S.mnz * (long long)sizeof(T));
    c.mcol = clampi(c.z - S.mz_off, S.mz_lo, S.mz_hi);
    c.mx_off = S.mx_off;
    c.mx_lo = S.mx_lo;
    c.mx_hi = S.mx_hi;
    c.mnz = S.mnz;
    c.xl = (const T *)S.xl;
    c.xr = (const T *)S.xr;
    c.nxl = S.nxl;
    c.nxr = S.nxr;
    c.nzl = S.nzl;
    c.nzr = S.nzr;
    c.general_phi = P.general_phi;
    c.has_rhs = P.rhs != nullptr;
    c.rhs = make_rsrc(P.rhs, c.has_rhs ? nbytes : 0);
    // halo lanes: 0..R-1 hold z0+64+lane, 64-R..63 hold z0+lane-64
    c.hz = lane < R ? z0 + 64 + lane : (lane >= 64 - R ? z0 + lane - 64 : -1);
    c.hin = c.hz >= 0 && c.hz < c.nz;
#pragma unroll
    for (int i = 1; i <= R; ++i) {
        c.srcp[i] = ((lane + i) & 63) << 2;
        c.srcm[i] = ((lane - i) & 63) << 2;
        c.inp[i] = lane + i <= 63;
        c.inm[i] = lane - i >= 0;
    }
    const T *zl = (const T *)S.zl, *zr = (const T *)S.zr;
    c.szk = c.szm = T(0);
    if (full && c.zin) {
        c.szk = (c.z <= c.nz - 2) ? sigma_kernel<T>(zl, S.nzl, zr, S.nzr, c.z, c.nz) : T(0);
        c.szm = sigma_mesh<T>(zl, S.nzl, zr, S.nzr, c.z, c.nz);
    }
    c.img = (MODE == 2) && (k % B.imaging_period == 0) && B.dw;
    const bool store = (MODE == 1) && (k % B.imaging_period == 0) && B.dw;
    const long long soff = (long long)(slot - seg_base) * B.total_n + off;
    const bool dwin = (c.img || MODE == 3) && B.dw;
    c.dwi = make_rsrc(dwin ? (const T *)B.dw + soff : nullptr, dwin ? nbytes : 0);
    c.has_dwo = store;
    c.dwo = make_rsrc(store ? (T *)B.dw + soff : nullptr, store ? nbytes : 0);
    c.has_fst = B.fstore != nullptr;
    c.fst = make_rsrc(c.has_fst ? (T *)B.fstore + soff : nullptr, c.has_fst ? nbytes : 0);
    c.ic = make_rsrc(MODE == 2 ? B.ic + off : nullptr,
                     MODE == 2 ? (long long)c.nx * c.nz * (long long)sizeof(double) : 0);
    c.m1 = make_rsrc(MODE == 3 ? B.born_m1 : nullptr,
                     MODE == 3 ? (long long)(S.mx_hi + 1) * S.mnz * (long long)sizeof(T) : 0);

    // sources of this tile, gathered into LDS (row-major TX x 64): one lane
    // per destination node, each summing its entries in the reference's order
    c.has_inj = false;
    c.finj = lds_f;
    if (B.inj_ptr) {
        const int n0 = B.inj_ptr[tile], n1 = B.inj_ptr[tile + 1];
        if (n1 > n0) {
            const double *amp_k = S.amp + (long long)k * S.namp;
#pragma unroll
            for (int j = 0; j < TX; ++j) lds_f[j * 64 + lane] = T(0);
            wave_sync();
            for (int n = n0 + lane; n < n1; n += 64) {
                double acc = 0.0;
                for (int e = B.inj_nptr[n]; e < B.inj_nptr[n + 1]; ++e)
                    acc += B.inj_w[e] * amp_k[B.inj_chan[e]];
                lds_f[B.inj_node[n]] = (T)acc;
            }
            wave_sync();
            c.has_inj = true;
        }
    }

    if (full) {
        constexpr int NB = block_rows<T, MODE, true>();
#pragma unroll 1
        for (int jb = 0; jb < TX; jb += NB) rows2d<T, R, NB, MODE, true>(c, P, x0 + jb, jb);
    } else {
        constexpr int NB = block_rows<T, MODE, false>();
#pragma unroll 1
        for (int jb = 0; jb < TX; jb += NB) rows2d<T, R, NB, MODE, false>(c, P, x0 + jb, jb);
    }
}

// Receiver sampling: 16 lanes per receiver (4 receivers per wavefront); lane
// g sums entries g, g+16, ... of its receiver, then the group reduces with a
// fixed butterfly, so the result does not depend on scheduling.
template <typename T>
__device__ __forceinline__ void sample_group(const pysit_amd_batch &B, int slot, int rec, int g)
{
    const bool act = rec < B.n_rec;
    int4 rd = make_int4(0, 0, 0, 0);
    if (act) rd = reinterpret_cast<const int4 *>(B.recs)[rec];  // {shot, r, e0, e1}
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

// Grid: the first `rec_blocks` blocks sample receivers (dispatched first, so
// their short gather chains overlap the tiles), the rest own tiles.  SINGLE
// is the reference-contract step: one shot passed by value, every tile PML.
template <typename T, int R, int MODE, bool SINGLE>
__global__ __launch_bounds__(256) void k_step2d(const pysit_amd_batch B, const Step2DParams<T> P,
                                                int slot, int seg_base, int n_tile_waves,
                                                int n_rec_lanes, int rec_blocks)
{
    __shared__ T lds_f[4][Tile2D<T>::TX * 64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = blockIdx.x;
    if (MODE != 2 && b < rec_blocks) {
        const int rw = b * 4 + w;
        const int rec = rw * 4 + (lane >> 4);
        if (rw * 4 < n_rec_lanes) sample_group<T>(B, slot, rec < n_rec_lanes ? rec : B.n_rec, lane & 15);
        return;
    }
    const int gw = xcd_block(b - rec_blocks, gridDim.x - rec_blocks) * 4 + w;
    if (gw >= n_tile_waves) return;
    if (SINGLE) {
        tile2d<T, R, MODE>(B, P, P.shot0, gw, 1, gw, slot, seg_base, lane, lds_f[w]);
    } else {
        const int4 td = reinterpret_cast<const int4 *>(B.tiles)[gw];
        tile2d<T, R, MODE>(B, P, B.shots[td.x], gw, td.z, td.y, slot, seg_base, lane, lds_f[w]);
    }
}

template <typename T, int R, int MODE>
static int launch2d(const pysit_amd_batch &B, const Step2DParams<T> &P, int slot, int seg_base,
                    int n_tiles, int n_recs, hipStream_t st)
{
    const int recs = (MODE != 2) ? n_recs : 0;
    // 16 receivers per block, block count rounded to whole XCD rounds
    const int rec_blocks = recs ? ((recs + 15) / 16 + 7) / 8 * 8 : 0;
    const int tile_blocks = (n_tiles + 3) / 4;
    if (rec_blocks + tile_blocks == 0) return PYSIT_AMD_OK;
    if (P.single)
        hipLaunchKernelGGL((k_step2d<T, R, MODE, true>), dim3(rec_blocks + tile_blocks), dim3(256), 0, st,
                           B, P, slot, seg_base, n_tiles, recs, rec_blocks);
    else
        hipLaunchKernelGGL((k_step2d<T, R, MODE, false>), dim3(rec_blocks + tile_blocks), dim3(256), 0, st,
                           B, P, slot, seg_base, n_tiles, recs, rec_blocks);
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

int tile_rows_2d(int dtype) { return dtype ? Tile2D<double>::TX : Tile2D<float>::TX; }

template <typename T>
int run2d(const pysit_amd_batch &B, int mode, int slot_begin, int slot_end, int direction,
          int seg_base, const int *active_tiles, const int *active_recs, hipStream_t st)
{
    PA_CHECK(B.tile_x == Tile2D<T>::TX, "2D tiles are " + std::to_string(Tile2D<T>::TX) +
                                            " rows along x for this dtype (tile_x)");
    Step2DParams<T> P{};
    const int r = B.order / 2;
    P.cx = axis_coef<T>(r, B.dx);
    P.cz = axis_coef<T>(r, B.dz);
    P.tc = time_coef<T>(B.dt);
    P.rhs = nullptr;
    P.single = 0;
    P.general_phi = 0;
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
    B.tile_x = Tile2D<T>::TX;
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
    P.general_phi = 1;
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

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
// TX rows along x (16 in FP32, 8 in FP64).  The z tiles of a shot start at
// 0, 64, 128, ... except the last, which ends on the last node; the node
// ranges the tiles own are disjoint (a tile lane that another tile owns only
// loads), so the PML band at the far z side costs one PML tile, not two.
// A tile is processed in register blocks of NB rows: all loads of a block
// (the NB + 2r rows of u its x derivatives need, u_{k-1}, the model, the
// imaging accumulator) are issued up front, then the rows are computed
// fully unrolled.  z derivatives take their neighbours from other lanes through
// an LDS row per wavefront; lanes 0..r-1 and 64-r..63 also carry the r-node
// halo beyond each side of the tile.
//
// Addressing: every array is reached through a raw buffer resource with a
// per-lane byte offset fixed for the whole tile (z, or an out-of-range value
// for masked lanes) and a wave-uniform row offset in an SGPR, so the row loop
// spends no vector instructions on addresses or masks.  In a batch the six
// state buffers (u, Phix, Phiz, two time levels each) lie in one region and
// share one resource.
//
// Two row paths: tiles whose nodes and r-neighbourhoods contain no PML node,
// no edge node and nothing outside the grid (the host flags them) run the
// plain leapfrog update with no Phi traffic and no checks; the other tiles
// run the full PML update, reading and writing Phi only where it can be
// nonzero (PML nodes) except for the reference-contract step entry point,
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
    return Tile2D<T>::TX / ((FULL ? 2 : 1) * (MODE >= 2 ? 2 : 1));
}

// state buffers of one step: u_k, u_{k-1} (overwritten by u_{k+1}),
// Phix_k, Phix_{k+1}, Phiz_k, Phiz_{k+1}
enum { F_U = 0, F_UO, F_PX, F_PXO, F_PZ, F_PZO, NF2 };

template <bool PACKED> struct State2D;
// batch: one resource over the state region plus a byte offset per buffer
template <> struct State2D<true> {
    Rsrc r;
    int o[NF2];
    __device__ __forceinline__ Rsrc rs(int) const { return r; }
    __device__ __forceinline__ int so(int f) const { return o[f]; }
};
// reference-contract step: caller-owned arrays, one resource each
template <> struct State2D<false> {
    Rsrc r[NF2];
    __device__ __forceinline__ Rsrc rs(int f) const { return r[f]; }
    __device__ __forceinline__ int so(int) const { return 0; }
};

template <typename T>
struct Step2DParams {
    AxisCoef<T> cx, cz;
    TimeCoef<T> tc;
    const T *rhs;          // dense rhs (step entry point only), may be null
    int single;            // 1: one shot given by value, tiles by arithmetic
    int general_phi;       // 1: Phi may be nonzero anywhere (step entry point)
    const void *state;     // batch: base of the state region
    int state_bytes;       //   its extent
    int fo[3][2];          //   byte offsets of u, Phix, Phiz buffers 0 and 1
    pysit_amd_shot shot0;  // used when single
};

// Per-tile constants shared by the row blocks.
template <typename T, int R, bool PACKED>
struct Ctx2D {
    State2D<PACKED> s;
    Rsrc model, m1, dw, fst, ic, rhs;  // dw: read (adjoint, Born) or written (forward)
    bool has_fst, has_dwo, has_rhs, has_inj, img, general_phi;
    const T *finj;  // LDS row sources of this wave
    const T *xl, *xr;
    int nxl, nxr;
    int nx, nz, lane;
    int rowb;                         // bytes per grid row
    int shb, shb8;                    // shot offset in slot-major stores (T) and in the accumulator (FP64)
    int mx_off, mx_lo, mx_hi, mrowb;  // model row window, bytes per model row
    // per-lane byte offsets, kOOB where masked
    int zv;    // centre node inside the grid
    int hv;    // halo node inside the grid
    int ov;    // owned node: u_{k-1} and every store
    int ov8;   // owned node, FP64 arrays
    int mv;    // model column (edge padding by clamping)
    int zpv;   // centre node where Phi can be nonzero without a PML row
    int hpv;   // halo node likewise
    bool zedge;
    T szk, szm;  // lane z profile (kernel, mesh)
    T *zbuf;     // this wave's z-exchange rows (LDS), 4 x ZROW
    int hslot;   // where this lane's halo value goes in a row
};

// z-exchange row: slots 0..R-1 left halo, R..R+63 the lanes, R+64..2R+63
// right halo, then a parking slot per lane for lanes that hold no halo
template <int R> constexpr int zrow_len() { return 2 * R + 128; }

// z neighbours at distance 1..R of every lane: each lane writes its value
// (and halo lanes their halo value) into an LDS row, then reads the 2R
// neighbours at fixed offsets.  LDS operations of one wavefront execute in
// order; wave_sync keeps the compiler from moving the reads above the writes,
// and rows alternate between buffers so the next row's writes cannot move
// above this row's reads either.
template <typename T, int R, bool PACKED>
__device__ __forceinline__ void z_taps(const Ctx2D<T, R, PACKED> &c, int buf, T v, T h, T (&zp)[R + 1],
                                       T (&zm)[R + 1])
{
    T *row = c.zbuf + buf * zrow_len<R>();
    row[c.lane + R] = v;
    row[c.hslot] = h;
    wave_sync();
#pragma unroll
    for (int i = 1; i <= R; ++i) {
        zp[i] = row[c.lane + R + i];
        zm[i] = row[c.lane + R - i];
    }
}

// Interior rows: no PML node within r, no edge node, every row in the grid.
template <typename T, int R, int NB, int MODE, bool PACKED>
__device__ __forceinline__ void rows_int(const Ctx2D<T, R, PACKED> &c, const Step2DParams<T> &P, int xb, int jb)
{
    const int rb = c.rowb;
    T uc[NB + 2 * R];
#pragma unroll
    for (int j = 0; j < NB + 2 * R; ++j) uc[j] = buf_ld<T>(c.s.rs(F_U), c.zv, c.s.so(F_U) + (xb - R + j) * rb);
    T uh[NB], um[NB], cv[NB], fv[NB], dv[NB], mv[NB];
    double icv[NB];
    const bool rd = MODE == 3 || (MODE == 2 && c.img);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int x = xb + j;
        const int xr = x * rb;
        uh[j] = buf_ld<T>(c.s.rs(F_U), c.hv, c.s.so(F_U) + xr);
        um[j] = buf_ld<T>(c.s.rs(F_UO), c.ov, c.s.so(F_UO) + xr);
        const int mr = clampi(x - c.mx_off, c.mx_lo, c.mx_hi) * c.mrowb;
        cv[j] = buf_ld<T>(c.model, c.mv, mr);
        fv[j] = c.has_inj ? c.finj[(jb + j) * 64 + c.lane] : T(0);
        dv[j] = rd ? buf_ld<T>(c.dw, c.ov, c.shb + xr) : T(0);
        mv[j] = (MODE == 3) ? buf_ld<T>(c.m1, c.mv, mr) : T(0);
        icv[j] = (MODE == 2 && c.img) ? buf_ld<double>(c.ic, c.ov8, c.shb8 + x * c.nz * 8) : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int x = xb + j;
        const int xr = x * rb;
        const T v = uc[R + j];
        T f = fv[j];
        if (MODE == 3) f += mv[j] * (-dv[j]);  // Born source: m1_pad * (-dWaveOp0_k)
        T zp[R + 1], zm[R + 1];
        z_taps<T, R, PACKED>(c, 2 * (j & 1), v, uh[j], zp, zm);
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
        const T src = f + lap;
        const T up = fma(P.tc.dt2, c2 * src, T(2) * v - um[j]);
        if (MODE == 2 && c.img) buf_st(c.ic, c.ov8, c.shb8 + x * c.nz * 8, icv[j] + (double)v * (double)dv[j]);
        buf_st(c.s.rs(F_UO), c.ov, c.s.so(F_UO) + xr, up);
        if (MODE == 1 && c.has_dwo) buf_st(c.dw, c.ov, c.shb + xr, c2 * src);
        if (c.has_fst) buf_st(c.fst, c.ov, c.shb + xr, v);
    }
}

// Rows with the PML update, edges and rows beyond the grid.
template <typename T, int R, int NB, int MODE, bool PACKED>
__device__ __forceinline__ void rows_full(const Ctx2D<T, R, PACKED> &c, const Step2DParams<T> &P, int xb, int jb)
{
    const int rb = c.rowb, nx = c.nx;
    T uc[NB + 2 * R], pc[NB + 2 * R];
#pragma unroll
    for (int j = 0; j < NB + 2 * R; ++j) {
        const int xx = xb - R + j;
        const bool rin = xx >= 0 && xx < nx;
        const bool xp = xx < c.nxl || xx >= nx - c.nxr;
        const int xr = (rin ? xx : 0) * rb;
        uc[j] = buf_ld<T>(c.s.rs(F_U), rin ? c.zv : kOOB, c.s.so(F_U) + xr);
        pc[j] = buf_ld<T>(c.s.rs(F_PX), rin ? (xp ? c.zv : c.zpv) : kOOB, c.s.so(F_PX) + xr);
    }
    T uh[NB], um[NB], cv[NB], fv[NB], dv[NB], mv[NB], pv[NB], ph[NB];
    double icv[NB];
    const bool rd = MODE == 3 || (MODE == 2 && c.img);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int x = xb + j;
        const bool own = x < nx;
        const int xs = own ? x : nx - 1;
        const int xr = xs * rb;
        const bool xp = x < c.nxl || x >= nx - c.nxr;
        const int ovr = own ? c.ov : kOOB;
        uh[j] = buf_ld<T>(c.s.rs(F_U), own ? c.hv : kOOB, c.s.so(F_U) + xr);
        um[j] = buf_ld<T>(c.s.rs(F_UO), ovr, c.s.so(F_UO) + xr);
        const int mr = clampi(xs - c.mx_off, c.mx_lo, c.mx_hi) * c.mrowb;
        cv[j] = buf_ld<T>(c.model, c.mv, mr);
        T f = c.has_rhs ? buf_ld<T>(c.rhs, ovr, xr) : T(0);
        if (c.has_inj) f += c.finj[(jb + j) * 64 + c.lane];
        fv[j] = f;
        dv[j] = rd ? buf_ld<T>(c.dw, ovr, c.shb + xr) : T(0);
        mv[j] = (MODE == 3) ? buf_ld<T>(c.m1, c.mv, mr) : T(0);
        icv[j] = (MODE == 2 && c.img) ? buf_ld<double>(c.ic, own ? c.ov8 : kOOB, c.shb8 + xs * c.nz * 8) : 0.0;
        pv[j] = buf_ld<T>(c.s.rs(F_PZ), own ? (xp ? c.zv : c.zpv) : kOOB, c.s.so(F_PZ) + xr);
        ph[j] = buf_ld<T>(c.s.rs(F_PZ), own ? (xp ? c.hv : c.hpv) : kOOB, c.s.so(F_PZ) + xr);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int x = xb + j;
        const bool own = x < nx;
        const int xr = (own ? x : nx - 1) * rb;
        const int ovr = own ? c.ov : kOOB;
        const T v = uc[R + j];
        T f = fv[j];
        if (MODE == 3) f += mv[j] * (-dv[j]);
        T zp[R + 1], zm[R + 1], pzp[R + 1], pzm[R + 1];
        z_taps<T, R, PACKED>(c, 2 * (j & 1), v, uh[j], zp, zm);
        z_taps<T, R, PACKED>(c, 2 * (j & 1) + 1, pv[j], ph[j], pzp, pzm);
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
        const bool edge = (x == 0) | (x == nx - 1) | c.zedge;

        T dux = T(0), duz = T(0), dpx = T(0), dpz = T(0);
#pragma unroll
        for (int i = R; i >= 1; --i) {
            dux = fma(P.cx.d1[i - 1], uc[R + j + i] - uc[R + j - i], dux);
            duz = fma(P.cz.d1[i - 1], zp[i] - zm[i], duz);
            dpx = fma(P.cx.d1[i - 1], pc[R + j + i] - pc[R + j - i], dpx);
            dpz = fma(P.cz.d1[i - 1], pzp[i] - pzm[i], dpz);
        }
        const T src = ((f + lap) + dpx) + dpz;
        const int xs = own ? x : nx - 1;
        const T sx = sigma_kernel<T>(c.xl, c.nxl, c.xr, c.nxr, max(0, min(xs, nx - 2)), nx);
        const T sxmj = sigma_mesh<T>(c.xl, c.nxl, c.xr, c.nxr, xs, nx);
        const T szk = c.szk, szm = c.szm;
        const T dt = P.tc.dt;
        const T pxc = pc[R + j], pzc = pv[j];
        const bool pml = sx != T(0) || szk != T(0);
        // damped update (PML node) and plain update, selected per lane
        const T pxd = (pxc - dt * sx * pxc) + dt * (szk - sx) * dux;
        const T pzd = (pzc - dt * szk * pzc) + dt * (sx - szk) * duz;
        const T ss = sx + szk;
        const T fac1 = P.tc.two_dt2 / (T(2) + dt * ss);
        const T fac2 = c2 * src - (u_m - T(2) * v) * P.tc.inv_dt2 + ss * u_m * P.tc.inv_2dt - (sx * szk) * v;
        const T upd = fac1 * fac2;
        const T upi = fma(P.tc.dt2, c2 * src, T(2) * v - u_m);
        T pxn = pml ? pxd : pxc;
        T pzn = pml ? pzd : pzc;
        T up = pml ? upd : upi;
        bool phi_store = pml || c.general_phi;
        // dWaveOp with the mesh profile: the step identity gives
        // (u+ - 2u + u-)/dt^2 = c^2 src - sum(sk)(u+ - u-)/(2dt) - prod(sk) u
        T dw = c2 * src + ((sxmj + szm) - (sx + szk)) * (up - u_m) * P.tc.inv_2dt + (sxmj * szm - sx * szk) * v;
        if (edge) {
            up = T(0);
            pxn = T(0);
            pzn = T(0);
            dw = T(0);
            phi_store = true;
        }
        if (MODE == 2 && c.img)
            buf_st(c.ic, own ? c.ov8 : kOOB, c.shb8 + (own ? x : nx - 1) * c.nz * 8, icv[j] + (double)v * (double)dv[j]);
        buf_st(c.s.rs(F_UO), ovr, c.s.so(F_UO) + xr, up);
        const int op = phi_store ? ovr : kOOB;
        buf_st(c.s.rs(F_PXO), op, c.s.so(F_PXO) + xr, pxn);
        buf_st(c.s.rs(F_PZO), op, c.s.so(F_PZO) + xr, pzn);
        if (MODE == 1 && c.has_dwo) buf_st(c.dw, ovr, c.shb + xr, dw);
        if (c.has_fst) buf_st(c.fst, ovr, c.shb + xr, v);
    }
}

template <typename T, int R, int MODE, bool SINGLE>
__device__ __forceinline__ void tile2d(const pysit_amd_batch &B, const Step2DParams<T> &P,
                                       const pysit_amd_shot &S, int tile, int full, int tloc,
                                       int slot, int seg_base, int lane, T *lds_f, T *zbuf)
{
    constexpr bool PACKED = !SINGLE;
    constexpr int TX = Tile2D<T>::TX;
    constexpr int SZ = sizeof(T);
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    Ctx2D<T, R, PACKED> c;
    const int nx = S.nx, nz = S.nz;
    c.nx = nx;
    c.nz = nz;
    // z tiles: 0, 64, ... and a last one ending on node nz-1; owned ranges disjoint
    const int ntz = S.ntz;
    const int tz = tloc % ntz, tx = tloc / ntz;
    const int zlast = nz > 64 ? nz - 64 : 0;
    const int z0 = (tz == ntz - 1) ? zlast : tz * 64;
    const int own_hi = (tz == ntz - 1) ? nz : min(z0 + 64, zlast);
    const int x0 = tx * TX;
    c.lane = lane;
    const int z = z0 + lane;
    const bool zin = z < nz;
    const bool own = z < own_hi;
    c.zv = zin ? z * SZ : kOOB;
    c.ov = own ? z * SZ : kOOB;
    c.ov8 = own ? z * 8 : kOOB;
    // halo lanes: 0..R-1 hold z0+64+lane, 64-R..63 hold z0+lane-64
    const int hz = lane < R ? z0 + 64 + lane : (lane >= 64 - R ? z0 + lane - 64 : -1);
    const bool hin = hz >= 0 && hz < nz;
    c.hv = hin ? hz * SZ : kOOB;
    const bool zp = P.general_phi || z < S.nzl || z >= nz - S.nzr;
    const bool hzp = P.general_phi || hz < S.nzl || hz >= nz - S.nzr;
    c.zpv = zp ? c.zv : kOOB;
    c.hpv = hzp ? c.hv : kOOB;
    c.zedge = z == 0 || z >= nz - 1;
    // which buffer holds time level k: parity of the steps this propagation took
    const int taken = (MODE == 2) ? (S.nt - 1 - k) : k;
    const int cur = taken & 1;
    c.rowb = nz * SZ;
    c.shb = (int)(S.off * SZ);
    c.shb8 = (int)(S.off * 8);
    if constexpr (PACKED) {
        c.s.r = make_rsrc(P.state, P.state_bytes);
        c.s.o[F_U] = P.fo[0][cur] + c.shb;
        c.s.o[F_UO] = P.fo[0][cur ^ 1] + c.shb;
        c.s.o[F_PX] = P.fo[1][cur] + c.shb;
        c.s.o[F_PXO] = P.fo[1][cur ^ 1] + c.shb;
        c.s.o[F_PZ] = P.fo[2][cur] + c.shb;
        c.s.o[F_PZO] = P.fo[2][cur ^ 1] + c.shb;
    } else {
        const long long nb = (long long)nx * nz * SZ;
        c.s.r[F_U] = make_rsrc(B.u[cur], nb);
        c.s.r[F_UO] = make_rsrc(B.u[cur ^ 1], nb);
        c.s.r[F_PX] = make_rsrc(B.phix[cur], nb);
        c.s.r[F_PXO] = make_rsrc(B.phix[cur ^ 1], nb);
        c.s.r[F_PZ] = make_rsrc(B.phiz[cur], nb);
        c.s.r[F_PZO] = make_rsrc(B.phiz[cur ^ 1], nb);
    }
    c.model = make_rsrc(S.model, (long long)(S.mx_hi + 1) * S.mnz * SZ);
    c.mv = clampi(z - S.mz_off, S.mz_lo, S.mz_hi) * SZ;
    c.mx_off = S.mx_off;
    c.mx_lo = S.mx_lo;
    c.mx_hi = S.mx_hi;
    c.mrowb = S.mnz * SZ;
    c.xl = (const T *)S.xl;
    c.xr = (const T *)S.xr;
    c.nxl = S.nxl;
    c.nxr = S.nxr;
    c.general_phi = P.general_phi;
    c.has_rhs = P.rhs != nullptr;
    c.rhs = make_rsrc(P.rhs, c.has_rhs ? (long long)nx * nz * SZ : 0);
    c.zbuf = zbuf;
    c.hslot = lane < R ? R + 64 + lane : (lane >= 64 - R ? lane - (64 - R) : 2 * R + 64 + lane);
    const T *zl = (const T *)S.zl, *zr = (const T *)S.zr;
    c.szk = c.szm = T(0);
    if (full && zin) {
        c.szk = (z <= nz - 2) ? sigma_kernel<T>(zl, S.nzl, zr, S.nzr, z, nz) : T(0);
        c.szm = sigma_mesh<T>(zl, S.nzl, zr, S.nzr, z, nz);
    }
    c.img = (MODE == 2) && (k % B.imaging_period == 0) && B.dw;
    const bool store = (MODE == 1) && (k % B.imaging_period == 0) && B.dw;
    const long long slot_el = (long long)(slot - seg_base) * B.total_n;
    const long long slot_bytes = B.total_n * SZ;
    const bool dwon = ((c.img || MODE == 3) && B.dw) || store;
    c.dw = make_rsrc(dwon ? (T *)B.dw + slot_el : nullptr, dwon ? slot_bytes : 0);
    c.has_dwo = store;
    c.has_fst = B.fstore != nullptr;
    c.fst = make_rsrc(c.has_fst ? (T *)B.fstore + slot_el : nullptr, c.has_fst ? slot_bytes : 0);
    c.ic = make_rsrc(MODE == 2 ? B.ic : nullptr, MODE == 2 ? B.total_n * 8 : 0);
    c.m1 = make_rsrc(MODE == 3 ? B.born_m1 : nullptr, MODE == 3 ? (long long)(S.mx_hi + 1) * S.mnz * SZ : 0);

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
                for (int e = B.inj_nptr[n]; e < B.inj_nptr[n + 1]; ++e) acc += B.inj_w[e] * amp_k[B.inj_chan[e]];
                lds_f[B.inj_node[n]] = (T)acc;
            }
            wave_sync();
            c.has_inj = true;
        }
    }

    if (full) {
        constexpr int NB = block_rows<T, MODE, true>();
#pragma unroll 1
        for (int jb = 0; jb < TX; jb += NB) rows_full<T, R, NB, MODE, PACKED>(c, P, x0 + jb, jb);
    } else {
        constexpr int NB = block_rows<T, MODE, false>();
#pragma unroll 1
        for (int jb = 0; jb < TX; jb += NB) rows_int<T, R, NB, MODE, PACKED>(c, P, x0 + jb, jb);
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
    __shared__ T lds_z[4][4 * zrow_len<R>()];
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
        tile2d<T, R, MODE, SINGLE>(B, P, P.shot0, gw, 1, gw, slot, seg_base, lane, lds_f[w], lds_z[w]);
    } else {
        const int4 td = reinterpret_cast<const int4 *>(B.tiles)[gw];
        tile2d<T, R, MODE, SINGLE>(B, P, B.shots[td.x], gw, td.z & 1, td.y, slot, seg_base, lane, lds_f[w], lds_z[w]);
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
    if (P.single) {
        if constexpr (MODE == 0) {
            hipLaunchKernelGGL((k_step2d<T, R, MODE, true>), dim3(rec_blocks + tile_blocks), dim3(256), 0, st,
                               B, P, slot, seg_base, n_tiles, recs, rec_blocks);
        } else {
            set_error("the single-shot step only propagates");
            return PYSIT_AMD_EINVAL;
        }
    } else {
        hipLaunchKernelGGL((k_step2d<T, R, MODE, false>), dim3(rec_blocks + tile_blocks), dim3(256), 0, st,
                           B, P, slot, seg_base, n_tiles, recs, rec_blocks);
    }
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
    // the six state buffers share one resource: they must lie in one region
    {
        const char *f[6] = {(const char *)B.u[0], (const char *)B.u[1], (const char *)B.phix[0],
                            (const char *)B.phix[1], (const char *)B.phiz[0], (const char *)B.phiz[1]};
        const char *lo = f[0], *hi = f[0];
        for (int i = 0; i < 6; ++i) {
            PA_CHECK(f[i] != nullptr, "2D batch state buffers (u, phix, phiz) must all be given");
            lo = f[i] < lo ? f[i] : lo;
            hi = f[i] > hi ? f[i] : hi;
        }
        const long long ext = (long long)(hi - lo) + B.total_n * (long long)sizeof(T);
        PA_CHECK(ext <= kMaxRsrcBytes, "2D batch state buffers must lie in one region of at most " +
                                           std::to_string(kMaxRsrcBytes) + " bytes");
        for (int i = 0; i < 6; ++i) {
            PA_CHECK((f[i] - lo) % (long long)sizeof(T) == 0, "misaligned 2D batch state buffer");
            P.fo[i / 2][i % 2] = (int)(f[i] - lo);
        }
        P.state = lo;
        P.state_bytes = (int)ext;
    }
    PA_CHECK(B.total_n * 8 <= kMaxRsrcBytes, "batch too large for one launch (total_n)");
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
    PA_CHECK(N * (long long)sizeof(T) <= kMaxRsrcBytes, "grid too large for one step launch");
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

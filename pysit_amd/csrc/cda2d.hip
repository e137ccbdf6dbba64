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
// TX = 8, 4 or 2 rows along x.  The z tiles of a shot start at
// 0, 64, 128, ... except the last, which ends on the last node; the node
// ranges the tiles own are disjoint (a tile lane that another tile owns only
// loads), so the PML band at the far z side costs one PML tile, not two.
// The rows of a tile are computed in order as a software-pipelined march:
// each u row is loaded once, the loads of the next rows are in flight while
// the current row is computed.  z derivatives take their neighbours from
// other lanes through an LDS row per wavefront; lanes 0..r-1 and 64-r..63
// also carry the r-node halo beyond each side of the tile.
//
// Addressing: every array is reached through a raw buffer resource with a
// per-lane byte offset (z, or an out-of-range value for masked lanes, plus
// the row) and a wave-uniform base offset in an SGPR, so the row loop spends
// one vector add per row on addresses and none on masks.  In a batch the six
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

#include <algorithm>
#include <type_traits>
#include <vector>

// No implicit FP contraction: every fused multiply-add below is written out,
// so all kernel modes (propagate, store dWaveOp, adjoint, Born) evaluate the
// step with the same roundings and a checkpoint recompute reproduces the
// stored history bit for bit.
#pragma clang fp contract(off)

namespace pa {

// Tile rows along x: every 2D step kernel is built for kTileRows2D[0] (the
// default), [1] and [2] rows; a batch picks one (tile_x).  Short tiles give
// more wavefronts and shorter marches (small grids: PySIT's 111 x 91
// reflector case runs 8.1, 5.9 and 5.0 us per step at 8, 4 and 2 rows), tall
// ones re-read fewer halo rows.  The persistent loop is built for [0] and [1].
constexpr int kTileRows2D[3] = {8, 4, 2};

// state buffers of one step: u_k, the other u buffer (batch: u_{k+1} is
// written there; reference-contract step: it holds u_{k-1} and u_{k+1}
// overwrites it), Phix_k, Phix_{k+1}, Phiz_k, Phiz_{k+1}, and (batch only)
// the increment du_k = u_k - u_{k-1}, updated in place
enum { F_U = 0, F_UO, F_PX, F_PXO, F_PZ, F_PZO, F_DU, NF2 };

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
    int fdu;               //   byte offset of the du buffer
    pysit_amd_shot shot0;  // used when single
};

// Per-tile constants shared by the row blocks.
template <typename T, int R, bool PACKED>
struct Ctx2D {
    State2D<PACKED> s;
    Rsrc model, m1, dw, fst, ic, rhs;  // dw: read (adjoint, Born) or written (forward)
    Rsrc dwo, volm, vold;              // second-order modes: dWaveOp out, volume source factors
    bool has_fst, has_dwo, has_rhs, has_inj, img, general_phi;
    bool dw_exact;  // store the kernel-profile dWaveOp (C^2 src) in PML tiles too
    const T *finj;  // this lane's LDS source column (or a zero slot)
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
    int ovd;   // owned node where the dWaveOp store is read or written this launch
    int ovi;   // owned node where the imaging accumulator is updated this launch (FP64)
    int ovf;   // owned node where the field store is written
    int ovo;   // owned node where dw_out is written (modes 5, 6)
    int ovv;   // owned node where the volume source is read (mode 6)
    int fstride;  // LDS row stride of the tile's sources: 64, or 0 (all rows read zslot)
    int mv;    // model column (edge padding by clamping)
    int zpv;   // centre node where Phi can be nonzero without a PML row
    int hpv;   // halo node likewise
    bool zedge;
    // PML profiles, each as the two halves of a masked lookup (summed only
    // after the first rows' loads are issued, so waiting for them costs nothing)
    T szk2[2], szm2[2];  // lane z profile (kernel, mesh)
    T vsig2[2];          // x profile of the tile rows: lane r < TX kernel, lane TX + r mesh
    T *zbuf;     // this wave's z-exchange rows (LDS), 4 x ZROW
    int hslot;   // where this lane's halo value goes in a row
};

// value of lane `l` (wave-uniform l) in an SGPR
__device__ __forceinline__ float lane_bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// PML profile value at node i (i < 0: none) of an axis of n nodes from the
// left array l (nl values, by node) and the right array r (nr values): the
// kernel form takes r one node shifted towards the edge and never covers the
// edge node itself (constant_density_acoustic_time_scalar_2D.h:137-157), the
// mesh form lays r on the last nr nodes (StructuredPML, mesh.py:770-828).
// Two masked buffer loads, at most one in range: the value is their sum.
template <typename T>
__device__ __forceinline__ void sigma_lane(Rsrc l, int nl, Rsrc r, int nr, int i, int n, bool kernel, T (&out)[2])
{
    constexpr int SZ = sizeof(T);
    const bool inl = i >= 0 && i < nl;
    const bool inr = !inl && i >= 0 && i >= n - nr && (!kernel || i <= n - 2);
    const int ri = kernel ? nr - ((n - 1) - i) : i - (n - nr);
    out[0] = buf_ld<T>(l, inl ? i * SZ : kOOB, 0);
    out[1] = buf_ld<T>(r, inr ? ri * SZ : kOOB, 0);
}

// z-exchange row: slots 0..R-1 left halo, R..R+63 the lanes, R+64..2R+63
// right halo, then a parking slot per lane for lanes that hold no halo
template <int R> constexpr int zrow_len() { return 2 * R + 128; }

// z neighbours at distance 1..R of every lane: each lane writes its value
// (and halo lanes their halo value) into an LDS row, then reads the 2R
// neighbours at fixed offsets.  LDS operations of one wavefront execute in
// order; wave_sync keeps the compiler from moving the reads above the writes,
// and rows alternate between buffers so the next row's writes cannot move
// above this row's reads either.
#ifndef PA2D_ZDPP
#define PA2D_ZDPP 0
#endif
// wavefront rotations by one lane (DPP): rol, lane l takes lane l+1 (mod 64);
// ror, lane l takes lane l-1
template <int CTRL>
__device__ __forceinline__ float wrot(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double wrot(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
constexpr int kDppRol1 = 0x134, kDppRor1 = 0x13C;

template <typename T, int R, bool PACKED>
__device__ __forceinline__ void z_taps(const Ctx2D<T, R, PACKED> &c, int buf, T v, T h, T (&zp)[R + 1],
                                       T (&zm)[R + 1])
{
#if PA2D_ZDPP
    // lane l's neighbour at z+i is lane l+i, or past the last lane the right
    // halo value that lane l+i-64 holds in h; at z-i likewise lane l-i, or the
    // left halo that lane l-i+64 holds
    T a = v, b = v, ha = h, hb = h;
#pragma unroll
    for (int i = 1; i <= R; ++i) {
        a = wrot<kDppRol1>(a);
        ha = wrot<kDppRol1>(ha);
        b = wrot<kDppRor1>(b);
        hb = wrot<kDppRor1>(hb);
        zp[i] = c.lane + i < 64 ? a : ha;
        zm[i] = c.lane - i >= 0 ? b : hb;
    }
    (void)buf;
    return;
#endif
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

// Kernel modes: 0 propagate, 1 propagate + store dWaveOp, 2 adjoint with the
// FP64 imaging accumulator, 3 Born, 4 adjoint with the FP32 window
// accumulator (flushed into the FP64 one by the host loop); the full
// Hessian's second-order runs: 5 Born that also stores its own dWaveOp
// (dw_out), 6 adjoint with the FP64 accumulator that also stores its own
// dWaveOp (dw_out) and adds the volume source vol_m * vol_dw_k.
constexpr bool is_adjoint(int mode) { return mode == 2 || mode == 4 || mode == 6; }
constexpr bool is_born(int mode) { return mode == 3 || mode == 5; }
constexpr bool stores_out(int mode) { return mode == 5 || mode == 6; }

// the FP32 window accumulator's read-modify-write may go non-temporal
// (measured no different)
#ifndef PA2D_IC_NT
#define PA2D_IC_NT 0
#endif
template <typename T, int MODE> struct Acc2D { typedef double type; };
template <typename T> struct Acc2D<T, 4> { typedef float type; };

// imaging accumulator of the owned node (per-lane offset `o`, kOOB if masked)
// in row byte offset xr of the T-typed fields
template <typename T, int MODE, typename C>
__device__ __forceinline__ typename Acc2D<T, MODE>::type ic_load(const C &c, int o, int xr)
{
    if constexpr (MODE == 4) return PA2D_IC_NT ? buf_ld_nt<float>(c.ic, o + xr, c.shb) : buf_ld<float>(c.ic, o + xr, c.shb);
    else return buf_ld<double>(c.ic, o + xr * (8 / (int)sizeof(T)), c.shb8);
}
template <typename T, int MODE, typename C>
__device__ __forceinline__ void ic_store(const C &c, int o, int xr, typename Acc2D<T, MODE>::type acc, T q, T dw)
{
    if constexpr (MODE == 4) {
        if constexpr (PA2D_IC_NT) buf_st_nt(c.ic, o + xr, c.shb, acc + q * dw);
        else buf_st(c.ic, o + xr, c.shb, acc + q * dw);
    }
    else buf_st(c.ic, o + xr * (8 / (int)sizeof(T)), c.shb8, acc + (double)q * (double)dw);
}

// Rows of a tile are computed in order (a march along x).  Each u row is
// loaded once per tile; the loads of the rows `PA2D_PREFETCH` ahead are
// issued before the current row is computed, so every wave keeps a steady
// stream of loads in flight while it computes.  Row byte offsets go into the
// per-lane offset (one vector add per row), the buffer's base offset stays in
// an SGPR per array.
#ifndef PA2D_PREFETCH
#define PA2D_PREFETCH 2
#endif
#ifndef PA2D_PREFETCH_ADJ
#define PA2D_PREFETCH_ADJ 3
#endif
#ifndef PA2D_PREFETCH_FULL
#define PA2D_PREFETCH_FULL 2
#endif
// the dWaveOp history streams through (written once, read back thousands of
// steps later): non-temporal accesses keep it from evicting the fields
#ifndef PA2D_DW_NT
#define PA2D_DW_NT 1
#endif
// The increment du is read and written once per step by the node that owns
// it.  Adjoint steps access it non-temporally (PA2D_DU_NT_ADJ): the caches
// then keep the fields the imaging condition and the stencil reuse (adjoint
// -10 to -14 % on the 16-shot headline); forward steps keep it cached
// (PA2D_DU_NT_FWD; non-temporal there made them 16 % slower).
#ifndef PA2D_DU_NT_FWD
#define PA2D_DU_NT_FWD 0
#endif
#ifndef PA2D_DU_NT_ADJ
#define PA2D_DU_NT_ADJ 1
#endif
#ifndef PA2D_UO_NT_ADJ
#define PA2D_UO_NT_ADJ 0
#endif
template <int MODE> __device__ __forceinline__ void uo_st(Rsrc r, int voff, int soff, float v)
{
    if constexpr (is_adjoint(MODE) && PA2D_UO_NT_ADJ) buf_st_nt(r, voff, soff, v);
    else buf_st(r, voff, soff, v);
}
template <int MODE> __device__ __forceinline__ void uo_st(Rsrc r, int voff, int soff, double v)
{
    if constexpr (is_adjoint(MODE) && PA2D_UO_NT_ADJ) buf_st_nt(r, voff, soff, v);
    else buf_st(r, voff, soff, v);
}
template <int MODE> constexpr bool du_nt() { return is_adjoint(MODE) ? PA2D_DU_NT_ADJ : PA2D_DU_NT_FWD; }
template <typename T, int MODE>
__device__ __forceinline__ T du_ld(Rsrc r, int voff, int soff)
{
    if constexpr (du_nt<MODE>()) return buf_ld_nt<T>(r, voff, soff);
    else return buf_ld<T>(r, voff, soff);
}
template <typename T, int MODE>
__device__ __forceinline__ void du_st(Rsrc r, int voff, int soff, T v)
{
    if constexpr (du_nt<MODE>()) buf_st_nt(r, voff, soff, v);
    else buf_st(r, voff, soff, v);
}
template <typename T>
__device__ __forceinline__ T dw_ld(Rsrc r, int voff, int soff)
{
    if constexpr (PA2D_DW_NT) return buf_ld_nt<T>(r, voff, soff);
    else return buf_ld<T>(r, voff, soff);
}
template <typename T>
__device__ __forceinline__ void dw_st(Rsrc r, int voff, int soff, T v)
{
    if constexpr (PA2D_DW_NT) buf_st_nt(r, voff, soff, v);
    else buf_st(r, voff, soff, v);
}

// Laplacian of a node from its x row queue (ux[R] is the node, ux[R +- i]
// its x neighbours) and its z neighbours, with difference-form taps:
// sum_i w_i ((u_{+i} - u) + (u_{-i} - u)) per axis, the centre weight
// -2 sum_i w_i being implied.  Neighbour differences of the smooth fields a
// step sees round far less than the terms of the plain form
// sum_i w_i (u_{+i} + u_{-i}) + w_0 u, which cancel to a value about (kh)^2
// times smaller than u: over thousands of FP32 steps the difference form
// (with the increment form of the update) keeps the traces near the FP64
// ones.
template <typename T, int R>
__device__ __forceinline__ T laplacian(const Step2DParams<T> &P, const T *ux, T v, const T (&zp)[R + 1],
                                       const T (&zm)[R + 1])
{
    T lx = T(0), lz = T(0);
#pragma unroll
    for (int i = R; i >= 1; --i) {
        lx = fma(P.cx.d2[i], (ux[R + i] - v) + (ux[R - i] - v), lx);
        lz = fma(P.cz.d2[i], (zp[i] - v) + (zm[i] - v), lz);
    }
    return lx + lz;
}

// Plain tiles: no PML node within r of an owned node, no edge node, every
// row of the march inside the grid.
template <typename T, int TX, int R, int MODE, bool PACKED>
__device__ __forceinline__ void march_int(const Ctx2D<T, R, PACKED> &c, const Step2DParams<T> &P, int x0)
{
    constexpr int D = is_adjoint(MODE) ? PA2D_PREFETCH_ADJ : PA2D_PREFETCH;
    constexpr int NU = TX + 2 * R;  // u rows x0-R .. x0+TX+R-1
    const int rb = c.rowb;
    const int so_u = c.s.so(F_U), so_uo = c.s.so(F_UO), so_du = c.s.so(F_DU);
    // um: du_k (batch, increment form) or u_{k-1} (reference-contract step)
    T u[NU], uh[TX], um[TX], cv[TX], fv[TX], dv[TX], mv[TX], wv[TX];
    typename Acc2D<T, MODE>::type icv[TX];
    const int zv = c.zv, hv = c.hv, ov = c.ov, ovd = c.ovd, ovi = c.ovi, ovf = c.ovf, mvo = c.mv;
    const int ovo = c.ovo, ovv = c.ovv;
#define PA_LD_U(s_) u[s_] = buf_ld<T>(c.s.rs(F_U), zv + (x0 - R + (s_)) * rb, so_u)
#define PA_LD_ROW(r_)                                                                               \
    do {                                                                                            \
        const int r = (r_), x = x0 + r, xr = x * rb;                                                \
        uh[r] = buf_ld<T>(c.s.rs(F_U), hv + xr, so_u);                                              \
        um[r] = PACKED ? du_ld<T, MODE>(c.s.rs(F_DU), ov + xr, so_du)                                     \
                       : buf_ld<T>(c.s.rs(F_UO), ov + xr, so_uo);                                   \
        const int mr = clampi(x - c.mx_off, c.mx_lo, c.mx_hi) * c.mrowb;                           \
        cv[r] = buf_ld<T>(c.model, mvo, mr);                                                        \
        fv[r] = c.finj[r * c.fstride];                                                              \
        dv[r] = (MODE >= 2) ? dw_ld<T>(c.dw, ovd + xr, c.shb) : T(0);                              \
        mv[r] = is_born(MODE) ? buf_ld<T>(c.m1, mvo, mr)                                            \
              : (MODE == 6 ? buf_ld<T>(c.volm, ovv + xr, c.shb) : T(0));                           \
        wv[r] = (MODE == 6) ? buf_ld<T>(c.vold, ovv + xr, c.shb) : T(0);                            \
        if (is_adjoint(MODE)) icv[r] = ic_load<T, MODE>(c, ovi, xr);                                \
    } while (0)
#pragma unroll
    for (int q = 0; q < 2 * R + D && q < NU; ++q) PA_LD_U(q);
#pragma unroll
    for (int q = 0; q < D && q < TX; ++q) PA_LD_ROW(q);
    __builtin_amdgcn_sched_barrier(0);  // the prologue's loads go out before any row work
#pragma unroll
    for (int j = 0; j < TX; ++j) {
        if (j + 2 * R + D < NU) PA_LD_U(j + 2 * R + D);
        if (j + D < TX) PA_LD_ROW(j + D);
        __builtin_amdgcn_sched_barrier(0);  // prefetch issued before this row is computed
        const int x = x0 + j;
        const int xr = x * rb;
        const T v = u[R + j];
        T f = fv[j];
        if (is_born(MODE)) f += mv[j] * (-dv[j]);  // Born source: m1_pad * (-dWaveOp0_k)
        if (MODE == 6) f += mv[j] * wv[j];         // volume source: m1 * dWaveOpAdj_k
        T zp[R + 1], zm[R + 1];
        z_taps<T, R, PACKED>(c, 2 * (j & 1), v, uh[j], zp, zm);
        const T lap = laplacian<T, R>(P, u + j, v, zp, zm);
        const T c2 = cv[j] * cv[j];
        const T src = f + lap;
        const T S = c2 * src;
        if (is_adjoint(MODE)) ic_store<T, MODE>(c, ovi, xr, icv[j], v, dv[j]);
        if constexpr (PACKED) {
            // increment form: du_{k+1} = du_k + dt^2 S, u_{k+1} = u_k + du_{k+1}
            const T dn = fma(P.tc.dt2, S, um[j]);
            uo_st<MODE>(c.s.rs(F_UO), ov + xr, so_uo, v + dn);
            du_st<T, MODE>(c.s.rs(F_DU), ov + xr, so_du, dn);
        } else {
            buf_st(c.s.rs(F_UO), ov + xr, so_uo, fma(P.tc.dt2, S, T(2) * v - um[j]));
        }
        if (MODE == 1) dw_st<T>(c.dw, ovd + xr, c.shb, S);
        if (stores_out(MODE)) dw_st<T>(c.dwo, ovo + xr, c.shb, S);
        if (c.has_fst) buf_st(c.fst, ovf + xr, c.shb, v);
    }
#undef PA_LD_U
#undef PA_LD_ROW
}

// PML tiles: the full update, edges, rows beyond the grid.  XC: no row of
// the tile or its R-row reach is an x PML row or an x edge (the host flags
// such tiles), so every row takes the form for rows outside the x PML and
// needs no row checks.
template <typename T, int TX, int R, int MODE, bool PACKED, bool XC>
__device__ __forceinline__ void march_full(const Ctx2D<T, R, PACKED> &c, const Step2DParams<T> &P, int x0)
{
    constexpr int D = PA2D_PREFETCH_FULL;
    constexpr int NU = TX + 2 * R;
    const int rb = c.rowb, nx = c.nx;
    const int so_u = c.s.so(F_U), so_uo = c.s.so(F_UO), so_px = c.s.so(F_PX), so_pz = c.s.so(F_PZ);
    const int so_du = c.s.so(F_DU);
    // values, not fields: a conditional on two fields would select their
    // addresses and keep the context in memory
    const int zv = c.zv, zpv = c.zpv, hv = c.hv, hpv = c.hpv, ov = c.ov, ovd = c.ovd, ovi = c.ovi, ovf = c.ovf;
    const int ovo = c.ovo, ovv = c.ovv;
    T u[NU], px[NU], uh[TX], um[TX], cv[TX], fv[TX], dv[TX], mv[TX], wv[TX], pv[TX], ph[TX];
    typename Acc2D<T, MODE>::type icv[TX];
    const int mvo = c.mv;
#define PA_LD_U(s_)                                                                                 \
    do {                                                                                            \
        const int xx = x0 - R + (s_);                                                               \
        const bool rin = XC || (xx >= 0 && xx < nx);                                                \
        const bool xp = !XC && (xx < c.nxl || xx >= nx - c.nxr);                                    \
        const int xr = (rin ? xx : 0) * rb;                                                         \
        u[s_] = buf_ld<T>(c.s.rs(F_U), rin ? zv + xr : kOOB, so_u);                                 \
        px[s_] = buf_ld<T>(c.s.rs(F_PX), rin ? (xp ? zv : zpv) + xr : kOOB, so_px);                 \
    } while (0)
#define PA_LD_ROW(r_)                                                                               \
    do {                                                                                            \
        const int r = (r_), x = x0 + r;                                                             \
        const bool own = XC || x < nx;                                                              \
        const int xs = own ? x : nx - 1;                                                            \
        const int xr = xs * rb;                                                                     \
        const bool xp = !XC && (x < c.nxl || x >= nx - c.nxr);                                      \
        const int ovr = own ? ov + xr : kOOB;                                                       \
        uh[r] = buf_ld<T>(c.s.rs(F_U), own ? hv + xr : kOOB, so_u);                                 \
        um[r] = PACKED ? du_ld<T, MODE>(c.s.rs(F_DU), ovr, so_du) : buf_ld<T>(c.s.rs(F_UO), ovr, so_uo);  \
        const int mr = clampi(xs - c.mx_off, c.mx_lo, c.mx_hi) * c.mrowb;                          \
        cv[r] = buf_ld<T>(c.model, mvo, mr);                                                        \
        T f = PACKED ? T(0) : buf_ld<T>(c.rhs, ovr, 0); /* dense rhs: reference-contract step */   \
        f += c.finj[r * c.fstride];                                                                 \
        fv[r] = f;                                                                                  \
        dv[r] = (MODE >= 2) ? dw_ld<T>(c.dw, own ? ovd + xr : kOOB, c.shb) : T(0);                 \
        mv[r] = is_born(MODE) ? buf_ld<T>(c.m1, mvo, mr)                                            \
              : (MODE == 6 ? buf_ld<T>(c.volm, own ? ovv + xr : kOOB, c.shb) : T(0));              \
        wv[r] = (MODE == 6) ? buf_ld<T>(c.vold, own ? ovv + xr : kOOB, c.shb) : T(0);               \
        if (is_adjoint(MODE)) icv[r] = ic_load<T, MODE>(c, own ? ovi : kOOB, xr);                  \
        pv[r] = buf_ld<T>(c.s.rs(F_PZ), own ? (xp ? zv : zpv) + xr : kOOB, so_pz);                  \
        ph[r] = buf_ld<T>(c.s.rs(F_PZ), own ? (xp ? hv : hpv) + xr : kOOB, so_pz);                  \
    } while (0)
#pragma unroll
    for (int q = 0; q < 2 * R + D && q < NU; ++q) PA_LD_U(q);
#pragma unroll
    for (int q = 0; q < D && q < TX; ++q) PA_LD_ROW(q);
    __builtin_amdgcn_sched_barrier(0);  // the prologue's loads go out before any row work
    const T szk = c.szk2[0] + c.szk2[1], szm = c.szm2[0] + c.szm2[1];
    const T vsig = XC ? T(0) : c.vsig2[0] + c.vsig2[1];
    // per-lane factors of the rows outside the x PML
    const T dtszk = P.tc.dt * szk;
    const T fac1z = P.tc.two_dt2 / (T(2) + dtszk);
    const T dsz = szm - szk;
    const T szk_dt = szk * P.tc.inv_dt;
#pragma unroll
    for (int j = 0; j < TX; ++j) {
        if (j + 2 * R + D < NU) PA_LD_U(j + 2 * R + D);
        if (j + D < TX) PA_LD_ROW(j + D);
        __builtin_amdgcn_sched_barrier(0);  // prefetch issued before this row is computed
        const int x = x0 + j;
        const bool own = XC || x < nx;
        const int xs = own ? x : nx - 1;
        const int xr = xs * rb;
        const int ovr = own ? ov + xr : kOOB;
        const T v = u[R + j];
        T f = fv[j];
        if (is_born(MODE)) f += mv[j] * (-dv[j]);
        if (MODE == 6) f += mv[j] * wv[j];
        T zp[R + 1], zm[R + 1], pzp[R + 1], pzm[R + 1];
        z_taps<T, R, PACKED>(c, 2 * (j & 1), v, uh[j], zp, zm);
        z_taps<T, R, PACKED>(c, 2 * (j & 1) + 1, pv[j], ph[j], pzp, pzm);
        const T lap = laplacian<T, R>(P, u + j, v, zp, zm);
        const T c2 = cv[j] * cv[j];
        const T dm = um[j];
        const bool edge = XC ? c.zedge : ((x == 0) | (x == nx - 1) | c.zedge);

        T dux = T(0), duz = T(0), dpx = T(0), dpz = T(0);
#pragma unroll
        for (int i = R; i >= 1; --i) {
            dux = fma(P.cx.d1[i - 1], u[R + j + i] - u[R + j - i], dux);
            duz = fma(P.cz.d1[i - 1], zp[i] - zm[i], duz);
            dpx = fma(P.cx.d1[i - 1], px[R + j + i] - px[R + j - i], dpx);
            dpz = fma(P.cz.d1[i - 1], pzp[i] - pzm[i], dpz);
        }
        const T src = ((f + lap) + dpx) + dpz;
        const T S = c2 * src;
        const T sx = XC ? T(0) : lane_bcast(vsig, j);
        const T sxmj = XC ? T(0) : lane_bcast(vsig, TX + j);
        const T dt = P.tc.dt;
        const T pxc = px[R + j], pzc = pv[j];
        // undamped update: increment form (batch) or the reference's form
        const T dni = fma(P.tc.dt2, S, dm);
        const T upi = PACKED ? v + dni : fma(P.tc.dt2, S, T(2) * v - dm);
        T pxn, pzn, up, dn = T(0), dw;
        bool pml;
        if (XC || (sx == T(0) && sxmj == T(0) && x > 0 && x < nx - 1)) {
            // a row outside the x PML (wave-uniform): the damped update below
            // with sx = 0 and the per-lane division hoisted; the same values
            // (zero terms dropped: at most the sign of a zero differs)
            pml = szk != T(0);
            const T pxd = pxc + dtszk * dux;
            const T pzd = (pzc - dtszk * pzc) + (-dtszk) * duz;
            pxn = pml ? pxd : pxc;
            pzn = pml ? pzd : pzc;
            if constexpr (PACKED) {
                dn = pml ? dm + fac1z * (S - szk_dt * dm) : dni;
                up = v + dn;
                dw = S + dsz * (dn + dm) * P.tc.inv_2dt;
            } else {
                const T fac2 = (S - (dm - T(2) * v) * P.tc.inv_dt2) + szk * dm * P.tc.inv_2dt;
                up = pml ? fac1z * fac2 : upi;
                dw = S + dsz * (up - dm) * P.tc.inv_2dt;
            }
        } else {
            pml = sx != T(0) || szk != T(0);
            // damped update (PML node) and plain update, selected per lane
            const T pxd = (pxc - dt * sx * pxc) + dt * (szk - sx) * dux;
            const T pzd = (pzc - dt * szk * pzc) + dt * (sx - szk) * duz;
            const T ss = sx + szk;
            const T pp = sx * szk;
            const T fac1 = P.tc.two_dt2 / (T(2) + dt * ss);
            pxn = pml ? pxd : pxc;
            pzn = pml ? pzd : pzc;
            // dWaveOp with the mesh profile: the step identity gives
            // (u+ - 2u + u-)/dt^2 = S - sum(sk)(u+ - u-)/(2dt) - prod(sk) u
            if constexpr (PACKED) {
                // du_{k+1} = du_k + fac1 (S - prod(sk) u - sum(sk) du_k / dt):
                // the reference's damped update with u_{k-1} = u_k - du_k
                dn = pml ? dm + fac1 * ((S - pp * v) - ss * dm * P.tc.inv_dt) : dni;
                up = v + dn;
                dw = S + ((sxmj + szm) - ss) * (dn + dm) * P.tc.inv_2dt + (sxmj * szm - pp) * v;
            } else {
                const T fac2 = S - (dm - T(2) * v) * P.tc.inv_dt2 + ss * dm * P.tc.inv_2dt - pp * v;
                up = pml ? fac1 * fac2 : upi;
                dw = S + ((sxmj + szm) - ss) * (up - dm) * P.tc.inv_2dt + (sxmj * szm - pp) * v;
            }
        }
        if (c.dw_exact) dw = S;
        bool phi_store = pml || c.general_phi;
        if (edge) {
            up = T(0);
            dn = T(0);
            pxn = T(0);
            pzn = T(0);
            dw = T(0);
            phi_store = true;
        }
        if (is_adjoint(MODE)) ic_store<T, MODE>(c, own ? ovi : kOOB, xr, icv[j], v, dv[j]);
        uo_st<MODE>(c.s.rs(F_UO), ovr, so_uo, up);
        if constexpr (PACKED) du_st<T, MODE>(c.s.rs(F_DU), ovr, so_du, dn);
        const int op = phi_store ? ovr : kOOB;
        buf_st(c.s.rs(F_PXO), op, c.s.so(F_PXO), pxn);
        buf_st(c.s.rs(F_PZO), op, c.s.so(F_PZO), pzn);
        if (MODE == 1) dw_st<T>(c.dw, own ? ovd + xr : kOOB, c.shb, dw);
        if (stores_out(MODE)) dw_st<T>(c.dwo, own ? ovo + xr : kOOB, c.shb, dw);
        if (c.has_fst) buf_st(c.fst, own ? ovf + xr : kOOB, c.shb, v);
    }
#undef PA_LD_U
#undef PA_LD_ROW
}

template <typename T, int TX, int R, int MODE, bool SINGLE>
__device__ __forceinline__ void tile2d(const pysit_amd_batch &B, const Step2DParams<T> &P,
                                       const pysit_amd_shot &S, int tile, int flags, int tloc, int tx,
                                       int slot, int seg_base, int lane, T *lds_f, T *zbuf)
{
    constexpr bool PACKED = !SINGLE;
    constexpr int SZ = sizeof(T);
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    Ctx2D<T, R, PACKED> c;
    const int nx = S.nx, nz = S.nz;
    c.nx = nx;
    c.nz = nz;
    // z tiles: 0, 64, ... and a last one ending on node nz-1; owned ranges disjoint
    const int ntz = S.ntz;
    const int tz = tloc - tx * ntz;
    const int zlast = z_last_tile(nz, z_align<T>());
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
    const int taken = is_adjoint(MODE) ? (S.nt - 1 - k) : k;
    const int cur = taken & 1;
    // disabled arrays get an empty resource on a valid base: every access is
    // out of range, loads return 0 and stores are dropped
    const void *vbase = B.u[0];
    c.rowb = S.zs * SZ;
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
        c.s.o[F_DU] = P.fdu + c.shb;
    } else {
        const long long nb = (long long)nx * S.zs * SZ;
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
    c.nxl = S.nxl;
    c.nxr = S.nxr;
    c.general_phi = P.general_phi;
    c.dw_exact = B.dw_form == 1;
    c.has_rhs = P.rhs != nullptr;
    c.rhs = make_rsrc(c.has_rhs ? (const void *)P.rhs : vbase, c.has_rhs ? (long long)nx * nz * SZ : 0);
    c.zbuf = zbuf;
    c.hslot = lane < R ? R + 64 + lane : (lane >= 64 - R ? lane - (64 - R) : 2 * R + 64 + lane);
    const bool imaged = B.imaging_period == 1 || k % B.imaging_period == 0;
    c.img = is_adjoint(MODE) && imaged && B.dw;
    const bool store = (MODE == 1) && imaged && B.dw;
    const long long slot_el = (long long)(slot - seg_base) * B.total_n;
    const long long slot_bytes = B.total_n * SZ;
    const bool dwon = ((c.img || is_born(MODE)) && B.dw) || store;
    c.dw = make_rsrc(dwon ? (const void *)((T *)B.dw + slot_el) : vbase, dwon ? slot_bytes : 0);
    c.has_dwo = store;
    c.has_fst = B.fstore != nullptr;
    c.fst = make_rsrc(c.has_fst ? (const void *)((T *)B.fstore + slot_el) : vbase, c.has_fst ? slot_bytes : 0);
    if constexpr (MODE == 4) c.ic = make_rsrc(B.ic32, B.total_n * 4);
    else c.ic = make_rsrc((MODE == 2 || MODE == 6) ? (const void *)B.ic : vbase,
                          (MODE == 2 || MODE == 6) ? B.total_n * 8 : 0);
    c.ovd = dwon ? c.ov : kOOB;
    c.ovi = c.img ? (MODE == 4 ? c.ov : c.ov8) : kOOB;
    c.ovf = c.has_fst ? c.ov : kOOB;
    c.m1 = make_rsrc(is_born(MODE) ? B.born_m1 : vbase, is_born(MODE) ? (long long)(S.mx_hi + 1) * S.mnz * SZ : 0);
    c.ovo = kOOB;
    c.ovv = kOOB;
    if constexpr (stores_out(MODE)) {
        const bool out = B.dw_out != nullptr;
        c.dwo = make_rsrc(out ? (const void *)((T *)B.dw_out + slot_el) : vbase, out ? slot_bytes : 0);
        c.ovo = out ? c.ov : kOOB;
    }
    if constexpr (MODE == 6) {
        const bool vol = B.vol_m && B.vol_dw;
        c.volm = make_rsrc(vol ? B.vol_m : vbase, vol ? slot_bytes : 0);
        c.vold = make_rsrc(vol ? (const void *)((const T *)B.vol_dw + slot_el) : vbase, vol ? slot_bytes : 0);
        c.ovv = vol ? c.ov : kOOB;
    }

    // sources of this tile, gathered into LDS (row-major TX x 64): one lane
    // per destination node, each summing its entries in the reference's order
    // no sources: every row reads one zero slot (stride 0), so the row loop
    // has no branch on the injection
    T *zslot = zbuf + 4 * zrow_len<R>() - 1;  // the last parking slot of the last buffer, never written by rows
    c.has_inj = false;
    c.finj = zslot;
    c.fstride = 0;
    if (B.inj_ptr) {
        const int n0 = B.inj_ptr[tile], n1 = B.inj_ptr[tile + 1];
        if (n1 > n0) {
            const double *amp_k = S.amp + (long long)k * S.namp;
#pragma unroll
            for (int j = 0; j < TX; ++j) lds_f[j * 64 + lane] = T(0);
            wave_sync();
            if (B.inj_pre) {
                // values evaluated for the whole run before its first step
                const T *pre = (const T *)B.inj_pre + (long long)(slot - seg_base) * B.n_inj_nodes;
                for (int n = n0 + lane; n < n1; n += 64) lds_f[B.inj_node[n]] = pre[n];
            } else {
                for (int n = n0 + lane; n < n1; n += 64) {
                    const int dst = B.inj_node[n];
                    lds_f[dst] = (T)inj_sum(B.inj_chan, B.inj_w, amp_k, B.inj_nptr[n], B.inj_nptr[n + 1]);
                }
            }
            wave_sync();
            c.has_inj = true;
            c.finj = lds_f + lane;
            c.fstride = 64;
        }
    }
    if (!c.has_inj) {
        *zslot = T(0);
        wave_sync();
    }

    if (flags & 1) {
        // PML profiles as masked buffer loads (no branches, so the loads stay
        // ahead of the march in the wait-count order): lane z profile, and the
        // x profile of the tile's rows, one lane per value (lane r < TX kernel
        // form, lane TX + r mesh form), read back per row with readlane; rows
        // past the grid repeat the last row.  Plain tiles need none.
        static_assert(2 * TX <= 64, "the x profile of a tile must fit one wavefront");
        const void *vbase = B.u[0];
        const Rsrc rzl = make_rsrc(S.nzl ? S.zl : vbase, (long long)S.nzl * SZ);
        const Rsrc rzr = make_rsrc(S.nzr ? S.zr : vbase, (long long)S.nzr * SZ);
        const Rsrc rxl = make_rsrc(S.nxl ? S.xl : vbase, (long long)S.nxl * SZ);
        const Rsrc rxr = make_rsrc(S.nxr ? S.xr : vbase, (long long)S.nxr * SZ);
        sigma_lane<T>(rzl, S.nzl, rzr, S.nzr, zin && z <= nz - 2 ? z : -1, nz, true, c.szk2);
        sigma_lane<T>(rzl, S.nzl, rzr, S.nzr, zin ? z : -1, nz, false, c.szm2);
        if (is_adjoint(MODE) && (flags & 2)) {
            // rows clear of the x PML and the x edges: no x profile needed
            // (adjoint only: the forward measured 1 % slower with it)
            march_full<T, TX, R, MODE, PACKED, true>(c, P, x0);
        } else {
            const int xs = min(x0 + (lane < TX ? lane : lane - TX), nx - 1);
            const int xi = lane < 2 * TX ? (lane < TX ? max(0, min(xs, nx - 2)) : xs) : -1;
            sigma_lane<T>(rxl, S.nxl, rxr, S.nxr, xi, nx, lane < TX, c.vsig2);
            march_full<T, TX, R, MODE, PACKED, false>(c, P, x0);
        }
    } else {
        march_int<T, TX, R, MODE, PACKED>(c, P, x0);
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
        acc = rec_gather<T>(B.rec_idx, B.rec_w, u, rd.z, rd.w, g);
    }
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 16);
    if (live && g == 0) S->trace[(long long)k * S->nrec + rd.y] = acc;
}

// Grid: the tile blocks, then `rec_blocks` blocks that sample receivers
// (u_k, which no block of the launch writes).  Dispatched last, the short
// receiver gathers fill the tail of the launch, where tile blocks leave CUs
// idle (first, they cost 10 % of a forward step).  SINGLE is the
// reference-contract step: one shot passed by value, every tile PML.
#ifndef PA2D_REC_LAST
#define PA2D_REC_LAST 1
#endif
#ifndef PA2D_WAVES_PER_EU
#define PA2D_WAVES_PER_EU 1
#endif
// One block of a step's grid: `b` of `nblocks` (tile blocks, then
// `rec_blocks` receiver blocks); each of its 4 wavefronts marches one tile.
template <typename T, int TX, int R, int MODE, bool SINGLE>
__device__ __forceinline__ void step2d_block(const pysit_amd_batch &B, const Step2DParams<T> &P, int slot,
                                             int seg_base, int n_tile_waves, int n_rec_lanes, int rec_blocks,
                                             int b, int nblocks, T (&lds_f)[4][TX * 64],
                                             T (&lds_z)[4][4 * zrow_len<R>()])
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile_blocks = nblocks - rec_blocks;
#if PA2D_REC_LAST
    const int rb = b - tile_blocks, tb = b;
#else
    const int rb = b, tb = b - rec_blocks;
#endif
    if (!is_adjoint(MODE) && rb >= 0 && rb < rec_blocks) {
        const int rw = rb * 4 + w;
        const int rec = rw * 4 + (lane >> 4);
        if (rw * 4 < n_rec_lanes) sample_group<T>(B, slot, rec < n_rec_lanes ? rec : B.n_rec, lane & 15);
        return;
    }
    const int gw = xcd_block(tb, tile_blocks) * 4 + w;
    if (gw >= n_tile_waves) return;
    if (SINGLE) {
        tile2d<T, TX, R, MODE, SINGLE>(B, P, P.shot0, gw, 1, gw, gw / P.shot0.ntz, slot, seg_base, lane, lds_f[w],
                                       lds_z[w]);
    } else {
        const int4 td = reinterpret_cast<const int4 *>(B.tiles)[gw];
        tile2d<T, TX, R, MODE, SINGLE>(B, P, B.shots[td.x], gw, td.z & 3, td.y, td.w, slot, seg_base, lane,
                                       lds_f[w], lds_z[w]);
    }
}

// With a step cursor (graph launches) the slot is cursor[0] + slot and the
// dWaveOp store base is cursor[1]; the host checks the slot range.
template <typename T, int TX, int R, int MODE, bool SINGLE>
__global__ __launch_bounds__(256, PA2D_WAVES_PER_EU) void k_step2d(const pysit_amd_batch B, const Step2DParams<T> P,
                                                int slot, int seg_base, int n_tile_waves,
                                                int n_rec_lanes, int rec_blocks, const int *cursor)
{
    __shared__ T lds_f[4][TX * 64];
    __shared__ T lds_z[4][4 * zrow_len<R>()];
    if (cursor) {
        slot += cursor[0];
        seg_base = cursor[1];
    }
    step2d_block<T, TX, R, MODE, SINGLE>(B, P, slot, seg_base, n_tile_waves, n_rec_lanes, rec_blocks, blockIdx.x,
                                         gridDim.x, lds_f, lds_z);
}

// ---- the persistent time loop ----------------------------------------------
// One cooperative launch runs every step of a call: each block takes the
// step's blocks b = blockIdx.x, blockIdx.x + gridDim.x, ... (the grid is a
// multiple of 8, so b lands on the XCD the per-step launch would use), then
// all blocks meet at a grid barrier before the next step reads the new time
// level.  The barrier is an arrival counter and a generation word in device
// memory, with device-scope release and acquire fences around them; a block
// that waits longer than `spin_limit` polls sets `err` and every block leaves
// the loop at its next barrier, so the launch always ends.
struct Persist2D {
    int slot_begin, n_steps, direction, seg_base, window;
    const int *counts;   // per step of the call: active tiles, active receivers
    unsigned *bar;       // [arrived, generation], zero at launch
    int *err;            // set on a barrier time-out
    long long spin_limit;
};

__device__ __forceinline__ bool grid_barrier(const Persist2D &A, unsigned &gen, int &ok)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned g = gen;
        if (atomicAdd(&A.bar[0], 1u) + 1u == gridDim.x) {
            __hip_atomic_store(&A.bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&A.bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long spins = 0;
            while (__hip_atomic_load(&A.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++spins > A.spin_limit || __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    atomicExch(A.err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        gen = g + 1;
        __threadfence();
        // one decision per block, read by all its threads after the barrier
        ok = __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    __syncthreads();
    return ok != 0;
}

template <typename T, int TX, int R, int MODE>
__global__ __launch_bounds__(256, PA2D_WAVES_PER_EU) void k_persist2d(const pysit_amd_batch B,
                                                                      const Step2DParams<T> P, const Persist2D A)
{
    __shared__ T lds_f[4][TX * 64];
    __shared__ T lds_z[4][4 * zrow_len<R>()];
    __shared__ int ok;
    unsigned gen = 0;
    for (int i = 0; i < A.n_steps; ++i) {
        const int slot = A.direction > 0 ? A.slot_begin + i : A.slot_begin + A.n_steps - 1 - i;
        const int nt = A.counts[2 * i];
        const int nr = is_adjoint(MODE) ? 0 : A.counts[2 * i + 1];
        const int rec_blocks = nr ? ((nr + 15) / 16 + 7) / 8 * 8 : 0;
        const int total = (nt + 3) / 4 + rec_blocks;
        for (int b = blockIdx.x; b < total; b += gridDim.x)
            step2d_block<T, TX, R, MODE, false>(B, P, slot, A.seg_base, nt, nr, rec_blocks, b, total, lds_f, lds_z);
        if (!grid_barrier(A, gen, ok)) return;
        if constexpr (MODE == 4) {
            if (A.window && window_end(slot, A.direction, A.window)) {
                // the FP32 window's flush: ic += ic32, ic32 = 0
                const long long n = B.total_n;
                for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
                     k += (long long)gridDim.x * blockDim.x) {
                    B.ic[k] += (double)B.ic32[k];
                    B.ic32[k] = 0.0f;
                }
                if (!grid_barrier(A, gen, ok)) return;
            }
        }
    }
}

template <typename T, int TX, int R, int MODE>
static int persist2d(const pysit_amd_batch &B, const Step2DParams<T> &P, const Persist2D &A, int max_blocks,
                     hipStream_t st)
{
    auto k = k_persist2d<T, TX, R, MODE>;
    int per_cu = 0, dev = 0;
    PA_HIP(hipGetDevice(&dev));
    PA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0));
    int cus = 0;
    PA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // whole rounds of 8 blocks (one per XCD), so that logical block b stays
    // on one XCD from step to step; a block beyond a step's count waits at
    // the barrier
    const int cap = per_cu * cus / 8 * 8;
    PA_CHECK(cap >= 8, "the persistent loop does not fit the device");
    const int grid = std::min(cap, (max_blocks + 7) / 8 * 8);
    void *args[] = {(void *)&B, (void *)&P, (void *)&A};
    PA_HIP(hipLaunchCooperativeKernel((const void *)k, dim3(grid), dim3(256), args, 0, st));
    return PYSIT_AMD_OK;
}

template <typename T, int TX, int R>
static int persist2d_mode(int kmode, const pysit_amd_batch &B, const Step2DParams<T> &P, const Persist2D &A,
                          int max_blocks, hipStream_t st)
{
    switch (kmode) {
    case 0: return persist2d<T, TX, R, 0>(B, P, A, max_blocks, st);
    case 1: return persist2d<T, TX, R, 1>(B, P, A, max_blocks, st);
    case 2: return persist2d<T, TX, R, 2>(B, P, A, max_blocks, st);
    case 4:
        if constexpr (std::is_same<T, float>::value) return persist2d<T, TX, R, 4>(B, P, A, max_blocks, st);
        break;
    default: break;
    }
    set_error("the persistent loop runs modes 0, 1 and 2");
    return PYSIT_AMD_EINVAL;
}

template <typename T, int TX>
static int persist2d_r(int r, int kmode, const pysit_amd_batch &B, const Step2DParams<T> &P, const Persist2D &A,
                       int max_blocks, hipStream_t st)
{
    switch (r) {
    case 1: return persist2d_mode<T, TX, 1>(kmode, B, P, A, max_blocks, st);
    case 2: return persist2d_mode<T, TX, 2>(kmode, B, P, A, max_blocks, st);
    case 3: return persist2d_mode<T, TX, 3>(kmode, B, P, A, max_blocks, st);
    case 4: return persist2d_mode<T, TX, 4>(kmode, B, P, A, max_blocks, st);
    }
    set_error("order must be 2, 4, 6 or 8");
    return PYSIT_AMD_EINVAL;
}

template <typename T, int TX, int R, int MODE>
static int launch2d(const pysit_amd_batch &B, const Step2DParams<T> &P, int slot, int seg_base,
                    int n_tiles, int n_recs, hipStream_t st, const int *cursor)
{
    const int recs = is_adjoint(MODE) ? 0 : n_recs;
    // 16 receivers per block, block count rounded to whole XCD rounds
    const int rec_blocks = recs ? ((recs + 15) / 16 + 7) / 8 * 8 : 0;
    const int tile_blocks = (n_tiles + 3) / 4;
    if (rec_blocks + tile_blocks == 0) return PYSIT_AMD_OK;
    if (P.single) {
        if constexpr (MODE == 0) {
            hipLaunchKernelGGL((k_step2d<T, TX, R, MODE, true>), dim3(rec_blocks + tile_blocks), dim3(256), 0, st,
                               B, P, slot, seg_base, n_tiles, recs, rec_blocks, cursor);
        } else {
            set_error("the single-shot step only propagates");
            return PYSIT_AMD_EINVAL;
        }
    } else {
        hipLaunchKernelGGL((k_step2d<T, TX, R, MODE, false>), dim3(rec_blocks + tile_blocks), dim3(256), 0, st,
                           B, P, slot, seg_base, n_tiles, recs, rec_blocks, cursor);
    }
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

template <typename T, int TX, int MODE>
static int launch2d_tx(int r, const pysit_amd_batch &B, const Step2DParams<T> &P, int slot, int seg_base,
                       int n_tiles, int n_recs, hipStream_t st, const int *cursor)
{
    switch (r) {
    case 1: return launch2d<T, TX, 1, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case 2: return launch2d<T, TX, 2, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case 3: return launch2d<T, TX, 3, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case 4: return launch2d<T, TX, 4, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    }
    set_error("order must be 2, 4, 6 or 8");
    return PYSIT_AMD_EINVAL;
}

template <typename T, int MODE>
static int launch2d_r(int r, const pysit_amd_batch &B, const Step2DParams<T> &P, int slot,
                      int seg_base, int n_tiles, int n_recs, hipStream_t st, const int *cursor = nullptr)
{
    switch (B.tile_x) {
    case kTileRows2D[0]: return launch2d_tx<T, kTileRows2D[0], MODE>(r, B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case kTileRows2D[1]: return launch2d_tx<T, kTileRows2D[1], MODE>(r, B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case kTileRows2D[2]: return launch2d_tx<T, kTileRows2D[2], MODE>(r, B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    }
    set_error("order must be 2, 4, 6 or 8");
    return PYSIT_AMD_EINVAL;
}

int tile_rows_2d(int) { return kTileRows2D[0]; }
bool tile_rows_2d_supported(int tx) { return tx == kTileRows2D[0] || tx == kTileRows2D[1] || tx == kTileRows2D[2]; }

// ic += ic32; ic32 = 0 (the FP32 window accumulator's flush)
__global__ void k_ic_flush(float *__restrict__ ic32, double *__restrict__ ic, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ic[i] += (double)ic32[i];
    ic32[i] = 0.0f;
}

int ic_flush(const pysit_amd_batch &B, hipStream_t st)
{
    const long long n = B.total_n;
    if (n == 0) return PYSIT_AMD_OK;
    hipLaunchKernelGGL(k_ic_flush, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B.ic32, B.ic, n);
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

template <typename T>
struct Sink2D final : StepSink {
    const pysit_amd_batch &B;
    const Step2DParams<T> &P;
    int r, mode;
    bool window;
    Sink2D(const pysit_amd_batch &B_, const Step2DParams<T> &P_, int r_, int mode_, bool window_)
        : B(B_), P(P_), r(r_), mode(mode_), window(window_) {}
    int step(int slot, int seg_base, int nt, int nr, const int *cursor, hipStream_t st) override
    {
        const bool second = B.dw_out || B.vol_dw;
        switch (mode) {
        case 0: return launch2d_r<T, 0>(r, B, P, slot, seg_base, nt, nr, st, cursor);
        case 1: return launch2d_r<T, 1>(r, B, P, slot, seg_base, nt, nr, st, cursor);
        case 3:
            if (second) return launch2d_r<T, 5>(r, B, P, slot, seg_base, nt, nr, st, cursor);
            return launch2d_r<T, 3>(r, B, P, slot, seg_base, nt, nr, st, cursor);
        default: break;
        }
        if (second) return launch2d_r<T, 6>(r, B, P, slot, seg_base, nt, 0, st, cursor);
        if (!window) return launch2d_r<T, 2>(r, B, P, slot, seg_base, nt, 0, st, cursor);
        if constexpr (std::is_same<T, float>::value)
            return launch2d_r<T, 4>(r, B, P, slot, seg_base, nt, 0, st, cursor);
        set_error("the FP32 window accumulator needs an FP32 batch");
        return PYSIT_AMD_EINVAL;
    }
    int flush(hipStream_t st) override { return ic_flush(B, st); }
};

// the batch steps' parameters (and the checks of the state region)
template <typename T>
static int params2d(const pysit_amd_batch &B, Step2DParams<T> &P)
{
    PA_CHECK(tile_rows_2d_supported(B.tile_x),
             "2D tiles are " + std::to_string(kTileRows2D[0]) + ", " + std::to_string(kTileRows2D[1]) + " or " +
                 std::to_string(kTileRows2D[2]) +
                 " rows along x (tile_x)");
    P = Step2DParams<T>{};
    const int r = B.order / 2;
    P.cx = axis_coef<T>(r, B.dx);
    P.cz = axis_coef<T>(r, B.dz);
    P.tc = time_coef<T>(B.dt);
    P.rhs = nullptr;
    P.single = 0;
    P.general_phi = 0;
    // the seven state buffers share one resource: they must lie in one region
    {
        const char *f[7] = {(const char *)B.u[0], (const char *)B.u[1], (const char *)B.phix[0],
                            (const char *)B.phix[1], (const char *)B.phiz[0], (const char *)B.phiz[1],
                            (const char *)B.du};
        const char *lo = f[0], *hi = f[0];
        for (int i = 0; i < 7; ++i) {
            PA_CHECK(f[i] != nullptr, "2D batch state buffers (u, phix, phiz, du) must all be given");
            lo = f[i] < lo ? f[i] : lo;
            hi = f[i] > hi ? f[i] : hi;
        }
        const long long ext = (long long)(hi - lo) + B.total_n * (long long)sizeof(T);
        PA_CHECK(ext <= kMaxRsrcBytes, "2D batch state buffers must lie in one region of at most " +
                                           std::to_string(kMaxRsrcBytes) + " bytes");
        for (int i = 0; i < 7; ++i) {
            PA_CHECK((f[i] - lo) % (long long)sizeof(T) == 0, "misaligned 2D batch state buffer");
            if (i < 6) P.fo[i / 2][i % 2] = (int)(f[i] - lo);
        }
        P.fdu = (int)(f[6] - lo);
        P.state = lo;
        P.state_bytes = (int)ext;
    }
    PA_CHECK(B.total_n * 8 <= kMaxRsrcBytes, "batch too large for one launch (total_n)");
    return PYSIT_AMD_OK;
}

template <typename T>
int run2d(const pysit_amd_batch &B, int mode, int slot_begin, int slot_end, int direction,
          int seg_base, const int *active_tiles, const int *active_recs, pysit_amd_plan *plan,
          hipStream_t st)
{
    Step2DParams<T> P;
    const int rc = params2d<T>(B, P);
    if (rc) return rc;
    const int r = B.order / 2;
    const bool window = mode == 2 && B.ic32 != nullptr;
    Sink2D<T> S(B, P, r, mode, window);
    const LoopSpec L{mode, slot_begin, slot_end, direction, seg_base, active_tiles, active_recs,
                     window ? B.ic_window : 0};
    return time_loop(plan, B, L, S, st);
}

template int run2d<float>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                          const int *, pysit_amd_plan *, hipStream_t);
template int run2d<double>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                           const int *, pysit_amd_plan *, hipStream_t);

// The whole call as one persistent cooperative launch (k_persist2d).  The
// per-step active counts go to a device buffer allocated on the stream; the
// call waits for the launch to read the barrier's time-out flag.
template <typename T>
int run2d_persistent(const pysit_amd_batch &B, int mode, int slot_begin, int slot_end, int direction,
                     int seg_base, const int *active_tiles, const int *active_recs, long long spin_limit,
                     hipStream_t st)
{
    Step2DParams<T> P;
    int rc = params2d<T>(B, P);
    if (rc) return rc;
    PA_CHECK(mode == 0 || mode == 1 || mode == 2, "the persistent loop runs modes 0, 1 and 2");
    PA_CHECK(!B.dw_out && !B.vol_dw, "the persistent loop has no second-order terms");
    PA_CHECK(B.tile_x == kTileRows2D[0] || B.tile_x == kTileRows2D[1],
             "the persistent loop is built for " + std::to_string(kTileRows2D[0]) + "- and " +
                 std::to_string(kTileRows2D[1]) + "-row tiles");
    const int n = slot_end - slot_begin;
    if (n <= 0) return PYSIT_AMD_OK;
    const bool window = mode == 2 && B.ic32 != nullptr;
    const int kmode = mode == 2 ? (window ? 4 : 2) : mode;
    std::vector<int> host(2 * (size_t)n + 4, 0);
    int max_blocks = 8;
    for (int i = 0; i < n; ++i) {
        const int j = direction > 0 ? slot_begin + i : slot_end - 1 - i;
        const int nt = active_tiles ? active_tiles[j] : B.n_tiles;
        const int nr = mode == 2 ? 0 : (active_recs ? active_recs[j] : B.n_rec);
        host[2 * i] = nt;
        host[2 * i + 1] = nr;
        const int rec_blocks = nr ? ((nr + 15) / 16 + 7) / 8 * 8 : 0;
        max_blocks = std::max(max_blocks, (nt + 3) / 4 + rec_blocks);
    }
    void *buf = nullptr;
    PA_HIP(hipMallocAsync(&buf, host.size() * sizeof(int), st));
    int *dcounts = (int *)buf;
    unsigned *bar = (unsigned *)(dcounts + 2 * n);
    int *err = (int *)(bar + 2);
    rc = PYSIT_AMD_OK;
    if (hipMemcpyAsync(buf, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error("copy of the persistent loop's step counts failed");
        rc = PYSIT_AMD_EDEVICE;
    }
    if (!rc) {
        Persist2D A{};
        A.slot_begin = slot_begin;
        A.n_steps = n;
        A.direction = direction;
        A.seg_base = seg_base;
        A.window = window ? B.ic_window : 0;
        A.counts = dcounts;
        A.bar = bar;
        A.err = err;
        A.spin_limit = spin_limit;
        rc = B.tile_x == kTileRows2D[0] ? persist2d_r<T, kTileRows2D[0]>(B.order / 2, kmode, B, P, A, max_blocks, st)
                                        : persist2d_r<T, kTileRows2D[1]>(B.order / 2, kmode, B, P, A, max_blocks, st);
    }
    int flag = 0;
    if (!rc && hipMemcpyAsync(&flag, err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) {
        set_error("read of the persistent loop's barrier flag failed");
        rc = PYSIT_AMD_EDEVICE;
    }
    const hipError_t se = hipStreamSynchronize(st);
    (void)hipFreeAsync(buf, st);
    if (!rc && se != hipSuccess) {
        set_error(std::string("persistent loop: ") + hipGetErrorString(se));
        rc = PYSIT_AMD_EDEVICE;
    }
    if (!rc && flag) {
        set_error("persistent loop: a grid barrier timed out");
        rc = PYSIT_AMD_EDEVICE;
    }
    return rc;
}

template int run2d_persistent<float>(const pysit_amd_batch &, int, int, int, int, int, const int *, const int *,
                                     long long, hipStream_t);
template int run2d_persistent<double>(const pysit_amd_batch &, int, int, int, int, int, const int *, const int *,
                                      long long, hipStream_t);

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
    B.tile_x = kTileRows2D[0];
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
    S.zs = nz;
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

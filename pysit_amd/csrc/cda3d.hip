// cda3d.hip -- 3D constant-density acoustic leapfrog + PML step for gfx950.
//
// Semantics: one time step of pysit/solvers/constant_density_acoustic/time/
// scalar/constant_density_acoustic_time_scalar_3D.h:46-236 (u, Phix, Phiy,
// Phiz, psi; taps beyond the grid read 0; edge nodes held at 0; psi += dt*u
// on every updated node; right-side PML index quirk as in 2D).
//
// Mapping to CDNA: a 512-thread workgroup owns 64 z (contiguous) x 8 y x 16 x
// nodes; wavefront w owns row y0 + w.  The z tiles of a shot start at 0, 64,
// ... except the last, which ends on the last node (owned ranges disjoint).
// The workgroup marches through its 16 x planes: each lane keeps a register
// queue of its own row along x (u, and in PML tiles Phix and psi) for the x
// derivatives, and the current y-z plane is staged in a double-buffered LDS
// plane for the y and z derivatives (one barrier per plane).  Every wave
// loads its own row, the z halo of its own row (lanes 0..r-1 and 64-r..63),
// and waves 0..2r-1 one y-halo row each.  The loads of the planes ahead are
// issued before the current plane is staged, and the barrier waits on LDS
// only, so they stay in flight.
//
// Addressing: per-field raw buffer resources based at the shot's region; a
// per-lane byte offset (masked lanes: out of range, reads 0, stores dropped)
// plus the plane offset.  Tiles farther than r nodes from every PML node,
// edge node and the grid boundary run the plain update and never touch Phi or
// psi (Phi is zero off the PML, psi only enters through sigma-weighted terms
// and PML-node derivatives); the reference-contract step marks every tile as
// PML so that it returns all five fields exactly as the reference defines them.
#include "common.h"

#include <type_traits>

// No implicit FP contraction (see cda2d.hip): all modes round identically.
#pragma clang fp contract(off)

namespace pa {

// Kernel modes: 0 propagate, 1 + dWaveOp store, 2 adjoint with the FP64
// imaging accumulator, 3 Born, 4 adjoint with the FP32 window accumulator
// (flushed into the FP64 one by the host loop, as in 2D); the full Hessian's
// second-order runs: 5 Born that also stores its own dWaveOp (dw_out), 6
// adjoint with the FP64 accumulator that also stores its own dWaveOp and adds
// the volume source vol_m * vol_dw_k.
constexpr bool adj3(int mode) { return mode == 2 || mode == 4 || mode == 6; }
constexpr bool born3(int mode) { return mode == 3 || mode == 5; }
constexpr bool out3(int mode) { return mode == 5 || mode == 6; }
constexpr bool ic64_3(int mode) { return mode == 2 || mode == 6; }
template <int MODE> struct Acc3D { typedef double type; };
template <> struct Acc3D<4> { typedef float type; };

// y rows per workgroup (one wavefront each): the waves 0..2r-1 also load the
// 2r y-halo rows.  16 rows in FP32 (1024-thread workgroups, one per CU, half
// the y-halo reads per node) ran both 3D steps 11 % slower than 8.
#ifndef PA3D_TY_F32
#define PA3D_TY_F32 8
#endif
template <typename T> struct Tile3 { static constexpr int TY = sizeof(T) == 4 ? PA3D_TY_F32 : 8; };
// x planes per tile: 16 runs the 3D steps 7-14 % faster than 8 (the x halo
// of 2r planes and each tile's start are spread over twice the planes); the
// injection block in LDS grows to TX3 * 8 * 64 values (65 KB per FP32
// workgroup, still two per CU)
#ifndef PA3D_TX
#define PA3D_TX 16
#endif
constexpr int TX3 = PA3D_TX;

int tile_rows_3d(int dtype, int axis) { return axis == 0 ? TX3 : (dtype ? Tile3<double>::TY : Tile3<float>::TY); }

// x planes whose loads are in flight ahead of the plane being computed: 1
// in PML tiles (their march is at the 128-VGPR limit; 2 spills), 2 in plain
// tiles (2 % faster steps than 1; 3 and 4 no faster).
#ifndef PA3D_PREFETCH
#define PA3D_PREFETCH 1
#endif
// x planes in flight in PML tiles that touch at most one PML band (they
// keep no psi, below): 1 (2 measured 2-6 % slower on the 3D steps)
#ifndef PA3D_PREFETCH_NOPSI
#define PA3D_PREFETCH_NOPSI 1
#endif
#ifndef PA3D_PSI_SKIP
#define PA3D_PSI_SKIP 1
#endif
#ifndef PA3D_PREFETCH_INT
#define PA3D_PREFETCH_INT 2
#endif

template <typename T>
struct Step3DParams {
    AxisCoef<T> cx, cy, cz;
    TimeCoef<T> tc;
    const T *rhs;
    int single;
    pysit_amd_shot shot0;
};

// LDS of a workgroup: double-buffered planes (u and psi with y and z halos,
// Phiy with y halo rows, Phiz with the z halo), the injection block, and a
// write-off slot per lane for lanes that hold no halo value.
template <typename T, int R>
struct Lds3D {
    static constexpr int LZ = 64 + 2 * R;
    static constexpr int TY = Tile3<T>::TY;
    static constexpr int LY = TY + 2 * R;
    T u[2][LY * LZ];
    T ps[2][LY * LZ];
    T py[2][LY * 64];
    T pz[2][TY * LZ];
    T f[TX3 * TY * 64];
    T junk[4][64];
};

__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T, int R>
struct Ctx3D {
    // u: u_k; uo: u_{k+1} (increment form) or u_{k-1}, overwritten by u_{k+1}
    // (the reference-contract step); du: the increment du_k = u_k - u_{k-1}
    Rsrc u, uo, du, px, pxo, py, pyo, pz, pzo, ps, pso, model, m1, dw, fst, ic, rhs;
    Rsrc dwo, volm, vold;    // second-order modes: dWaveOp out, volume source factors
    bool has_fst, has_rhs;
    int nx, ny, nz, lane, w, y;
    int plane;               // bytes per x plane of the shot grid
    int co, cho, hro;        // own row centre, own row z-halo value, y-halo row centre (kOOB if none)
    int ov, ov8, ovd, ovi, ovf;
    int ovo, ovv;            // owned node where dw_out is written / the volume source read (kOOB if not)
    int mv, mplane, mx_off, mx_lo, mx_hi;
    int shb, shb8;
    int hl;                  // LDS index (in a plane buffer) of this lane's z-halo slot in its own row, or -1
    int hrow;                // LDS row of this wave's y-halo row, or -1
    bool yzedge;             // y or z on the grid edge
    // Phi (x, y, z) is nonzero only at PML nodes: a full tile reads and
    // writes it only where the node lies in a PML band (every node for the
    // reference-contract step, whose Phi input is arbitrary).  Lane masks:
    // own node, y-halo row node, z-halo node in the y or z band; the x band
    // is a per-row test.
    bool gen, bp, bph, bpz;
    int nxl, nxr;
    T syk, sym, szk, szm;    // lane y/z profiles (kernel, mesh)
    T vsig;                  // x profile of the tile planes: lane p < TX kernel, lane TX + p mesh
    const T *finj;
    int fstride;
};

// LDS store of this lane's z-halo value (into the write-off slot when the
// lane holds none)
template <typename T, int R>
__device__ __forceinline__ void st_halo(T *plane, T *junk, int hl, int lane, T v)
{
    T *dst = hl >= 0 ? plane + hl : junk + lane;
    *dst = v;
}

// PML profile value at node i (i < 0: none), kernel or mesh form, as two
// masked buffer loads (see cda2d.hip)
template <typename T>
__device__ __forceinline__ T sigma3(Rsrc l, int nl, Rsrc r, int nr, int i, int n, bool kernel)
{
    constexpr int SZ = sizeof(T);
    const bool inl = i >= 0 && i < nl;
    const bool inr = !inl && i >= 0 && i >= n - nr && (!kernel || i <= n - 2);
    const int ri = kernel ? nr - ((n - 1) - i) : i - (n - nr);
    return buf_ld<T>(l, inl ? i * SZ : kOOB, 0) + buf_ld<T>(r, inr ? ri * SZ : kOOB, 0);
}

__device__ __forceinline__ float bcast3(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast3(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The increment du: adjoint steps may access it non-temporally
// (PA3D_DU_NT_ADJ), as the 2D adjoint does; forward steps keep it cached.
#ifndef PA3D_DU_NT_ADJ
#define PA3D_DU_NT_ADJ 1
#endif
template <typename T, int MODE>
__device__ __forceinline__ T du_ld3(Rsrc r, int voff)
{
    if constexpr (adj3(MODE) && PA3D_DU_NT_ADJ) return buf_ld_nt<T>(r, voff, 0);
    else return buf_ld<T>(r, voff, 0);
}
template <typename T, int MODE>
__device__ __forceinline__ void du_st3(Rsrc r, int voff, T v)
{
    if constexpr (adj3(MODE) && PA3D_DU_NT_ADJ) buf_st_nt(r, voff, 0, v);
    else buf_st(r, voff, 0, v);
}

// byte offset `o` plus plane offset `p`, kept out of range when o is
__device__ __forceinline__ int at(int o, int p) { return o == kOOB ? kOOB : o + p; }

// Laplacian with difference-form taps, sum_i w_i ((u_{+i} - u) + (u_{-i} - u))
// per axis (the centre weight implied), from the x queue (qx[R] is the node)
// and the LDS plane (own: the node; rows LZ apart): neighbour differences of
// smooth fields round far less than the plain form's terms, which cancel
// (see cda2d.hip laplacian)
template <typename T, int R, int LZ>
__device__ __forceinline__ T laplacian3(const Step3DParams<T> &P, const T *qx, const T *lu, int own, T v)
{
    T lx = T(0), lyy = T(0), lz = T(0);
#pragma unroll
    for (int i = R; i >= 1; --i) {
        lx = fma(P.cx.d2[i], (qx[R + i] - v) + (qx[R - i] - v), lx);
        lyy = fma(P.cy.d2[i], (lu[own + i * LZ] - v) + (lu[own - i * LZ] - v), lyy);
        lz = fma(P.cz.d2[i], (lu[own + i] - v) + (lu[own - i] - v), lz);
    }
    return (lx + lyy) + lz;
}

// ---- plain tiles -------------------------------------------------------------
template <typename T, int R, int MODE, bool INC>
__device__ __forceinline__ void march3_int(const Ctx3D<T, R> &c, const Step3DParams<T> &P, Lds3D<T, R> &L,
                                           int x0)
{
    constexpr int TX = TX3, D = PA3D_PREFETCH_INT, NQ = TX + 2 * R, LZ = Lds3D<T, R>::LZ;
    constexpr int SZ = sizeof(T);
    const int pl = c.plane;
    const int co = c.co, cho = c.cho, hro = c.hro, ov = c.ov, mv = c.mv;
    const int ovd = c.ovd, ovi = c.ovi, ovf = c.ovf, ovo = c.ovo, ovv = c.ovv;
    T q[NQ], uz[TX], yh[TX], um[TX], cv[TX], fv[TX], dv[TX], mm[TX], wv[TX];
    typename Acc3D<MODE>::type icv[TX];
#define PA_Q(s_) q[s_] = buf_ld<T>(c.u, at(co, (x0 - R + (s_)) * pl), 0)
#define PA_ROW(r_)                                                                                  \
    do {                                                                                            \
        const int r = (r_), xp = (x0 + r) * pl;                                                     \
        uz[r] = buf_ld<T>(c.u, at(cho, xp), 0);                                                     \
        yh[r] = buf_ld<T>(c.u, at(hro, xp), 0);                                                     \
        um[r] = INC ? du_ld3<T, MODE>(c.du, at(ov, xp)) : buf_ld<T>(c.uo, at(ov, xp), 0);           \
        const int mp = clampi(x0 + r - c.mx_off, c.mx_lo, c.mx_hi) * c.mplane;                      \
        cv[r] = buf_ld<T>(c.model, mv, mp);                                                         \
        fv[r] = c.finj[r * c.fstride];                                                              \
        dv[r] = (MODE >= 2) ? buf_ld<T>(c.dw, at(ovd, xp), c.shb) : T(0);                           \
        mm[r] = born3(MODE) ? buf_ld<T>(c.m1, mv, mp)                                               \
              : (MODE == 6 ? buf_ld<T>(c.volm, at(ovv, xp), c.shb) : T(0));                         \
        wv[r] = (MODE == 6) ? buf_ld<T>(c.vold, at(ovv, xp), c.shb) : T(0);                         \
        if constexpr (ic64_3(MODE)) icv[r] = buf_ld<double>(c.ic, at(ovi, xp * (8 / SZ)), c.shb8);  \
        if constexpr (MODE == 4) icv[r] = buf_ld<float>(c.ic, at(ovi, xp), c.shb);                    \
    } while (0)
#pragma unroll
    for (int s = 0; s < 2 * R + D && s < NQ; ++s) PA_Q(s);
#pragma unroll
    for (int s = 0; s < D && s < TX; ++s) PA_ROW(s);
    __builtin_amdgcn_sched_barrier(0);
    const int own = (R + c.w) * LZ + R + c.lane;
#pragma unroll
    for (int j = 0; j < TX; ++j) {
        if (j + 2 * R + D < NQ) PA_Q(j + 2 * R + D);
        if (j + D < TX) PA_ROW(j + D);
        __builtin_amdgcn_sched_barrier(0);
        T *lu = L.u[j & 1];
        const T v = q[R + j];
        lu[own] = v;
        st_halo<T, R>(lu, L.junk[0], c.hl, c.lane, uz[j]);
        if (c.hrow >= 0) lu[c.hrow * LZ + R + c.lane] = yh[j];
        lds_barrier();
        const int x = x0 + j, xp = x * pl;
        T f = fv[j];
        if (born3(MODE)) f += mm[j] * (-dv[j]);
        if (MODE == 6) f += mm[j] * wv[j];
        const T lap = laplacian3<T, R, LZ>(P, q + j, lu, own, v);
        const T c2 = cv[j] * cv[j];
        const T src = f + lap;
        const T S = c2 * src;
        if constexpr (ic64_3(MODE)) buf_st(c.ic, at(ovi, xp * (8 / SZ)), c.shb8, icv[j] + (double)v * (double)dv[j]);
        if constexpr (MODE == 4) buf_st(c.ic, at(ovi, xp), c.shb, icv[j] + v * dv[j]);
        if constexpr (INC) {
            // increment form: du_{k+1} = du_k + dt^2 S, u_{k+1} = u_k + du_{k+1}
            const T dn = fma(P.tc.dt2, S, um[j]);
            buf_st(c.uo, at(ov, xp), 0, v + dn);
            du_st3<T, MODE>(c.du, at(ov, xp), dn);
        } else {
            buf_st(c.uo, at(ov, xp), 0, fma(P.tc.dt2, S, T(2) * v - um[j]));
        }
        if (MODE == 1) buf_st(c.dw, at(ovd, xp), c.shb, S);
        if (out3(MODE)) buf_st(c.dwo, at(ovo, xp), c.shb, S);
        if (c.has_fst) buf_st(c.fst, at(ovf, xp), c.shb, v);
    }
#undef PA_Q
#undef PA_ROW
}

// ---- PML tiles -----------------------------------------------------------------
// XC: no plane of the tile or within R planes of it is an x PML plane or an
// x edge (the host flags such tiles): sigma_x is zero on every plane, the
// terms it multiplies drop out and the damping factor is one per lane.
// PSI: some node of the tile or within R nodes of it lies in two or more PML
// bands.  psi only enters multiplied by a product of two or three sigmas:
// the node's own psi by sigma_x sigma_y sigma_z, and the psi derivative
// along one axis by the product of the other two axes' sigmas at the node
// taking it, whose stencil nodes share those two coordinates, hence those
// two bands.  So psi is read only at nodes in two or more bands, and a tile
// without such a node (PSI false) neither reads nor advances it: its psi
// stays zero, every term it would enter is an exact zero, and the results
// equal the full form's (at most the sign of a zero differs).
template <typename T, int R, int MODE, bool XC, bool INC, bool PSI>
__device__ __forceinline__ void march3_full(const Ctx3D<T, R> &c, const Step3DParams<T> &P, Lds3D<T, R> &L,
                                            int x0, int hl_pz)
{
    constexpr int TX = TX3, D = PSI ? PA3D_PREFETCH : PA3D_PREFETCH_NOPSI, NQ = TX + 2 * R;
    constexpr int LZ = Lds3D<T, R>::LZ;
    constexpr int SZ = sizeof(T);
    const int pl = c.plane, nx = c.nx;
    const int co = c.co, cho = c.cho, hro = c.hro, ov = c.ov, mv = c.mv;
    const int ovd = c.ovd, ovi = c.ovi, ovf = c.ovf, ovo = c.ovo, ovv = c.ovv;
    T q[NQ], qx[NQ], qs[NQ];
    T uz[TX], yh[TX], sz[TX], sh[TX], pyc[TX], pyh[TX], pzc[TX], pzz[TX];
    T um[TX], cv[TX], fv[TX], dv[TX], mm[TX], wv[TX];
    typename Acc3D<MODE>::type icv[TX];
#define PA_Q(s_)                                                                                    \
    do {                                                                                            \
        const int xx = x0 - R + (s_);                                                               \
        const bool in = xx >= 0 && xx < nx;                                                         \
        const int o = in ? at(co, xx * pl) : kOOB;                                                  \
        const bool xb = c.gen || xx < c.nxl || xx >= nx - c.nxr;                                    \
        q[s_] = buf_ld<T>(c.u, o, 0);                                                               \
        qx[s_] = buf_ld<T>(c.px, (xb || c.bp) ? o : kOOB, 0);                                       \
        qs[s_] = PSI ? buf_ld<T>(c.ps, o, 0) : T(0);                                                \
    } while (0)
#define PA_ROW(r_)                                                                                  \
    do {                                                                                            \
        const int r = (r_), x = x0 + r;                                                             \
        const bool in = x < nx;                                                                     \
        const int xs = in ? x : nx - 1, xp = xs * pl;                                              \
        const int oc = in ? at(co, xp) : kOOB, oz = in ? at(cho, xp) : kOOB;                        \
        const int oh = in ? at(hro, xp) : kOOB, oo = in ? at(ov, xp) : kOOB;                        \
        const bool xb = c.gen || x < c.nxl || x >= nx - c.nxr;                                      \
        uz[r] = buf_ld<T>(c.u, oz, 0);                                                              \
        yh[r] = buf_ld<T>(c.u, oh, 0);                                                              \
        sz[r] = PSI ? buf_ld<T>(c.ps, oz, 0) : T(0);                                                \
        sh[r] = PSI ? buf_ld<T>(c.ps, oh, 0) : T(0);                                                \
        pyc[r] = buf_ld<T>(c.py, (xb || c.bp) ? oc : kOOB, 0);                                      \
        pyh[r] = buf_ld<T>(c.py, (xb || c.bph) ? oh : kOOB, 0);                                     \
        pzc[r] = buf_ld<T>(c.pz, (xb || c.bp) ? oc : kOOB, 0);                                      \
        pzz[r] = buf_ld<T>(c.pz, (xb || c.bpz) ? oz : kOOB, 0);                                     \
        um[r] = INC ? du_ld3<T, MODE>(c.du, oo) : buf_ld<T>(c.uo, oo, 0);                           \
        const int mp = clampi(xs - c.mx_off, c.mx_lo, c.mx_hi) * c.mplane;                         \
        cv[r] = buf_ld<T>(c.model, mv, mp);                                                         \
        T f_ = c.has_rhs ? buf_ld<T>(c.rhs, oo, 0) : T(0);                                          \
        f_ += c.finj[r * c.fstride];                                                                \
        fv[r] = f_;                                                                                 \
        dv[r] = (MODE >= 2) ? buf_ld<T>(c.dw, in ? at(ovd, xp) : kOOB, c.shb) : T(0);               \
        mm[r] = born3(MODE) ? buf_ld<T>(c.m1, mv, mp)                                               \
              : (MODE == 6 ? buf_ld<T>(c.volm, in ? at(ovv, xp) : kOOB, c.shb) : T(0));            \
        wv[r] = (MODE == 6) ? buf_ld<T>(c.vold, in ? at(ovv, xp) : kOOB, c.shb) : T(0);             \
        if constexpr (ic64_3(MODE)) icv[r] = buf_ld<double>(c.ic, in ? at(ovi, xp * (8 / SZ)) : kOOB, c.shb8); \
        if constexpr (MODE == 4) icv[r] = buf_ld<float>(c.ic, in ? at(ovi, xp) : kOOB, c.shb);         \
    } while (0)
#pragma unroll
    for (int s = 0; s < 2 * R + D && s < NQ; ++s) PA_Q(s);
#pragma unroll
    for (int s = 0; s < D && s < TX; ++s) PA_ROW(s);
    __builtin_amdgcn_sched_barrier(0);
    const T syk = c.syk, sym = c.sym, szk = c.szk, szm = c.szm, vsig = c.vsig;
    const int own = (R + c.w) * LZ + R + c.lane;     // u, psi planes
    const int owny = (R + c.w) * 64 + c.lane;        // Phiy plane
    const int ownz = c.w * LZ + R + c.lane;          // Phiz plane
#pragma unroll
    for (int j = 0; j < TX; ++j) {
        if (j + 2 * R + D < NQ) PA_Q(j + 2 * R + D);
        if (j + D < TX) PA_ROW(j + D);
        __builtin_amdgcn_sched_barrier(0);
        const int b = j & 1;
        T *lu = L.u[b], *lps = L.ps[b], *lpy = L.py[b], *lpz = L.pz[b];
        const T uc = q[R + j], psc = qs[R + j];
        lu[own] = uc;
        if constexpr (PSI) lps[own] = psc;
        lpy[owny] = pyc[j];
        lpz[ownz] = pzc[j];
        st_halo<T, R>(lu, L.junk[0], c.hl, c.lane, uz[j]);
        if constexpr (PSI) st_halo<T, R>(lps, L.junk[1], c.hl, c.lane, sz[j]);
        st_halo<T, R>(lpz, L.junk[2], hl_pz, c.lane, pzz[j]);
        if (c.hrow >= 0) {
            lu[c.hrow * LZ + R + c.lane] = yh[j];
            if constexpr (PSI) lps[c.hrow * LZ + R + c.lane] = sh[j];
            lpy[c.hrow * 64 + c.lane] = pyh[j];
        }
        lds_barrier();
        const int x = x0 + j;
        const bool inx = x < nx;
        const int xp = (inx ? x : nx - 1) * pl;
        T f = fv[j];
        if (born3(MODE)) f += mm[j] * (-dv[j]);
        if (MODE == 6) f += mm[j] * wv[j];
        const T lap = laplacian3<T, R, LZ>(P, q + j, lu, own, uc);
        const T c2 = cv[j] * cv[j];
        const T u_m = um[j];
        const bool edge = XC ? c.yzedge : ((x == 0) | (x == nx - 1) | c.yzedge);
        T dux = T(0), duy = T(0), duz = T(0), dpx = T(0), dpy = T(0), dpz = T(0);
        T dsx = T(0), dsy = T(0), dsz = T(0);
#pragma unroll
        for (int i = R; i >= 1; --i) {
            const T ax = P.cx.d1[i - 1], ay = P.cy.d1[i - 1], az = P.cz.d1[i - 1];
            dux = fma(ax, q[R + j + i] - q[R + j - i], dux);
            dpx = fma(ax, qx[R + j + i] - qx[R + j - i], dpx);
            duy = fma(ay, lu[own + i * LZ] - lu[own - i * LZ], duy);
            dpy = fma(ay, lpy[owny + i * 64] - lpy[owny - i * 64], dpy);
            duz = fma(az, lu[own + i] - lu[own - i], duz);
            dpz = fma(az, lpz[ownz + i] - lpz[ownz - i], dpz);
            if constexpr (PSI) {
                dsx = fma(ax, qs[R + j + i] - qs[R + j - i], dsx);
                dsy = fma(ay, lps[own + i * LZ] - lps[own - i * LZ], dsy);
                dsz = fma(az, lps[own + i] - lps[own - i], dsz);
            }
        }
        const T src = (((f + lap) + dpx) + dpy) + dpz;
        const T dt = P.tc.dt;
        T psn = PSI ? psc + dt * uc : T(0);
        const T pxc = qx[R + j], pycj = pyc[j], pzcj = pzc[j];
        T s1k, s2k, s1m, s2m, srcp, pxd, pyd, pzd, dws;
        bool pml;
        if constexpr (XC) {
            // sigma_x = 0: the form below with its zero terms dropped (the
            // same values; at most the sign of a zero differs), and every
            // factor of the lane's own profiles, the division included, the
            // same on every plane
            s1k = syk + szk;
            s2k = syk * szk;
            s1m = sym + szm;
            s2m = sym * szm;
            pml = syk != T(0) || szk != T(0);
            srcp = src;
            pxd = pxc + dt * (syk + szk) * dux;
            if constexpr (PSI) pxd = pxd + dt * (syk * szk * dsx);
            pyd = (pycj - dt * syk * pycj) + dt * (szk - syk) * duy;
            pzd = (pzcj - dt * szk * pzcj) + dt * (syk - szk) * duz;
            dws = T(0);
        } else {
            const T sxk = bcast3(vsig, j), sxm = bcast3(vsig, TX + j);
            s1k = (sxk + syk) + szk;
            s2k = (sxk * syk + syk * szk) + szk * sxk;
            const T s3k = (sxk * syk) * szk;
            pml = sxk != T(0) || syk != T(0) || szk != T(0);
            pxd = (pxc - dt * sxk * pxc) + dt * ((syk + szk) - sxk) * dux;
            pyd = (pycj - dt * syk * pycj) + dt * ((szk + sxk) - syk) * duy;
            pzd = (pzcj - dt * szk * pzcj) + dt * ((sxk + syk) - szk) * duz;
            s1m = (sxm + sym) + szm;
            s2m = (sxm * sym + sxm * szm) + sym * szm;
            if constexpr (PSI) {
                srcp = src - s3k * psc;
                pxd = pxd + dt * (syk * szk * dsx);
                pyd = pyd + dt * (szk * sxk * dsy);
                pzd = pzd + dt * (sxk * syk * dsz);
                const T s3m = (sxm * sym) * szm;
                dws = s3m * psc;
            } else {
                (void)s3k;
                srcp = src;
                dws = T(0);
            }
        }
        const T fac1 = P.tc.two_dt2 / (T(2) + dt * s1k);
        const T S = c2 * srcp;
        T pxn = pml ? pxd : pxc;
        T pyn = pml ? pyd : pycj;
        T pzn = pml ? pzd : pzcj;
        T up, dn = T(0), dw;
        if constexpr (INC) {
            // du_{k+1} = du_k + fac1 (S - s2 u - s1 du_k / dt): the reference's
            // damped update with u_{k-1} = u_k - du_k; undamped du_k + dt^2 S
            const T dnd = u_m + fac1 * ((S - s2k * uc) - s1k * u_m * P.tc.inv_dt);
            const T dni = fma(P.tc.dt2, c2 * src, u_m);
            dn = pml ? dnd : dni;
            up = uc + dn;
            dw = S + (s1m - s1k) * (dn + u_m) * P.tc.inv_2dt + (s2m - s2k) * uc;
        } else {
            const T fac2 = S - (u_m - T(2) * uc) * P.tc.inv_dt2 + s1k * u_m * P.tc.inv_2dt - s2k * uc;
            const T upd = fac1 * fac2;
            const T upi = fma(P.tc.dt2, c2 * src, T(2) * uc - u_m);
            up = pml ? upd : upi;
            dw = S + (s1m - s1k) * (up - u_m) * P.tc.inv_2dt + (s2m - s2k) * uc;
        }
        if constexpr (!XC && PSI) dw = dw + dws;
        if (edge) {
            up = T(0);
            dn = T(0);
            dw = T(0);
            pxn = T(0);
            pyn = T(0);
            pzn = T(0);
            psn = T(0);
        }
        const int oo = inx ? at(ov, xp) : kOOB;
        // Phi keeps its value (zero) where the node is not a PML node
        const int op = (c.gen || pml || edge) ? oo : kOOB;
        if constexpr (ic64_3(MODE))
            buf_st(c.ic, inx ? at(ovi, xp * (8 / SZ)) : kOOB, c.shb8, icv[j] + (double)uc * (double)dv[j]);
        if constexpr (MODE == 4) buf_st(c.ic, inx ? at(ovi, xp) : kOOB, c.shb, icv[j] + uc * dv[j]);
        buf_st(c.uo, oo, 0, up);
        if constexpr (INC) du_st3<T, MODE>(c.du, oo, dn);
        buf_st(c.pxo, op, 0, pxn);
        buf_st(c.pyo, op, 0, pyn);
        buf_st(c.pzo, op, 0, pzn);
        if constexpr (PSI) buf_st(c.pso, oo, 0, psn);
        if (MODE == 1) buf_st(c.dw, inx ? at(ovd, xp) : kOOB, c.shb, dw);
        if (out3(MODE)) buf_st(c.dwo, inx ? at(ovo, xp) : kOOB, c.shb, dw);
        if (c.has_fst) buf_st(c.fst, inx ? at(ovf, xp) : kOOB, c.shb, uc);
    }
#undef PA_Q
#undef PA_ROW
}

template <typename T, int R, int MODE, bool INC, bool FST>
__device__ __forceinline__ void tile3d(const pysit_amd_batch &B, const Step3DParams<T> &P,
                                       const pysit_amd_shot &S, int tile, int full, int xclear, int psi, int tloc,
                                       int slot,
                                       int seg_base, Lds3D<T, R> &L)
{
    constexpr int SZ = sizeof(T), TX = TX3, LZ = Lds3D<T, R>::LZ;
    const int k = slot - S.kshift;
    if (k < 0 || k >= S.nt) return;
    Ctx3D<T, R> c;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    c.nx = nx;
    c.ny = ny;
    c.nz = nz;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    c.lane = lane;
    c.w = w;
    // tiles: z fastest, then y, then x; z tiles as in 2D (the last ends on nz-1)
    const int ntz = S.ntz;
    const int tz = tloc % ntz, tyi = (tloc / ntz) % S.nty, tx = tloc / (ntz * S.nty);
    const int zlast = z_last_tile(nz, z_align<T>());
    const int zs = S.zs;
    const int z0 = (tz == ntz - 1) ? zlast : tz * 64;
    const int own_hi = (tz == ntz - 1) ? nz : min(z0 + 64, zlast);
    constexpr int TY3 = Tile3<T>::TY;
    const int y0 = tyi * TY3, x0 = tx * TX;
    const int z = z0 + lane, y = y0 + w;
    c.y = y;
    const bool zin = z < nz, yin = y < ny;
    const bool own = zin && z < own_hi && yin;
    c.plane = ny * zs * SZ;
    c.co = (zin && yin) ? (y * zs + z) * SZ : kOOB;
    c.ov = own ? (y * zs + z) * SZ : kOOB;
    c.ov8 = own ? (y * zs + z) * 8 : kOOB;
    const int hz = lane < R ? z0 + 64 + lane : (lane >= 64 - R ? z0 + lane - 64 : -1);
    const bool hin = hz >= 0 && hz < nz;
    c.cho = (hin && yin) ? (y * zs + hz) * SZ : kOOB;
    // LDS column of the z-halo value, own row in the u/psi planes and the Phiz plane
    const int hcol = lane < R ? R + 64 + lane : (lane >= 64 - R ? lane - (64 - R) : -1);
    c.hl = hcol >= 0 ? (R + w) * LZ + hcol : -1;
    const int hl_pz = hcol >= 0 ? w * LZ + hcol : -1;
    // y-halo rows: waves 0..r-1 above the tile, waves r..2r-1 below it
    c.hrow = w < R ? w : (w < 2 * R ? TY3 + w : -1);
    const int hy = y0 - R + c.hrow;
    c.hro = (c.hrow >= 0 && hy >= 0 && hy < ny && zin) ? (hy * zs + z) * SZ : kOOB;
    c.yzedge = y == 0 || y >= ny - 1 || z == 0 || z >= nz - 1;
    {
        auto yb = [&](int yy) { return yy < S.nyl || yy >= ny - S.nyr; };
        auto zb = [&](int zz) { return zz < S.nzl || zz >= nz - S.nzr; };
        // the batch steps (INC) never take the reference-contract step's
        // general Phi input or dense rhs: compile-time false there, which
        // frees their scalar registers
        c.gen = !INC && P.single != 0;
        c.bp = yb(y) || zb(z);
        c.bph = yb(hy) || zb(z);
        c.bpz = yb(y) || zb(hz);
        c.nxl = S.nxl;
        c.nxr = S.nxr;
    }
    const int taken = adj3(MODE) ? (S.nt - 1 - k) : k;
    const int cur = taken & 1;
    const long long off = S.off;
    const long long nb = (long long)nx * ny * zs * SZ;
    const T *vb = (const T *)B.u[0];
    c.u = make_rsrc((const T *)B.u[cur] + off, nb);
    c.uo = make_rsrc((T *)B.u[cur ^ 1] + off, nb);
    c.du = make_rsrc(INC ? (const void *)((T *)B.du + off) : (const void *)vb, INC ? nb : 0);
    c.px = make_rsrc((const T *)B.phix[cur] + off, full ? nb : 0);
    c.pxo = make_rsrc((T *)B.phix[cur ^ 1] + off, full ? nb : 0);
    c.py = make_rsrc((const T *)B.phiy[cur] + off, full ? nb : 0);
    c.pyo = make_rsrc((T *)B.phiy[cur ^ 1] + off, full ? nb : 0);
    c.pz = make_rsrc((const T *)B.phiz[cur] + off, full ? nb : 0);
    c.pzo = make_rsrc((T *)B.phiz[cur ^ 1] + off, full ? nb : 0);
    c.ps = make_rsrc((const T *)B.psi[cur] + off, full ? nb : 0);
    c.pso = make_rsrc((T *)B.psi[cur ^ 1] + off, full ? nb : 0);
    c.model = make_rsrc(S.model, (long long)(S.mx_hi + 1) * S.mny * S.mnz * SZ);
    c.mv = (clampi(y - S.my_off, S.my_lo, S.my_hi) * S.mnz + clampi(z - S.mz_off, S.mz_lo, S.mz_hi)) * SZ;
    c.mplane = S.mny * S.mnz * SZ;
    c.mx_off = S.mx_off;
    c.mx_lo = S.mx_lo;
    c.mx_hi = S.mx_hi;
    c.has_rhs = !INC && P.rhs != nullptr;
    c.rhs = make_rsrc(c.has_rhs ? (const void *)P.rhs : vb, c.has_rhs ? nb : 0);
    c.shb = (int)(off * SZ);
    c.shb8 = (int)(off * 8);
    const bool img = adj3(MODE) && (k % B.imaging_period == 0) && B.dw;
    const bool store = (MODE == 1) && (k % B.imaging_period == 0) && B.dw;
    const long long slot_el = (long long)(slot - seg_base) * B.total_n;
    const long long slot_bytes = B.total_n * SZ;
    const bool dwon = ((img || born3(MODE)) && B.dw) || store;
    c.dw = make_rsrc(dwon ? (const void *)((const T *)B.dw + slot_el) : vb, dwon ? slot_bytes : 0);
    // FST: the launch stores u_k of every step (B.fstore); kernels built
    // without it hold no resource for that store
    c.has_fst = FST && B.fstore != nullptr;
    c.fst = make_rsrc(c.has_fst ? (const void *)((const T *)B.fstore + slot_el) : vb, c.has_fst ? slot_bytes : 0);
    if constexpr (MODE == 4) c.ic = make_rsrc(B.ic32, B.total_n * 4);
    else c.ic = make_rsrc(ic64_3(MODE) ? (const void *)B.ic : vb, ic64_3(MODE) ? B.total_n * 8 : 0);
    c.m1 = make_rsrc(born3(MODE) ? B.born_m1 : vb,
                     born3(MODE) ? (long long)(S.mx_hi + 1) * S.mny * S.mnz * SZ : 0);
    c.ovd = dwon ? c.ov : kOOB;
    c.ovi = img ? (MODE == 4 ? c.ov : c.ov8) : kOOB;
    c.ovf = c.has_fst ? c.ov : kOOB;
    c.ovo = kOOB;
    c.ovv = kOOB;
    if constexpr (out3(MODE)) {
        const bool out = B.dw_out != nullptr;
        c.dwo = make_rsrc(out ? (const void *)((T *)B.dw_out + slot_el) : vb, out ? slot_bytes : 0);
        c.ovo = out ? c.ov : kOOB;
    }
    if constexpr (MODE == 6) {
        const bool vol = B.vol_m && B.vol_dw;
        c.volm = make_rsrc(vol ? B.vol_m : vb, vol ? slot_bytes : 0);
        c.vold = make_rsrc(vol ? (const void *)((const T *)B.vol_dw + slot_el) : vb, vol ? slot_bytes : 0);
        c.ovv = vol ? c.ov : kOOB;
    }
    // PML profiles: lane y/z values and the tile planes' x profile (lane p:
    // kernel form of plane x0+p, lane TX+p: mesh form), masked buffer loads
    static_assert(2 * TX <= 64, "the x profile of a tile must fit one wavefront");
    {
        const Rsrc rxl = make_rsrc(S.nxl ? S.xl : vb, (long long)S.nxl * SZ);
        const Rsrc rxr = make_rsrc(S.nxr ? S.xr : vb, (long long)S.nxr * SZ);
        const Rsrc ryl = make_rsrc(S.nyl ? S.yl : vb, (long long)S.nyl * SZ);
        const Rsrc ryr = make_rsrc(S.nyr ? S.yr : vb, (long long)S.nyr * SZ);
        const Rsrc rzl = make_rsrc(S.nzl ? S.zl : vb, (long long)S.nzl * SZ);
        const Rsrc rzr = make_rsrc(S.nzr ? S.zr : vb, (long long)S.nzr * SZ);
        const bool f = full && zin && yin;
        c.syk = sigma3<T>(ryl, S.nyl, ryr, S.nyr, f && y <= ny - 2 ? y : -1, ny, true);
        c.sym = sigma3<T>(ryl, S.nyl, ryr, S.nyr, f ? y : -1, ny, false);
        c.szk = sigma3<T>(rzl, S.nzl, rzr, S.nzr, f && z <= nz - 2 ? z : -1, nz, true);
        c.szm = sigma3<T>(rzl, S.nzl, rzr, S.nzr, f ? z : -1, nz, false);
        const int xs = min(x0 + (lane < TX ? lane : lane - TX), nx - 1);
        const int xi = (full && lane < 2 * TX) ? (lane < TX ? max(0, min(xs, nx - 2)) : xs) : -1;
        c.vsig = sigma3<T>(rxl, S.nxl, rxr, S.nxr, xi, nx, lane < TX);
    }
    // sources of this tile gathered into LDS (plane, row, lane): one thread
    // per destination node, entries summed in the reference's order; without
    // sources every plane reads a zero slot
    c.finj = &L.junk[3][0];
    c.fstride = 0;
    bool has_inj = false;
    if (B.inj_ptr) {
        const int n0 = B.inj_ptr[tile], n1 = B.inj_ptr[tile + 1];
        if (n1 > n0) has_inj = true;
        if (has_inj) {
            const double *amp_k = S.amp + (long long)k * S.namp;
            for (int i = threadIdx.x; i < TX3 * TY3 * 64; i += TY3 * 64) L.f[i] = T(0);
            lds_barrier();
            if (B.inj_pre) {
                // values evaluated for the whole run before its first step
                const T *pre = (const T *)B.inj_pre + (long long)(slot - seg_base) * B.n_inj_nodes;
                for (int n = n0 + (int)threadIdx.x; n < n1; n += TY3 * 64) L.f[B.inj_node[n]] = pre[n];
            } else {
                for (int n = n0 + (int)threadIdx.x; n < n1; n += TY3 * 64) {
                    const int dst = B.inj_node[n];
                    L.f[dst] = (T)inj_sum(B.inj_chan, B.inj_w, amp_k, B.inj_nptr[n], B.inj_nptr[n + 1]);
                }
            }
            c.finj = &L.f[w * 64 + lane];
            c.fstride = TY3 * 64;
        }
    }
    if (!has_inj && threadIdx.x == 0) L.junk[3][0] = T(0);
    lds_barrier();
    if (full && xclear && psi) march3_full<T, R, MODE, true, INC, true>(c, P, L, x0, hl_pz);
    else if (full && xclear) march3_full<T, R, MODE, true, INC, false>(c, P, L, x0, hl_pz);
    else if (full && psi) march3_full<T, R, MODE, false, INC, true>(c, P, L, x0, hl_pz);
    else if (full) march3_full<T, R, MODE, false, INC, false>(c, P, L, x0, hl_pz);
    else march3_int<T, R, MODE, INC>(c, P, L, x0);
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
        acc = rec_gather<T>(B.rec_idx, B.rec_w, u, rd.z, rd.w, g);
    }
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 16);
    if (live && g == 0) S->trace[(long long)k * S->nrec + rd.y] = acc;
}

// at most 128 VGPRs (FP32): two 8-wave workgroups per CU
#ifndef PA3D_WAVES_PER_EU
#define PA3D_WAVES_PER_EU 4
#endif
// INC: the batch's increment form (du); false: the reference-contract step
// (u_{k-1} in the second u buffer)
template <typename T, int R, int MODE, bool INC, bool FST>
__global__ __launch_bounds__(Tile3<T>::TY * 64)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? PA3D_WAVES_PER_EU : 2)))
void k_step3d(const pysit_amd_batch B, const Step3DParams<T> P,
                                                int slot, int seg_base, int n_tile_blocks,
                                                int n_rec_lanes, int rec_blocks, const int *cursor)
{
    __shared__ Lds3D<T, R> L;
    // with a step cursor (graph launches): slot offset from cursor[0], store base cursor[1]
    if (cursor) {
        slot += cursor[0];
        seg_base = cursor[1];
    }
    // the tile blocks, then the receiver blocks (they fill the launch's tail)
    constexpr int TY3 = Tile3<T>::TY;
    const int b = blockIdx.x;
    const int tile_blocks = gridDim.x - rec_blocks;
    if (!adj3(MODE) && b >= tile_blocks) {
        const int rec = (b - tile_blocks) * (TY3 * 4) + (threadIdx.x >> 4);
        sample_group3<T>(B, slot, rec < n_rec_lanes ? rec : B.n_rec, threadIdx.x & 15);
        return;
    }
    const int blk = xcd_block(b, tile_blocks);
    if (blk >= n_tile_blocks) return;
    if (!INC) {
        tile3d<T, R, MODE, INC, FST>(B, P, P.shot0, blk, 1, 0, 1, blk, slot, seg_base, L);
    } else {
        const int4 td = reinterpret_cast<const int4 *>(B.tiles)[blk];
        // flags bit2: the tile and its reach touch at most one PML band (no psi)
        const int psi = PA3D_PSI_SKIP ? ((td.z >> 2) & 1) ^ 1 : 1;
        tile3d<T, R, MODE, INC, FST>(B, P, B.shots[td.x], blk, td.z & 1, (td.z >> 1) & 1, psi, td.y, slot, seg_base,
                                     L);
    }
}

template <typename T, int R, int MODE>
static int launch3d(const pysit_amd_batch &B, const Step3DParams<T> &P, int slot, int seg_base,
                    int n_tiles, int n_recs, hipStream_t st, const int *cursor = nullptr)
{
    constexpr int TY3 = Tile3<T>::TY;
    const int recs = adj3(MODE) ? 0 : n_recs;
    const int rec_blocks = recs ? (recs + TY3 * 4 - 1) / (TY3 * 4) : 0;
    const int blocks = n_tiles + rec_blocks;
    if (blocks == 0) return PYSIT_AMD_OK;
    if (P.single) {
        if constexpr (MODE == 0) {
            hipLaunchKernelGGL((k_step3d<T, R, MODE, false, false>), dim3(blocks), dim3(TY3 * 64), 0, st, B, P,
                               slot, seg_base, n_tiles, recs, rec_blocks, cursor);
        } else {
            set_error("the single-shot step only propagates");
            return PYSIT_AMD_EINVAL;
        }
    } else if (B.fstore) {
        hipLaunchKernelGGL((k_step3d<T, R, MODE, true, true>), dim3(blocks), dim3(TY3 * 64), 0, st, B, P, slot,
                           seg_base, n_tiles, recs, rec_blocks, cursor);
    } else {
        hipLaunchKernelGGL((k_step3d<T, R, MODE, true, false>), dim3(blocks), dim3(TY3 * 64), 0, st, B, P, slot,
                           seg_base, n_tiles, recs, rec_blocks, cursor);
    }
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

template <typename T, int MODE>
static int launch3d_r(int r, const pysit_amd_batch &B, const Step3DParams<T> &P, int slot,
                      int seg_base, int n_tiles, int n_recs, hipStream_t st, const int *cursor = nullptr)
{
    switch (r) {
    case 1: return launch3d<T, 1, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case 2: return launch3d<T, 2, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case 3: return launch3d<T, 3, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    case 4: return launch3d<T, 4, MODE>(B, P, slot, seg_base, n_tiles, n_recs, st, cursor);
    }
    set_error("order must be 2, 4, 6 or 8");
    return PYSIT_AMD_EINVAL;
}

int ic_flush(const pysit_amd_batch &B, hipStream_t st);

template <typename T>
struct Sink3D final : StepSink {
    const pysit_amd_batch &B;
    const Step3DParams<T> &P;
    int r, mode;
    bool window;
    Sink3D(const pysit_amd_batch &B_, const Step3DParams<T> &P_, int r_, int mode_, bool window_)
        : B(B_), P(P_), r(r_), mode(mode_), window(window_) {}
    int step(int slot, int seg_base, int nt, int nr, const int *cursor, hipStream_t st) override
    {
        const bool second = B.dw_out || B.vol_dw;
        switch (mode) {
        case 0: return launch3d_r<T, 0>(r, B, P, slot, seg_base, nt, nr, st, cursor);
        case 1: return launch3d_r<T, 1>(r, B, P, slot, seg_base, nt, nr, st, cursor);
        case 2:
            if (second) return launch3d_r<T, 6>(r, B, P, slot, seg_base, nt, 0, st, cursor);
            if (!window) return launch3d_r<T, 2>(r, B, P, slot, seg_base, nt, 0, st, cursor);
            if constexpr (std::is_same<T, float>::value)
                return launch3d_r<T, 4>(r, B, P, slot, seg_base, nt, 0, st, cursor);
            set_error("the FP32 window accumulator needs an FP32 batch");
            return PYSIT_AMD_EINVAL;
        default:
            if (second) return launch3d_r<T, 5>(r, B, P, slot, seg_base, nt, nr, st, cursor);
            return launch3d_r<T, 3>(r, B, P, slot, seg_base, nt, nr, st, cursor);
        }
    }
    int flush(hipStream_t st) override { return ic_flush(B, st); }
};

template <typename T>
int run3d(const pysit_amd_batch &B, int mode, int slot_begin, int slot_end, int direction,
          int seg_base, const int *active_tiles, const int *active_recs, pysit_amd_plan *plan,
          hipStream_t st)
{
    constexpr int TY3 = Tile3<T>::TY;
    PA_CHECK(B.tile_x == TX3 && B.tile_y == TY3,
             "3D tiles are " + std::to_string(TX3) + " x " + std::to_string(TY3) + " (tile_x, tile_y)");
    // every per-batch array (the FP64 accumulator at 8 B per node the
    // largest) is reached through 32-bit buffer offsets
    PA_CHECK(B.total_n * 8 <= kMaxRsrcBytes, "batch too large for one launch (total_n)");
    PA_CHECK(B.u[0] && B.u[1] && B.du && B.phix[0] && B.phix[1] && B.phiy[0] && B.phiy[1] && B.phiz[0] &&
                 B.phiz[1] && B.psi[0] && B.psi[1],
             "3D batch state buffers (u, phix, phiy, phiz, psi, du) must all be given");
    Step3DParams<T> P{};
    const int r = B.order / 2;
    P.cx = axis_coef<T>(r, B.dx);
    P.cy = axis_coef<T>(r, B.dy);
    P.cz = axis_coef<T>(r, B.dz);
    P.tc = time_coef<T>(B.dt);
    P.rhs = nullptr;
    P.single = 0;
    const bool window = mode == 2 && B.ic32 != nullptr;
    Sink3D<T> S(B, P, r, mode, window);
    const LoopSpec L{mode, slot_begin, slot_end, direction, seg_base, active_tiles, active_recs,
                     window ? B.ic_window : 0};
    return time_loop(plan, B, L, S, st);
}

template int run3d<float>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                          const int *, pysit_amd_plan *, hipStream_t);
template int run3d<double>(const pysit_amd_batch &, int, int, int, int, int, const int *,
                           const int *, pysit_amd_plan *, hipStream_t);

template <typename T>
int step3d_entry(int order, const T *km1_u, const T *k_phix, const T *k_phiy, const T *k_phiz,
                 const T *k_psi, const T *k_u, const T *C, const T *rhs, const T *xl, int nxl,
                 const T *xr, int nxr, const T *yl, int nyl, const T *yr, int nyr, const T *zl,
                 int nzl, const T *zr, int nzr, double dt, double dx, double dy, double dz, int nx,
                 int ny, int nz, T *kp1_phix, T *kp1_phiy, T *kp1_phiz, T *kp1_psi, T *kp1_u,
                 hipStream_t st)
{
    constexpr int TY3 = Tile3<T>::TY;
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
    PA_CHECK(N * (long long)sizeof(T) <= kMaxRsrcBytes, "grid too large for one step launch");
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
    S.zs = nz;
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

// common.h -- shared device helpers and host-side error plumbing for the
// pysit_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/pysit_amd.h"

namespace pa {

// ---- errors (thread-local message behind pysit_amd_last_error) ----------
void set_error(const std::string &msg);
void clear_error();

#define PA_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::pa::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
            return PYSIT_AMD_EDEVICE;                                             \
        }                                                                         \
    } while (0)

#define PA_CHECK(cond, msg)                                                       \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ::pa::set_error(msg);                                                 \
            return PYSIT_AMD_EINVAL;                                              \
        }                                                                         \
    } while (0)

// ---- centred finite-difference weights, accuracy 2r (r = 1..4) ----------
// first derivative at offset +j (offset -j has the opposite sign)
__host__ __device__ constexpr double fd1(int r, int j)
{
    return r == 1 ? (j == 1 ? 1. / 2. : 0.)
         : r == 2 ? (j == 1 ? 2. / 3. : j == 2 ? -1. / 12. : 0.)
         : r == 3 ? (j == 1 ? 3. / 4. : j == 2 ? -3. / 20. : j == 3 ? 1. / 60. : 0.)
         : (j == 1 ? 4. / 5. : j == 2 ? -1. / 5. : j == 3 ? 4. / 105. : j == 4 ? -1. / 280. : 0.);
}
// second derivative at offset +-j (j = 0 is the centre)
__host__ __device__ constexpr double fd2(int r, int j)
{
    return r == 1 ? (j == 0 ? -2.0 : j == 1 ? 1.0 : 0.)
         : r == 2 ? (j == 0 ? -5. / 2. : j == 1 ? 4. / 3. : j == 2 ? -1. / 12. : 0.)
         : r == 3 ? (j == 0 ? -49. / 18. : j == 1 ? 3. / 2. : j == 2 ? -3. / 20. : j == 3 ? 1. / 90. : 0.)
         : (j == 0 ? -205. / 72. : j == 1 ? 8. / 5. : j == 2 ? -1. / 5. : j == 3 ? 8. / 315. : j == 4 ? -1. / 560. : 0.);
}

// Per-axis stencil coefficients with the grid spacing folded in.
template <typename T>
struct AxisCoef {
    T d1[4];  // fd1(r, j+1) / h
    T d2[5];  // fd2(r, j) / h^2
};

template <typename T>
inline AxisCoef<T> axis_coef(int r, double h)
{
    AxisCoef<T> c{};
    for (int j = 0; j < 4; ++j) c.d1[j] = (T)(fd1(r, j + 1) / h);
    for (int j = 0; j < 5; ++j) c.d2[j] = (T)(fd2(r, j) / (h * h));
    return c;
}

template <typename T>
struct TimeCoef {
    T dt, dt2, inv_dt2, inv_2dt, two_dt2, inv_dt;
};

template <typename T>
inline TimeCoef<T> time_coef(double dt)
{
    TimeCoef<T> c;
    c.dt = (T)dt;
    c.dt2 = (T)(dt * dt);
    c.inv_dt2 = (T)(1.0 / (dt * dt));
    c.inv_2dt = (T)(1.0 / (2.0 * dt));
    c.two_dt2 = (T)(2.0 * dt * dt);
    c.inv_dt = (T)(1.0 / dt);
    return c;
}

// Block index remap so that the blocks the dispatcher sends to one XCD
// (round robin over 8 XCDs) cover one contiguous range of logical blocks:
// neighbouring tiles (and the shots of one tile) then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    constexpr int NXCD = 8;
    const int xcd = b % NXCD, q = b / NXCD;
    const int full = nb / NXCD, rem = nb % NXCD;
    return (xcd < rem) ? xcd * (full + 1) + q : rem * (full + 1) + (xcd - rem) * full + q;
}

// LDS exchange inside one wavefront: order the LDS stores of all lanes before
// the loads that follow (compiler barrier plus wavefront-scope fences).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// z tiles of an axis of nz nodes (2D and 3D): 64 nodes each, starting at 0,
// 64, 128, ... except the last, which starts at the first multiple of
// `align` nodes at or after nz - 64 (so it still reaches node nz - 1, and
// with rows stored at an aligned stride its accesses start on a cache line);
// each tile owns the nodes up to the next tile's start.  The host builds the
// same tiles (pysit_amd/device.py z_tiles).
__host__ __device__ constexpr int z_last_tile(int nz, int align)
{
    return nz > 64 ? ((nz - 64 + align - 1) / align) * align : 0;
}
// Receiver gather of one 16-lane group: lane g sums the entries g, g+16,
// g+32, ... of [e0, e1) in that order, w[e] * u[idx[e]] in FP64.  The index
// and weight loads of up to four entries go out before their field loads,
// so a lane waits for two rounds of loads, not two per entry.
template <typename T>
__device__ __forceinline__ double rec_gather(const long long *__restrict__ idx, const double *__restrict__ w,
                                             const T *__restrict__ u, int e0, int e1, int g)
{
    constexpr int U = 4;
    double acc = 0.0;
    for (int base = e0 + g; base < e1; base += 16 * U) {
        long long ix[U];
        double wv[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = base + 16 * q;
            ix[q] = e < e1 ? idx[e] : 0;
            wv[q] = e < e1 ? w[e] : 0.0;
        }
        T uv[U];
#pragma unroll
        for (int q = 0; q < U; ++q) uv[q] = u[ix[q]];
#pragma unroll
        for (int q = 0; q < U; ++q)
            if (base + 16 * q < e1) acc += wv[q] * (double)uv[q];
    }
    return acc;
}

// Injection value of one destination node: the sum over its entries
// [e0, e1) of w[e] * amp_k[chan[e]] in entry order (FP64, each product
// rounded before it is added).  The channel and weight loads of up to
// kInjU entries go out together, then their amplitude loads, so a node
// waits for two rounds of loads per kInjU entries instead of two per entry
// (a Gaussian receiver spreads over 7 nodes per axis: 7 entries per node).
constexpr int kInjU = 8;
__device__ __forceinline__ double inj_sum(const int *__restrict__ chan, const double *__restrict__ w,
                                          const double *__restrict__ amp_k, int e0, int e1)
{
#pragma clang fp contract(off)
    double acc = 0.0;
    for (int e = e0; e < e1; e += kInjU) {
        int ch[kInjU];
        double wv[kInjU], av[kInjU];
#pragma unroll
        for (int q = 0; q < kInjU; ++q) {
            const bool ok = e + q < e1;
            ch[q] = ok ? chan[e + q] : 0;
            wv[q] = ok ? w[e + q] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kInjU; ++q) av[q] = e + q < e1 ? amp_k[ch[q]] : 0.0;
#pragma unroll
        for (int q = 0; q < kInjU; ++q)
            if (e + q < e1) acc += wv[q] * av[q];
    }
    return acc;
}

// row alignment in nodes: one 128-byte cache line
template <typename T> constexpr int z_align() { return 128 / (int)sizeof(T); }

// ---- raw buffer access -----------------------------------------------------
// A buffer resource covers `bytes` from `base`.  An access names a per-lane
// byte offset (VGPR) and a wave-uniform byte offset (SGPR, e.g. the row and
// field); a lane whose per-lane offset is kOOB lies beyond every resource, so
// its load returns 0 and its store is dropped by the hardware range check:
// masked lanes need no branches.  Resources are wave-uniform (SGPRs).
// Resource sizes stay below kOOB (kMaxRsrcBytes); the host checks that every
// region it hands to a kernel fits.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kOOB = 0x7ffffff0;
constexpr long long kMaxRsrcBytes = 0x7fff0000LL;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer, 32-bit data format

__device__ __forceinline__ Rsrc make_rsrc(const void *base, long long bytes)
{
    // the size is wave-uniform; the clamp to [0, kMaxRsrcBytes] works on the
    // two 32-bit halves (scalar instructions), and readfirstlane keeps the
    // result (and so the whole resource) in SGPRs, where buffer instructions
    // need it
    const int hi = (int)(bytes >> 32);
    const unsigned lo = (unsigned)bytes;
    const int n = __builtin_amdgcn_readfirstlane(
        hi < 0 ? 0 : (hi > 0 || lo > (unsigned)kMaxRsrcBytes ? (int)kMaxRsrcBytes : (int)lo));
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, n, kRsrcWord3);
}

template <typename T> __device__ __forceinline__ T buf_ld(Rsrc r, int voff, int soff);
template <> __device__ __forceinline__ float buf_ld<float>(Rsrc r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <> __device__ __forceinline__ double buf_ld<double>(Rsrc r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_st(Rsrc r, int voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
}
__device__ __forceinline__ void buf_st(Rsrc r, int voff, int soff, double v)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
}

// Streaming (non-temporal) forms for data that is not reused soon, such as
// the dWaveOp history: the caches keep the fields that are
template <typename T> __device__ __forceinline__ T buf_ld_nt(Rsrc r, int voff, int soff);
template <> __device__ __forceinline__ float buf_ld_nt<float>(Rsrc r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}
template <> __device__ __forceinline__ double buf_ld_nt<double>(Rsrc r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}
__device__ __forceinline__ void buf_st_nt(Rsrc r, int voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 2);
}
__device__ __forceinline__ void buf_st_nt(Rsrc r, int voff, int soff, double v)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 2);
}

// Kernel-side PML profile value at node i of an axis of n nodes (reference
// indexing: left array by node, right array offset by one towards the edge;
// the edge node itself is never updated, callers pass i <= n-2).
template <typename T>
__device__ __forceinline__ T sigma_kernel(const T *l, int nl, const T *r, int nr, int i, int n)
{
    if (nl > 0 && i < nl) return l[i];
    if (nr > 0 && i >= n - nr && i <= n - 2) return r[nr - ((n - 1) - i)];
    return T(0);
}
// Mesh PML profile (StructuredPML.eval_on_mesh + build_sigma): the right
// array is laid on the last nr nodes without the shift.
template <typename T>
__device__ __forceinline__ T sigma_mesh(const T *l, int nl, const T *r, int nr, int i, int n)
{
    if (nl > 0 && i < nl) return l[i];
    if (nr > 0 && i >= n - nr) return r[i - (n - nr)];
    return T(0);
}

// ---- batch time loop -------------------------------------------------------
// One dimension's step launcher as the loop driver sees it: `step` launches
// the step of `slot` (with a cursor: the offset from the cursor's slot),
// `flush` the FP32 window accumulator's flush.
struct StepSink {
    virtual int step(int slot, int seg_base, int n_tiles, int n_recs, const int *cursor, hipStream_t st) = 0;
    virtual int flush(hipStream_t st) = 0;
    virtual ~StepSink() {}
};

struct LoopSpec {
    int mode, slot_begin, slot_end, direction, seg_base;
    const int *active_tiles, *active_recs;
    int window;  // FP32 imaging window: flush after each slot that ends one (0: never)
};

// The FP32 imaging windows follow the batch's slots, not the calls: a window
// ends after slot j when j % w == 0 in a backward run ((j + 1) % w == 0 in a
// forward one), so a sweep split into calls (checkpoint segments, binomial
// replays) groups its sums exactly as one call does.  A backward sweep ends
// on slot 0, which ends a window.
__host__ __device__ inline bool window_end(int j, int direction, int w)
{
    return direction < 0 ? j % w == 0 : (j + 1) % w == 0;
}

// Runs the steps of `L` on `st`: one launch per step, or, with a plan, in
// graph chunks of the plan's step count (see plan.hip).
int time_loop(pysit_amd_plan *plan, const pysit_amd_batch &B, const LoopSpec &L, StepSink &S, hipStream_t st);

}  // namespace pa

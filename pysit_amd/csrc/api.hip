// api.hip -- the extern "C" surface of libpysit_amd (declared in
// include/pysit_amd.h): argument checks, dispatch on dimension and dtype,
// error strings, and the PML fold kernel.
#include "common.h"

#include <cstdlib>
#include <mutex>

namespace pa {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
void clear_error() { g_err.clear(); }

template <typename T>
int run2d(const pysit_amd_batch &, int, int, int, int, int, const int *, const int *, pysit_amd_plan *,
          hipStream_t);
template <typename T>
int run3d(const pysit_amd_batch &, int, int, int, int, int, const int *, const int *, pysit_amd_plan *,
          hipStream_t);
template <typename T>
int run2d_persistent(const pysit_amd_batch &, int, int, int, int, int, const int *, const int *, long long,
                     hipStream_t);

int ic_flush(const pysit_amd_batch &B, hipStream_t st);
int tile_rows_2d(int dtype);
bool tile_rows_2d_supported(int tx);
int tile_rows_3d(int dtype, int axis);

template <typename T>
int step2d_entry(int, const T *, const T *, const T *, const T *, const T *, const T *, const T *,
                 int, const T *, int, const T *, int, const T *, int, double, double, double, int,
                 int, T *, T *, T *, hipStream_t);
template <typename T>
int step3d_entry(int, const T *, const T *, const T *, const T *, const T *, const T *, const T *,
                 const T *, const T *, int, const T *, int, const T *, int, const T *, int,
                 const T *, int, const T *, int, double, double, double, double, int, int, int, T *,
                 T *, T *, T *, T *, hipStream_t);

// ---- fold (add_padding) ---------------------------------------------------
// One thread per unpadded output node; the PML slabs that fold onto it are
// summed x first, then y, then z, as the reference's successive passes do.
__global__ void k_fold(int nx, int ny, int nz, int lx, int rx, int ly, int ry, int lz, int rz,
                       const double *__restrict__ in, double scale, double *__restrict__ out,
                       int out_ny, int out_nz, int accumulate)
{
    const int ox = nx - lx - rx, oy = ny - ly - ry, oz = nz - lz - rz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ox * oy * oz) return;
    const int ko = (int)(t % oz);
    const int jo = (int)((t / oz) % oy);
    const int io = (int)(t / ((long long)oz * oy));
    auto A = [&](int i, int j, int k) { return in[((long long)i * ny + j) * nz + k]; };
    auto X = [&](int j, int k) {
        double v = A(io + lx, j, k);
        if (io == 0) {
            double s = 0.0;
            for (int i = 0; i < lx; ++i) s += A(i, j, k);
            v = v + s;
        }
        if (io == ox - 1) {
            double s = 0.0;
            for (int i = lx + ox; i < nx; ++i) s += A(i, j, k);
            v = v + s;
        }
        return v;
    };
    auto Y = [&](int k) {
        double v = X(jo + ly, k);
        if (jo == 0) {
            double s = 0.0;
            for (int j = 0; j < ly; ++j) s += X(j, k);
            v = v + s;
        }
        if (jo == oy - 1) {
            double s = 0.0;
            for (int j = ly + oy; j < ny; ++j) s += X(j, k);
            v = v + s;
        }
        return v;
    };
    double v = Y(ko + lz);
    if (ko == 0) {
        double s = 0.0;
        for (int k = 0; k < lz; ++k) s += Y(k);
        v = v + s;
    }
    if (ko == oz - 1) {
        double s = 0.0;
        for (int k = lz + oz; k < nz; ++k) s += Y(k);
        v = v + s;
    }
    double *o = out + ((long long)io * out_ny + jo) * out_nz + ko;
    *o = accumulate ? *o + scale * v : scale * v;
}

// ---- precomputed injection -------------------------------------------------
// Every injection node's value for slots [slot_begin, slot_end): one thread
// per node and kInjPreSlots slots, the value inj_sum gives (zero where the
// node's shot is inactive), stored at (slot - seg_base) * n_inj_nodes + n
// (consecutive threads, consecutive nodes).  A node of at most kInjU entries
// keeps its channels and weights in registers and loads the amplitudes of
// kInjPreU slots together; the sum runs in entry order as in inj_sum.
constexpr int kInjPreSlots = 32;
constexpr int kInjPreU = 4;
template <typename T>
__global__ void k_inj_pre(const pysit_amd_batch B, int slot_begin, int slot_end, int seg_base)
{
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= B.n_inj_nodes) return;
    const int s0 = slot_begin + (int)blockIdx.y * kInjPreSlots;
    const int s1 = min(slot_end, s0 + kInjPreSlots);
    const pysit_amd_shot &S = B.shots[B.inj_node_shot[n]];
    const int e0 = B.inj_nptr[n], e1 = B.inj_nptr[n + 1];
    const int ne = e1 - e0;
    T *out = (T *)B.inj_pre + n;
    const long long stride = B.n_inj_nodes;
    if (ne > kInjU) {
        for (int slot = s0; slot < s1; ++slot) {
            const int k = slot - S.kshift;
            const double v =
                (k >= 0 && k < S.nt) ? inj_sum(B.inj_chan, B.inj_w, S.amp + (long long)k * S.namp, e0, e1) : 0.0;
            out[(slot - seg_base) * stride] = (T)v;
        }
        return;
    }
    int ch[kInjU];
    double wv[kInjU];
#pragma unroll
    for (int q = 0; q < kInjU; ++q) {
        ch[q] = q < ne ? B.inj_chan[e0 + q] : 0;
        wv[q] = q < ne ? B.inj_w[e0 + q] : 0.0;
    }
    for (int slot = s0; slot < s1; slot += kInjPreU) {
        double av[kInjPreU][kInjU];
        bool act[kInjPreU];
#pragma unroll
        for (int j = 0; j < kInjPreU; ++j) {
            const int k = slot + j - S.kshift;
            act[j] = slot + j < s1 && k >= 0 && k < S.nt;
            const double *amp_k = S.amp + (long long)(act[j] ? k : 0) * S.namp;
#pragma unroll
            for (int q = 0; q < kInjU; ++q) av[j][q] = (act[j] && q < ne) ? amp_k[ch[q]] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kInjPreU; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < kInjU; ++q)
                if (q < ne) acc += wv[q] * av[j][q];
            if (slot + j < s1) out[(slot + j - seg_base) * stride] = (T)(act[j] ? acc : 0.0);
        }
    }
}

// out[i] += sum over slots s = s0, s0 + step, ... < s1 of store[s * stride + i]^2
// (FP64), i < n: one thread per node, the slots summed in order.
template <typename T>
__global__ void k_slot_square_sum(const T *__restrict__ store, long long stride, int s0, int s1, int step,
                                  long long n, double *__restrict__ out)
{
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int s = s0; s < s1; s += step) {
        const double v = (double)store[(long long)s * stride + i];
        acc += v * v;
    }
    out[i] += acc;
}

static int inj_precompute(const pysit_amd_batch &B, int slot_begin, int slot_end, int seg_base, hipStream_t st)
{
    if (slot_end <= slot_begin || B.n_inj_nodes == 0) return PYSIT_AMD_OK;
    const dim3 grid((unsigned)((B.n_inj_nodes + 255) / 256),
                    (unsigned)((slot_end - slot_begin + kInjPreSlots - 1) / kInjPreSlots));
    if (B.dtype)
        hipLaunchKernelGGL(k_inj_pre<double>, grid, dim3(256), 0, st, B, slot_begin, slot_end, seg_base);
    else
        hipLaunchKernelGGL(k_inj_pre<float>, grid, dim3(256), 0, st, B, slot_begin, slot_end, seg_base);
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

const char *pysit_amd_last_error(void) { return g_err.c_str(); }

const char *pysit_amd_version(void) { return "pysit_amd 0.1 (gfx950)"; }

int pysit_amd_tile_rows(int dim, int dtype, int axis)
{
    clear_error();
    if ((dim != 2 && dim != 3) || (dtype != 0 && dtype != 1) || axis < 0 || axis > 1) {
        set_error("tile_rows: dim must be 2 or 3, dtype 0 or 1, axis 0 (x) or 1 (y)");
        return PYSIT_AMD_EINVAL;
    }
    if (dim == 2) return axis == 0 ? pa::tile_rows_2d(dtype) : 1;
    return pa::tile_rows_3d(dtype, axis);
}

int pysit_amd_tile_rows_supported(int dim, int dtype, int tile_x)
{
    clear_error();
    if ((dim != 2 && dim != 3) || (dtype != 0 && dtype != 1)) {
        set_error("tile_rows_supported: dim must be 2 or 3, dtype 0 or 1");
        return PYSIT_AMD_EINVAL;
    }
    if (dim == 2) return pa::tile_rows_2d_supported(tile_x) ? 1 : 0;
    return tile_x == pa::tile_rows_3d(dtype, 0) ? 1 : 0;
}

int pysit_amd_abi_check(int shot_size, int batch_size)
{
    clear_error();
    if (shot_size != (int)sizeof(pysit_amd_shot) || batch_size != (int)sizeof(pysit_amd_batch)) {
        set_error("struct size mismatch: library has shot " + std::to_string(sizeof(pysit_amd_shot)) +
                  ", batch " + std::to_string(sizeof(pysit_amd_batch)));
        return PYSIT_AMD_EINVAL;
    }
    return PYSIT_AMD_OK;
}

int pysit_amd_cda_step2d_f32(int order, const float *km1_u, const float *k_phix,
                             const float *k_phiz, const float *k_u, const float *C,
                             const float *rhs, const float *xlpml, int n_xlpml,
                             const float *xrpml, int n_xrpml, const float *zlpml, int n_zlpml,
                             const float *zrpml, int n_zrpml, double dt, double dx, double dz,
                             int nx, int nz, float *kp1_phix, float *kp1_phiz, float *kp1_u,
                             void *stream)
{
    clear_error();
    return step2d_entry<float>(order, km1_u, k_phix, k_phiz, k_u, C, rhs, xlpml, n_xlpml, xrpml,
                               n_xrpml, zlpml, n_zlpml, zrpml, n_zrpml, dt, dx, dz, nx, nz,
                               kp1_phix, kp1_phiz, kp1_u, (hipStream_t)stream);
}

int pysit_amd_cda_step2d_f64(int order, const double *km1_u, const double *k_phix,
                             const double *k_phiz, const double *k_u, const double *C,
                             const double *rhs, const double *xlpml, int n_xlpml,
                             const double *xrpml, int n_xrpml, const double *zlpml, int n_zlpml,
                             const double *zrpml, int n_zrpml, double dt, double dx, double dz,
                             int nx, int nz, double *kp1_phix, double *kp1_phiz, double *kp1_u,
                             void *stream)
{
    clear_error();
    return step2d_entry<double>(order, km1_u, k_phix, k_phiz, k_u, C, rhs, xlpml, n_xlpml, xrpml,
                                n_xrpml, zlpml, n_zlpml, zrpml, n_zrpml, dt, dx, dz, nx, nz,
                                kp1_phix, kp1_phiz, kp1_u, (hipStream_t)stream);
}

int pysit_amd_cda_step3d_f32(int order, const float *km1_u, const float *k_phix,
                             const float *k_phiy, const float *k_phiz, const float *k_psi,
                             const float *k_u, const float *C, const float *rhs,
                             const float *xlpml, int n_xlpml, const float *xrpml, int n_xrpml,
                             const float *ylpml, int n_ylpml, const float *yrpml, int n_yrpml,
                             const float *zlpml, int n_zlpml, const float *zrpml, int n_zrpml,
                             double dt, double dx, double dy, double dz, int nx, int ny, int nz,
                             float *kp1_phix, float *kp1_phiy, float *kp1_phiz, float *kp1_psi,
                             float *kp1_u, void *stream)
{
    clear_error();
    return step3d_entry<float>(order, km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xlpml,
                               n_xlpml, xrpml, n_xrpml, ylpml, n_ylpml, yrpml, n_yrpml, zlpml,
                               n_zlpml, zrpml, n_zrpml, dt, dx, dy, dz, nx, ny, nz, kp1_phix,
                               kp1_phiy, kp1_phiz, kp1_psi, kp1_u, (hipStream_t)stream);
}

int pysit_amd_cda_step3d_f64(int order, const double *km1_u, const double *k_phix,
                             const double *k_phiy, const double *k_phiz, const double *k_psi,
                             const double *k_u, const double *C, const double *rhs,
                             const double *xlpml, int n_xlpml, const double *xrpml, int n_xrpml,
                             const double *ylpml, int n_ylpml, const double *yrpml, int n_yrpml,
                             const double *zlpml, int n_zlpml, const double *zrpml, int n_zrpml,
                             double dt, double dx, double dy, double dz, int nx, int ny, int nz,
                             double *kp1_phix, double *kp1_phiy, double *kp1_phiz,
                             double *kp1_psi, double *kp1_u, void *stream)
{
    clear_error();
    return step3d_entry<double>(order, km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xlpml,
                                n_xlpml, xrpml, n_xrpml, ylpml, n_ylpml, yrpml, n_yrpml, zlpml,
                                n_zlpml, zrpml, n_zrpml, dt, dx, dy, dz, nx, ny, nz, kp1_phix,
                                kp1_phiy, kp1_phiz, kp1_psi, kp1_u, (hipStream_t)stream);
}

}  // extern "C"

namespace pa {

// pysit_amd_batch_run's checks and dispatch, with an optional plan
static int batch_run(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end, int direction,
                     int seg_base, const int *active_tiles, const int *active_recs, pysit_amd_plan *plan,
                     void *stream, long long persistent_spins = 0)
{
    PA_CHECK(b != nullptr, "null batch");
    const pysit_amd_batch &B = *b;
    PA_CHECK(B.dim == 2 || B.dim == 3, "dim must be 2 or 3");
    PA_CHECK(B.dtype == 0 || B.dtype == 1, "dtype must be 0 (f32) or 1 (f64)");
    PA_CHECK(B.order == 2 || B.order == 4 || B.order == 6 || B.order == 8,
             "order must be 2, 4, 6 or 8");
    PA_CHECK(mode >= 0 && mode <= 3, "mode must be 0, 1, 2 or 3");
    PA_CHECK(mode != 3 || (B.born_m1 && B.imaging_period == 1),
             "Born mode needs m1 and imaging_period 1");
    PA_CHECK(direction == 1 || direction == -1, "direction must be +1 or -1");
    PA_CHECK(slot_begin >= 0 && slot_end >= slot_begin, "bad slot range");
    PA_CHECK(B.tile_x >= 1 && (B.dim == 2 || B.tile_y >= 1), "tile sizes must be positive");
    PA_CHECK(B.n_shots >= 1 && B.shots && B.n_tiles >= 0 && (B.n_tiles == 0 || B.tiles),
             "shot or tile table missing");
    PA_CHECK(B.u[0] && B.u[1] && B.phix[0] && B.phix[1] && B.phiz[0] && B.phiz[1] && B.du,
             "state buffers missing (u, phix, phiz and du)");
    PA_CHECK(B.dim == 2 || (B.phiy[0] && B.phiy[1] && B.psi[0] && B.psi[1]),
             "3D state buffers missing");
    PA_CHECK(B.imaging_period >= 1, "imaging_period must be >= 1");
    PA_CHECK(mode != 2 || !B.dw || B.ic, "imaging needs the FP64 accumulator");
    PA_CHECK(!B.ic32 || (B.dtype == 0 && B.ic && B.ic_window >= 1),
             "the FP32 window accumulator needs an FP32 batch, ic and ic_window >= 1");
    PA_CHECK(mode == 0 || mode == 2 || B.dw, "modes 1 and 3 need the dWaveOp store");
    PA_CHECK(!B.vol_m == !B.vol_dw, "the volume source needs both vol_m and vol_dw");
    if (B.dw_out || B.vol_dw) {
        PA_CHECK(mode == 2 || (mode == 3 && !B.vol_dw),
                 "dw_out applies to modes 2 and 3, the volume source to mode 2");
        PA_CHECK(B.imaging_period == 1, "the second-order terms need imaging_period 1");
        PA_CHECK(!B.ic32, "an adjoint run with second-order terms sums into ic directly (ic32 must be NULL)");
        PA_CHECK(B.dw_out != B.dw && (!B.vol_dw || B.vol_dw != B.dw_out),
                 "dw_out must not alias dw or vol_dw");
    }
    PA_CHECK(B.n_rec == 0 || (B.recs && B.rec_idx && B.rec_w), "receiver tables missing");
    PA_CHECK(seg_base >= 0 && seg_base <= slot_begin, "seg_base must not exceed slot_begin");
    if (active_tiles || active_recs) {
        for (int j = slot_begin; j < slot_end; ++j) {
            PA_CHECK(!active_tiles || (active_tiles[j] >= 0 && active_tiles[j] <= B.n_tiles),
                     "active tile count out of range");
            PA_CHECK(mode == 2 || !active_recs || (active_recs[j] >= 0 && active_recs[j] <= B.n_rec),
                     "active receiver count out of range");
        }
    }
    PA_CHECK(!B.inj_pre || (B.inj_ptr && B.inj_node_shot && B.n_inj_nodes >= 0 && B.inj_pre_slots >= 0 &&
                            slot_end - seg_base <= B.inj_pre_slots &&
                            (slot_end - slot_begin + kInjPreSlots - 1) / kInjPreSlots <= 65535),
             "precomputed injection: tables missing, or the run does not fit inj_pre_slots");
    hipStream_t st = (hipStream_t)stream;
    if (B.inj_pre) {
        const int rc = inj_precompute(B, slot_begin, slot_end, seg_base, st);
        if (rc) return rc;
    }
    if (persistent_spins > 0) {
        PA_CHECK(B.dim == 2, "the persistent loop is built for 2D batches");
        return B.dtype ? run2d_persistent<double>(B, mode, slot_begin, slot_end, direction, seg_base, active_tiles,
                                                  active_recs, persistent_spins, st)
                       : run2d_persistent<float>(B, mode, slot_begin, slot_end, direction, seg_base, active_tiles,
                                                 active_recs, persistent_spins, st);
    }
    if (B.dim == 2)
        return B.dtype ? run2d<double>(B, mode, slot_begin, slot_end, direction, seg_base,
                                       active_tiles, active_recs, plan, st)
                       : run2d<float>(B, mode, slot_begin, slot_end, direction, seg_base,
                                      active_tiles, active_recs, plan, st);
    return B.dtype ? run3d<double>(B, mode, slot_begin, slot_end, direction, seg_base,
                                   active_tiles, active_recs, plan, st)
                   : run3d<float>(B, mode, slot_begin, slot_end, direction, seg_base,
                                  active_tiles, active_recs, plan, st);
}

int plan_device(const pysit_amd_plan *plan);

// ---- the run without a caller workspace (pysit_amd_batch_run_static) ------
// The FP32 window accumulator of an adjoint run is partial sums: scratch
// memory beyond the caller's inputs and outputs.  pysit_amd_batch_run takes it
// from the caller (b->ic32); the static run keeps it in this file-scope
// device array instead, so it allocates nothing.  Each GPU the library runs
// on holds its own copy until the process ends.  In host code `scratch` names
// a host-side stand-in, not device memory: host code reaches the array only
// through the device address hipGetSymbolAddress returns (scratch_ptr).
constexpr long long kScratchBytes = 64LL << 20;
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_dev[kMaxDevices];  // device address of `scratch` per device (null until first use)

// the current device's copy of `scratch`: looked up once per device, with
// that device current, on the first static run there
static int scratch_ptr(float **out)
{
    int dev = -1;
    PA_HIP(hipGetDevice(&dev));
    PA_CHECK(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        PA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PA_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = (float *)p;
    }
    *out = g_scratch_dev[dev];
    return PYSIT_AMD_OK;
}

// zero n floats (a kernel, not a memset node, so the run can be captured)
__global__ void k_zero_f32(float *__restrict__ p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace pa

extern "C" {

int pysit_amd_batch_run(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end,
                        int direction, int seg_base, const int *active_tiles,
                        const int *active_recs, void *stream)
{
    clear_error();
    return batch_run(b, mode, slot_begin, slot_end, direction, seg_base, active_tiles, active_recs, nullptr,
                     stream);
}

int pysit_amd_batch_run_persistent(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end,
                                   int direction, int seg_base, const int *active_tiles,
                                   const int *active_recs, void *stream)
{
    clear_error();
    // a barrier wait longer than 2^24 polls of about 0.1 us (seconds, where a
    // step takes microseconds) ends the launch with an error;
    // PYSIT_AMD_PERSIST_SPIN_LIMIT (>= 1) sets another bound, e.g. a tiny one
    // that makes the time-out path run
    long long spins = 1LL << 24;
    if (const char *env = std::getenv("PYSIT_AMD_PERSIST_SPIN_LIMIT")) {
        const long long v = std::atoll(env);
        PA_CHECK(v >= 1, "PYSIT_AMD_PERSIST_SPIN_LIMIT must be an integer >= 1");
        spins = v;
    }
    return batch_run(b, mode, slot_begin, slot_end, direction, seg_base, active_tiles, active_recs, nullptr,
                     stream, spins);
}

int pysit_amd_batch_run_static(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end,
                               int direction, int seg_base, const int *active_tiles,
                               const int *active_recs, void *stream)
{
    clear_error();
    PA_CHECK(b != nullptr, "null batch");
    PA_CHECK(!b->ic32 && !b->inj_pre,
             "the static run takes no workspace: ic32 and inj_pre must be NULL (its FP32 window "
             "accumulator is the library's scratch array; steps gather their sources themselves)");
    pysit_amd_batch B = *b;
    const bool window = mode == 2 && B.dtype == 0 && B.ic_window >= 1 && B.ic && B.dw && !B.dw_out && !B.vol_dw;
    if (window) {
        PA_CHECK(B.total_n >= 0 && B.total_n * (long long)sizeof(float) <= kScratchBytes,
                 "static run: the FP32 window accumulator of " + std::to_string(B.total_n) +
                     " nodes exceeds the scratch array (at most " +
                     std::to_string(kScratchBytes / (long long)sizeof(float)) +
                     " nodes); use pysit_amd_batch_run with a caller workspace (ic32)");
        float *p = nullptr;
        const int rc = scratch_ptr(&p);
        if (rc) return rc;
        // the array keeps the last call's contents: the window sums start at 0
        if (B.total_n > 0) {
            hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((B.total_n + 255) / 256)), dim3(256), 0,
                               (hipStream_t)stream, p, B.total_n);
            PA_HIP(hipGetLastError());
        }
        B.ic32 = p;
    }
    int rc = batch_run(&B, mode, slot_begin, slot_end, direction, seg_base, active_tiles, active_recs, nullptr,
                       stream);
    // the next call clears the array: a run that stops inside a window adds
    // its open window into ic now
    const int last = direction > 0 ? slot_end - 1 : slot_begin;
    if (!rc && window && slot_end > slot_begin && !window_end(last, direction, B.ic_window))
        rc = ic_flush(B, (hipStream_t)stream);
    return rc;
}

int pysit_amd_plan_run(pysit_amd_plan *plan, const pysit_amd_batch *b, int mode, int slot_begin,
                       int slot_end, int direction, int seg_base, const int *active_tiles,
                       const int *active_recs, void *stream)
{
    clear_error();
    PA_CHECK(plan != nullptr, "null plan");
    int dev = -1;
    PA_HIP(hipGetDevice(&dev));
    PA_CHECK(dev == plan_device(plan), "the plan belongs to another device");
    return batch_run(b, mode, slot_begin, slot_end, direction, seg_base, active_tiles, active_recs, plan,
                     stream);
}

int pysit_amd_slot_square_sum(const void *store, int dtype, long long slot_stride, int slot_begin, int slot_end,
                              int slot_step, long long n, double *out, void *stream)
{
    clear_error();
    PA_CHECK(store && out, "null array argument");
    PA_CHECK(dtype == 0 || dtype == 1, "dtype must be 0 (f32) or 1 (f64)");
    PA_CHECK(slot_step >= 1 && slot_begin >= 0 && slot_end >= slot_begin && n >= 0 && slot_stride >= n,
             "bad slot range, step or stride");
    if (n == 0 || slot_end == slot_begin) return PYSIT_AMD_OK;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (dtype)
        hipLaunchKernelGGL(k_slot_square_sum<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double *)store, slot_stride, slot_begin, slot_end, slot_step, n, out);
    else
        hipLaunchKernelGGL(k_slot_square_sum<float>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const float *)store, slot_stride, slot_begin, slot_end, slot_step, n, out);
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

int pysit_amd_fold(int dim, int nx, int ny, int nz, int lx, int rx, int ly, int ry, int lz, int rz,
                   const double *in, double scale, double *out, int out_ny, int out_nz,
                   int accumulate, void *stream)
{
    clear_error();
    PA_CHECK(dim == 2 || dim == 3, "dim must be 2 or 3");
    if (dim == 2) {
        ny = 1;
        ly = ry = 0;
        out_ny = 1;
    }
    PA_CHECK(in && out, "null array argument");
    PA_CHECK(lx >= 0 && rx >= 0 && ly >= 0 && ry >= 0 && lz >= 0 && rz >= 0, "negative padding");
    const int ox = nx - lx - rx, oy = ny - ly - ry, oz = nz - lz - rz;
    PA_CHECK(ox >= 1 && oy >= 1 && oz >= 1, "padding leaves an empty grid");
    PA_CHECK(out_nz >= oz && out_ny >= oy, "output strides smaller than the unpadded block");
    const long long n = (long long)ox * oy * oz;
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_fold, dim3(blocks), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, lx,
                       rx, ly, ry, lz, rz, in, scale, out, out_ny, out_nz, accumulate);
    PA_HIP(hipGetLastError());
    return PYSIT_AMD_OK;
}

}  // extern "C"

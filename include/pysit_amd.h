/*
 * pysit_amd.h -- C ABI of the MI355X (gfx950) library behind the drop-in
 * constant-density acoustic time solver.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative code on error;
 *     pysit_amd_last_error() then returns a message (thread-local).  No C++
 *     exception crosses the ABI.
 *   - Array arguments are DEVICE pointers unless stated otherwise.  Grids are
 *     the reference's padded grids with the reference's flat layout:
 *     2D idx = i*nz + k, 3D idx = (i*ny + j)*nz + k (z contiguous).
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Calls on one stream are ordered; a batch descriptor must not be used
 *     from two threads at once.
 *   - dtype codes: 0 = float32, 1 = float64.
 */
#ifndef PYSIT_AMD_H
#define PYSIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYSIT_AMD_OK 0
#define PYSIT_AMD_EINVAL (-1)   /* bad argument (shape, order, dtype, pointer) */
#define PYSIT_AMD_EDEVICE (-2)  /* a HIP runtime call failed */

/* Message of the last error on this thread ("" if none). */
const char *pysit_amd_last_error(void);

/* Library version string. */
const char *pysit_amd_version(void);

/* Default tile geometry: rows along x (axis 0) or y (axis 1, 3D) of one
 * tile, for dim 2 or 3 and dtype 0 (f32) or 1 (f64).  A batch's tile_x /
 * tile_y carry these values or others pysit_amd_tile_rows_supported accepts
 * (2D: 8, 4 or 2 rows along x; 3D: 16 x planes by 8 y rows). */
int pysit_amd_tile_rows(int dim, int dtype, int axis);

/* 1 if a batch of this dim and dtype may use tile_x rows along x, else 0
 * (negative on bad arguments). */
int pysit_amd_tile_rows_supported(int dim, int dtype, int tile_x);

/* ABI self-check for bindings: returns 0 when the caller's sizes of
 * pysit_amd_shot and pysit_amd_batch match the library's, else -1. */
int pysit_amd_abi_check(int shot_size, int batch_size);

/* ------------------------------------------------------------------------ */
/* Step entry points: one leapfrog+PML step with the reference kernel        */
/* contract.  Replace the SWIG-exported                                      */
/*   constant_density_acoustic_time_scalar_2D_{2,4,6,8}os                     */
/* (pysit/solvers/constant_density_acoustic/time/scalar/                     */
/*  constant_density_acoustic_time_scalar.i:57-84, C++ signature             */
/*  constant_density_acoustic_time_scalar_2D.h:44-62) with the same argument */
/* order minus the (rows, cols) pairs: inputs km1_u, k_Phix, k_Phiz, k_u, C, */
/* rhs (padded nx*nz), PML arrays with lengths, dt, dx, dz, nx, nz, and the  */
/* caller-owned outputs kp1_Phix, kp1_Phiz, kp1_u.  Outputs must not alias  */
/* inputs.  order in {2,4,6,8}.                                              */
/* ------------------------------------------------------------------------ */
int pysit_amd_cda_step2d_f32(int order, const float *km1_u, const float *k_phix,
                             const float *k_phiz, const float *k_u, const float *C,
                             const float *rhs, const float *xlpml, int n_xlpml,
                             const float *xrpml, int n_xrpml, const float *zlpml,
                             int n_zlpml, const float *zrpml, int n_zrpml, double dt,
                             double dx, double dz, int nx, int nz, float *kp1_phix,
                             float *kp1_phiz, float *kp1_u, void *stream);
int pysit_amd_cda_step2d_f64(int order, const double *km1_u, const double *k_phix,
                             const double *k_phiz, const double *k_u, const double *C,
                             const double *rhs, const double *xlpml, int n_xlpml,
                             const double *xrpml, int n_xrpml, const double *zlpml,
                             int n_zlpml, const double *zrpml, int n_zrpml, double dt,
                             double dx, double dz, int nx, int nz, double *kp1_phix,
                             double *kp1_phiz, double *kp1_u, void *stream);

/* 3D counterparts of constant_density_acoustic_time_scalar_3D_{2,4,6,8}os
 * (constant_density_acoustic_time_scalar.i:78-84, C++ signature
 * constant_density_acoustic_time_scalar_3D.h:47-72). */
int pysit_amd_cda_step3d_f32(int order, const float *km1_u, const float *k_phix,
                             const float *k_phiy, const float *k_phiz, const float *k_psi,
                             const float *k_u, const float *C, const float *rhs,
                             const float *xlpml, int n_xlpml, const float *xrpml,
                             int n_xrpml, const float *ylpml, int n_ylpml,
                             const float *yrpml, int n_yrpml, const float *zlpml,
                             int n_zlpml, const float *zrpml, int n_zrpml, double dt,
                             double dx, double dy, double dz, int nx, int ny, int nz,
                             float *kp1_phix, float *kp1_phiy, float *kp1_phiz,
                             float *kp1_psi, float *kp1_u, void *stream);
int pysit_amd_cda_step3d_f64(int order, const double *km1_u, const double *k_phix,
                             const double *k_phiy, const double *k_phiz,
                             const double *k_psi, const double *k_u, const double *C,
                             const double *rhs, const double *xlpml, int n_xlpml,
                             const double *xrpml, int n_xrpml, const double *ylpml,
                             int n_ylpml, const double *yrpml, int n_yrpml,
                             const double *zlpml, int n_zlpml, const double *zrpml,
                             int n_zrpml, double dt, double dx, double dy, double dz,
                             int nx, int ny, int nz, double *kp1_phix, double *kp1_phiy,
                             double *kp1_phiz, double *kp1_psi, double *kp1_u,
                             void *stream);

/* ------------------------------------------------------------------------ */
/* Batched shot loops.  Replace the per-step Python loops of                */
/* TemporalModeling.forward_model (pysit/modeling/temporal_modeling.py:      */
/* 60-169) and adjoint_model (:228-387) for many shots at once: one kernel   */
/* launch advances every active shot of the batch by one time step, with    */
/* source injection, receiver sampling, the dWaveOp store (forward) and the */
/* FP64 imaging condition (adjoint) fused into the step.                    */
/*                                                                           */
/* A batch is described by device tables the host builds once per geometry  */
/* (see pysit_amd/device.py, which fills them):                              */
/* ------------------------------------------------------------------------ */

/* Per-shot descriptor (device array, one per shot). */
typedef struct {
    int nx, ny, nz;        /* padded grid of this shot (ny = 1 in 2D)            */
    int zs;                /* row stride along z in the batch's arrays, >= nz: node
                            * (i,j,k) sits at off + (i*ny + j)*zs + k            */
    int nt;                /* number of time steps                               */
    int kshift;            /* local step k = slot - kshift                       */
    int ntx, nty, ntz;     /* tiles of this shot along x, y, z                   */
    long long off;         /* element offset of this shot in the batch's arrays  */
    /* model (wave speed, dtype of the batch): value at padded (i,j,k) is
     * model[((clamp(i-mx_off,mx_lo,mx_hi))*mny + clamp(j-my_off,..))*mnz
     *        + clamp(k-mz_off,mz_lo,mz_hi)]  -- edge padding by clamping,
     * the 'edge' pad of solver_base.py:107-117.                              */
    const void *model;
    int mny, mnz;
    int mx_off, mx_lo, mx_hi, my_off, my_lo, my_hi, mz_off, mz_lo, mz_hi;
    /* PML arrays in the reference form (dtype of the batch) with lengths    */
    const void *xl, *xr, *yl, *yr, *zl, *zr;
    int nxl, nxr, nyl, nyr, nzl, nzr;
    /* injection amplitudes: amp[k * namp + chan] (FP64)                     */
    const double *amp;
    int namp;
    int nrec;              /* receivers of this shot                             */
    double *trace;         /* trace[k * nrec + r] (FP64), forward sampling       */
} pysit_amd_shot;

typedef struct {
    int dtype;             /* 0 f32, 1 f64                                       */
    int dim;               /* 2 or 3                                             */
    int order;             /* 2, 4, 6, 8                                         */
    int tile_x;            /* rows along x per wave tile                         */
    int tile_y;            /* rows along y per wave tile (3D)                    */
    double dt, dx, dy, dz;
    int n_shots;
    const pysit_amd_shot *shots;
    /* tiles: int4 {shot, tile index t within the shot, flags, x tile index
     * t / (tiles per x row)}; flags bit0 = the tile is within `order/2` nodes
     * of a PML node (full path); bit1 = no row (2D) or x plane (3D) of the
     * tile or within `order/2` of it is an x PML row/plane or an x edge;
     * 3D bit2 = no node of the tile or within `order/2` nodes of it lies in
     * two or more PML bands: the full path then neither reads nor advances
     * psi, which enters only multiplied by products of two or three sigmas
     * (a bit that is wrongly set gives wrong results; clear is always
     * correct).  Shots sorted so that, for every slot, the active tiles form
     * a prefix.                                                              */
    int n_tiles;
    const int *tiles;
    /* injection, two-level CSR: tile t owns destination nodes
     * [inj_ptr[t], inj_ptr[t+1]); node n sits at local position inj_node[n]
     * (2D: row_in_tile*64 + z - z0; 3D: (row_in_tile*8 + y - y0)*64 + z - z0)
     * and sums entries [inj_nptr[n], inj_nptr[n+1]) in order:
     * rhs += inj_w[e] * amp[k * namp + inj_chan[e]]                        */
    const int *inj_ptr;
    const int *inj_node;
    const int *inj_nptr;
    const int *inj_chan;
    const double *inj_w;
    /* receiver sampling: rec {shot, r, e0, e1}; entries rec_idx (flat index
     * in the shot's padded grid) and rec_w (FP64, weight * prod(deltas));
     * 16 lanes share one receiver and reduce in a fixed order              */
    int n_rec;
    const int *recs;
    const long long *rec_idx;
    const double *rec_w;
    /* state: each field has two buffers; the local step parity picks which
     * holds time level k.  2D uses u, phix, phiz; 3D adds phiy and psi.
     * du (one buffer, required) holds the increment du_k = u_k - u_{k-1},
     * updated in place: a step computes du_{k+1} = du_k + (update) and
     * u_{k+1} = u_k + du_{k+1} (the reference's leapfrog in increment form;
     * in FP32 it keeps the low bits of the time difference that
     * 2u_k - u_{k-1} loses).  A batch at rest has u = du = 0.               */
    void *u[2], *phix[2], *phiy[2], *phiz[2], *psi[2];
    void *du;
    long long total_n;     /* sum of shot grid sizes                             */
    /* dWaveOp store (dtype of the batch): slot-major, dw[(slot - seg_base) *
     * total_n + off + idx]; NULL = not stored                               */
    void *dw;
    /* imaging-condition accumulator (FP64, total_n), adjoint only           */
    double *ic;
    int imaging_period;
    /* optional field store (dtype of the batch), same layout as dw: when not
     * NULL every launch writes the field of time level k it reads (u_k
     * forward, q_k adjoint) -- the 'wavefield' / 'adjointfield' returns     */
    void *fstore;
    /* Born source (mode 3): model-grid perturbation m1 (dtype of the batch,
     * same shape and window as the model); rhs_k = -m1 * dWaveOp0_k with
     * dWaveOp0 read from dw                                                 */
    const void *born_m1;
    /* two-level imaging accumulator (FP32 batches, adjoint): when ic32 is
     * not NULL the per-step products are summed into ic32 (FP32, total_n,
     * zero at the start of a sweep) and ic32 is added into ic (FP64) and
     * cleared after every slot j with j % ic_window == 0 (backward runs).
     * The windows follow the slots, not the calls: a sweep split into calls
     * (checkpoint segments) sums exactly as one call, and a call that stops
     * inside a window leaves it open in ic32 for the next call.  A backward
     * sweep ends on slot 0, so ic holds the complete sum after it.
     * NULL: every step adds into ic                                         */
    float *ic32;
    int ic_window;
    /* precomputed injection (optional, any mode): when inj_pre is not NULL,
     * a run first evaluates the injection of every node of the CSR above for
     * each of its slots, inj_pre[(slot - seg_base) * n_inj_nodes + n] (dtype
     * of the batch, the same value the step would sum: FP64 sum in entry
     * order, rounded once), and its steps read that value instead of summing
     * the node's entries.  inj_node_shot[n] is the shot of node n; the
     * buffer holds inj_pre_slots slots (slot_end - seg_base must fit).     */
    void *inj_pre;
    const int *inj_node_shot;
    int n_inj_nodes;
    int inj_pre_slots;
    /* dWaveOp form (what mode 1 stores and mode 3 reads):
     *   0  PySIT's (constant_density_acoustic_base.py:24-40): the step's time
     *      operator with the MESH PML profile, whose right side sits one node
     *      off the kernel's (the reference's index quirk);
     *   1  exact (2D only): the step's time operator with the KERNEL profile,
     *      i.e. C^2 (rhs + L u + Dx Phix + Dz Phiz) by the step identity --
     *      the derivative of the discrete step with respect to m = C^-2, so
     *      the imaging condition against the adjoint recursion is the exact
     *      gradient of the discrete misfit (that recursion is the step's
     *      transpose after the change of variables Phi' = -dt (sz - sx) mu).
     * Both forms agree wherever the two profiles do (off the right PML).   */
    int dw_form;
    /* second-order terms of the full Hessian (apply_hessian 'full',
     * temporal_least_squares.py:224-239); imaging_period must be 1:
     *   dw_out: when not NULL, mode 3 also stores the dWaveOp of the Born
     *     field u1 (linear_forward_model's 'dWaveOp1', temporal_modeling.py:
     *     521-522) and mode 2 the dWaveOp of the adjoint field q
     *     (adjoint_model's 'dWaveOpAdj', :340-343: the form above applied to
     *     the adjoint's own three time levels), both in dw's layout;
     *   vol_m, vol_dw: when not NULL (both), mode 2 adds the volume source
     *     vol_m[off + idx] * vol_dw[(slot - seg_base) * total_n + off + idx]
     *     to rhs_k (_setup_adjoint_rhs, :216-226: operand_model, zero-padded,
     *     times operand_dWaveOpAdj[k]); vol_m is laid out like the state.
     * A mode 2 run with any of them sums into ic directly (ic32 must be
     * NULL).                                                                */
    void *dw_out;
    const void *vol_m;
    const void *vol_dw;
} pysit_amd_batch;

/* Run slots [slot_begin, slot_end) forward (direction = +1) or
 * (slot_end-1 .. slot_begin) backward (direction = -1).
 *   mode 0: propagate only (+ receiver sampling when recs are given)
 *   mode 1: propagate and store dWaveOp for k % imaging_period == 0
 *   mode 2: adjoint: ic += q_k * dWaveOp_k (k % imaging_period == 0), then
 *           propagate; receiver sampling is skipped (with ic32, the sum goes
 *           through the FP32 window accumulator described above)
 *   mode 3: Born (linear_forward_model, temporal_modeling.py:389-546):
 *           propagate with the extra source -m1 * dWaveOp0_k read from dw
 *           (imaging_period must be 1), recording traces
 * seg_base: slot index of the dWaveOp store's first slot.
 * active_tiles[j], active_recs[j] (HOST arrays indexed by slot, may be NULL =
 * all): how many leading tiles / receivers are active at slot j.
 * A backward run starts with q at the state the previous call left (both
 * zero for a fresh accumulation). */
int pysit_amd_batch_run(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end,
                        int direction, int seg_base, const int *active_tiles,
                        const int *active_recs, void *stream);

/* pysit_amd_batch_run as one persistent launch (2D, modes 0, 1 and 2, no
 * second-order terms).  The reference steps its time loop from Python, one
 * step call per time level (temporal_modeling.py:120-142, :313-350); here a
 * cooperative launch runs every step of the call: each block takes its
 * share of every step's tile and receiver blocks and all blocks meet at a
 * device-wide barrier between steps (and around the FP32 window's flushes).
 * The results equal pysit_amd_batch_run's bit for bit.  On MI355X the
 * barrier (2-4 us at 8-24 blocks) costs more than the launch boundary it
 * replaces, so the graph loop of pysit_amd_plan_run stays the faster
 * choice for small batches.  The call allocates a
 * small device buffer on the stream (the per-step counts and the barrier)
 * and waits for the launch to end; a barrier that waits for seconds (2^24
 * polls, or the environment's PYSIT_AMD_PERSIST_SPIN_LIMIT) ends the launch
 * with PYSIT_AMD_EDEVICE. */
int pysit_amd_batch_run_persistent(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end,
                                   int direction, int seg_base, const int *active_tiles,
                                   const int *active_recs, void *stream);

/* pysit_amd_batch_run without a caller workspace.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid allocation).  It allocates nothing: no hipMalloc, hipFree or
 * allocating library routine.  Same arguments and results as
 * pysit_amd_batch_run (bit for bit), except:
 *   - b->ic32 and b->inj_pre must be NULL (no workspace is taken); the steps
 *     gather their sources themselves, which gives the same values;
 *   - the FP32 window accumulator of an FP32 imaging run (mode 2,
 *     ic_window >= 1, ic and dw given) lives in the library's static device
 *     array of 64 MiB per GPU, so such a run holds at most 16 Mi nodes
 *     (total_n); larger batches are refused with PYSIT_AMD_EINVAL;
 *   - every call on one GPU shares that GPU's array: calls that overlap on
 *     one GPU (two streams, two threads) race.  Each call clears the part it
 *     uses before its first step (with a kernel, not a memset), and a call
 *     that stops inside an FP32 window adds that window into ic at its end
 *     (pysit_amd_batch_run leaves it open in ic32 for the next call).     */
int pysit_amd_batch_run_static(const pysit_amd_batch *b, int mode, int slot_begin, int slot_end,
                               int direction, int seg_base, const int *active_tiles,
                               const int *active_recs, void *stream);

/* ------------------------------------------------------------------------ */
/* Plans: the batch time loop as replayed hipGraphs.                         */
/*                                                                           */
/* The reference steps its time loop from Python, one step call per time    */
/* level (temporal_modeling.py:120-142 forward, :313-350 adjoint).  A plan   */
/* is the opaque handle that runs pysit_amd_batch_run's loop as graphs: it   */
/* owns a device step cursor, a capture stream and a cache of instantiated   */
/* graphs.  A graph holds `graph_steps` consecutive step launches (and the   */
/* FP32 window flushes among them) that read their slot from the cursor,    */
/* then a node that advances the cursor, so one graph serves every chunk of */
/* a run with the same batch contents, mode, direction and active counts.   */
/* Chunks whose active tile or receiver counts change inside them, and the  */
/* remainder of a run, are launched step by step.  Results are identical to  */
/* pysit_amd_batch_run's.                                                    */
/*   graph_steps: steps per graph (>= 1; with the FP32 window accumulator a */
/*                multiple of ic_window, else that run goes step by step)   */
/*   max_graphs:  graphs kept (least recently used evicted), >= 1           */
/* A plan belongs to the device current at creation; its calls must not    */
/* overlap (one stream at a time).  Graphs capture the batch by value: a    */
/* batch whose tables or buffers change is a new cache entry.               */
/* ------------------------------------------------------------------------ */
typedef struct pysit_amd_plan pysit_amd_plan;

int pysit_amd_plan_create(int graph_steps, int max_graphs, pysit_amd_plan **out);

/* pysit_amd_batch_run through the plan (same arguments and checks). */
int pysit_amd_plan_run(pysit_amd_plan *plan, const pysit_amd_batch *b, int mode, int slot_begin,
                       int slot_end, int direction, int seg_base, const int *active_tiles,
                       const int *active_recs, void *stream);

/* stats[0] graph launches, [1] graph captures, [2] steps launched one by
 * one, [3] graphs cached now */
int pysit_amd_plan_stats(const pysit_amd_plan *plan, long long *stats);

/* Waits for the device, then frees the plan's graphs, cursor and stream. */
int pysit_amd_plan_destroy(pysit_amd_plan *plan);

/* Sum of squares over time slots of a slot-major store (the dWaveOp store):
 * out[i] += sum over s = slot_begin, slot_begin + slot_step, ... < slot_end of
 * store[s * slot_stride + i]^2 in FP64, for i < n (store: dtype 0 f32 or
 * 1 f64, device).  The per-shot term of the pseudo-Hessian diagonal,
 * TemporalLeastSquares._pseudo_hessian_diagonal_component_shot
 * (pysit/objective_functions/temporal_least_squares.py:121-139), before its
 * crop to the physical grid and its imaging_period * dt factor. */
int pysit_amd_slot_square_sum(const void *store, int dtype, long long slot_stride, int slot_begin,
                              int slot_end, int slot_step, long long n, double *out, void *stream);

/* Fold (ModelPerturbationBase.add_padding, pysit/solvers/model_parameter.py:
 * 537-593) of one padded FP64 image onto its unpadded grid: the PML slabs are
 * summed onto the first/last physical plane of each axis (x, then y, then z),
 * the result is scaled and written (accumulate = 0) or added (accumulate = 1)
 * into `out`, which points at the window origin inside a larger array whose
 * row strides are out_ny * out_nz (x) and out_nz (y).  nx, ny, nz are the
 * padded sizes, lx/rx/... the padding widths (2D: ny = 1, ly = ry = 0). */
int pysit_amd_fold(int dim, int nx, int ny, int nz, int lx, int rx, int ly, int ry,
                   int lz, int rz, const double *in, double scale, double *out,
                   int out_ny, int out_nz, int accumulate, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PYSIT_AMD_H */

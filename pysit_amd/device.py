"""Device-side shot batches: tables, state, and the fused time loops.

A :class:`ShotGeometry` describes one propagation in the reference's terms
(padded grid, the kernel PML arrays, the edge-padded model window, point
operators as sparse entry lists, the wavelet amplitudes).  A
:class:`DeviceBatch` packs any number of them for one launch per time step of
``pysit_amd_batch_run`` and runs

* ``forward``      -- record traces (and optionally keep dWaveOp),
* ``migrate``      -- the time-reversed adjoint with the FP64 imaging
  condition, recomputing dWaveOp from state checkpoints when the store does
  not fit the memory budget,
* ``fold``         -- the PML fold of each shot's image onto the model grid.

Time slots: all shots of a batch end on the same slot; shot ``s`` with
``nt_s`` steps starts at slot ``kshift_s = max_nt - nt_s``.  Shots are sorted
by ``kshift`` so that the active tiles (and receivers) of every slot form a
prefix of the tables and each launch covers exactly the active work.

Device memory comes from torch (plumbing only); every kernel is ours.
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as N

TILE_Z = 64      # z nodes per wavefront (the contiguous axis)


def z_last_tile(nz, align):
    """Start of the last z tile: the first multiple of ``align`` nodes at or
    after nz - 64 (0 when nz <= 64), common.h z_last_tile."""
    return -(-(nz - TILE_Z) // align) * align if nz > TILE_Z else 0


def z_align(itemsize):
    """Row alignment in nodes: one 128-byte cache line."""
    return 128 // itemsize


# Largest region a launch reaches through 32-bit buffer offsets (the
# kernels' kMaxRsrcBytes, common.h)
MAX_REGION_BYTES = 0x7FFF0000


def max_batch_nodes(dim, itemsize):
    """Most padded nodes one batch may hold: the FP64 imaging accumulator
    (8 B per node) and, in 2D, the seven state buffers, kept in one region,
    must each fit one buffer resource."""
    return MAX_REGION_BYTES // (max(8, 7 * itemsize) if dim == 2 else 8)


def shot_nodes(shape, itemsize):
    """Padded nodes of one shot of grid ``shape`` as a batch stores it (rows
    along z at a whole number of cache lines)."""
    a = z_align(itemsize)
    return int(np.prod(shape[:-1])) * (-(-int(shape[-1]) // a) * a)


def shots_per_batch(shape, itemsize):
    """Most shots of grid ``shape`` that fit one batch (at least 1; a shot
    too large on its own is refused when its batch is built)."""
    return max(1, max_batch_nodes(len(shape), itemsize) // shot_nodes(shape, itemsize))


def z_tiles_2d(nz, align=1):
    """z tiles (2D and 3D): (start, owned end) per tile.  Tiles start at 0, 64,
    128, ... except the last, which starts at z_last_tile(nz, align) and
    reaches node nz-1; each node is owned by exactly one tile (the kernels'
    tiling, cda2d.hip tile2d, cda3d.hip tile3d)."""
    ntz = -(-nz // TILE_Z)
    zlast = z_last_tile(nz, align)
    out = []
    for tz in range(ntz):
        if tz == ntz - 1:
            out.append((zlast, nz))
        else:
            out.append((tz * TILE_Z, min(tz * TILE_Z + TILE_Z, zlast)))
    return out


def _dev():
    if not torch.cuda.is_available():
        raise N.NativeError("no ROCm device visible: the pysit_amd solver path runs on the GPU only")
    return torch.device("cuda", torch.cuda.current_device())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


_PLANS = {}


def shared_plan(graph_steps, device=None):
    """The graph plan of this process for a device and a graph step count,
    shared by every batch there.  A graph is keyed by the whole batch
    descriptor (every buffer address and size) and the run, so it replays
    only with the kernel arguments it was captured with; a new batch whose
    descriptor equals an earlier one's (per-shot batches get the same
    buffers back from torch's allocator) replays the earlier graphs instead
    of capturing its own.  Batches use it one at a time (one thread)."""
    dev = torch.cuda.current_device() if device is None else int(device)
    key = (dev, int(graph_steps))
    plan = _PLANS.get(key)
    if plan is None:
        with torch.cuda.device(dev):
            plan = _PLANS[key] = N.Plan(graph_steps, max_graphs=32)
    return plan


class ShotGeometry:
    """One propagation, in padded-grid terms.

    Parameters
    ----------
    shape : padded grid ``(nx, nz)`` or ``(nx, ny, nz)``
    kpml : per axis ``(left, right)`` kernel PML arrays (the reference's
        ``mesh.<axis>.{lbc,rbc}.sigma``; empty arrays for non-PML sides)
    window : per axis ``(off, lo, hi)``; the model value of padded node ``i``
        is ``model[clamp(i - off, lo, hi)]`` (edge padding by clamping)
    nt : number of time steps
    src : ``(pos, chan, w)`` injection entries, ``pos`` (n, dim) padded
        integer coordinates, in the reference's accumulation order
    amp : ``(nt, n_src_chan)`` FP64 amplitudes, ``amp[k, c] = intensity*w(k dt)``
    rec : ``(idx, chan, w)`` sampling entries: flat padded index, receiver,
        weight (already times prod(deltas)), grouped by receiver
    nrec : receivers
    adj : ``(pos, chan, w)`` adjoint injection entries (weights without
        prod(deltas)), ``chan`` = receiver
    """

    def __init__(self, shape, kpml, window, nt, src, amp, rec, nrec, adj=None):
        self.shape = tuple(int(s) for s in shape)
        self.dim = len(self.shape)
        self.kpml = [(np.asarray(l, dtype=np.float64).reshape(-1), np.asarray(r, dtype=np.float64).reshape(-1))
                     for l, r in kpml]
        self.window = [tuple(int(v) for v in w) for w in window]
        self.nt = int(nt)
        self.src = _as_entries(src, self.dim)
        self.amp = np.ascontiguousarray(amp, dtype=np.float64)
        if self.amp.ndim != 2 or self.amp.shape[0] < self.nt:
            raise ValueError("amp must be (>= nt, n_src_chan)")
        self.rec = (np.asarray(rec[0], dtype=np.int64), np.asarray(rec[1], dtype=np.int64),
                    np.asarray(rec[2], dtype=np.float64))
        self.nrec = int(nrec)
        self.adj = _as_entries(adj, self.dim) if adj is not None else None
        if len(self.kpml) != self.dim or len(self.window) != self.dim:
            raise ValueError("kpml and window need one entry per axis")

    @property
    def npts(self):
        return int(np.prod(self.shape))


def _as_entries(e, dim):
    pos, chan, w = e
    pos = np.asarray(pos, dtype=np.int64).reshape(-1, dim)
    return pos, np.asarray(chan, dtype=np.int64).reshape(-1), np.asarray(w, dtype=np.float64).reshape(-1)


def checkpoint_levels(nt, per_slot, ckpt, budget, max_levels=4, state=None):
    """Segment lengths [K_1, ..., K_L] of a nested checkpoint schedule for a
    run of ``nt`` slots, or None if none fits ``budget`` bytes.

    Level 1 keeps a state checkpoint every K_1 slots of the forward pass;
    level l > 1 replays one level l-1 segment from its checkpoint and keeps a
    checkpoint every K_l slots of it; the last level replays a K_L-slot
    segment storing its dWaveOp (per_slot bytes per slot), which the adjoint
    then consumes.  Memory: K_L per_slot plus, per level, the checkpoints of
    one segment of the level above, plus one working state (ckpt bytes each).
    Each level below the first costs one more forward sweep, so the fewest
    levels that fit win, and among those the longest stored segment.  The
    intermediate lengths split the run evenly: K_l ~ nt / r^l with
    r^L = nt / K_L.  ``state`` is the working state's size (default: a
    checkpoint's)."""
    state = ckpt if state is None else state
    for L in range(1, max_levels + 1):
        for KL in range(nt, 0, -1):
            r = (nt / KL) ** (1.0 / L)
            Ks = [max(KL, -(-nt // int(round(r ** l))) if r ** l >= 1 else nt) for l in range(1, L)] + [KL]
            Ks = [min(nt, k) for k in Ks]
            if any(Ks[i] < Ks[i + 1] for i in range(len(Ks) - 1)):
                continue
            n = [-(-nt // Ks[0])] + [-(-Ks[i - 1] // Ks[i]) for i in range(1, L)]
            if KL * per_slot + sum(n) * ckpt + state <= budget:
                return Ks
            if sum(n) * ckpt + state > budget and KL == 1:
                break
    return None


def _beta(s, t):
    """C(s + t, s): the most steps s snapshots reverse with t sweeps."""
    from math import comb
    return comb(s + t, s)


def binomial_split(n, s):
    """Steps to advance before taking the next snapshot when reversing ``n``
    steps from a stored state with ``s`` free snapshots: the binomial
    (Revolve) schedule of Griewank (1992) and Griewank & Walther (2000).
    With S = s + 1 snapshots counting the stored state and the repetition
    number t, the least with C(S + t, S) >= n, an advance m is optimal when
    the part after the new snapshot (n - m steps, S - 1 snapshots) needs t
    sweeps, C(S - 1 + t - 1, S - 1) <= n - m <= C(S - 1 + t, S - 1), and the
    part before it (m steps, S snapshots) t - 1, C(S + t - 2, S) <= m <=
    C(S + t - 1, S).  The total is then t n - C(S + t, S + 1) forward steps,
    the least any schedule needs (``binomial_cost``; the CPU tests check it
    against a brute-force search over every split)."""
    if n <= 1:
        return 0
    if s <= 0:
        return n - 1
    S = s + 1
    t = 0
    while _beta(S, t) < n:
        t += 1
    lo = max(1, n - _beta(S - 1, t), _beta(S, t - 2) if t >= 2 else 1)
    hi = min(n - 1, n - _beta(S - 1, t - 1), _beta(S, t - 1))
    return max(1, min(n - 1, (lo + hi) // 2))


def binomial_cost(n, s):
    """Forward steps the binomial schedule advances to reverse n steps with s
    free snapshots (the one-step stores of the reverse steps not counted)."""
    if n <= 1:
        return 0
    if s <= 0:
        return n * (n - 1) // 2
    m = binomial_split(n, s)
    return m + binomial_cost(n - m, s - 1) + binomial_cost(m, s)


GRAPH_MAX_NODES = 1 << 20
# FP32 imaging window (slots summed in FP32 before each FP64 flush) and the
# default graph chunk, a multiple of it so that graph chunks hold whole windows
IC_WINDOW = 64
GRAPH_STEPS = 64


class DeviceBatch:
    """Tables and state of a shot batch on one GPU."""

    # adjoint sources evaluated per run before its first step (False: each
    # step gathers them; the values are the same)
    precompute_injection = True

    def __init__(self, shots, model, deltas, dt, order, dtype=np.float32, tile_x=None,
                 imaging_period=1, memory_budget=None, ic_window=None, tile_order="shot", graph_steps=None,
                 adjoint="pysit"):
        if not shots:
            raise ValueError("empty shot batch")
        self.dim = shots[0].dim
        if any(s.dim != self.dim for s in shots):
            raise ValueError("all shots of a batch share the dimension")
        # adjoint: "pysit" stores PySIT's dWaveOp (mesh PML profile), the
        # imaging factor of migrate_shot; "exact" stores the step's own
        # derivative with respect to m (kernel profile), with which the
        # adjoint recursion gives the exact gradient of the discrete misfit
        # (2D; in 3D only without PML, where the two agree)
        if adjoint not in ("pysit", "exact"):
            raise ValueError("adjoint must be 'pysit' or 'exact'")
        if adjoint == "exact" and self.dim == 3 and any(l.size or r.size for s in shots for l, r in s.kpml):
            raise ValueError("the exact adjoint is built for 2D; 3D batches with PML use adjoint='pysit'")
        self.adjoint = adjoint
        self.dev = _dev()
        if order not in (2, 4, 6, 8):
            raise ValueError("spatial_accuracy_order must be 2, 4, 6 or 8")
        self.order = int(order)
        self.R = self.order // 2
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be float32 or float64")
        self.tdtype = torch.float32 if self.dtype == np.float32 else torch.float64
        self.deltas = tuple(float(d) for d in deltas)
        self.dt = float(dt)
        self.imaging_period = int(imaging_period)
        # tile rows along x: the library's default, or another height its
        # kernels are built for (2D: 8, 4 or 2)
        code = 1 if self.dtype == np.float64 else 0
        if tile_x:
            if not N.tile_rows_supported(self.dim, code, int(tile_x)):
                raise ValueError(f"{self.dim}D {self.dtype} batches do not support tile_x={tile_x}")
            self.tile_x = int(tile_x)
        else:
            self.tile_x = N.tile_rows(self.dim, code, 0)
        self.tile_y = N.tile_rows(self.dim, 1 if self.dtype == np.float64 else 0, 1) if self.dim == 3 else 1
        self.zalign = z_align(self.dtype.itemsize)
        self.memory_budget = memory_budget
        # checkpoint schedule: "auto" stores what fits and nests segment
        # checkpoints (falling back to the binomial schedule when even four
        # levels do not fit); "binomial" always reverses from rest with state
        # snapshots (max_snaps caps their number)
        self.schedule = "auto"
        self.max_snaps = None
        self.snaps = None
        # FP32 adjoints sum the imaging condition in FP32 over windows of
        # `ic_window` slots and in FP64 across windows (0: FP64 every step)
        if ic_window is None:
            ic_window = IC_WINDOW if self.dtype == np.float32 else 0
        if ic_window and self.dtype != np.float32:
            raise ValueError("the FP32 window accumulator is for FP32 batches")
        self.ic_window = int(ic_window)
        # tile table order: "shot" (all tiles of a shot together: halo rows of
        # neighbouring tiles are reused in L2) or "interleave" (the shots of
        # one tile together: the model window is reused; alike shots only)
        if tile_order not in ("shot", "interleave"):
            raise ValueError("tile_order must be 'shot' or 'interleave'")
        self.tile_order = tile_order
        # time loop as replayed hipGraphs of `graph_steps` steps (0: one launch
        # per step); the plan is made on the first run.  Default: graphs for
        # batches of fewer than GRAPH_MAX_NODES nodes, whose steps are short
        # enough for launch gaps to show; larger batches launch step by step
        # (a graph step reads its slot from device memory first, which costs
        # them 1-3 %).
        self._graph_steps_arg = graph_steps
        if graph_steps is not None and int(graph_steps) < 0:
            raise ValueError("graph_steps must be >= 0")
        self.plan = None
        # 2D time loops without second-order terms (8- or 4-row tiles) as one
        # persistent cooperative launch per run
        # (pysit_amd_batch_run_persistent) in place of step launches or
        # graphs.  The results are the same; its
        # device-wide barrier between steps costs more than a graph's launch
        # boundary on MI355X, hence off by default.
        self.persistent = False

        if isinstance(model, torch.Tensor):
            self.model = model.detach().to(device=self.dev, dtype=self.tdtype).contiguous()
        else:
            self.model = torch.from_numpy(np.ascontiguousarray(model, dtype=self.dtype)).to(self.dev)
        if self.model.dim() != self.dim:
            raise ValueError("model must have the grid's dimension")
        self.model_shape = tuple(self.model.shape)
        # optional per-mode launch timing: {mode: [(start, end, n_launches)]}
        self.timer = None

        self.max_nt = max(s.nt for s in shots)
        kshift = [self.max_nt - s.nt for s in shots]
        self.order_idx = sorted(range(len(shots)), key=lambda i: kshift[i])  # stable
        self.shots = [shots[i] for i in self.order_idx]
        self.kshift = [kshift[i] for i in self.order_idx]
        self.n_shots = len(self.shots)
        self._build_tables()
        self._alloc_state()
        gs = self._graph_steps_arg
        self.graph_steps = int(gs) if gs is not None else (GRAPH_STEPS if self.total_n < GRAPH_MAX_NODES else 0)

    # ------------------------------------------------------------------ tables
    def _tiles_of(self, s):
        """(ntx, nty, ntz) and the per-tile PML flag array (tloc order)."""
        R = self.R
        if self.dim == 2:
            # a tile takes the plain path only if no node within R of the
            # nodes it owns is a PML node, an edge node or outside the grid
            nx, nz = s.shape
            zt = z_tiles_2d(nz, self.zalign)
            ntz = len(zt)
            ntx = -(-nx // self.tile_x)
            flags = np.zeros(ntx * ntz, dtype=np.int32)
            (xl, xr), (zl, zr) = s.kpml
            for tx in range(ntx):
                x0, x1 = tx * self.tile_x, min(nx, (tx + 1) * self.tile_x) - 1
                xp = (x0 - R < max(xl.size, 1)) or (x1 + R > nx - 1 - max(xr.size, 1))
                for tz, (z0, zhi) in enumerate(zt):
                    z1 = zhi - 1
                    zp = (z0 - R < max(zl.size, 1)) or (z1 + R > nz - 1 - max(zr.size, 1))
                    # bit0: the full path; bit1: its rows need no x PML form
                    flags[tx * ntz + tz] = (1 if (xp or zp) else 0) | (0 if xp else 2)
            return (ntx, 1, ntz), flags
        nx, ny, nz = s.shape
        zt = z_tiles_2d(nz, self.zalign)
        ntz = len(zt)
        nty = -(-ny // self.tile_y)
        ntx = -(-nx // self.tile_x)
        (xl, xr), (yl, yr), (zl, zr) = s.kpml

        def band(lo, hi, n, left, right):
            # the node range [lo - R, hi + R] meets a PML band of the axis
            return (left.size > 0 and lo - R < left.size) or (right.size > 0 and hi + R > n - 1 - right.size)

        flags = np.zeros(ntx * nty * ntz, dtype=np.int32)
        for tx in range(ntx):
            x0, x1 = tx * self.tile_x, min(nx, (tx + 1) * self.tile_x) - 1
            xp = (x0 - R < max(xl.size, 1)) or (x1 + R > nx - 1 - max(xr.size, 1))
            xb = band(x0, x1, nx, xl, xr)
            for ty in range(nty):
                y0, y1 = ty * self.tile_y, min(ny, (ty + 1) * self.tile_y) - 1
                yp = (y0 - R < max(yl.size, 1)) or (y1 + R > ny - 1 - max(yr.size, 1))
                yb = band(y0, y1, ny, yl, yr)
                for tz, (z0, zhi) in enumerate(zt):
                    zp = (z0 - R < max(zl.size, 1)) or (zhi - 1 + R > nz - 1 - max(zr.size, 1))
                    zb = band(z0, zhi - 1, nz, zl, zr)
                    # bit0: the full path; bit1: its planes need no x PML
                    # form; bit2: the tile's reach meets at most one PML band,
                    # so no node it reads or writes lies in two (the box it
                    # reaches is a product of ranges, the bands are slabs):
                    # psi is neither read nor advanced there
                    flags[(tx * nty + ty) * ntz + tz] = ((1 if (xp or yp or zp) else 0) | (0 if xp else 2)
                                                         | (4 if int(xb) + int(yb) + int(zb) <= 1 else 0))
        return (ntx, nty, ntz), flags

    def _tile_of_pos(self, s, s_tiles, pos):
        """(tile within the shot, position within the tile) of grid nodes."""
        ntx, nty, ntz = s_tiles
        if self.dim == 2:
            nz = s.shape[1]
            zlast = z_last_tile(nz, self.zalign)
            z = pos[:, 1]
            tz = np.where(z >= zlast, ntz - 1, z // TILE_Z)
            z0 = np.where(tz == ntz - 1, zlast, tz * TILE_Z)
            return (pos[:, 0] // self.tile_x) * ntz + tz, (pos[:, 0] % self.tile_x) * TILE_Z + (z - z0)
        ty = self.tile_y
        nz = s.shape[2]
        zlast = z_last_tile(nz, self.zalign)
        z = pos[:, 2]
        tz = np.where(z >= zlast, ntz - 1, z // TILE_Z)
        z0 = np.where(tz == ntz - 1, zlast, tz * TILE_Z)
        t = ((pos[:, 0] // self.tile_x) * nty + pos[:, 1] // ty) * ntz + tz
        local = ((pos[:, 0] % self.tile_x) * ty + pos[:, 1] % ty) * TILE_Z + (z - z0)
        return t, local

    def _build_tables(self):
        dev = self.dev
        T = self.dtype
        # storage: every shot's rows along z at a stride of whole cache lines
        # (zs >= nz; the nodes past nz are never read or written)
        self.offs, self.zs, self.snpts = [], [], []
        off = 0
        tiles_info = []
        for s in self.shots:
            nz = s.shape[-1]
            zs = -(-nz // self.zalign) * self.zalign
            sn = int(np.prod(s.shape[:-1])) * zs
            self.offs.append(off)
            self.zs.append(zs)
            self.snpts.append(sn)
            off += sn
            tiles_info.append(self._tiles_of(s))
        self.total_n = off
        limit = max_batch_nodes(self.dim, self.dtype.itemsize)
        if off > limit:
            raise ValueError(f"{len(self.shots)} shots of {self.dim}D grids hold {off} padded nodes; one batch "
                             f"holds at most {limit}: split the shots (batch_size=)")
        self.tiles_info = tiles_info

        # tile table: interleave the shots of one tile when they are alike,
        # else shot-major (active prefix by kshift)
        same = all(s.shape == self.shots[0].shape and s.window == self.shots[0].window
                   for s in self.shots) and len(set(self.kshift)) == 1 and self.tile_order == "interleave"
        rows = []
        # column 3: the x tile index, t // (tiles per x row)
        per_x = [ti[0][1] * ti[0][2] for ti in tiles_info]
        if same:
            ntile = len(tiles_info[0][1])
            for t in range(ntile):
                for si in range(self.n_shots):
                    rows.append((si, t, int(tiles_info[si][1][t]), t // per_x[si]))
        else:
            for si in range(self.n_shots):
                fl = tiles_info[si][1]
                for t in range(len(fl)):
                    rows.append((si, t, int(fl[t]), t // per_x[si]))
        tiles = np.asarray(rows, dtype=np.int32).reshape(-1, 4)
        self.n_tiles = tiles.shape[0]
        # global tile id of (shot, tile within the shot): one lookup array per shot
        self._tile_lookup = []
        for si in range(self.n_shots):
            look = np.full(len(tiles_info[si][1]), -1, dtype=np.int64)
            mine = tiles[:, 0] == si
            look[tiles[mine, 1]] = np.nonzero(mine)[0]
            self._tile_lookup.append(look)
        self.tiles = torch.from_numpy(tiles.reshape(-1).copy()).to(dev)

        self._same = same
        self._active_prefix()

        # PML arrays and model windows (shared tensors where equal)
        self._keep = []
        self.fwd_inj = self._injection_tables([s.src for s in self.shots])
        has_adj = all(s.adj is not None for s in self.shots)
        self.adj_inj = self._injection_tables([s.adj for s in self.shots]) if has_adj else None

        # receivers
        recs, ridx, rw = [], [], []
        e = 0
        for si, s in enumerate(self.shots):
            idx, chan, w = s.rec
            nz = s.shape[-1]
            idx = (np.asarray(idx, dtype=np.int64) // nz) * self.zs[si] + np.asarray(idx, dtype=np.int64) % nz
            order = np.argsort(chan, kind="stable")
            idx, chan, w = idx[order], chan[order], w[order]
            counts = np.bincount(chan, minlength=s.nrec)
            starts = np.concatenate([[0], np.cumsum(counts)])
            for r in range(s.nrec):
                recs.append((si, r, e + starts[r], e + starts[r + 1]))
            ridx.append(idx)
            rw.append(w)
            e += idx.size
        self.n_rec = len(recs)
        self.recs = torch.from_numpy(np.asarray(recs, dtype=np.int32).reshape(-1)).to(dev)
        self.rec_idx = torch.from_numpy(np.concatenate(ridx).astype(np.int64) if ridx else np.zeros(0, np.int64)).to(dev)
        self.rec_w = torch.from_numpy(np.concatenate(rw) if rw else np.zeros(0)).to(dev)

        # per-shot device arrays
        self.pml_dev = []
        for s in self.shots:
            arrs = []
            for l, r in s.kpml:
                arrs.append(torch.from_numpy(l.astype(T)).to(dev) if l.size else None)
                arrs.append(torch.from_numpy(r.astype(T)).to(dev) if r.size else None)
            self.pml_dev.append(arrs)
        self.amp_dev = [torch.from_numpy(s.amp).to(dev) for s in self.shots]
        self.traces = [torch.zeros((s.nt, s.nrec), dtype=torch.float64, device=dev) for s in self.shots]
        self._model_layout()
        self.shots_fwd = self._shot_table(self.amp_dev, [s.amp.shape[1] for s in self.shots])
        self.residual_dev = None
        self.shots_adj = None

    def _active_prefix(self):
        """Active tiles and receivers per slot (shots start at their kshift)."""
        shot_tiles = np.array([len(ti[1]) for ti in self.tiles_info])
        shot_recs = np.array([s.nrec for s in self.shots])
        ks = np.array(self.kshift)
        slots = np.arange(self.max_nt)
        active = ks[None, :] <= slots[:, None]
        if self._same:
            self.active_tiles = np.full(self.max_nt, self.n_tiles, dtype=np.int32)
        else:
            self.active_tiles = (active * shot_tiles[None, :]).sum(1).astype(np.int32)
        self.active_recs = (active * shot_recs[None, :]).sum(1).astype(np.int32)

    def retime(self, shots, dt):
        """Take a new time axis in place: ``shots`` are the batch's shots (in
        the caller's order) with new step counts and amplitudes at step
        ``dt``, and everything else equal (grids, PML, model windows,
        injection and receiver entries).  The tables, the launch
        configuration and the state are kept; the histories of the old time
        axis (dWaveOp stores, checkpoints, traces) are dropped.  A trial
        model of an inversion changes dt with its largest speed: retiming
        costs what the new amplitudes cost instead of a rebuilt batch.
        Raises ValueError when the shots differ otherwise or their order by
        length changes (build a new batch then)."""
        if len(shots) != self.n_shots:
            raise ValueError("retime needs the batch's shots")
        max_nt = max(s.nt for s in shots)
        kshift = [max_nt - s.nt for s in shots]
        order = sorted(range(len(shots)), key=lambda i: kshift[i])
        if order != self.order_idx or (self._same and len(set(kshift)) != 1):
            raise ValueError("the shots' order by length changes")
        new = [shots[i] for i in order]

        def same_entries(a, b):
            return a is b or (a is not None and b is not None and all(np.array_equal(x, y) for x, y in zip(a, b)))

        for a, b in zip(new, self.shots):
            if (a.shape != b.shape or a.window != b.window or a.nrec != b.nrec
                    or not all(np.array_equal(x, y) for p, q in zip(a.kpml, b.kpml) for x, y in zip(p, q))
                    or not same_entries(a.src, b.src) or not same_entries(a.rec, b.rec)
                    or not same_entries(a.adj, b.adj)):
                raise ValueError("retime changes only the time axis of the same shots")
        self.generation += 1
        self.dt = float(dt)
        self.max_nt = max_nt
        self.kshift = [kshift[i] for i in order]
        self.shots = new
        self._active_prefix()
        self.amp_dev = [torch.from_numpy(s.amp).to(self.dev) for s in self.shots]
        self.traces = [torch.zeros((s.nt, s.nrec), dtype=torch.float64, device=self.dev) for s in self.shots]
        self.shots_fwd = self._shot_table(self.amp_dev, [s.amp.shape[1] for s in self.shots])
        self.residual_dev = None
        self.shots_adj = None
        # histories of the old time axis (the store is kept for reuse: it is
        # indexed by slot, so a store of at least the slots a run needs serves)
        self.ckpts = []
        self._dw_seg = None
        self.have_dw = False
        self.dw1 = None
        self.dw_adj = None
        self.field_store = None
        self.adjoint_store = None
        self.K = self.max_nt

    def _model_layout(self):
        """The kernels' copy of the model.  When every shot takes the same z
        window of the model, the copy lays each model row along the padded z
        axis at the batch's row stride (edge values by clamping, as the
        window does), so model rows start on the same cache-line offsets as
        the field rows; otherwise it is the model itself."""
        zw = {tuple(int(v) for v in s.window[-1]) for s in self.shots}
        nzs = {s.shape[-1] for s in self.shots}
        self._mz = None
        if len(zw) == 1 and len(nzs) == 1:
            off, lo, hi = zw.pop()
            nz = nzs.pop()
            zs = -(-nz // self.zalign) * self.zalign
            idx = torch.clamp(torch.arange(zs, device=self.dev) - off, lo, hi)
            self._mz = (idx, zs)
        self.model_k = self._kernel_layout(self.model)

    def _kernel_layout(self, m):
        if self._mz is None:
            return m
        return m.index_select(m.dim() - 1, self._mz[0]).contiguous()

    def set_model(self, model):
        """Take a new model of the same shape (host array or device tensor)
        in place; the tables keep pointing at it."""
        m = torch.as_tensor(np.ascontiguousarray(model) if not isinstance(model, torch.Tensor) else model)
        self.model.copy_(m.reshape(self.model.shape).to(device=self.dev, dtype=self.tdtype))
        if self._mz is not None:
            self.model_k.copy_(self._kernel_layout(self.model))

    def _injection_tables(self, entries):
        """Two-level CSR: tiles -> destination nodes -> entries.  Entries of
        one node keep the order they were given in (the order the reference's
        sparse product accumulates in).  Also the shot of every node (for the
        precomputed injection)."""
        node_tile, node_pos, node_key, e_chan, e_w, e_shot = [], [], [], [], [], []
        for si, (s, e) in enumerate(zip(self.shots, entries)):
            pos, chan, w = e
            if pos.shape[0] == 0:
                continue
            if np.any(pos < 0) or np.any(pos >= np.asarray(s.shape)[None, :]):
                raise ValueError("injection entry outside the padded grid")
            tloc, local = self._tile_of_pos(s, self.tiles_info[si][0], pos)
            tglob = self._tile_lookup[si][np.asarray(tloc, dtype=np.int64)]
            node_tile.append(tglob)
            node_pos.append(local)
            e_chan.append(chan)
            e_w.append(w)
            e_shot.append(np.full(tglob.size, si, dtype=np.int32))
        if node_tile:
            tg = np.concatenate(node_tile)
            lp = np.concatenate(node_pos)
            ch = np.concatenate(e_chan)
            wt = np.concatenate(e_w)
            es = np.concatenate(e_shot)
            order = np.lexsort((np.arange(tg.size), lp, tg))  # stable within a node
            tg, lp, ch, wt, es = tg[order], lp[order], ch[order], wt[order], es[order]
            key = tg * (1 << 20) + lp
            starts = np.concatenate([[True], key[1:] != key[:-1]])
            node_first = np.nonzero(starts)[0]
            n_tile = tg[node_first]
            n_pos = lp[node_first].astype(np.int32)
            n_shot = es[node_first]
            nptr = np.concatenate([node_first, [tg.size]]).astype(np.int32)
            tcount = np.bincount(n_tile, minlength=self.n_tiles)
        else:
            n_pos = np.zeros(0, np.int32)
            n_shot = np.zeros(0, np.int32)
            nptr = np.zeros(1, np.int32)
            ch = np.zeros(0, np.int64)
            wt = np.zeros(0)
            tcount = np.zeros(self.n_tiles, np.int64)
        ptr = np.concatenate([[0], np.cumsum(tcount)]).astype(np.int32)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(self.dev)
        return (t(ptr, np.int32), t(n_pos, np.int32), t(nptr, np.int32), t(ch, np.int32), t(wt, np.float64),
                t(n_shot, np.int32))

    def _shot_table(self, amps, namps):
        arr = (N.Shot * self.n_shots)()
        for si, s in enumerate(self.shots):
            d = arr[si]
            d.nx = s.shape[0]
            d.ny = s.shape[1] if self.dim == 3 else 1
            d.nz = s.shape[-1]
            d.zs = self.zs[si]
            d.nt = s.nt
            d.kshift = self.kshift[si]
            ntx, nty, ntz = self.tiles_info[si][0]
            d.ntx, d.nty, d.ntz = ntx, nty, ntz
            d.off = self.offs[si]
            d.model = self.model_k.data_ptr()
            d.mny = self.model_shape[1] if self.dim == 3 else 1
            wx = s.window[0]
            wz = s.window[-1]
            wy = s.window[1] if self.dim == 3 else (0, 0, 0)
            if self._mz is not None:
                # model rows already laid along the padded z axis
                d.mnz = self._mz[1]
                wz = (0, 0, self._mz[1] - 1)
            else:
                d.mnz = self.model_shape[-1]
            d.mx_off, d.mx_lo, d.mx_hi = wx
            d.my_off, d.my_lo, d.my_hi = wy
            d.mz_off, d.mz_lo, d.mz_hi = wz
            pm = self.pml_dev[si]
            if self.dim == 2:
                xl, xr, zl, zr = pm
                yl = yr = None
                (nxl, nxr), (nzl, nzr) = [(l.size, r.size) for l, r in s.kpml]
                nyl = nyr = 0
            else:
                xl, xr, yl, yr, zl, zr = pm
                (nxl, nxr), (nyl, nyr), (nzl, nzr) = [(l.size, r.size) for l, r in s.kpml]
            for name, t in (("xl", xl), ("xr", xr), ("yl", yl), ("yr", yr), ("zl", zl), ("zr", zr)):
                setattr(d, name, t.data_ptr() if t is not None else None)
            d.nxl, d.nxr, d.nyl, d.nyr, d.nzl, d.nzr = nxl, nxr, nyl, nyr, nzl, nzr
            d.amp = amps[si].data_ptr() if amps[si] is not None else None
            d.namp = namps[si]
            d.nrec = s.nrec
            d.trace = self.traces[si].data_ptr()
        raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
        # pinned staging: the copy does not wait for the work queued before it
        return torch.from_numpy(raw).pin_memory().to(self.dev, non_blocking=True)

    # ------------------------------------------------------------------- state
    def _fields(self):
        return ("u", "phix", "phiz") if self.dim == 2 else ("u", "phix", "phiy", "phiz", "psi")

    def _new_state(self):
        """Both time levels of every field and the increment du, in one
        allocation (the 2D kernel reaches all of them through one buffer
        resource).  At rest every buffer is zero."""
        fields = self._fields()
        block = torch.zeros((2 * len(fields) + 1, self.total_n), dtype=self.tdtype, device=self.dev)
        st = {f: block[2 * i:2 * i + 2] for i, f in enumerate(fields)}
        st["du"] = block[2 * len(fields):]
        return st

    def _alloc_state(self):
        # forward() calls so far: a consumer of the stores (a pending
        # backward) checks that no later forward replaced them
        self.generation = 0
        self.fstate = self._new_state()
        self.qstate = None
        self.dw = None
        self.dw_slots = 0
        self.ic = None
        self.dw1 = None     # dWaveOp of the Born field (born(store_dw=True))
        self.dw_adj = None  # dWaveOp of the adjoint field (migrate(store_adjoint_dw=True))

    def state_bytes(self):
        return (len(self._fields()) * 2 + 1) * self.total_n * self.dtype.itemsize

    def ckpt_bytes(self):
        """A state checkpoint: the current time level of every field and du."""
        return (len(self._fields()) + 1) * self.total_n * self.dtype.itemsize

    def _levels_at(self, slot):
        """Per shot, the buffer that holds the current time level of a
        forward run about to take step ``slot`` (the local step's parity; a
        shot that has not started is at rest in both)."""
        return [max(0, slot - ks) & 1 for ks in self.kshift]

    def _ckpt_take(self, state, slot):
        """Checkpoint of a forward state before step ``slot``.  A step reads
        only the current level of each field (and du) and rewrites the other
        level wherever it can be nonzero, so the current level is the whole
        state."""
        cur = self._levels_at(slot)
        out = {"du": state["du"].clone()}
        for f in self._fields():
            if len(set(cur)) == 1:
                out[f] = state[f][cur[0]].clone()
            else:
                t = torch.empty(self.total_n, dtype=self.tdtype, device=self.dev)
                for pos, c in enumerate(cur):
                    a, n = self.offs[pos], self.snpts[pos]
                    t[a:a + n].copy_(state[f][c][a:a + n])
                out[f] = t
        return out

    def _ckpt_restore(self, ckpt, state, slot):
        cur = self._levels_at(slot)
        state["du"].copy_(ckpt["du"])
        for f in self._fields():
            if len(set(cur)) == 1:
                state[f][cur[0]].copy_(ckpt[f])
            else:
                for pos, c in enumerate(cur):
                    a, n = self.offs[pos], self.snpts[pos]
                    state[f][c][a:a + n].copy_(ckpt[f][a:a + n])

    def _batch(self, state, adjoint, ic=None, dw=None):
        b = N.Batch()
        b.dtype = 1 if self.dtype == np.float64 else 0
        b.dim = self.dim
        b.order = self.order
        b.tile_x = self.tile_x
        b.tile_y = self.tile_y
        b.dt = self.dt
        b.dx = self.deltas[0]
        b.dy = self.deltas[1] if self.dim == 3 else 1.0
        b.dz = self.deltas[-1]
        b.n_shots = self.n_shots
        b.shots = (self.shots_adj if adjoint else self.shots_fwd).data_ptr()
        b.n_tiles = self.n_tiles
        b.tiles = self.tiles.data_ptr()
        inj = self.adj_inj if adjoint else self.fwd_inj
        b.inj_ptr, b.inj_node, b.inj_nptr, b.inj_chan, b.inj_w = [x.data_ptr() for x in inj[:5]]
        b.n_rec = 0 if adjoint else self.n_rec
        b.recs = self.recs.data_ptr() if self.n_rec else None
        b.rec_idx = self.rec_idx.data_ptr() if self.n_rec else None
        b.rec_w = self.rec_w.data_ptr() if self.n_rec else None
        for f in ("u", "phix", "phiy", "phiz", "psi"):
            arr = getattr(b, f)
            if f in state:
                arr[0] = state[f][0].data_ptr()
                arr[1] = state[f][1].data_ptr()
        b.du = state["du"].data_ptr()
        b.total_n = self.total_n
        b.dw_form = 1 if (self.adjoint == "exact" and self.dim == 2) else 0
        b.dw = dw.data_ptr() if dw is not None else None
        b.ic = ic.data_ptr() if ic is not None else None
        b.imaging_period = self.imaging_period
        return b

    def _run(self, b, mode, j0, j1, direction, seg_base, record=True):
        at = np.ascontiguousarray(self.active_tiles)
        ar = np.ascontiguousarray(self.active_recs if record else np.zeros_like(self.active_recs))
        ev = None
        if self.timer is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        atp, arp = at.ctypes.data_as(ctypes.c_void_p), ar.ctypes.data_as(ctypes.c_void_p)
        if (self.persistent and self.dim == 2 and mode in (0, 1, 2) and not (b.dw_out or b.vol_dw)
                and self.tile_x in (8, 4)):
            N.call("pysit_amd_batch_run_persistent", ctypes.byref(b), mode, j0, j1, direction, seg_base, atp, arp,
                   _stream())
        elif self.graph_steps:
            if self.plan is None:
                self.plan = shared_plan(self.graph_steps, self.dev.index)
            self.plan.run(b, mode, j0, j1, direction, seg_base, atp, arp, _stream())
        else:
            N.call("pysit_amd_batch_run", ctypes.byref(b), mode, j0, j1, direction, seg_base, atp, arp, _stream())
        if ev is not None:
            ev[1].record()
            self.timer.setdefault(mode, []).append((ev[0], ev[1], j1 - j0))

    # ---------------------------------------------------------- checkpointing
    def plan_segments(self, keep_dw):
        """Slots per dWaveOp segment K (K = max_nt means every step is stored).

        The store needs K * total_n values, the adjoint's precomputed
        sources K values per injection node; each checkpoint keeps the whole
        state.  K is the largest value whose store plus checkpoints fit the
        memory budget (default: 90% of the memory available to this batch:
        free device memory, memory torch holds unused, and this batch's own
        store and checkpoints, which a new run reuses or releases)."""
        self.snaps = None
        if not keep_dw:
            return self.max_nt
        if self.schedule not in ("auto", "binomial"):
            raise ValueError("schedule must be 'auto' or 'binomial'")
        budget = self.memory_budget
        item = self.dtype.itemsize
        n_inj = int(self.adj_inj[2].numel()) - 1 if self.adj_inj is not None else 0
        per_slot = (self.total_n + n_inj) * item
        ckpt = self.ckpt_bytes()
        state = self.state_bytes()
        if budget is None:
            # everything this batch holds from an earlier gradient counts as
            # available (a new one reuses or releases it), so repeated calls
            # plan the same schedule and keep their buffers
            free, _ = torch.cuda.mem_get_info(self.dev)
            unused = torch.cuda.memory_reserved(self.dev) - torch.cuda.memory_allocated(self.dev)
            mine = [getattr(self, "dw", None), getattr(self, "_inj_pre", None), getattr(self, "ic", None),
                    getattr(self, "_ic32", None)]
            mine += [v for c in getattr(self, "ckpts", []) for v in c.values()]
            held = sum(v.numel() * v.element_size() for v in mine if v is not None)
            held += state if getattr(self, "qstate", None) is not None else 0
            budget = int(0.9 * (free + unused + held))
        # what migrate allocates beside the store: the adjoint state, the
        # working state of the replays and the FP64 and FP32 imaging
        # accumulators
        budget -= state + 12 * self.total_n
        self.levels = [self.max_nt]
        self.snaps = None
        if self.schedule == "binomial":
            return self._plan_binomial(budget, per_slot, ckpt, state)
        if self.max_nt * per_slot <= budget:
            return self.max_nt
        levels = checkpoint_levels(self.max_nt, per_slot, ckpt, budget, state=state)
        if levels is None:
            # fewer bytes than the nested schedule needs: state snapshots
            # alone, reversed step by step (binomial schedule)
            return self._plan_binomial(budget, per_slot, ckpt, state)
        held = getattr(self, "dw_slots", 0) if getattr(self, "dw", None) is not None else 0
        if (self.memory_budget is None and len(levels) == 1 and held and held == getattr(self, "K", 0)
                and abs(levels[0] - held) <= max(2, levels[0] // 20)):
            # the earlier gradient ran this schedule with the store it still
            # holds: keep it rather than reallocate for a few slots of
            # difference in the free-memory estimate (a reallocation of a
            # store of hundreds of GB takes seconds)
            levels = [held]
        self.levels = levels
        return levels[0]

    def _plan_binomial(self, budget, per_slot, ckpt, state):
        """The binomial schedule: the forward pass only records traces; the
        adjoint reverses the run from rest with as many state snapshots as
        fit (a working state and a one-slot dWaveOp store beside them), each
        reverse step recomputing its own dWaveOp.  Returns K = 1."""
        snaps = (budget - state - per_slot) // ckpt
        if snaps < 0:
            raise MemoryError("memory budget too small for the checkpoint schedule")
        snaps = int(min(snaps, self.max_nt))
        if self.max_snaps is not None:
            snaps = min(snaps, int(self.max_snaps))
        self.snaps = snaps
        self.levels = [1]
        return 1

    def _ensure_dw(self, K):
        # zeroed once per allocation: no step writes the row padding past nz
        # (or the slots before an uneven shot starts), and sums over a
        # shot's whole region (the dWaveOp squares) read it
        if self.dw is None or self.dw_slots < K:
            self.dw = None
            self.dw_slots = 0
            try:
                self.dw = torch.zeros((K, self.total_n), dtype=self.tdtype, device=self.dev)
            except torch.cuda.OutOfMemoryError:
                # the released store may still sit in torch's cache
                torch.cuda.empty_cache()
                self.dw = torch.zeros((K, self.total_n), dtype=self.tdtype, device=self.dev)
            self.dw_slots = K

    # ----------------------------------------------------------------- forward
    def forward(self, keep_dw=False, segment=None, store_field=False):
        """Run every shot forward from rest, recording traces.

        With ``keep_dw`` the state is prepared for :meth:`migrate`: either the
        whole dWaveOp history is stored (one segment) or checkpoints are kept
        at segment boundaries.  ``store_field`` also keeps u_k of every step
        (``field_store``).  Returns the per-shot traces (device FP64, in the
        caller's shot order)."""
        self.generation += 1
        for t in self.traces:
            t.zero_()
        for f in self.fstate.values():
            f.zero_()
        # histories derived from the previous forward state
        self.dw1 = None
        self.dw_adj = None
        if segment:
            K = int(segment)
            self.levels = [K]
        else:
            K = self.plan_segments(keep_dw)
        self.K = K
        self.ckpts = []
        self._dw_seg = None
        self.field_store = self._field_store() if store_field else None
        if keep_dw and self.snaps is not None and not segment:
            # binomial schedule: the adjoint reverses the run from rest; the
            # forward pass only records the traces
            if store_field:
                raise ValueError("storing the wavefield needs the whole dWaveOp history in memory")
            self._run(self._batch(self.fstate, False), 0, 0, self.max_nt, 1, 0)
            self.have_dw = True
            return self.traces_in_caller_order()
        if segment:
            self.snaps = None
        if keep_dw and K >= self.max_nt:
            self._ensure_dw(self.max_nt)
            b = self._batch(self.fstate, False, dw=self.dw)
            b.fstore = self.field_store.data_ptr() if store_field else None
            self._run(b, 1, 0, self.max_nt, 1, 0)
        elif keep_dw:
            if store_field:
                raise ValueError("storing the wavefield needs the whole dWaveOp history in memory")
            if self.dw is not None and self.dw_slots > K:
                # the budget counted a larger store held from an earlier call
                # as available: release it before the checkpoints are taken
                self.dw = None
                self.dw_slots = 0
                self._inj_pre = None
                torch.cuda.empty_cache()
            b = self._batch(self.fstate, False)
            one_level = len(getattr(self, "levels", [K])) == 1
            self._ensure_dw(self.levels[-1] if not one_level else K)
            last = (self.max_nt - 1) // K * K if one_level else -1
            for j0 in range(0, self.max_nt, K):
                self.ckpts.append(self._ckpt_take(self.fstate, j0))
                if j0 == last:
                    # the last segment stores its dWaveOp as it goes: migrate
                    # starts there and needs no recompute for it
                    bs = self._batch(self.fstate, False, dw=self.dw)
                    self._run(bs, 1, j0, self.max_nt, 1, j0)
                else:
                    self._run(b, 0, j0, min(self.max_nt, j0 + K), 1, j0)
            if one_level:
                self._dw_seg = (last, self.max_nt)
        else:
            b = self._batch(self.fstate, False)
            b.fstore = self.field_store.data_ptr() if store_field else None
            self._run(b, 0, 0, self.max_nt, 1, 0)
        self.have_dw = keep_dw
        return self.traces_in_caller_order()

    def _field_store(self):
        # zeroed: no step writes the slots before an uneven shot starts or
        # the row padding past nz, and the store is returned whole
        return torch.zeros((self.max_nt, self.total_n), dtype=self.tdtype, device=self.dev)

    def load_dw(self, per_shot_images):
        """Use host-provided dWaveOp (per shot, caller order: a sequence of
        padded arrays, one per imaged step) as the stored history."""
        p = self.imaging_period
        self.generation += 1
        self._ensure_dw(self.max_nt)
        self.dw.zero_()
        for pos, orig in enumerate(self.order_idx):
            s = self.shots[pos]
            imgs = per_shot_images[orig]
            for i, a in enumerate(imgs):
                k = i * p
                if k >= s.nt:
                    break
                slot = k + self.kshift[pos]
                src = torch.as_tensor(np.ascontiguousarray(np.asarray(a).reshape(s.shape), dtype=self.dtype))
                self._shot_view(self.dw[slot], pos).copy_(src.to(self.dev))
        self.K = self.max_nt
        self.ckpts = []
        self.have_dw = True

    def _shot_view(self, flat, pos):
        """The padded grid of shot ``pos`` inside a flat batch array (a strided
        view: rows along z are stored at the stride zs >= nz)."""
        s = self.shots[pos]
        zs = self.zs[pos]
        grid = flat[self.offs[pos]:self.offs[pos] + self.snpts[pos]].view(tuple(s.shape[:-1]) + (zs,))
        return grid[..., :s.shape[-1]]

    def shot_field(self, store, si_caller, k):
        """Padded field of time level k of one shot from a field store."""
        pos = self.order_idx.index(si_caller)
        slot = k + self.kshift[pos]
        return self._shot_view(store[slot], pos)

    def shot_dw(self, si_caller, k, store="dw"):
        """Stored dWaveOp of step k of one shot (store-all histories only):
        the forward's (``"dw"``), the Born field's (``"dw1"``, from
        ``born(store_dw=True)``) or the adjoint field's (``"dw_adj"``, from
        ``migrate(store_adjoint_dw=True)``)."""
        if store == "dw" and self.ckpts:
            raise RuntimeError("dWaveOp was checkpointed, not stored")
        hist = getattr(self, store, None)
        if hist is None:
            raise RuntimeError(f"no {store} history is held")
        return self.shot_field(hist, si_caller, k)

    def padded_operand(self, si_caller, grid):
        """A flat batch array (dtype of the batch, laid out like the state)
        holding ``grid`` (the padded grid of one shot) in that shot's region
        and zeros elsewhere."""
        pos = self.order_idx.index(si_caller)
        out = torch.zeros(self.total_n, dtype=self.tdtype, device=self.dev)
        g = torch.as_tensor(np.ascontiguousarray(np.asarray(grid, dtype=self.dtype).reshape(self.shots[pos].shape)))
        self._shot_view(out, pos).copy_(g.to(self.dev))
        return out

    def history_from(self, si_caller, images):
        """A (max_nt, total_n) device history holding host arrays of one shot
        (one padded grid per step from step 0, imaging period 1) in that
        shot's slots and zeros elsewhere."""
        hist = torch.zeros((self.max_nt, self.total_n), dtype=self.tdtype, device=self.dev)
        pos = self.order_idx.index(si_caller)
        s = self.shots[pos]
        for k, a in enumerate(images):
            if k >= s.nt:
                break
            src = torch.as_tensor(np.ascontiguousarray(np.asarray(a).reshape(s.shape), dtype=self.dtype))
            self._shot_view(hist[k + self.kshift[pos]], pos).copy_(src.to(self.dev))
        return hist

    def born(self, m1, store_dw=False, store_field=False):
        """Traces of the Born wavefield u1 driven by -m1 * dWaveOp0_k (mode 3);
        needs the whole background dWaveOp history from forward(keep_dw=True).
        ``m1`` has the model's shape.  With ``store_dw`` the dWaveOp of u1 of
        every step is kept in ``self.dw1`` (linear_forward_model's 'dWaveOp1'),
        with ``store_field`` u1 of every step in ``field_store``.  Returns
        per-shot traces (caller order)."""
        if not getattr(self, "have_dw", False) or self.ckpts or self.imaging_period != 1:
            raise RuntimeError("born needs forward(keep_dw=True) with every step stored (imaging_period 1)")
        m1 = torch.as_tensor(np.ascontiguousarray(m1, dtype=self.dtype)).to(self.dev)
        if tuple(m1.shape) != tuple(self.model_shape):
            raise ValueError(f"m1 shape {tuple(m1.shape)} != model shape {self.model_shape}")
        for t in self.traces:
            t.zero_()
        state = self._new_state()
        b = self._batch(state, False, dw=self.dw)
        b.inj_ptr = b.inj_node = b.inj_nptr = b.inj_chan = b.inj_w = None  # u1: Born source alone
        m1 = self._kernel_layout(m1)
        b.born_m1 = m1.data_ptr()
        self._born_keep = m1
        if store_dw:
            # zeroed: slots before an uneven shot starts and the row padding
            # are never written
            self.dw1 = None
            self.dw1 = torch.zeros((self.max_nt, self.total_n), dtype=self.tdtype, device=self.dev)
            b.dw_out = self.dw1.data_ptr()
        self.field_store = self._field_store() if store_field else None
        b.fstore = self.field_store.data_ptr() if store_field else None
        self._run(b, 3, 0, self.max_nt, 1, 0)
        return self.traces_in_caller_order()

    def traces_in_caller_order(self):
        out = [None] * self.n_shots
        for pos, orig in enumerate(self.order_idx):
            out[orig] = self.traces[pos]
        return out

    # ----------------------------------------------------------------- migrate
    def migrate(self, residuals, store_field=False, image=True, dw_squares=False, image_store="dw",
                store_adjoint_dw=False, volume=None):
        """Imaging condition sum_k q_k dWaveOp_k of every shot (padded, FP64,
        concatenated at ``offs``; the caller applies -dt*p when folding).
        ``residuals`` are per-shot (nt, nrec) operands in the caller's shot
        order (numpy or device tensors).  ``store_field`` keeps q_k of every
        step in ``adjoint_store``; ``image=False`` only propagates q.
        ``dw_squares`` also sums dWaveOp_k^2 over the imaged steps of every
        shot into ``self.dw_sq`` (padded, FP64, laid out like the images).

        The full Hessian's terms (every step stored, imaging period 1):
        ``image_store`` names the history imaged against ("dw", or "dw1" of
        born(store_dw=True)); ``store_adjoint_dw`` keeps the dWaveOp of q of
        every step in ``self.dw_adj``; ``volume`` = (vol_m, history) adds the
        source vol_m * history_k (vol_m: a flat batch array, see
        padded_operand; history: a (max_nt, total_n) store)."""
        second = image_store != "dw" or store_adjoint_dw or volume is not None
        binomial = image and self.snaps is not None
        if second and (self.imaging_period != 1 or (image and (self.ckpts or binomial))):
            raise RuntimeError("an adjoint with second-order terms needs imaging_period 1 and, when it images, "
                               "every step of the forward stored")
        if store_field and image and (self.ckpts or binomial):
            raise ValueError("storing the adjoint field needs the whole dWaveOp history (no checkpoints)")
        if image and not getattr(self, "have_dw", False):
            raise RuntimeError("migrate needs a forward(keep_dw=True) first")
        if self.adj_inj is None:
            raise RuntimeError("the shots carry no adjoint injection entries")
        img_hist = getattr(self, image_store, None) if image else None
        recomputed = image_store == "dw" and (self.ckpts or self.snaps is not None)
        if image and img_hist is None and not recomputed:
            raise RuntimeError(f"no {image_store} history is held")
        res = [None] * self.n_shots
        for pos, orig in enumerate(self.order_idx):
            r = residuals[orig]
            s = self.shots[pos]
            r = torch.as_tensor(r, dtype=torch.float64, device=self.dev).contiguous()
            if tuple(r.shape) != (s.nt, s.nrec):
                raise ValueError(f"residual of shot {orig} has shape {tuple(r.shape)}, expected {(s.nt, s.nrec)}")
            res[pos] = r
        self.residual_dev = res
        self.shots_adj = self._shot_table(res, [s.nrec for s in self.shots])
        if self.qstate is None:
            self.qstate = self._new_state()
        for f in self.qstate.values():
            f.zero_()
        self.ic = torch.zeros(self.total_n, dtype=torch.float64, device=self.dev)
        self.dw_sq = torch.zeros(self.total_n, dtype=torch.float64, device=self.dev) if dw_squares else None
        ic32 = torch.zeros(self.total_n, dtype=torch.float32, device=self.dev) \
            if (self.ic_window and not second) else None
        K = getattr(self, "K", self.max_nt)
        if not hasattr(self, "ckpts"):
            self.ckpts = []
        qb = self._batch(self.qstate, True, ic=self.ic, dw=None)
        if ic32 is not None:
            qb.ic32 = ic32.data_ptr()
            qb.ic_window = self.ic_window
            self._ic32 = ic32
        # the adjoint sources of each run (the whole loop, or one checkpoint
        # segment) are evaluated before its first step: a step then reads one
        # value per node instead of gathering the node's receiver entries
        n_nodes = int(self.adj_inj[2].numel()) - 1
        if n_nodes > 0 and self.precompute_injection:
            slots = K if ((self.ckpts or binomial) and image) else self.max_nt
            qb.inj_pre = self._inj_pre_buffer(slots, n_nodes).data_ptr()
            qb.inj_node_shot = self.adj_inj[5].data_ptr()
            qb.n_inj_nodes = n_nodes
            qb.inj_pre_slots = slots
        self.adjoint_store = None
        if store_field:
            self.adjoint_store = self._field_store()
            qb.fstore = self.adjoint_store.data_ptr()
        if store_adjoint_dw:
            self.dw_adj = None
            self.dw_adj = torch.zeros((self.max_nt, self.total_n), dtype=self.tdtype, device=self.dev)
            qb.dw_out = self.dw_adj.data_ptr()
        if volume is not None:
            vm, vd = volume
            if tuple(vm.shape) != (self.total_n,) or tuple(vd.shape) != (self.max_nt, self.total_n) \
                    or vm.dtype != self.tdtype or vd.dtype != self.tdtype:
                raise ValueError("volume source arrays do not match the batch layout")
            qb.vol_m = vm.data_ptr()
            qb.vol_dw = vd.data_ptr()
            self._vol_keep = (vm, vd)
        if not image:
            qb.dw = None
            self._run(qb, 2, 0, self.max_nt, -1, 0, record=False)
        elif binomial:
            self._ensure_dw(1)
            qb.dw = self.dw.data_ptr()
            work = self._new_state()
            fb = self._batch(work, False, dw=self.dw)
            wb = self._batch(work, False)
            self.binomial_advances = 0
            self._treeverse(0, self.max_nt, self.snaps, None, work, fb, wb, qb)
        elif not self.ckpts:
            qb.dw = img_hist.data_ptr()
            self._dw_squares(0, self.max_nt, 0)
            self._run(qb, 2, 0, self.max_nt, -1, 0, record=False)
        else:
            levels = getattr(self, "levels", [K])
            self._ensure_dw(levels[-1])
            qb.dw = self.dw.data_ptr()
            work = self._new_state()
            fb = self._batch(work, False, dw=self.dw)
            wb = self._batch(work, False)
            for si in range(len(self.ckpts) - 1, -1, -1):
                j0 = si * K
                self._reverse(j0, min(self.max_nt, j0 + K), self.ckpts[si], 1, levels, work, fb, wb, qb)
        return self.ic

    def _treeverse(self, a, b, snaps, ckpt, work, fb, wb, qb):
        """The adjoint over slots [a, b) with the binomial schedule
        (``binomial_split``): the state at slot a is ``ckpt`` (None: rest, at
        slot 0) and ``snaps`` more snapshots may be taken.  Each reverse step
        restores or advances the state to its slot, recomputes that slot's
        dWaveOp into the one-slot store and runs the adjoint step on it."""
        while b - a > 0:
            if b - a == 1 or snaps == 0:
                k = b - 1
                self._restore_or_rest(ckpt, work, a)
                if k > a:
                    self._run(wb, 0, a, k, 1, a, record=False)
                    self.binomial_advances += k - a
                self._run(fb, 1, k, k + 1, 1, k, record=False)
                self._dw_squares(k, k + 1, k)
                self._run(qb, 2, k, k + 1, -1, k, record=False)
                b = k
                continue
            m = a + binomial_split(b - a, snaps)
            self._restore_or_rest(ckpt, work, a)
            self._run(wb, 0, a, m, 1, a, record=False)
            self.binomial_advances += m - a
            snap = self._ckpt_take(work, m)
            self._treeverse(m, b, snaps - 1, snap, work, fb, wb, qb)
            del snap
            b = m

    def _restore_or_rest(self, ckpt, work, slot):
        if ckpt is None:
            for f in work.values():
                f.zero_()
        else:
            self._ckpt_restore(ckpt, work, slot)

    def _reverse(self, j0, j1, ckpt, lvl, levels, work, fb, wb, qb):
        """The adjoint over slots [j0, j1) from the state checkpoint taken at
        j0: at the last level recompute the segment's dWaveOp into the store
        (unless it holds it) and run the adjoint over it; above, replay the
        segment keeping a checkpoint every levels[lvl] slots and reverse
        those sub-segments last to first."""
        if lvl == len(levels):
            if getattr(self, "_dw_seg", None) != (j0, j1):
                self._ckpt_restore(ckpt, work, j0)
                self._run(fb, 1, j0, j1, 1, j0, record=False)
                self._dw_seg = (j0, j1)
            self._dw_squares(j0, j1, j0)
            self._run(qb, 2, j0, j1, -1, j0, record=False)
            return
        Kn = levels[lvl]
        self._ckpt_restore(ckpt, work, j0)
        subs = []
        for s0 in range(j0, j1, Kn):
            subs.append((s0, min(j1, s0 + Kn), self._ckpt_take(work, s0)))
            if s0 + Kn < j1:
                self._run(wb, 0, s0, s0 + Kn, 1, s0, record=False)
        for s0, s1, c in reversed(subs):
            self._reverse(s0, s1, c, lvl + 1, levels, work, fb, wb, qb)
            c.clear()

    def _dw_squares(self, j0, j1, seg_base):
        """dw_sq += dWaveOp^2 of the imaged slots in [j0, j1) that the store
        holds from seg_base on (no-op unless migrate asked for it)."""
        if self.dw_sq is None:
            return
        p = self.imaging_period
        item = self.dtype.itemsize
        for pos in range(self.n_shots):
            ks = self.kshift[pos]
            first = max(j0, ks)
            s0 = first + (-(first - ks)) % p  # the first imaged slot at or after `first`
            if s0 >= j1:
                continue
            off = self.offs[pos]
            N.call("pysit_amd_slot_square_sum", ctypes.c_void_p(self.dw.data_ptr() + off * item),
                   1 if self.dtype == np.float64 else 0, self.total_n, s0 - seg_base, j1 - seg_base, p,
                   self.snpts[pos], ctypes.c_void_p(self.dw_sq.data_ptr() + off * 8), _stream())

    def dw_squares_unpadded(self, si_caller, pads):
        """Sum over the imaged steps of dWaveOp^2 of one shot, cropped to the
        physical grid (device FP64, the unpadded grid's shape)."""
        v = self._shot_view(self.dw_sq, self.order_idx.index(si_caller))
        sl = tuple(slice(l, n - r) for (l, r), n in zip(pads, v.shape))
        return v[sl]

    def _inj_pre_buffer(self, slots, n_nodes):
        """Device buffer for the precomputed injection of ``slots`` slots
        (kept between calls)."""
        need = slots * n_nodes
        buf = getattr(self, "_inj_pre", None)
        if buf is None or buf.numel() < need:
            self._inj_pre = None
            buf = self._inj_pre = torch.empty(need, dtype=self.tdtype, device=self.dev)
        return buf

    def probe_step_time(self, slots=16, reps=2):
        """Seconds per slot of the forward step storing dWaveOp (mode 1 with
        sources) plus the adjoint imaging step, without receiver sampling and
        adjoint sources, over the last ``slots`` slots (where every shot is
        active), from rest; the best of ``reps`` timed passes after one
        warm-up.  Works on scratch state: the batch's state, traces and
        stores are untouched."""
        S = max(1, min(int(slots), self.max_nt))
        j0 = self.max_nt - S
        state = self._new_state()
        qstate = self._new_state()
        dw = torch.empty((S, self.total_n), dtype=self.tdtype, device=self.dev)
        ic = torch.zeros(self.total_n, dtype=torch.float64, device=self.dev)
        fb = self._batch(state, False, dw=dw)
        fb.n_rec = 0
        qb = self._batch(qstate, False, ic=ic, dw=dw)
        qb.n_rec = 0
        qb.inj_ptr = qb.inj_node = qb.inj_nptr = qb.inj_chan = qb.inj_w = None
        ic32 = None
        if self.ic_window:
            ic32 = torch.zeros(self.total_n, dtype=torch.float32, device=self.dev)
            qb.ic32 = ic32.data_ptr()
            qb.ic_window = self.ic_window
        timer, self.timer = self.timer, None
        best = None
        try:
            for rep in range(int(reps) + 1):
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                torch.cuda.synchronize()
                ev[0].record()
                self._run(fb, 1, j0, self.max_nt, 1, j0, record=False)
                self._run(qb, 2, j0, self.max_nt, -1, j0, record=False)
                ev[1].record()
                torch.cuda.synchronize()
                if rep:
                    t = ev[0].elapsed_time(ev[1]) * 1e-3 / S
                    best = t if best is None else min(best, t)
        finally:
            self.timer = timer
        del state, qstate, dw, ic, ic32
        return best

    def image(self, si_caller, scale=1.0):
        """Padded FP64 image of one shot (caller order), a view of the accumulator."""
        return self._shot_view(self.ic, self.order_idx.index(si_caller))

    def fold_sum(self, pads, out, scale=1.0):
        """Add the fold of every shot's image onto ``out`` (the sum over the
        shots of :meth:`fold`).  Shots of one padded shape are summed on the
        padded grid first and folded once (the fold is linear)."""
        s0 = self.shots[0]
        if self.n_shots > 1 and all(s.shape == s0.shape for s in self.shots):
            img = self.ic.view(self.n_shots, -1).sum(dim=0)
            img = img.view(tuple(s0.shape[:-1]) + (self.zs[0],))[..., :s0.shape[-1]].contiguous()
            self._fold_grid(img, s0.shape, pads, out, None, scale, True)
        else:
            for i in range(self.n_shots):
                self.fold(i, pads, out, scale=scale, accumulate=True)

    def fold(self, si_caller, pads, out, out_origin=None, scale=1.0, accumulate=True):
        """Fold one shot's image onto ``out`` (device FP64 tensor holding the
        full unpadded model grid); ``out_origin`` is the shot window's first
        node in ``out`` (default 0)."""
        pos = self.order_idx.index(si_caller)
        img = self._shot_view(self.ic, pos).contiguous()  # the fold reads a dense padded grid
        self._fold_grid(img, self.shots[pos].shape, pads, out, out_origin, scale, accumulate)

    def _fold_grid(self, img, shape, pads, out, out_origin, scale, accumulate):
        origin = out_origin or (0,) * self.dim
        if self.dim == 2:
            (lx, rx), (lz, rz) = pads
            ly = ry = 0
            onz = out.shape[-1]
            ony = 1
            o = out.reshape(-1)[origin[0] * onz + origin[1]:]
            nx, nz = shape
            ny = 1
        else:
            (lx, rx), (ly, ry), (lz, rz) = pads
            ony, onz = out.shape[1], out.shape[2]
            o = out.reshape(-1)[(origin[0] * ony + origin[1]) * onz + origin[2]:]
            nx, ny, nz = shape
        N.call("pysit_amd_fold", self.dim, nx, ny, nz, lx, rx, ly, ry, lz, rz, _ptr(img), scale,
               _ptr(o), ony, onz, 1 if accumulate else 0, _stream())

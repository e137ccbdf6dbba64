"""ctypes binding of ``libpysit_amd.so`` (the C ABI in ``include/pysit_amd.h``).

The library is built in-tree by ``pysit_amd._build`` (``__graft_entry__.build``
does it).  Loading fails loudly when it is missing: there is no CPU fallback
for the solver path.
"""
import ctypes
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpysit_amd.so")

_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong
_d = ctypes.c_double


class Shot(ctypes.Structure):
    """``pysit_amd_shot``."""
    _fields_ = [
        ("nx", _i), ("ny", _i), ("nz", _i), ("zs", _i), ("nt", _i), ("kshift", _i),
        ("ntx", _i), ("nty", _i), ("ntz", _i),
        ("off", _ll),
        ("model", _vp), ("mny", _i), ("mnz", _i),
        ("mx_off", _i), ("mx_lo", _i), ("mx_hi", _i),
        ("my_off", _i), ("my_lo", _i), ("my_hi", _i),
        ("mz_off", _i), ("mz_lo", _i), ("mz_hi", _i),
        ("xl", _vp), ("xr", _vp), ("yl", _vp), ("yr", _vp), ("zl", _vp), ("zr", _vp),
        ("nxl", _i), ("nxr", _i), ("nyl", _i), ("nyr", _i), ("nzl", _i), ("nzr", _i),
        ("amp", _vp), ("namp", _i), ("nrec", _i), ("trace", _vp),
    ]


class Batch(ctypes.Structure):
    """``pysit_amd_batch``."""
    _fields_ = [
        ("dtype", _i), ("dim", _i), ("order", _i), ("tile_x", _i), ("tile_y", _i),
        ("dt", _d), ("dx", _d), ("dy", _d), ("dz", _d),
        ("n_shots", _i), ("shots", _vp),
        ("n_tiles", _i), ("tiles", _vp),
        ("inj_ptr", _vp), ("inj_node", _vp), ("inj_nptr", _vp), ("inj_chan", _vp), ("inj_w", _vp),
        ("n_rec", _i), ("recs", _vp), ("rec_idx", _vp), ("rec_w", _vp),
        ("u", _vp * 2), ("phix", _vp * 2), ("phiy", _vp * 2), ("phiz", _vp * 2), ("psi", _vp * 2),
        ("du", _vp),
        ("total_n", _ll),
        ("dw", _vp), ("ic", _vp),
        ("imaging_period", _i),
        ("fstore", _vp), ("born_m1", _vp),
        ("ic32", _vp), ("ic_window", _i),
        ("inj_pre", _vp), ("inj_node_shot", _vp), ("n_inj_nodes", _i), ("inj_pre_slots", _i),
        ("dw_form", _i),
        ("dw_out", _vp), ("vol_m", _vp), ("vol_dw", _vp),
    ]


class NativeError(RuntimeError):
    pass


_lib = None

# name -> (argtypes) ; all return int status
_STEP2D = [_i] + [_vp] * 6 + [_vp, _i] * 4 + [_d] * 3 + [_i] * 2 + [_vp] * 3 + [_vp]
_STEP3D = [_i] + [_vp] * 8 + [_vp, _i] * 6 + [_d] * 4 + [_i] * 3 + [_vp] * 5 + [_vp]
SIGNATURES = {
    "pysit_amd_abi_check": [_i, _i],
    "pysit_amd_tile_rows": [_i, _i, _i],
    "pysit_amd_tile_rows_supported": [_i, _i, _i],
    "pysit_amd_cda_step2d_f32": _STEP2D,
    "pysit_amd_cda_step2d_f64": _STEP2D,
    "pysit_amd_cda_step3d_f32": _STEP3D,
    "pysit_amd_cda_step3d_f64": _STEP3D,
    "pysit_amd_batch_run": [ctypes.POINTER(Batch), _i, _i, _i, _i, _i, _vp, _vp, _vp],
    "pysit_amd_batch_run_static": [ctypes.POINTER(Batch), _i, _i, _i, _i, _i, _vp, _vp, _vp],
    "pysit_amd_batch_run_persistent": [ctypes.POINTER(Batch), _i, _i, _i, _i, _i, _vp, _vp, _vp],
    "pysit_amd_fold": [_i] * 10 + [_vp, _d, _vp, _i, _i, _i, _vp],
    "pysit_amd_slot_square_sum": [_vp, _i, _ll, _i, _i, _i, _ll, _vp, _vp],
    "pysit_amd_plan_create": [_i, _i, ctypes.POINTER(_vp)],
    "pysit_amd_plan_run": [_vp, ctypes.POINTER(Batch), _i, _i, _i, _i, _i, _vp, _vp, _vp],
    "pysit_amd_plan_stats": [_vp, _vp],
    "pysit_amd_plan_destroy": [_vp],
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is missing: build it with `python -m pysit_amd._build` "
                "(there is no CPU fallback for the HIP solver path)")
        L = ctypes.CDLL(LIB_PATH)
        L.pysit_amd_last_error.restype = ctypes.c_char_p
        L.pysit_amd_last_error.argtypes = []
        L.pysit_amd_version.restype = ctypes.c_char_p
        L.pysit_amd_version.argtypes = []
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = _i
            f.argtypes = args
        if L.pysit_amd_abi_check(ctypes.sizeof(Shot), ctypes.sizeof(Batch)) != 0:
            raise NativeError(L.pysit_amd_last_error().decode())
        _lib = L
    return _lib


def tile_rows(dim, dtype_code, axis=0):
    """Tile rows along x (axis 0) or y (axis 1) the library was built with."""
    n = lib().pysit_amd_tile_rows(dim, dtype_code, axis)
    if n <= 0:
        raise NativeError(lib().pysit_amd_last_error().decode())
    return n


def tile_rows_supported(dim, dtype_code, tile_x):
    """Whether a batch may use ``tile_x`` rows along x."""
    n = lib().pysit_amd_tile_rows_supported(dim, dtype_code, int(tile_x))
    if n < 0:
        raise NativeError(lib().pysit_amd_last_error().decode())
    return n == 1


def call(name, *args):
    """Call an entry point and raise NativeError with the library's message on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != 0:
        raise NativeError(f"{name} failed ({rc}): {L.pysit_amd_last_error().decode()}")
    return rc


def exported_symbols():
    """Entry points declared in include/pysit_amd.h."""
    return ["pysit_amd_last_error", "pysit_amd_version"] + list(SIGNATURES)


class Plan(object):
    """A pysit_amd_plan: the batch time loop as replayed hipGraphs of
    ``graph_steps`` steps (at most ``max_graphs`` kept).  Freed with the
    object (or by ``close``)."""

    def __init__(self, graph_steps, max_graphs=16):
        h = _vp()
        call("pysit_amd_plan_create", int(graph_steps), int(max_graphs), ctypes.byref(h))
        self.graph_steps = int(graph_steps)
        self._h = h
        self._fin = weakref.finalize(self, _plan_destroy, h.value)

    def run(self, b, mode, j0, j1, direction, seg_base, active_tiles, active_recs, stream):
        call("pysit_amd_plan_run", self._h, ctypes.byref(b), mode, j0, j1, direction, seg_base, active_tiles,
             active_recs, stream)

    def stats(self):
        """{'graph_launches', 'captures', 'direct_steps', 'cached'}"""
        out = (ctypes.c_longlong * 4)()
        call("pysit_amd_plan_stats", self._h, ctypes.cast(out, _vp))
        return dict(zip(("graph_launches", "captures", "direct_steps", "cached"), list(out)))

    def close(self):
        self._fin()


def _plan_destroy(handle):
    if handle and _lib is not None:
        _lib.pysit_amd_plan_destroy(handle)

"""ctypes binding of the C ABI in include/ssegpu.h (libssegpu.so).

This is the Python stand-in for the Julia ``ccall`` layer: it marshals a
FlatModel and walker parameters into the plain structs of the ABI and loads
the in-tree HIP library.  There is no CPU fallback: if the library is not
built, or no gfx950 device is present, calls fail with an error.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libssegpu.so")

SSE_MODE_WAVE = 0
SSE_MODE_LANE = 1
SSE_OPT_GRAPH = 1
SSE_OPT_TIMING = 2
SSE_OPT_WORM_RECORDS = 3
SSE_OPT_PERSISTENT = 4
SSE_OPT_STREAMS = 5
SSE_OPT_K2_LDS = 6
SSE_OPT_ORDER = 7
SSE_OPT_CHASE = 8
SSE_OPT_TUNE = 9
SSE_OPT_FUSE_MEASURE = 10
SSE_OPT_LANE_WIDTH = 11
SSE_OPT_GENERIC = 12

STATUS = {0: "SSE_OK", 1: "SSE_ERR_INVALID", 2: "SSE_ERR_HIP", 3: "SSE_ERR_OOM",
          4: "SSE_ERR_SCATTER", 5: "SSE_ERR_NO_DEVICE", 6: "SSE_ERR_STATE"}


MAX_BOND_SITES = 4


class VertexDataDesc(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("dims", C.c_int32 * MAX_BOND_SITES),
                ("energy_offset", C.c_double),
                ("n_vertices", C.c_int32), ("weights", C.POINTER(C.c_double)),
                ("signs", C.POINTER(C.c_int8)), ("leg_states", C.POINTER(C.c_uint8)),
                ("diagonal_vertices", C.POINTER(C.c_int32)), ("max_worm_count", C.c_int32),
                ("transition_offset", C.POINTER(C.c_int32)),
                ("transition_count", C.POINTER(C.c_int32)), ("n_exits", C.c_int32),
                ("exit_cumprob", C.POINTER(C.c_double)), ("exit_target", C.POINTER(C.c_int32)),
                ("exit_leg", C.POINTER(C.c_uint8)), ("exit_worm", C.POINTER(C.c_uint8))]


class ModelDesc(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("site_dims", C.POINTER(C.c_uint8)),
                ("n_bonds", C.c_int32), ("bond_n_sites", C.c_int32),
                ("bond_type", C.POINTER(C.c_int32)),
                ("bond_sites", C.POINTER(C.c_int32)), ("n_types", C.c_int32),
                ("types", C.POINTER(VertexDataDesc)), ("energy_offset", C.c_double),
                ("normalization_site_count", C.c_double), ("n_estimators", C.c_int32),
                ("est_max_dim", C.c_int32), ("est_coef", C.POINTER(C.c_double)),
                ("est_mval", C.POINTER(C.c_double))]


class WalkerDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("diagonal_warmup_sweeps", C.c_int32),
                ("T", C.c_double), ("seed", C.c_uint64), ("init_opstring_cutoff", C.c_int64),
                ("init_num_worms", C.c_double), ("num_worms_attenuation_factor", C.c_double),
                ("target_worm_length_fraction", C.c_double), ("thermalization", C.c_int64)]


class WalkerScalars(C.Structure):
    _fields_ = [("cutoff", C.c_int64), ("num_operators", C.c_int64),
                ("avg_worm_length", C.c_double), ("num_worms", C.c_double),
                ("rng_counter", C.c_uint64), ("sweeps", C.c_int64),
                ("last_scatter_calls", C.c_int64), ("last_total_worm_length", C.c_double),
                ("last_worm_length_fraction", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class ModelDescs:
    """Array of ModelDesc built from FlatModels; holds every numpy buffer the
    structs point into."""

    def __init__(self, flats):
        self.flats = list(flats)
        self.keep = []
        self.array = (ModelDesc * len(self.flats))()
        for k, fm in enumerate(self.flats):
            self.array[k] = self._model(fm)

    def _arr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.size == 0:
            a = np.zeros(1, dtype=dtype)
        self.keep.append(a)
        return a

    def _model(self, fm):
        types = (VertexDataDesc * len(fm.types))()
        self.keep.append(types)
        for t, d in enumerate(fm.types):
            vd = types[t]
            vd.n_sites = len(d["dims"])
            for k, dk in enumerate(d["dims"]):
                vd.dims[k] = dk
            vd.energy_offset = d["energy_offset"]
            vd.n_vertices = d["n_vertices"]
            vd.weights = _ptr(self._arr(d["weights"], np.float64), C.c_double)
            vd.signs = _ptr(self._arr(d["signs"], np.int8), C.c_int8)
            vd.leg_states = _ptr(self._arr(d["leg_states"], np.uint8), C.c_uint8)
            vd.diagonal_vertices = _ptr(self._arr(d["diagonal_vertices"], np.int32), C.c_int32)
            vd.max_worm_count = d["max_worm_count"]
            vd.transition_offset = _ptr(self._arr(d["transition_offset"], np.int32), C.c_int32)
            vd.transition_count = _ptr(self._arr(d["transition_count"], np.int32), C.c_int32)
            vd.n_exits = d["n_exits"]
            vd.exit_cumprob = _ptr(self._arr(d["exit_cumprob"], np.float64), C.c_double)
            vd.exit_target = _ptr(self._arr(d["exit_target"], np.int32), C.c_int32)
            vd.exit_leg = _ptr(self._arr(d["exit_leg"], np.uint8), C.c_uint8)
            vd.exit_worm = _ptr(self._arr(d["exit_worm"], np.uint8), C.c_uint8)
        m = ModelDesc()
        m.n_sites = fm.n_sites
        m.site_dims = _ptr(self._arr(fm.site_dims, np.uint8), C.c_uint8)
        m.n_bonds = fm.n_bonds
        m.bond_n_sites = fm.bond_n_sites
        m.bond_type = _ptr(self._arr(fm.bond_type, np.int32), C.c_int32)
        m.bond_sites = _ptr(self._arr(fm.bond_sites, np.int32), C.c_int32)
        m.n_types = len(fm.types)
        m.types = C.cast(types, C.POINTER(VertexDataDesc))
        m.energy_offset = fm.energy_offset
        m.normalization_site_count = fm.normalization_site_count
        m.n_estimators = fm.n_estimators
        m.est_max_dim = fm.est_max_dim
        m.est_coef = _ptr(self._arr(fm.est_coef, np.float64), C.c_double)
        m.est_mval = _ptr(self._arr(fm.est_mval, np.float64), C.c_double)
        return m


def walker_desc(model=0, T=1.0, seed=0, init_opstring_cutoff=-1, diagonal_warmup_sweeps=5,
                init_num_worms=5.0, num_worms_attenuation_factor=0.01,
                target_worm_length_fraction=2.0, thermalization=0):
    w = WalkerDesc()
    w.model = model
    w.T = T
    w.seed = seed
    w.init_opstring_cutoff = init_opstring_cutoff
    w.diagonal_warmup_sweeps = diagonal_warmup_sweeps
    w.init_num_worms = init_num_worms
    w.num_worms_attenuation_factor = num_worms_attenuation_factor
    w.target_worm_length_fraction = target_worm_length_fraction
    w.thermalization = thermalization
    return w


def walker_array(walkers):
    arr = (WalkerDesc * len(walkers))()
    for k, w in enumerate(walkers):
        arr[k] = w
    return arr


def default_cutoff(n_sites, T):
    """round(Int, N*T) with ties to even (sse.jl:50-51)."""
    return int(round(n_sites * T))


_LIB = None


def load_library(path=None):
    """Load libssegpu.so from the package directory and declare signatures."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libssegpu.so not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(p)
    vp = C.c_void_p
    lib.sse_batch_create.argtypes = [C.POINTER(ModelDesc), C.c_int32, C.POINTER(WalkerDesc),
                                     C.c_int32, C.c_int32, C.POINTER(vp)]
    lib.sse_batch_destroy.argtypes = [vp]
    lib.sse_init.argtypes = [vp]
    lib.sse_sweep.argtypes = [vp, C.c_int32]
    lib.sse_measure.argtypes = [vp, C.POINTER(C.c_double)]
    lib.sse_measure_width.argtypes = [vp, C.c_int32]
    lib.sse_measure_width.restype = C.c_int32
    lib.sse_run.argtypes = [vp, C.c_int32, C.c_int32]
    lib.sse_run_budgets.argtypes = [vp, C.POINTER(C.c_int64), C.c_int32]
    lib.sse_run_budgets_noalloc.argtypes = [vp, C.POINTER(C.c_int64), C.c_int32]
    lib.sse_tuning.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                               C.POINTER(C.c_int32)]
    lib.sse_read_accumulators.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int64)]
    lib.sse_reset_accumulators.argtypes = [vp]
    lib.sse_read_worm_fraction.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_int64)]
    lib.sse_get_walker.argtypes = [vp, C.c_int32, C.POINTER(WalkerScalars),
                                   C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_uint8)]
    lib.sse_set_walker.argtypes = [vp, C.c_int32, C.POINTER(WalkerScalars),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)]
    lib.sse_sweep_phases.argtypes = [vp, C.c_int32]
    lib.sse_get_vertex_list.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.c_int64,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.sse_phase_times.argtypes = [vp, C.POINTER(C.c_double)]
    lib.sse_kernel_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.sse_call_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.sse_lds_bytes.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int32)]
    lib.sse_work_counters.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.sse_opcode_to_reference.argtypes = [C.c_uint32]
    lib.sse_opcode_to_reference.restype = C.c_uint64
    lib.sse_opcode_from_reference.argtypes = [C.c_uint64]
    lib.sse_opcode_from_reference.restype = C.c_uint32
    lib.sse_add_model.argtypes = [vp, C.POINTER(ModelDesc), C.POINTER(C.c_int32)]
    lib.sse_slot_start.argtypes = [vp, C.c_int32, C.POINTER(WalkerDesc), C.c_int64]
    lib.sse_slot_stop.argtypes = [vp, C.c_int32]
    lib.sse_set_option.argtypes = [vp, C.c_int32, C.c_int64]
    lib.sse_slot_poll.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]
    lib.sse_pool_usage.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.sse_pt_log_weight_ratio.argtypes = [vp, C.c_int32, C.c_double, C.POINTER(C.c_double)]
    lib.sse_pt_change_parameter.argtypes = [vp, C.c_int32, C.c_double]
    lib.sse_pt_get_parameter.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    lib.sse_pt_set_ladders.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_uint64,
                                       C.c_int32]
    lib.sse_pt_exchange.argtypes = [vp]
    lib.sse_pt_get_ladders.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    i32p, f64p, i64p = C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    lib.sse_vertex_transitions.argtypes = [C.c_int32, i32p, C.c_int32, C.POINTER(C.c_uint8),
                                           f64p, C.c_double, C.c_double, i32p, i32p, C.c_int64,
                                           f64p, i32p, i32p, i32p, i64p, i32p]
    lib.sse_last_error.restype = C.c_char_p
    lib.sse_abi_version.restype = C.c_int32
    if path is None:
        _LIB = lib
    return lib


EXPORTED_SYMBOLS = (
    "sse_batch_create", "sse_batch_destroy", "sse_init", "sse_sweep", "sse_measure",
    "sse_measure_width", "sse_run", "sse_read_accumulators", "sse_reset_accumulators",
    "sse_get_walker", "sse_set_walker", "sse_sweep_phases", "sse_get_vertex_list",
    "sse_phase_times", "sse_work_counters", "sse_opcode_to_reference",
    "sse_opcode_from_reference", "sse_last_error", "sse_abi_version", "sse_add_model",
    "sse_slot_start", "sse_slot_stop", "sse_slot_poll", "sse_pool_usage", "sse_set_option",
    "sse_call_time", "sse_lds_bytes", "sse_pt_log_weight_ratio", "sse_pt_change_parameter",
    "sse_pt_get_parameter", "sse_pt_set_ladders", "sse_pt_exchange", "sse_pt_get_ladders",
    "sse_vertex_transitions", "sse_read_worm_fraction", "sse_run_budgets",
    "sse_run_budgets_noalloc", "sse_tuning", "sse_kernel_times")


def vertex_transitions(dims, leg_states, weights, tolerance, lp_tolerance):
    """Native transition tables (sse_vertex_transitions).  Returns
    (offset, length, cumprobs, targets, out_leg, out_worm, flags) with
    offset/length shaped [n_vertices][max_worm][2 * n_sites]."""
    lib = load_library()
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    ls = np.ascontiguousarray(leg_states, dtype=np.uint8)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    ns, nv = len(dims), len(w)
    mw = max(1, int(dims.max()) - 1)
    legs = 2 * ns
    n_steps = sum(max(0, int(d) - 1) for d in dims) * 2
    cap = max(1, nv * n_steps ** 3)  # <= nv * n_steps LPs, each <= n_steps^2 exits
    off = np.zeros((nv, mw, legs), dtype=np.int32)
    ln = np.zeros((nv, mw, legs), dtype=np.int32)
    cum = np.zeros(cap)
    tg, ol, ow = (np.zeros(cap, dtype=np.int32) for _ in range(3))
    n_exits, flags = C.c_int64(), C.c_int32()
    check(lib, lib.sse_vertex_transitions(
        ns, _ptr(dims, C.c_int32), nv, _ptr(ls, C.c_uint8), _ptr(w, C.c_double), tolerance,
        lp_tolerance, _ptr(off, C.c_int32), _ptr(ln, C.c_int32), cap, _ptr(cum, C.c_double),
        _ptr(tg, C.c_int32), _ptr(ol, C.c_int32), _ptr(ow, C.c_int32), C.byref(n_exits),
        C.byref(flags)))
    k = n_exits.value
    return off, ln, cum[:k], tg[:k], ol[:k], ow[:k], flags.value


class SSEError(RuntimeError):
    pass


def check(lib, status):
    if status != 0:
        msg = lib.sse_last_error()
        raise SSEError(f"{STATUS.get(status, status)}: {msg.decode() if msg else ''}")
    return status


class Batch:
    """A batch of walkers resident on the GPU (owns one sse_batch handle)."""

    def __init__(self, flats, walkers, mode=SSE_MODE_WAVE, lib=None):
        self.lib = lib or load_library()
        self.models = ModelDescs(flats)
        self.walkers = list(walkers)
        self.flats = self.models.flats
        warr = walker_array(self.walkers)
        h = C.c_void_p()
        check(self.lib, self.lib.sse_batch_create(self.models.array, len(self.flats), warr,
                                                  len(self.walkers), mode, C.byref(h)))
        self.handle = h
        self.mode = mode
        self._slot_widths = [self.flats[w.model].measure_width for w in self.walkers]

    def close(self):
        if self.handle:
            self.lib.sse_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_walkers(self):
        return len(self.walkers)

    def init(self):
        check(self.lib, self.lib.sse_init(self.handle))

    def sweep(self, n=1):
        check(self.lib, self.lib.sse_sweep(self.handle, n))

    def sweep_phases(self, phases):
        check(self.lib, self.lib.sse_sweep_phases(self.handle, phases))

    def width(self, w=0):
        return self.lib.sse_measure_width(self.handle, w)

    def measure(self):
        widths = self._slot_widths
        total = sum(widths)
        out = np.zeros(total, dtype=np.float64)
        check(self.lib, self.lib.sse_measure(self.handle, _ptr(out, C.c_double)))
        res, k = [], 0
        for w, wd in enumerate(widths):
            res.append(out[k:k + self.flats[self.walkers[w].model].measure_width])
            k += wd
        return res

    def run(self, n_sweeps, measure=True):
        check(self.lib, self.lib.sse_run(self.handle, n_sweeps, int(bool(measure))))

    def run_budgets(self, sweeps, measure=True):
        """Walker w advances by sweeps[w] sweeps in one persistent launch."""
        sw = np.ascontiguousarray(sweeps, dtype=np.int64)
        if sw.shape != (len(self.walkers),):
            raise ValueError("one sweep count per walker")
        check(self.lib, self.lib.sse_run_budgets(self.handle, _ptr(sw, C.c_int64),
                                                 int(bool(measure))))

    def run_budgets_noalloc(self, sweeps, measure=True):
        """run_budgets through the entry point that allocates no device
        memory (static per-GPU scratch; at most SSE_NOALLOC_MAX_WALKERS)."""
        sw = np.ascontiguousarray(sweeps, dtype=np.int64)
        if sw.shape != (len(self.walkers),):
            raise ValueError("one sweep count per walker")
        check(self.lib, self.lib.sse_run_budgets_noalloc(self.handle, _ptr(sw, C.c_int64),
                                                         int(bool(measure))))

    def accumulators(self):
        widths = self._slot_widths
        total = sum(widths)
        s = np.zeros(total)
        s2 = np.zeros(total)
        cnt = np.zeros(len(self.walkers), dtype=np.int64)
        check(self.lib, self.lib.sse_read_accumulators(self.handle, _ptr(s, C.c_double),
                                                       _ptr(s2, C.c_double),
                                                       _ptr(cnt, C.c_int64)))
        out, k = [], 0
        for w, wd in enumerate(widths):
            cw = self.flats[self.walkers[w].model].measure_width
            out.append((s[k:k + cw], s2[k:k + cw], int(cnt[w])))
            k += wd
        return out

    def accumulators_raw(self, with_sumsq=True):
        """(sums, sums of squares or None, counts) as flat arrays: walker w's
        row starts at slot_offsets()[w]."""
        total = sum(self._slot_widths)
        s = np.zeros(total)
        s2 = np.zeros(total) if with_sumsq else None
        cnt = np.zeros(len(self.walkers), dtype=np.int64)
        check(self.lib, self.lib.sse_read_accumulators(
            self.handle, _ptr(s, C.c_double), _ptr(s2, C.c_double) if with_sumsq else None,
            _ptr(cnt, C.c_int64)))
        return s, s2, cnt

    def slot_offsets(self):
        return np.concatenate([[0], np.cumsum(self._slot_widths)[:-1]]).astype(np.int64)

    def worm_fractions_raw(self):
        n = len(self.walkers)
        s, s2 = np.zeros(n), np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        check(self.lib, self.lib.sse_read_worm_fraction(self.handle, _ptr(s, C.c_double),
                                                        _ptr(s2, C.c_double),
                                                        _ptr(cnt, C.c_int64)))
        return s, s2, cnt

    def worm_fractions(self):
        """Per walker (sum, sum of squares, count) of WormLengthFraction over
        the measured sweeps since the last reset (sse.jl:200-202)."""
        n = len(self.walkers)
        s, s2 = np.zeros(n), np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        check(self.lib, self.lib.sse_read_worm_fraction(self.handle, _ptr(s, C.c_double),
                                                        _ptr(s2, C.c_double),
                                                        _ptr(cnt, C.c_int64)))
        return [(float(s[k]), float(s2[k]), int(cnt[k])) for k in range(n)]

    def reset_accumulators(self):
        check(self.lib, self.lib.sse_reset_accumulators(self.handle))

    def get_walker(self, w, with_ops=True, with_spins=True):
        s = WalkerScalars()
        check(self.lib, self.lib.sse_get_walker(self.handle, w, C.byref(s), None, 0, None))
        ops = np.zeros(max(1, s.cutoff), dtype=np.uint32) if with_ops else None
        n_sites = self.flats[self.walkers[w].model].n_sites
        spins = np.zeros(n_sites, dtype=np.uint8) if with_spins else None
        check(self.lib, self.lib.sse_get_walker(
            self.handle, w, C.byref(s),
            _ptr(ops, C.c_uint32) if ops is not None else None, len(ops) if ops is not None else 0,
            _ptr(spins, C.c_uint8) if spins is not None else None))
        if ops is not None:
            ops = ops[:s.cutoff]
        return s.as_dict(), ops, spins

    def set_walker(self, w, scalars, ops, spins):
        s = WalkerScalars()
        for k, _ in WalkerScalars._fields_:
            if k in scalars:
                setattr(s, k, scalars[k])
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        spins = np.ascontiguousarray(spins, dtype=np.uint8)
        s.cutoff = len(ops)
        check(self.lib, self.lib.sse_set_walker(self.handle, w, C.byref(s), _ptr(ops, C.c_uint32),
                                                _ptr(spins, C.c_uint8)))

    def vertex_list(self, w):
        sc, _, _ = self.get_walker(w, with_ops=False, with_spins=False)
        L = sc["cutoff"]
        flat = self.flats[self.walkers[w].model]
        n_sites, legs = flat.n_sites, 2 * flat.bond_n_sites
        links = np.zeros(max(1, legs * L), dtype=np.int32)
        first = np.zeros(n_sites, dtype=np.int32)
        last = np.zeros(n_sites, dtype=np.int32)
        check(self.lib, self.lib.sse_get_vertex_list(self.handle, w, _ptr(links, C.c_int32),
                                                     len(links), _ptr(first, C.c_int32),
                                                     _ptr(last, C.c_int32)))
        return links[:legs * L], first, last

    # ---- serving slots
    def add_model(self, flat):
        descs = ModelDescs([flat])
        self._extra_models = getattr(self, "_extra_models", [])
        self._extra_models.append(descs)
        idx = C.c_int32()
        check(self.lib, self.lib.sse_add_model(self.handle, descs.array, C.byref(idx)))
        self.flats.append(flat)
        return idx.value

    def slot_start(self, slot, walker, n_sweeps):
        check(self.lib, self.lib.sse_slot_start(self.handle, slot, C.byref(walker), int(n_sweeps)))
        self.walkers[slot] = walker

    def set_option(self, option, value):
        check(self.lib, self.lib.sse_set_option(self.handle, option, int(value)))

    def slot_stop(self, slot):
        check(self.lib, self.lib.sse_slot_stop(self.handle, slot))

    def slot_poll(self):
        buf = (C.c_int32 * self.n_walkers)()
        n = C.c_int32()
        check(self.lib, self.lib.sse_slot_poll(self.handle, buf, self.n_walkers, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def slot_result(self, slot):
        """(sums, sums of squares, measurement count) of one slot."""
        return self.accumulators()[slot]

    def pool_usage(self):
        u, t = C.c_int64(), C.c_int64()
        check(self.lib, self.lib.sse_pool_usage(self.handle, C.byref(u), C.byref(t)))
        return u.value, t.value

    def phase_times(self):
        t = np.zeros(3)
        check(self.lib, self.lib.sse_phase_times(self.handle, _ptr(t, C.c_double)))
        return t

    def kernel_times(self):
        """(ms sums, launch counts) per kernel of the last call with
        SSE_OPT_TIMING 2: diag_vlist, worm, measure, persistent."""
        ms = np.zeros(4)
        n = np.zeros(4, dtype=np.int64)
        check(self.lib, self.lib.sse_kernel_times(self.handle, _ptr(ms, C.c_double),
                                                  _ptr(n, C.c_int64)))
        return ms, n

    def call_time(self):
        """(device ms, sweeps) of the last sweep/run call (HIP events)."""
        ms, sw = C.c_double(), C.c_int64()
        check(self.lib, self.lib.sse_call_time(self.handle, C.byref(ms), C.byref(sw)))
        return ms.value, sw.value

    def tuning(self, which=0):
        """The autotuner's choice for sse_run (0) or sse_run_budgets (1) as a
        dict, or None while it is still measuring."""
        p, g, ch, cand = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        lds, rate = C.c_int64(), C.c_double()
        st = self.lib.sse_tuning(self.handle, which, C.byref(p), C.byref(g), C.byref(ch),
                                 C.byref(lds), C.byref(rate), C.byref(cand))
        if st == 6:
            return None
        check(self.lib, st)
        return {"persistent": bool(p.value), "groups": g.value, "chase": ch.value,
                "k2_lds": lds.value, "vertex_updates_per_s": rate.value,
                "candidates_measured": cand.value}

    def lds_bytes(self):
        """(K1 bytes, K2 bytes, K2 walkers per CU) of the current launches."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int32()
        check(self.lib, self.lib.sse_lds_bytes(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    # ---- parallel tempering in T (src/sse.jl:390-405 + the exchange loop)
    def pt_log_weight_ratio(self, walker, new_T):
        """Carlo.parallel_tempering_log_weight_ratio(mc, :T, new_T): -n log(new_T/T)."""
        r = C.c_double()
        check(self.lib, self.lib.sse_pt_log_weight_ratio(self.handle, walker, new_T, C.byref(r)))
        return r.value

    def pt_change_parameter(self, walker, new_T):
        """Carlo.parallel_tempering_change_parameter!(mc, :T, new_T)."""
        check(self.lib, self.lib.sse_pt_change_parameter(self.handle, walker, new_T))

    def pt_get_parameter(self, walker):
        t = C.c_double()
        check(self.lib, self.lib.sse_pt_get_parameter(self.handle, walker, C.byref(t)))
        return t.value

    def pt_set_ladders(self, ladders, seed, interval=1):
        """ladders: n_ladders x ladder_len walker indices in temperature-slot
        order (an empty list removes the ladders)."""
        lad = np.ascontiguousarray(np.asarray(ladders, dtype=np.int32).reshape(len(ladders), -1)
                                   if len(ladders) else np.zeros((0, 0), np.int32))
        self._pt_shape = lad.shape
        check(self.lib, self.lib.sse_pt_set_ladders(self.handle, _ptr(lad, C.c_int32),
                                                    lad.shape[0], lad.shape[1] if lad.size else 0,
                                                    seed, interval))

    def pt_exchange(self):
        check(self.lib, self.lib.sse_pt_exchange(self.handle))

    def pt_ladders(self):
        """(slot_walkers, accepted, proposed) as n_ladders x ladder_len arrays,
        and the number of rounds run."""
        shape = getattr(self, "_pt_shape", (0, 0))
        sw = np.zeros(shape, dtype=np.int32)
        acc = np.zeros(shape, dtype=np.int64)
        prop = np.zeros(shape, dtype=np.int64)
        r = C.c_int64()
        check(self.lib, self.lib.sse_pt_get_ladders(self.handle, _ptr(sw, C.c_int32),
                                                    _ptr(acc, C.c_int64), _ptr(prop, C.c_int64),
                                                    C.byref(r)))
        return sw, acc, prop, r.value

    def work_counters(self):
        a = [C.c_int64() for _ in range(4)]
        check(self.lib, self.lib.sse_work_counters(self.handle, *[C.byref(x) for x in a]))
        return tuple(x.value for x in a)

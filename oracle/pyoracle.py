"""ctypes binding of oracle/libsseoracle.so (test infrastructure only).

The oracle is the checker: tests compare the GPU path against it, smoke()
runs one tiny case against it, bench.py times it as the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SSE_ORACLE_LIB selects another build of the same source (the sanitizer
# build of run_sanitized.sh)
LIB = os.environ.get("SSE_ORACLE_LIB") or os.path.join(HERE, "libsseoracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib(capi):
    """Load the oracle library; `capi` supplies the shared data-format structs."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.so_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]
    L.so_philox_word.argtypes = [C.c_uint64, C.c_uint64]
    L.so_philox_word.restype = C.c_uint64
    L.so_word_to_float.argtypes = [C.c_uint64]
    L.so_word_to_float.restype = C.c_double
    L.so_range.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), C.c_uint64]
    L.so_range.restype = C.c_uint64
    L.so_expm1.argtypes = [C.c_double]
    L.so_expm1.restype = C.c_double
    L.so_tanh.argtypes = [C.c_double]
    L.so_tanh.restype = C.c_double
    L.so_walker_create.argtypes = [C.POINTER(capi.ModelDesc), C.POINTER(capi.WalkerDesc)]
    L.so_walker_create.restype = vp
    L.so_walker_destroy.argtypes = [vp]
    L.so_init.argtypes = [vp]
    L.so_sweep.argtypes = [vp]
    L.so_worm_traverse_from.argtypes = [vp, C.c_int64, C.c_int, C.c_int]
    L.so_worm_traverse_from.restype = C.c_int64
    L.so_sweep_phases.argtypes = [vp, C.c_int]
    L.so_measure.argtypes = [vp, C.POINTER(C.c_double)]
    L.so_measure_width.argtypes = [vp]
    L.so_get_scalars.argtypes = [vp, C.POINTER(capi.WalkerScalars)]
    L.so_get_ops.argtypes = [vp, C.POINTER(C.c_uint32), C.c_int64]
    L.so_get_ops.restype = C.c_int64
    L.so_get_spins.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.so_set_state.argtypes = [vp, C.POINTER(capi.WalkerScalars), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint8)]
    L.so_get_vertex_list.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]
    L.so_make_vertex_list.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.so_run_timed.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    L.so_run_timed.restype = C.c_double
    L.so_run_series.argtypes = [vp, C.c_int64, C.POINTER(C.c_double)]
    L.so_run_series.restype = C.c_int64
    L.so_work_counters.argtypes = [vp, C.POINTER(C.c_int64)]
    L.so_log.argtypes = [C.c_double]
    L.so_log.restype = C.c_double
    L.so_pt_log_weight_ratio.argtypes = [vp, C.c_double]
    L.so_pt_log_weight_ratio.restype = C.c_double
    L.so_pt_change_parameter.argtypes = [vp, C.c_double]
    L.so_pt_get_T.argtypes = [vp]
    L.so_pt_get_T.restype = C.c_double
    L.so_pt_uniform.argtypes = [C.c_uint64, C.c_int32, C.c_int64, C.c_int32]
    L.so_pt_uniform.restype = C.c_double
    L.so_pt_exchange.argtypes = [C.POINTER(vp), C.POINTER(C.c_int32), C.c_int32, C.c_uint64,
                                 C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleWalker:
    """One walker of the CPU restatement (reference order of operations)."""

    def __init__(self, capi, flat, walker):
        self.capi = capi
        self.L = lib(capi)
        self.descs = capi.ModelDescs([flat])
        self.wd = walker
        self.flat = flat
        self.h = self.L.so_walker_create(self.descs.array, C.byref(self.wd))

    def __del__(self):
        try:
            if self.h:
                self.L.so_walker_destroy(self.h)
        except Exception:
            pass

    def init(self):
        self.L.so_init(self.h)

    def sweep(self, n=1):
        for _ in range(n):
            err = self.L.so_sweep(self.h)
            if err:
                raise RuntimeError(f"oracle sweep error {err}")

    def worm_traverse_from(self, p0, l0, wf0):
        """One worm from (slot p0, leg l0, worm wf0), all 0-based but the
        1-based worm, over a freshly built vertex list; returns its length."""
        return self.L.so_worm_traverse_from(self.h, p0, l0, wf0)

    def sweep_phases(self, phases):
        return self.L.so_sweep_phases(self.h, phases)

    def measure(self):
        out = np.zeros(self.flat.measure_width)
        self.L.so_measure(self.h, _p(out, C.c_double))
        return out

    def scalars(self):
        s = self.capi.WalkerScalars()
        self.L.so_get_scalars(self.h, C.byref(s))
        return s.as_dict()

    def ops(self):
        s = self.scalars()
        out = np.zeros(max(1, s["cutoff"]), dtype=np.uint32)
        self.L.so_get_ops(self.h, _p(out, C.c_uint32), len(out))
        return out[:s["cutoff"]]

    def spins(self):
        out = np.zeros(self.flat.n_sites, dtype=np.uint8)
        self.L.so_get_spins(self.h, _p(out, C.c_uint8))
        return out

    def state(self):
        return self.scalars(), self.ops(), self.spins()

    def set_state(self, scalars, ops, spins):
        s = self.capi.WalkerScalars()
        for k, _ in self.capi.WalkerScalars._fields_:
            if k in scalars:
                setattr(s, k, scalars[k])
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        spins = np.ascontiguousarray(spins, dtype=np.uint8)
        s.cutoff = len(ops)
        return self.L.so_set_state(self.h, C.byref(s), _p(ops, C.c_uint32), _p(spins, C.c_uint8))

    def vertex_list(self):
        L = self.scalars()["cutoff"]
        legs = 2 * self.flat.bond_n_sites
        links = np.zeros(max(1, legs * L), dtype=np.int64)
        first = np.zeros(self.flat.n_sites, dtype=np.int64)
        last = np.zeros(self.flat.n_sites, dtype=np.int64)
        self.L.so_get_vertex_list(self.h, _p(links, C.c_int64), _p(first, C.c_int64),
                                  _p(last, C.c_int64))
        return links[:legs * L], first, last

    def run_series(self, n_sweeps):
        out = np.zeros((n_sweeps, self.flat.measure_width))
        rows = self.L.so_run_series(self.h, n_sweeps, _p(out, C.c_double))
        if rows < 0:
            raise RuntimeError("oracle sweep error")
        return out[:rows]

    def work_counters(self):
        out = np.zeros(4, dtype=np.int64)
        self.L.so_work_counters(self.h, _p(out, C.c_int64))
        return tuple(int(x) for x in out)

    def pt_log_weight_ratio(self, new_T):
        """Carlo.parallel_tempering_log_weight_ratio (sse.jl:390-396)."""
        return self.L.so_pt_log_weight_ratio(self.h, new_T)

    def pt_change_parameter(self, new_T):
        """Carlo.parallel_tempering_change_parameter! (sse.jl:398-405)."""
        self.L.so_pt_change_parameter(self.h, new_T)

    @property
    def T(self):
        return self.L.so_pt_get_T(self.h)

    def run_timed(self, n_sweeps, measure=True):
        sums = np.zeros(self.flat.measure_width)
        t = self.L.so_run_timed(self.h, n_sweeps, int(measure), _p(sums, C.c_double))
        return t, sums


def make_vertex_list(capi, n_sites, bond_sites, ops):
    """make_vertex_list! of the oracle; bond_sites is [n_bonds][NSites]."""
    L_ = lib(capi)
    ops = np.ascontiguousarray(ops, dtype=np.uint32)
    bs = np.ascontiguousarray(bond_sites, dtype=np.int32)
    ns = bs.shape[1] if bs.ndim == 2 else 2
    legs = 2 * ns
    L = len(ops)
    links = np.zeros(max(1, legs * L), dtype=np.int64)
    first = np.zeros(n_sites, dtype=np.int64)
    last = np.zeros(n_sites, dtype=np.int64)
    L_.so_make_vertex_list(n_sites, ns, _p(bs, C.c_int32), _p(ops, C.c_uint32), L,
                           _p(links, C.c_int64), _p(first, C.c_int64), _p(last, C.c_int64))
    return links[:legs * L], first, last


def philox_block(capi, ctr, key):
    L_ = lib(capi)
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L_.so_philox4x32_10(c, k, o)
    return tuple(o)


def log(capi, x):
    """The series logarithm both sides use in the exchange test."""
    return lib(capi).so_log(x)


def pt_uniform(capi, seed, ladder, rnd, slot):
    return lib(capi).so_pt_uniform(seed, ladder, rnd, slot)


def pt_exchange(capi, walkers, slot_walker, seed, ladder, rnd, accepted, proposed):
    """One exchange round over one ladder; slot_walker, accepted and proposed
    (int32 / int64 numpy arrays) are updated in place."""
    L_ = lib(capi)
    hs = (C.c_void_p * len(walkers))(*[w.h for w in walkers])
    assert slot_walker.dtype == np.int32 and accepted.dtype == np.int64
    assert proposed.dtype == np.int64
    L_.so_pt_exchange(hs, _p(slot_walker, C.c_int32), len(slot_walker), seed, ladder, rnd,
                      _p(accepted, C.c_int64), _p(proposed, C.c_int64))

"""A stand-in for capi.Batch whose walkers are oracle walkers (CPU), with
the subset of the Batch interface that carlo_mc.SharedBatch uses.  It lets
the CPU suite check the host-side protocol of the shared batch (owed sweeps,
one launch per round, filing and resetting of accumulators, slot reuse,
checkpoints) without a GPU; GPU tests run the same protocol on the device."""
import math

import numpy as np


class OracleBatch:
    def __init__(self, oracle_mod, capi, flats, walkers, mode=0, lib=None):
        self.oracle_mod = oracle_mod
        self.capi = capi
        self.flats = list(flats)
        self.walkers = list(walkers)
        n = len(walkers)
        self.w = [None] * n
        self.acc = [None] * n
        self.wlf = [[0.0, 0.0, 0] for _ in range(n)]
        self.calls = 0
        self.base = {}
        # a slot's accumulator row has the width of the model it was created
        # with, as in the library
        self.widths = [self.flats[w.model].measure_width for w in walkers]

    def _width(self, s):
        return self.widths[s]

    def slot_stop(self, s):
        self.w[s] = None

    def add_model(self, flat):
        self.flats.append(flat)
        return len(self.flats) - 1

    def slot_start(self, slot, wd, n_sweeps):
        assert n_sweeps == -1
        own = self.capi.walker_desc(model=0, T=wd.T, seed=wd.seed,
                                    init_opstring_cutoff=wd.init_opstring_cutoff,
                                    diagonal_warmup_sweeps=wd.diagonal_warmup_sweeps,
                                    init_num_worms=wd.init_num_worms,
                                    num_worms_attenuation_factor=wd.num_worms_attenuation_factor,
                                    target_worm_length_fraction=wd.target_worm_length_fraction,
                                    thermalization=wd.thermalization)
        o = self.oracle_mod.OracleWalker(self.capi, self.flats[wd.model], own)
        o.init()
        self.w[slot] = o
        self.walkers[slot] = wd
        self.acc[slot] = [np.zeros(self._width(slot)), np.zeros(self._width(slot)), 0]
        self.wlf[slot] = [0.0, 0.0, 0]
        self.base[slot] = np.array(o.work_counters())

    def run(self, n_sweeps, measure=True):
        self.run_budgets([n_sweeps if o is not None else 0 for o in self.w], measure)

    def run_budgets(self, sweeps, measure=True):
        self.calls += 1
        for s, k in enumerate(sweeps):
            o = self.w[s]
            if o is None or k == 0:
                continue
            for _ in range(int(k)):
                o.sweep(1)
                sc = o.scalars()
                if measure and sc["sweeps"] > self.walkers[s].thermalization:
                    row = o.measure()
                    a = self.acc[s]
                    a[0][:len(row)] += row
                    a[1][:len(row)] += row * row
                    a[2] += 1
                    f = sc["last_worm_length_fraction"]
                    if not math.isnan(f):
                        self.wlf[s][0] += f
                        self.wlf[s][1] += f * f
                        self.wlf[s][2] += 1

    def accumulators(self):
        """Per slot (sums, sums of squares, count), rows of the slot's model
        width."""
        out = []
        for s, (a, b, c) in enumerate(self._full()):
            cw = self.flats[self.walkers[s].model].measure_width
            out.append((a[:cw], b[:cw], c))
        return out

    def _full(self):
        out = []
        for s in range(len(self.w)):
            a = self.acc[s]
            if a is None:
                out.append((np.zeros(self._width(s)), np.zeros(self._width(s)), 0))
            else:
                out.append((a[0].copy(), a[1].copy(), a[2]))
        return out

    def slot_offsets(self):
        return np.concatenate([[0], np.cumsum(self.widths)[:-1]]).astype(np.int64)

    def accumulators_raw(self, with_sumsq=True):
        acc = self._full()
        return (np.concatenate([a[0] for a in acc]), np.concatenate([a[1] for a in acc]),
                np.array([a[2] for a in acc], dtype=np.int64))

    def worm_fractions_raw(self):
        return (np.array([x[0] for x in self.wlf]), np.array([x[1] for x in self.wlf]),
                np.array([x[2] for x in self.wlf], dtype=np.int64))

    def worm_fractions(self):
        return [tuple(x) for x in self.wlf]

    def work_counters(self):
        tot = np.zeros(4, dtype=np.int64)
        for s, o in enumerate(self.w):
            if o is not None:
                tot += np.array(o.work_counters()) - self.base.get(s, 0)
        return tuple(int(x) for x in tot)

    def reset_accumulators(self):
        self.base = {s: np.array(o.work_counters()) for s, o in enumerate(self.w) if o is not None}
        for s in range(len(self.w)):
            if self.acc[s] is not None:
                self.acc[s] = [np.zeros(self._width(s)), np.zeros(self._width(s)), 0]
            self.wlf[s] = [0.0, 0.0, 0]

    def get_walker(self, s, with_ops=True, with_spins=True):
        sc, ops, spins = self.w[s].state()
        return sc, ops if with_ops else None, spins if with_spins else None

    def set_walker(self, s, scalars, ops, spins):
        rc = self.w[s].set_state(scalars, ops, spins)
        if rc:
            raise ValueError(f"oracle set_state {rc}")

    def pt_log_weight_ratio(self, s, new_T):
        return self.w[s].pt_log_weight_ratio(new_T)

    def pt_change_parameter(self, s, new_T):
        self.w[s].pt_change_parameter(new_T)

    def pt_get_parameter(self, s):
        return self.w[s].T

    def close(self):
        self.w = [None] * len(self.w)

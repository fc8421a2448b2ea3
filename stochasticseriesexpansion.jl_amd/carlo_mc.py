"""Carlo's AbstractMC methods over one process-wide batch of walkers.

The reference gives every Carlo run its own `MC` and Carlo calls
`init!`, `sweep!`, `measure!`, `write_checkpoint` and `read_checkpoint` on it
(src/sse.jl:26-107), one run at a time.  Here an `MC` is a set of slots of a
`SharedBatch`, a batch that many `MC` objects of one process share:

* `sweep!` only posts the sweep: the slot's count of owed sweeps goes up.
* The owed sweeps of every slot of the batch run in ONE device call
  (`sse_run_budgets`, one budget per slot; `sse_run` when every attached
  slot owes the same count) when any `MC` needs a result:
  `measure!`, a checkpoint, a parallel-tempering hook, or a second `sweep!`
  of an `MC` whose previous sweep has not run yet.
* The launch measures every thermalized slot on the device; its rows (and
  the WormLengthFraction the reference measures inside `sweep!`,
  src/sse.jl:200-202) are read once per launch and handed to each `MC`'s
  context, in sweep order.

So results do not depend on the call order: every walker follows its own
Philox stream, and what an `MC` reports equals the reference's per-sweep
`measure!` calls.  A driver that calls `sweep!` on all its `MC` objects
before it calls `measure!` on any (`step_all`) gets one launch for all of
them per sweep; a driver that calls `sweep!`, `measure!` per `MC` in turn
(Carlo's own loop over one run) gets one launch per `MC` and sweep.  An `MC`
may also carry R runs (`runs`): its R walkers always sweep in one launch and
`measure!` reports their mean, which Carlo bins like any other sample.

`MCContext` is the part of Carlo's context these methods use: the sweep and
thermalization counters (a run measures once `sweeps > thermalization`) and
`measure!(ctx, name, value)`.
"""
import hashlib

import numpy as np

from . import capi, carlo

WORM_LENGTH_FRACTION = carlo.WORM_LENGTH_FRACTION


class MCContext:
    """Carlo's MCContext as the MC methods see it: `sweeps` (incremented by
    the driver after each sweep!), `thermalization`, `is_thermalized()` and
    `measure(name, value)`, which keeps every sample per observable."""

    def __init__(self, thermalization):
        self.sweeps = 0
        self.thermalization = int(thermalization)
        self.samples = {}

    def is_thermalized(self):
        return self.sweeps > self.thermalization

    def measure(self, name, value):
        self.samples.setdefault(name, []).append(float(value))

    def series(self, name):
        return np.array(self.samples.get(name, []))


def flat_key(flat):
    """Content hash of a FlatModel: MC objects built from equal parameters
    share one device model."""
    h = hashlib.sha1()

    def add(a):
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())

    for a in (flat.site_dims, flat.bond_type, flat.bond_sites, flat.est_coef, flat.est_mval):
        add(a)
    h.update(repr((flat.bond_n_sites, flat.energy_offset, flat.normalization_site_count,
                   [e.prefix for e in flat.estimators])).encode())
    for t in flat.types:
        for k in sorted(t):
            v = t[k]
            if isinstance(v, np.ndarray):
                add(v)
            else:
                h.update(f"{k}={v!r};".encode())
    return h.hexdigest()


class SharedBatch:
    """A batch of `n_slots` walkers that MC objects of this process share.

    `proto` (a FlatModel) sizes the slots: no model attached later may have
    more sites or more estimators.  Attached walkers run for an unlimited
    number of sweeps until their MC releases them.  `batch_cls` replaces
    capi.Batch (same constructor and methods)."""

    def __init__(self, proto, n_slots, mode=capi.SSE_MODE_WAVE, lib=None, batch_cls=None):
        walkers = [capi.walker_desc(model=0, T=1.0, seed=0) for _ in range(n_slots)]
        self.batch = (batch_cls or capi.Batch)([proto], walkers, mode=mode, lib=lib)
        for s in range(n_slots):
            self.batch.slot_stop(s)
        self.n_slots = n_slots
        self.offsets = self.batch.slot_offsets()
        self.free = list(range(n_slots - 1, -1, -1))
        self.models = {flat_key(proto): 0}
        self.owed = np.zeros(n_slots, dtype=np.int64)
        self.attached = np.zeros(n_slots, dtype=bool)
        self.owners = {}        # id(MC) -> MC whose launches get filed
        self.launches = 0
        # sum over launches of the work counters (cutoff L, operators n,
        # scatter calls W, walker-sweeps)
        self.work = np.zeros(4, dtype=np.int64)

    # ---- slots
    def model_index(self, flat):
        key = flat_key(flat)
        if key not in self.models:
            self.models[key] = self.batch.add_model(flat)
        return self.models[key]

    def attach(self, flat, wd):
        """Start a walker with parameters `wd` (its model field is replaced by
        the device index of `flat`) in a free slot; returns the slot."""
        if not self.free:
            raise RuntimeError("SharedBatch: no free slot")
        mi = self.model_index(flat)
        slot = self.free.pop()
        w = capi.walker_desc(model=mi, T=wd.T, seed=wd.seed,
                             init_opstring_cutoff=wd.init_opstring_cutoff,
                             diagonal_warmup_sweeps=wd.diagonal_warmup_sweeps,
                             init_num_worms=wd.init_num_worms,
                             num_worms_attenuation_factor=wd.num_worms_attenuation_factor,
                             target_worm_length_fraction=wd.target_worm_length_fraction,
                             thermalization=wd.thermalization)
        try:
            self.batch.slot_start(slot, w, -1)
        except Exception:
            self.free.append(slot)
            raise
        self.attached[slot] = True
        return slot

    def release(self, slot):
        self.owed[slot] = 0
        self.attached[slot] = False
        self.batch.slot_stop(slot)
        self.free.append(slot)

    # ---- sweeps
    def post(self, slots, n):
        """Owe `n` more sweeps to each of `slots` (an index array, or one
        slot as an int); slots that still owe sweeps first have them run
        (with everyone else's)."""
        if isinstance(slots, int):
            if self.owed[slots] > 0:
                self.flush()
        elif np.any(self.owed[slots] > 0):
            self.flush()
        self.owed[slots] += n

    def flush(self):
        """Run every owed sweep in one launch and file the launch's
        measurements with every registered MC that took part."""
        if not np.any(self.owed > 0):
            return
        b = self.batch
        swept = self.owed > 0
        k = int(self.owed.max())
        if np.array_equal(swept, self.attached) and np.all(self.owed[swept] == k):
            # every attached slot owes the same count: sse_run (stopped slots
            # take no sweep) with the launch form the tuner picked for it
            b.run(k, measure=True)
        else:
            b.run_budgets(self.owed, measure=True)
        self.launches += 1
        self.owed[:] = 0
        sums, _, cnt = b.accumulators_raw(with_sumsq=False)
        ws, _, wc = b.worm_fractions_raw()
        self.work += np.array(b.work_counters(), dtype=np.int64)
        b.reset_accumulators()
        rec = (sums, cnt, ws, wc)
        for mc in self.owners.values():
            if swept[mc.slots[0]]:
                mc._filed.append(rec)

    def close(self):
        self.batch.close()


class MC:
    """An AbstractMC run (src/sse.jl:26-107) whose walkers live in a
    SharedBatch.

    params: "model" (a host model class called with params, or a model
    object with .flat(), or a FlatModel), "T", "seed", "thermalization",
    and optionally the reference's "init_opstring_cutoff", "init_num_worms",
    "num_worms_attenuation_factor", "target_worm_length_fraction",
    "diagonal_warmup_sweeps".  `runs` walkers (seeds seed .. seed+runs-1)
    form this MC; `sweeps_per_step` device sweeps make one Carlo sweep
    (measure! then reports their mean)."""

    def __init__(self, params, shared, runs=1, sweeps_per_step=1):
        model = params["model"]
        if isinstance(model, type):
            model = model(params)
        self.model = model
        self.flat = model.flat() if hasattr(model, "flat") else model
        self.shared = shared
        self.runs = int(runs)
        self.k = int(sweeps_per_step)
        if self.runs < 1 or self.k < 1:
            raise ValueError("runs and sweeps_per_step must be positive")
        self.names = list(self.flat.observable_names())
        self._filed = []      # launch records (sums, counts, wlf sums, wlf counts)
        cutoff = params.get("init_opstring_cutoff", -1)
        slots = []
        self.slots = np.zeros(0, dtype=np.int64)
        try:
            for r in range(self.runs):
                wd = capi.walker_desc(
                    T=float(params["T"]), seed=int(params["seed"]) + r,
                    init_opstring_cutoff=int(cutoff),
                    diagonal_warmup_sweeps=int(params.get("diagonal_warmup_sweeps", 5)),
                    init_num_worms=float(params.get("init_num_worms", 5)),
                    num_worms_attenuation_factor=float(
                        params.get("num_worms_attenuation_factor", 0.01)),
                    target_worm_length_fraction=float(
                        params.get("target_worm_length_fraction", 2.0)),
                    thermalization=self.k * int(params.get("thermalization", 0)))
                slots.append(shared.attach(self.flat, wd))
        except Exception:
            for s in slots:
                shared.release(s)
            raise
        self.slots = np.array(slots, dtype=np.int64)
        width = len(self.names)
        self._rows = shared.offsets[self.slots][:, None] + np.arange(width)[None, :]
        # one run: plain-int indexing on the host (no index arrays)
        self._one = int(slots[0]) if self.runs == 1 else None
        self._off = int(shared.offsets[slots[0]])
        shared.owners[id(self)] = self

    # Carlo.init! (sse.jl:47-60): the slots were started at construction
    # and are initialised on the device by the first launch that includes
    # them
    def init(self, ctx, params=None):
        return None

    # Carlo.sweep! (sse.jl:62-68), posted
    def sweep(self, ctx):
        self.shared.post(self._one if self._one is not None else self.slots, self.k)

    # Carlo.measure! (sse.jl:70-87)
    def measure(self, ctx):
        self._flush_own()
        if not self._report(ctx):
            raise RuntimeError("measure!: no measured sweep to report (the walker is not "
                               "thermalized on the device)")

    def _flush_own(self):
        owed = self.shared.owed
        if (owed[self._one] > 0) if self._one is not None else np.any(owed[self.slots] > 0):
            self.shared.flush()

    def _report(self, ctx):
        """Hand every filed launch's measurements to ctx, oldest first: the
        mean over runs (and the step's sweeps) of each observable, and the
        mean WormLengthFraction.  Returns the number of rows reported."""
        filed, self._filed = self._filed, []
        reported = 0
        if self._one is not None:
            s, width = self._one, len(self.names)
            for sums, cnt, ws, wc in filed:
                if cnt[s]:
                    if cnt[s] != self.k:
                        raise RuntimeError(
                            f"measure!: {int(cnt[s])} of {self.k} sweeps of the step measured")
                    row = sums[self._off:self._off + width] / self.k
                    for name, v in zip(self.names, row.tolist()):
                        ctx.measure(name, v)
                    reported += 1
                if wc[s]:
                    ctx.measure(WORM_LENGTH_FRACTION, ws[s] / wc[s])
            return reported
        for sums, cnt, ws, wc in filed:
            c = cnt[self.slots]
            if np.any(c != c[0]):
                raise RuntimeError("runs of one MC measured different sweep counts")
            if c[0]:
                if c[0] != self.k:
                    raise RuntimeError(
                        f"measure!: {int(c[0])} of {self.k} sweeps of the step measured")
                row = sums[self._rows].sum(axis=0) / (self.k * self.runs)
                for name, v in zip(self.names, row):
                    ctx.measure(name, v)
                reported += 1
            # WormLengthFraction: measured by the reference inside sweep! of
            # a thermalized sweep with n != 0 (sse.jl:200-202); here the mean
            # over the step's samples of all runs
            n = wc[self.slots].sum()
            if n:
                ctx.measure(WORM_LENGTH_FRACTION, ws[self.slots].sum() / n)
        return reported

    # Carlo.write_checkpoint / read_checkpoint (sse.jl:89-107): the
    # reference's five datasets per run plus this library's stream position
    def write_checkpoint(self, group):
        self._flush_own()
        b = self.shared.batch
        for r, s in enumerate(self.slots):
            g = group if self.runs == 1 else group.create_group(f"run{r}")
            try:
                carlo.write_checkpoint(b, int(s), g)
                sc, _, _ = b.get_walker(int(s), with_ops=False, with_spins=False)
                g.write("philox_counter", np.uint64(sc["rng_counter"]))
                g.write("sweeps", int(sc["sweeps"]))
            finally:
                if g is not group:
                    g.close()

    def read_checkpoint(self, group):
        self._flush_own()
        b = self.shared.batch
        for r, s in enumerate(self.slots):
            g = group if self.runs == 1 else group.group(f"run{r}")
            try:
                carlo.read_checkpoint(b, int(s), g, sweeps=int(g.read("sweeps")),
                                      rng_counter=int(g.read("philox_counter")))
            finally:
                if g is not group:
                    g.close()
        self._filed = []

    # Carlo's parallel-tempering hooks (sse.jl:390-405), one run per MC
    def parallel_tempering_log_weight_ratio(self, parameter, new_value):
        self._one_run(parameter)
        self._flush_own()
        return self.shared.batch.pt_log_weight_ratio(int(self.slots[0]), float(new_value))

    def parallel_tempering_change_parameter(self, parameter, new_value):
        self._one_run(parameter)
        self._flush_own()
        self.shared.batch.pt_change_parameter(int(self.slots[0]), float(new_value))

    def _one_run(self, parameter):
        if parameter != "T":
            raise ValueError(f"unsupported parallel tempering parameter {parameter}")
        if self.runs != 1:
            raise ValueError("parallel tempering exchanges one run per MC")

    def close(self):
        self.shared.owners.pop(id(self), None)
        for s in self.slots:
            self.shared.release(int(s))
        self.slots = np.zeros(0, dtype=np.int64)


def step_all(mcs, ctxs):
    """One Carlo sweep of every MC, grouped by phase: every sweep! is posted
    before any measure!, so all of them run in one launch.  Per MC this is
    Carlo's step (sweep!, sweeps += 1, measure! once thermalized)."""
    for mc, ctx in zip(mcs, ctxs):
        mc.sweep(ctx)
        ctx.sweeps += 1
    for mc, ctx in zip(mcs, ctxs):
        if ctx.is_thermalized():
            mc.measure(ctx)


def step_one(mc, ctx):
    """Carlo's step for one MC alone: sweep!, sweeps += 1, measure!."""
    mc.sweep(ctx)
    ctx.sweeps += 1
    if ctx.is_thermalized():
        mc.measure(ctx)

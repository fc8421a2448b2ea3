"""Carlo-compatible persistence and result analysis for batched walkers.

The reference runs one `MC` per Carlo run.  Carlo stores a run's state through
`write_checkpoint(mc, group)` / `read_checkpoint(mc, group)` (src/sse.jl:89-107)
and merges the internal bins of all runs of a task into `results.json`
(layout: docs/src/bani2v2o8.results.json, written by Carlo 0.2.2).  Here one
run is one walker of a `Batch`.

Checkpoints
  write_mc_group / read_mc_group   the reference's five datasets of one walker:
      num_operators (int64 scalar), avg_worm_length, num_worms (float64
      scalars), operators (compound {code: uint64}[L], the reference OperCode
      encoding of src/opercode.jl:37-47) and state (uint8[N], 1-based states)
  write_checkpoint / read_checkpoint   the same for walker w of a Batch
  save_run / load_run   a file holding that group ("simulation") and this
      library's stream position ("context": Philox key and counter, sweeps,
      thermalization, T), so a resumed walker continues bit for bit

Results
  bin_walkers       internal bins: means over `binsize` consecutive measured
                    sweeps of every walker (Carlo's accumulator bins), with
                    WormLengthFraction as the last column
  merge_observable  Carlo's merge of a task's internal bins: rebin count,
                    rebin length, mean, error, autocorrelation time
  jackknife         bias-corrected jackknife over the rebinned means
  evaluables        the reference's evaluables (src/sse.jl:111-134,
                    src/models/common/magnetization_estimator.jl:236-258)
  evaluate_task / write_results / run_tasks   results.json entries
"""
import json
import math

import numpy as np

from . import h5lite

VERSION = {"carlo_version": None, "mc_package": "StochasticSeriesExpansion.jl_amd",
           "mc_version": "0.1.2"}

_VIDX_BITS = 9        # device vertex index width (vertex index < 512)
_REF_VCODE_BITS = 25  # src/opercode.jl:2, vertex_code_maxbits = 8*3+1
_REF_BOND_SHIFT = 1 + _REF_VCODE_BITS


# ---------------------------------------------------------------- opcodes
def opcodes_to_reference(ops):
    """Device operator codes (uint32: 1 | vcode << 1 | bond << 11 with
    vcode = diag | vidx << 1, 0-based) to the reference's OperCode values
    (uint64: 1 | vcode' << 1 | bond' << 26, vcode' = diag | vidx' << 1 with
    1-based vidx' and bond'; src/opercode.jl:18-47).  Identity stays 0."""
    ops = np.asarray(ops, dtype=np.uint64)
    vc = (ops >> 1) & 1023
    vref = (vc & 1) | (((vc >> 1) + 1) << 1)
    code = 1 | (vref << 1) | (((ops >> 11) + 1) << _REF_BOND_SHIFT)
    return np.where(ops != 0, code, 0).astype(np.uint64)


def opcodes_from_reference(codes):
    """Inverse of opcodes_to_reference.  Raises ValueError for a code that is
    not an OperCode of a two-site bond operator this library can hold (bit 0
    clear on a non-zero code, vertex index 0 or above 512, bond 0 or above
    2^21)."""
    codes = np.asarray(codes, dtype=np.uint64)
    nz = codes != 0
    vref = (codes & ((1 << _REF_VCODE_BITS) - 1)) >> 1
    vidx1 = vref >> 1
    bond1 = codes >> _REF_BOND_SHIFT
    bad = nz & (((codes & 1) == 0) | (vidx1 < 1) | (vidx1 > (1 << _VIDX_BITS)) | (bond1 < 1)
                | (bond1 > (1 << 21)))
    if bad.any():
        k = int(np.nonzero(bad)[0][0])
        raise ValueError(f"operator {k}: 0x{int(codes[k]):x} is not a valid OperCode")
    vc = (vref & 1) | ((np.where(nz, vidx1, 1) - 1) << 1)
    dev = 1 | (vc << 1) | ((np.where(nz, bond1, 1) - 1) << 11)
    return np.where(nz, dev, 0).astype(np.uint32)


# ---------------------------------------------------------------- checkpoints
def write_mc_group(group, scalars, ops, spins):
    """Write one walker's state as the reference's checkpoint group
    (src/sse.jl:89-96).  `ops` are device operator codes."""
    group.write("num_operators", int(scalars["num_operators"]))
    group.write("avg_worm_length", float(scalars["avg_worm_length"]))
    group.write("num_worms", float(scalars["num_worms"]))
    group.write("operators", opcodes_to_reference(ops), compound_member="code")
    group.write("state", np.ascontiguousarray(spins, dtype=np.uint8))


def read_mc_group(group):
    """Read a checkpoint group written by write_mc_group or by the reference
    (src/sse.jl:98-107); `operators` may be the compound {code} form or plain
    uint64.  Returns (scalars, device ops, spins); raises ValueError when the
    stored operator count does not match the operator string."""
    ops = opcodes_from_reference(group.read("operators"))
    spins = np.ascontiguousarray(group.read("state"), dtype=np.uint8)
    scalars = {"num_operators": int(group.read("num_operators")),
               "avg_worm_length": float(group.read("avg_worm_length")),
               "num_worms": float(group.read("num_worms")),
               "cutoff": int(len(ops))}
    if scalars["num_operators"] != int(np.count_nonzero(ops)):
        raise ValueError(f"checkpoint: num_operators {scalars['num_operators']} but "
                         f"{int(np.count_nonzero(ops))} operators in the string")
    return scalars, ops, spins


def write_checkpoint(batch, walker, group):
    """Carlo.write_checkpoint for walker `walker` of a Batch."""
    sc, ops, spins = batch.get_walker(walker)
    write_mc_group(group, sc, ops, spins)


def read_checkpoint(batch, walker, group, sweeps=None, rng_counter=None):
    """Carlo.read_checkpoint for walker `walker` of a Batch.  The walker keeps
    its sweep count and stream position unless `sweeps` / `rng_counter` are
    given (the reference keeps both in Carlo's context, not in this group).
    The library rejects states outside the model (bond, vertex or local state
    out of range)."""
    now, _, _ = batch.get_walker(walker, with_ops=False, with_spins=False)
    sc, ops, spins = read_mc_group(group)
    merged = dict(now)
    merged.update(sc)
    if sweeps is not None:
        merged["sweeps"] = int(sweeps)
    if rng_counter is not None:
        merged["rng_counter"] = int(rng_counter)
    batch.set_walker(walker, merged, ops, spins)


def write_context(batch, walker, group):
    """This library's part of a run's context: the Philox key and counter of
    the walker's stream, its sweep count, thermalization and temperature."""
    sc, _, _ = batch.get_walker(walker, with_ops=False, with_spins=False)
    wd = batch.walkers[walker]
    group.write("philox_key", np.uint64(wd.seed))
    group.write("philox_counter", np.uint64(sc["rng_counter"]))
    group.write("sweeps", int(sc["sweeps"]))
    group.write("thermalization", int(wd.thermalization))
    group.write("T", float(batch.pt_get_parameter(walker)))


def save_run(path, batch, walker):
    """Write walker `walker` to an HDF5 file: group "simulation" holds the
    reference's checkpoint datasets, group "context" the stream position."""
    with h5lite.File(path, "w") as f:
        with f.create_group("simulation") as g:
            write_checkpoint(batch, walker, g)
        with f.create_group("context") as g:
            write_context(batch, walker, g)


def load_run(path, batch, walker, check_key=True):
    """Restore walker `walker` from a file written by save_run.  The walker
    must have been created with the same seed (Philox key) unless
    `check_key` is False; its temperature is set to the stored one."""
    with h5lite.File(path, "r") as f:
        with f.group("context") as g:
            key = int(g.read("philox_key"))
            ctr = int(g.read("philox_counter"))
            sweeps = int(g.read("sweeps"))
            T = float(g.read("T"))
        if check_key and key != int(batch.walkers[walker].seed):
            raise ValueError(f"run file belongs to Philox key {key}, walker {walker} has key "
                             f"{int(batch.walkers[walker].seed)}")
        if T != float(batch.pt_get_parameter(walker)):
            batch.pt_change_parameter(walker, T)
        with f.group("simulation") as g:
            read_checkpoint(batch, walker, g, sweeps=sweeps, rng_counter=ctr)


# ---------------------------------------------------------------- binning
WORM_LENGTH_FRACTION = "WormLengthFraction"


def binned_names(flat):
    """Column names of bin_walkers' bins: the measured observables, then
    WormLengthFraction."""
    return list(flat.observable_names()) + [WORM_LENGTH_FRACTION]


def bin_walkers(batch, binsize, n_bins):
    """Run n_bins * binsize sweeps and return every walker's internal bins: a
    list with one [bin][measure width + 1] array per walker (columns as
    binned_names), each bin the mean of `binsize` consecutive per-sweep
    measurements (accumulated in sweep order on the device, divided by the
    bin length, as Carlo's accumulator forms a bin).  The last column is the
    bin's mean WormLengthFraction, which the reference measures only in
    sweeps with n != 0 (sse.jl:200-202): its mean over the samples the bin
    has (NaN without any).  Every walker must be past thermalization; a walker
    that measured fewer sweeps in a bin raises ValueError.  The batch's
    accumulators are reset before the first bin and after every bin (they
    hold one bin at a time, so each bin's sum is formed in sweep order from
    zero)."""
    if binsize < 1 or n_bins < 1:
        raise ValueError("binsize and n_bins must be positive")
    out = [np.zeros((n_bins, batch.flats[w.model].measure_width + 1)) for w in batch.walkers]
    batch.reset_accumulators()
    for b in range(n_bins):
        batch.run(binsize, measure=True)
        wlf = batch.worm_fractions()
        for k, (s, _, cnt) in enumerate(batch.accumulators()):
            if cnt != binsize:
                raise ValueError(f"walker {k} measured {cnt} of {binsize} sweeps in bin {b}: "
                                 "thermalize the batch first")
            out[k][b, :-1] = s / binsize
            ws, _, wc = wlf[k]
            out[k][b, -1] = ws / wc if wc else float("nan")
        batch.reset_accumulators()
    return out


def calc_rebin_count(sample_count, min_bin_count=10):
    """Number of rebinned samples Carlo forms from `sample_count` internal
    bins: all of them up to min_bin_count, then min_bin_count plus the
    rounded cube root of the rest.  Checked against every task of the
    reference's results.json (rebin_count, rebin_len = count // rebin_count)."""
    if sample_count <= min_bin_count:
        return int(sample_count)
    return int(min_bin_count + round(float(np.cbrt(sample_count - min_bin_count))))


def _nan_to_none(x):
    """JSON has no NaN or infinity: such values are written as null."""
    if isinstance(x, float) and not math.isfinite(x):
        return None
    return x


def merge_observable(samples, internal_bin_len):
    """Merge the internal bins of one observable (all runs of a task in run
    order).  The leading rebin_count * rebin_len bins are used (the mean of
    the reference's OperatorCount times rebin_count * rebin_len *
    internal_bin_len is an integer in every task of its results.json); the
    error is the standard error of the rebinned means (sample variance with
    n - 1), and the autocorrelation time, in internal bins, is
    0.5 * rebin_len * var(rebinned) / var(internal bins).  Returns
    (entry, rebinned means)."""
    x = np.asarray(samples, dtype=np.float64)
    n = len(x)
    rc = calc_rebin_count(n)
    rl = n // rc if rc else 0
    used = x[:rc * rl]
    rebins = used.reshape(rc, rl).mean(axis=1) if rc else used
    mean = float(used.sum() / (rc * rl)) if rc else float("nan")
    if rc >= 2:
        vr = float(np.var(rebins, ddof=1))
        err = math.sqrt(vr / rc)
        vs = float(np.var(used, ddof=1))
        tau = 0.5 * rl * vr / vs if vs > 0 else float("nan")
    else:
        err = tau = float("nan")
    entry = {"mean": mean, "error": err, "autocorr_time": tau, "rebin_len": int(rl),
             "internal_bin_len": int(internal_bin_len), "rebin_count": int(rc)}
    return entry, rebins


def jackknife(func, *rebinned):
    """Bias-corrected jackknife of func over equally long rebinned series:
    returns (mean, error).  func is applied elementwise to numpy arrays."""
    sets = [np.asarray(s, dtype=np.float64) for s in rebinned]
    n = len(sets[0])
    if n < 2 or any(len(s) != n for s in sets):
        raise ValueError("jackknife needs >= 2 samples per input, equally many")
    sums = [s.sum() for s in sets]
    full = func(*[t / n for t in sums])
    jk = np.asarray(func(*[(t - s) / (n - 1) for t, s in zip(sums, sets)]), dtype=np.float64)
    jmean = jk.mean()
    mean = full - (n - 1) * (jmean - full)
    err = math.sqrt((n - 1) / n * float(((jk - jmean) ** 2).sum()))
    return float(mean), err


def _unsign(x, s):
    return x / s


def _binder(smag4, smag2, sign):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where((smag2 == 0) & (smag4 == 0), 0.0 * smag2, smag2 ** 2 / smag4 / sign)


def evaluables(names, normalization_site_count):
    """The reference's evaluables for a model measuring `names` (in the order
    FlatModel.observable_names() gives): per magnetization estimator with
    prefix p, {p}Mag, {p}AbsMag, {p}Mag2, {p}Mag4, {p}MagChi = Sign{p}X / Sign
    and {p}BinderRatio = Sign{p}Mag2^2 / Sign{p}Mag4 / Sign (0 when both are
    0) (magnetization_estimator.jl:236-258); then Energy = SignEnergy / Sign
    and SpecificHeat = (sn2/s - sn^2/s^2 - sn/s) / N (sse.jl:124-131).
    Returns a list of (name, input names, function)."""
    out = []
    for nm in names:
        if nm.startswith("Sign") and nm.endswith("Mag4"):
            p = nm[4:-4]
            for obs in ("Mag", "AbsMag", "Mag2", "Mag4", "MagChi"):
                out.append((f"{p}{obs}", (f"Sign{p}{obs}", "Sign"), _unsign))
            out.append((f"{p}BinderRatio", (f"Sign{p}Mag4", f"Sign{p}Mag2", "Sign"), _binder))
    nrm = float(normalization_site_count)
    out.append(("Energy", ("SignEnergy", "Sign"), _unsign))
    out.append(("SpecificHeat", ("SignOperatorCount2", "SignOperatorCount", "Sign"),
                lambda sn2, sn, s: (sn2 / s - sn * sn / s ** 2 - sn / s) / nrm))
    return out


def evaluate_task(names, runs, internal_bin_len, normalization_site_count):
    """results.json "results" of one task.  `runs` holds each run's internal
    bins ([bins][len(names)]), concatenated in run order as Carlo merges the
    run files of a task.  Observables get Carlo's merged entry; evaluables
    get {"mean": [m], "error": [e]} from the jackknife over the rebinned
    means, autocorr_time None, the inputs' rebin_len and rebin_count 0, as
    in the reference's results.json."""
    samples = np.concatenate([np.asarray(r, dtype=np.float64) for r in runs], axis=0)
    if samples.ndim != 2 or samples.shape[1] != len(names):
        raise ValueError("runs must be [bins][observables] arrays")
    res, rebinned = {}, {}
    for j, nm in enumerate(names):
        entry, rb = merge_observable(samples[:, j], internal_bin_len)
        res[nm] = {k: _nan_to_none(v) for k, v in entry.items()}
        rebinned[nm] = rb
    rl = res[names[0]]["rebin_len"]
    for nm, inputs, func in evaluables(names, normalization_site_count):
        if not all(i in rebinned for i in inputs):
            continue
        if len(rebinned[inputs[0]]) >= 2:
            m, e = jackknife(func, *[rebinned[i] for i in inputs])
        else:
            m = e = float("nan")
        res[nm] = {"mean": [_nan_to_none(m)], "error": [_nan_to_none(e)], "autocorr_time": None,
                   "rebin_len": rl, "internal_bin_len": int(internal_bin_len), "rebin_count": 0}
    return res


def _plain(x):
    if isinstance(x, dict):
        return {str(k): _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    if isinstance(x, np.ndarray):
        return [_plain(v) for v in x.tolist()]
    if isinstance(x, np.integer):
        return int(x)
    if isinstance(x, (float, np.floating)):
        return _nan_to_none(float(x))
    return x


def write_results(path, entries):
    """Write results.json: a list of {"task", "parameters", "results",
    "version"} objects (NaN written as null)."""
    out = []
    for e in entries:
        out.append({"task": e["task"], "parameters": _plain(e.get("parameters", {})),
                    "results": _plain(e["results"]), "version": dict(VERSION)})
    with open(path, "w") as f:
        json.dump(out, f, indent=1, allow_nan=False)


def run_tasks(tasks, runs_per_task=1, seed0=0, mode=0, batch_cls=None, progress=None,
              pt_interval=0, pt_seed=1):
    """Run Carlo-style tasks on the GPU, every run of every task in one batch
    per (thermalization, sweeps, binsize) group.

    tasks: dicts with "name", "model" (a host model with .flat()), "T",
    "thermalization", "sweeps", "binsize" and any further parameters (kept
    in the entry).  Tasks that pass the same model object share one device
    model.  Run r of task t uses seed seed0 + t * runs_per_task + r.

    pt_interval > 0 runs parallel tempering in T (the reference's hooks,
    sse.jl:390-405): the tasks of a group must share one model object; run r
    of every task forms ladder r (temperatures in increasing order) and an
    exchange round follows every pt_interval-th sweep, thermalization
    included.  A task's bins then belong to its temperature, whichever walker
    holds it.  Returns the results.json entries in task order."""
    from .capi import Batch, walker_desc
    batch_cls = batch_cls or Batch
    groups = {}
    for t, task in enumerate(tasks):
        key = (int(task["thermalization"]), int(task["sweeps"]), int(task["binsize"]))
        if key[1] % key[2]:
            raise ValueError(f"task {task['name']}: sweeps must be a multiple of binsize")
        groups.setdefault(key, []).append(t)
    entries = [None] * len(tasks)
    for (therm, sweeps, binsize), idx in groups.items():
        flats, model_of = [], {}
        for t in idx:
            mid = id(tasks[t]["model"])
            if mid not in model_of:
                model_of[mid] = len(flats)
                flats.append(tasks[t]["model"].flat())
        if pt_interval and len(flats) != 1:
            raise ValueError("parallel tempering: the tasks of a group must share one model object")
        walkers = []
        for t in idx:
            for r in range(runs_per_task):
                walkers.append(walker_desc(model=model_of[id(tasks[t]["model"])],
                                           T=float(tasks[t]["T"]),
                                           seed=seed0 + t * runs_per_task + r,
                                           thermalization=therm))
        batch = batch_cls(flats, walkers, mode=mode)
        try:
            batch.init()
            if pt_interval:
                order = sorted(range(len(idx)), key=lambda k: float(tasks[idx[k]]["T"]))
                batch.pt_set_ladders([[k * runs_per_task + r for k in order]
                                      for r in range(runs_per_task)], pt_seed, pt_interval)
            if therm:
                batch.run(therm, measure=False)
            n_bins = sweeps // binsize
            chunks = []
            done = 0
            while done < n_bins:
                step = min(n_bins - done, max(1, n_bins // 10))
                chunks.append(bin_walkers(batch, binsize, step))
                done += step
                if progress:
                    progress(f"tasks {idx}: {done}/{n_bins} bins")
        finally:
            batch.close()
        for k, t in enumerate(idx):
            flat = flats[model_of[id(tasks[t]["model"])]]
            runs = [np.concatenate([c[k * runs_per_task + r] for c in chunks], axis=0)
                    for r in range(runs_per_task)]
            params = {key: v for key, v in tasks[t].items() if key not in ("model", "name")}
            if pt_interval:
                params["parallel_tempering_interval"] = int(pt_interval)
            entries[t] = {"task": tasks[t]["name"], "parameters": params,
                          "results": evaluate_task(binned_names(flat), runs, binsize,
                                                   flat.normalization_site_count)}
    return entries

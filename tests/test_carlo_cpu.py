"""Carlo-compatible persistence and result analysis on the CPU.

Checkpoints: the HDF5 layout of the reference's write_checkpoint
(src/sse.jl:89-107) written and read through libhdf5, the OperCode
conversion against the C ABI's own map, and a chain resumed from a file equal
to the uninterrupted chain (oracle walkers).

Results: Carlo's merge rules against every task of the reference's published
results.json (tests/golden/carlo_results_meta.json), the jackknife against
closed forms, the autocorrelation time against an AR(1) process, and a task
merged from oracle bins against exact diagonalisation."""
import json
import math
import os

import numpy as np
import pytest

import ed
from fixtures import TS, job_magnet_square, square_heisenberg

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def carlo(sse_pkg):
    if not sse_pkg.h5lite.available():
        pytest.skip("libhdf5 not available")
    return sse_pkg.carlo


@pytest.fixture(scope="module")
def meta():
    return json.load(open(os.path.join(HERE, "golden", "carlo_results_meta.json")))["tasks"]


def _scratch(tmp_path, name):
    return str(tmp_path / name)


# ------------------------------------------------------------ HDF5 + checkpoints
def test_h5_roundtrip_and_layout(sse_pkg, carlo, tmp_path):
    h5 = sse_pkg.h5lite
    path = _scratch(tmp_path, "a.h5")
    codes = np.array([0, 1 | (5 << 1) | (3 << 26), 0, (1 << 40) | 7], dtype=np.uint64)
    with h5.File(path, "w") as f:
        with f.create_group("simulation") as g:
            g.write("num_operators", 7)
            g.write("num_worms", 2.25)
            g.write("operators", codes, compound_member="code")
            g.write("state", np.array([1, 2, 2, 1], dtype=np.uint8))
            g.write("num_worms", 3.5)  # replaces
            g.write("zero_d", np.array(4, dtype=np.uint64))
    with h5.File(path, "r") as f:
        assert "simulation" in f and "nothing" not in f
        with f.group("simulation") as g:
            assert g.read("num_operators") == 7 and isinstance(g.read("num_operators"), int)
            assert g.read("num_worms") == 3.5
            assert np.array_equal(g.read("operators"), codes)
            assert g.read("state").dtype == np.uint8
            assert g.dataset_info("num_operators") == (h5.H5T_INTEGER, 8, None, ())
            assert g.dataset_info("num_worms") == (h5.H5T_FLOAT, 8, None, ())
            assert g.dataset_info("operators") == (h5.H5T_COMPOUND, 8, "code", (4,))
            assert g.dataset_info("state") == (h5.H5T_INTEGER, 1, None, (4,))
            assert g.dataset_info("zero_d") == (h5.H5T_INTEGER, 8, None, ())
            assert g.read("zero_d") == 4
            with pytest.raises(KeyError):
                g.read("missing")


def test_opcode_conversion_matches_c_abi(sse_pkg, carlo):
    lib = sse_pkg.capi.load_library()
    rng = np.random.default_rng(5)
    dev = np.zeros(400, dtype=np.uint32)
    vidx = rng.integers(0, 512, 400)
    diag = rng.integers(0, 2, 400)
    bond = rng.integers(0, 1 << 20, 400)
    dev[:] = 1 | ((diag | (vidx << 1)) << 1) | (bond << 11)
    dev[::7] = 0
    ref = carlo.opcodes_to_reference(dev)
    for d, r in zip(dev, ref):
        assert lib.sse_opcode_to_reference(int(d)) == int(r)
    assert np.array_equal(carlo.opcodes_from_reference(ref), dev)
    for bad in (2, 1 | (0 << 2) | (1 << 26), 1 | (1 << 2) | (0 << 26), 1 | (600 << 2) | (1 << 26)):
        with pytest.raises(ValueError):
            carlo.opcodes_from_reference(np.array([0, bad], dtype=np.uint64))


def test_checkpoint_file_resumes_chain(sse_pkg, oracle_mod, carlo, tmp_path):
    """write_mc_group -> file -> read_mc_group restores a state from which the
    chain continues exactly as the uninterrupted one (oracle walkers)."""
    flat = square_heisenberg(sse_pkg, 4).flat()
    wd = sse_pkg.walker_desc(T=0.4, seed=77, thermalization=3)
    a = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    a.init()
    a.sweep(9)
    sc, ops, spins = a.state()
    path = _scratch(tmp_path, "run.h5")
    with sse_pkg.h5lite.File(path, "w") as f:
        with f.create_group("simulation") as g:
            carlo.write_mc_group(g, sc, ops, spins)
    with sse_pkg.h5lite.File(path, "r") as f:
        with f.group("simulation") as g:
            rsc, rops, rspins = carlo.read_mc_group(g)
            assert g.read("operators").dtype == np.uint64
    assert np.array_equal(rops, ops) and np.array_equal(rspins, spins)
    for k in ("num_operators", "avg_worm_length", "num_worms", "cutoff"):
        assert rsc[k] == sc[k], k
    b = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    b.init()
    full = dict(sc)
    full.update(rsc)
    b.set_state(full, rops, rspins)
    a.sweep(11)
    b.sweep(11)
    sa, oa, pa = a.state()
    sb, ob, pb = b.state()
    assert np.array_equal(oa, ob) and np.array_equal(pa, pb)
    assert sa["num_worms"] == sb["num_worms"] and sa["rng_counter"] == sb["rng_counter"]


def test_checkpoint_count_mismatch_rejected(sse_pkg, carlo, tmp_path):
    path = _scratch(tmp_path, "bad.h5")
    with sse_pkg.h5lite.File(path, "w") as f:
        with f.create_group("simulation") as g:
            g.write("num_operators", 2)
            g.write("avg_worm_length", 1.0)
            g.write("num_worms", 1.0)
            g.write("operators", np.array([0, 1 | (2 << 2) | (1 << 26)], dtype=np.uint64))
            g.write("state", np.array([1, 1], dtype=np.uint8))
    with sse_pkg.h5lite.File(path, "r") as f:
        with f.group("simulation") as g:
            with pytest.raises(ValueError):
                carlo.read_mc_group(g)


# ------------------------------------------------------------ merge rules
def test_rebin_count_matches_reference_results(carlo, meta):
    """Every task's (rebin_count, rebin_len) is what calc_rebin_count gives
    for the number of samples the task's OperatorCount mean implies, and that
    number is exactly rebin_count * rebin_len: the per-sweep operator counts
    are integers, so mean * samples * internal_bin_len is an integer for the
    samples the mean was taken over."""
    for t in meta:
        oc = t["results"]["OperatorCount"]
        rc, rl, b = oc["rebin_count"], oc["rebin_len"], oc["internal_bin_len"]
        assert b == t["binsize"]
        used = rc * rl
        total = oc["mean"] * used * b
        assert abs(total - round(total)) < 1e-6 * max(1.0, total) ** 0.5, t["T"]
        # some sample count n with n // rc == rl yields this rebin count
        ok = [n for n in range(used, used + rc) if carlo.calc_rebin_count(n) == rc and n // rc == rl]
        assert ok, (t["size"], t["T"], rc, rl)
        if t["sweeps"] // t["binsize"] in range(used, used + rc):  # complete single run
            assert carlo.calc_rebin_count(t["sweeps"] // t["binsize"]) == rc


def test_merge_uses_leading_rebins_and_integral_mean(carlo):
    rng = np.random.default_rng(1)
    per_sweep = rng.integers(100, 200, size=(800 * 100))
    bins = per_sweep.reshape(800, 100).mean(axis=1)
    e, rebins = carlo.merge_observable(bins, 100)
    assert (e["rebin_count"], e["rebin_len"]) == (19, 42)
    used = 19 * 42
    assert len(rebins) == 19
    tot = e["mean"] * used * 100
    assert abs(tot - per_sweep[:used * 100].sum()) < 1e-4
    assert abs(e["error"] - np.std(rebins, ddof=1) / math.sqrt(19)) < 1e-15
    for n in (1, 5, 10):
        assert carlo.calc_rebin_count(n) == n


def test_sign_one_relations_as_in_reference(sse_pkg, carlo, meta):
    """With Sign identically 1 the reference's results.json has Sign error 0
    and autocorr_time null, and each unsigned evaluable equals its signed
    observable in mean and error up to rounding; the merge reproduces that
    (it would not with a population-variance error: factor sqrt(n/(n-1)))."""
    for t in meta:
        r = t["results"]
        assert r["Sign"]["mean"] == 1.0 and r["Sign"]["error"] == 0.0
        assert r["Sign"]["autocorr_time"] is None
        for ev, ob in (("Energy", "SignEnergy"), ("MagChi", "SignMagChi")):
            assert r[ev]["rebin_count"] == 0 and r[ev]["autocorr_time"] is None
            assert r[ev]["rebin_len"] == r[ob]["rebin_len"]
            assert abs(r[ev]["mean"][0] - r[ob]["mean"]) <= 1e-10 * abs(r[ob]["mean"])
            assert abs(r[ev]["error"][0] - r[ob]["error"]) <= 1e-10 * r[ob]["error"]
    flat = square_heisenberg(sse_pkg, 4).flat()
    names = flat.observable_names()
    rng = np.random.default_rng(2)
    bins = rng.normal(size=(517, len(names)))
    bins[:, names.index("Sign")] = 1.0
    res = carlo.evaluate_task(names, [bins[:300], bins[300:]], 100, flat.normalization_site_count)
    assert res["Sign"]["error"] == 0.0 and res["Sign"]["autocorr_time"] is None
    for ev, ob in (("Energy", "SignEnergy"), ("MagChi", "SignMagChi"), ("StagMag2", "SignStagMag2")):
        assert res[ev]["rebin_count"] == 0 and res[ev]["autocorr_time"] is None
        assert abs(res[ev]["mean"][0] - res[ob]["mean"]) <= 1e-12 * abs(res[ob]["mean"])
        assert abs(res[ev]["error"][0] - res[ob]["error"]) <= 1e-12 * res[ob]["error"]
    keys = set(meta[0]["results"]["SignEnergy"].keys())
    assert set(res["SignEnergy"].keys()) == keys and set(res["Energy"].keys()) == keys


def test_jackknife_closed_forms(carlo):
    rng = np.random.default_rng(3)
    x = rng.normal(2.0, 1.5, size=40)
    n = len(x)
    m, e = carlo.jackknife(lambda a: a, x)
    assert abs(m - x.mean()) < 1e-12 and abs(e - x.std(ddof=1) / math.sqrt(n)) < 1e-12
    # bias-corrected jackknife of mean^2 is the unbiased xbar^2 - s^2/n
    m2, _ = carlo.jackknife(lambda a: a * a, x)
    assert abs(m2 - (x.mean() ** 2 - x.var(ddof=1) / n)) < 1e-10
    # ratio of two correlated series: error from the delta method to O(1/n)
    y = 5.0 + 0.1 * rng.normal(size=n)
    r, er = carlo.jackknife(lambda a, b: a / b, x, y)
    assert abs(r - x.mean() / y.mean()) < 0.05


def test_autocorrelation_time_ar1(carlo):
    """AR(1) with coefficient rho has integrated autocorrelation time
    0.5 (1 + rho) / (1 - rho) samples."""
    rng = np.random.default_rng(4)
    rho, n = 0.8, 400000
    eps = rng.normal(size=n)
    x = np.empty(n)
    x[0] = eps[0]
    for i in range(1, n):
        x[i] = rho * x[i - 1] + eps[i]
    e, _ = carlo.merge_observable(x, 1)
    tau = 0.5 * (1 + rho) / (1 - rho)
    assert abs(e["autocorr_time"] - tau) < 0.3 * tau, e


def test_binder_ratio_zero_case(carlo):
    ev = {n: f for n, _, f in carlo.evaluables(["Sign", "SignMag4", "SignMag2"], 4)}
    assert ev["BinderRatio"](np.array([0.0]), np.array([0.0]), np.array([1.0]))[0] == 0.0
    assert ev["BinderRatio"](np.array([2.0]), np.array([1.0]), np.array([1.0]))[0] == 0.5


def test_task_from_oracle_bins_matches_ed(sse_pkg, oracle_mod, carlo, meta, tmp_path):
    """One task of the reference's ED job (test/test_jobs.jl:4-29) as 2 runs
    of oracle walkers, binned by 100 sweeps and merged: every evaluable ED
    gives within 4 errors; results.json is written with the reference's
    entry and observable keys."""
    T = float(TS[2])
    model = job_magnet_square(sse_pkg, T)
    flat = model.flat()
    names = flat.observable_names()
    runs = []
    for r in range(2):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat,
                                    sse_pkg.walker_desc(T=T, seed=3100 + r, thermalization=2000))
        o.init()
        rows = o.run_series(2000 + 20000)
        assert len(rows) == 20000
        runs.append(rows.reshape(200, 100, -1).mean(axis=1))
    res = carlo.evaluate_task(names, runs, 100, flat.normalization_site_count)
    exact = ed.magnet_ed(job_magnet_square(sse_pkg, TS[0]), [T], sse_pkg.spin_operators)
    for name, ev in exact.items():
        mean, err = res[name]["mean"][0], res[name]["error"][0]
        z = abs(mean - ev[0]) / max(err, 1e-8)
        assert z <= 4, f"{name}: {mean}+-{err} ED {ev[0]} z={z:.2f}"
    path = _scratch(tmp_path, "results.json")
    carlo.write_results(path, [{"task": "task0001", "parameters": {"T": T, "binsize": 100},
                                "results": res}])
    out = json.load(open(path))
    assert sorted(out[0].keys()) == meta[0]["entry_keys"]
    assert set(out[0]["version"].keys()) == set(meta[0]["version"].keys())
    assert out[0]["results"]["Sign"]["mean"] == 1.0

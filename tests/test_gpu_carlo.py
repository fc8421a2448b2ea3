"""Carlo-compatible persistence and binning through the HIP path.

A walker saved to an HDF5 run file (the reference's checkpoint group plus
the Philox stream position) and loaded into a fresh batch continues bit for
bit; a checkpoint group written from an oracle state resumes on the GPU
exactly as the oracle continues; restored states outside the model are
rejected before any kernel runs; device bins equal the oracle's per-sweep
measurements binned; a small job merged into results.json agrees with exact
diagonalisation."""
import json
import sys

import numpy as np
import pytest

import ed
from fixtures import TS, job_magnet_square, square_heisenberg

pytestmark = pytest.mark.gpu

KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def carlo(sse_pkg):
    if not sse_pkg.h5lite.available():
        pytest.skip("libhdf5 not available")
    return sse_pkg.carlo


@pytest.mark.parametrize("mode", [0, 1])
def test_run_file_roundtrip(sse_pkg, carlo, tmp_path, mode):
    _say(f"carlo: run file round trip mode={mode}")
    flat = square_heisenberg(sse_pkg, 6).flat()
    wd = [sse_pkg.walker_desc(T=0.3, seed=60 + s, thermalization=5) for s in range(3)]
    a = sse_pkg.Batch([flat], wd, mode=mode)
    a.init()
    a.sweep(12)
    for k in range(3):
        carlo.save_run(str(tmp_path / f"run{k}.h5"), a, k)
    with sse_pkg.h5lite.File(str(tmp_path / "run1.h5"), "r") as f:
        with f.group("simulation") as g:
            sc, ops, spins = a.get_walker(1)
            assert g.read("num_operators") == sc["num_operators"]
            assert np.array_equal(g.read("operators"), carlo.opcodes_to_reference(ops))
            assert np.array_equal(g.read("state"), spins)
    b = sse_pkg.Batch([flat], wd, mode=mode)
    b.init()
    b.sweep(3)  # a different state and stream position, overwritten by the load
    for k in range(3):
        carlo.load_run(str(tmp_path / f"run{k}.h5"), b, k)
    a.sweep(15)
    b.sweep(15)
    for k in range(3):
        sa, oa, pa = a.get_walker(k)
        sb, ob, pb = b.get_walker(k)
        assert np.array_equal(oa, ob) and np.array_equal(pa, pb), k
        for key in KEYS:
            assert sa[key] == sb[key], key
    with pytest.raises(ValueError):
        carlo.load_run(str(tmp_path / "run0.h5"), b, 1)  # key of walker 0
    a.close()
    b.close()


def test_oracle_checkpoint_resumes_on_gpu(sse_pkg, oracle_mod, carlo, tmp_path):
    _say("carlo: oracle checkpoint -> GPU")
    flat = square_heisenberg(sse_pkg, 4).flat()
    wd = sse_pkg.walker_desc(T=0.5, seed=91, thermalization=4)
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    o.init()
    o.sweep(10)
    sc, ops, spins = o.state()
    path = str(tmp_path / "o.h5")
    with sse_pkg.h5lite.File(path, "w") as f:
        with f.create_group("simulation") as g:
            carlo.write_mc_group(g, sc, ops, spins)
    b = sse_pkg.Batch([flat], [wd])
    b.init()
    with sse_pkg.h5lite.File(path, "r") as f:
        with f.group("simulation") as g:
            carlo.read_checkpoint(b, 0, g, sweeps=sc["sweeps"], rng_counter=sc["rng_counter"])
    o.sweep(8)
    b.sweep(8)
    gs, gops, gspins = b.get_walker(0)
    os_, oops, ospins = o.state()
    assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins)
    for key in KEYS:
        assert gs[key] == os_[key], key
    b.close()


def test_restored_state_outside_model_rejected(sse_pkg):
    _say("carlo: invalid restored states")
    flat = square_heisenberg(sse_pkg, 4).flat()
    b = sse_pkg.Batch([flat], [sse_pkg.walker_desc(T=0.5, seed=1)])
    b.init()
    sc, ops, spins = b.get_walker(0)
    nv = flat.types[int(flat.bond_type[0])]["n_vertices"]
    bad_ops = ops.copy()
    k = int(np.nonzero(ops)[0][0]) if np.count_nonzero(ops) else 0
    if bad_ops[k] == 0:
        sc["num_operators"] += 1
    bad_ops[k] = 1 | ((1 | (nv << 1)) << 1)  # vertex index = vertex count, bond 0
    with pytest.raises(sse_pkg.capi.SSEError):
        b.set_walker(0, sc, bad_ops, spins)
    sc2, ops2, spins2 = b.get_walker(0)
    bad_spins = spins2.copy()
    bad_spins[3] = 3  # spin-1/2 sites hold states 1..2
    with pytest.raises(sse_pkg.capi.SSEError):
        b.set_walker(0, sc2, ops2, bad_spins)
    bad_spins[3] = 0
    with pytest.raises(sse_pkg.capi.SSEError):
        b.set_walker(0, sc2, ops2, bad_spins)
    b.sweep(2)  # the walker is unchanged and still runs
    b.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_bins_equal_oracle_rows(sse_pkg, oracle_mod, carlo, mode):
    _say(f"carlo: device bins vs oracle mode={mode}")
    flat = square_heisenberg(sse_pkg, 6).flat()
    wd = [sse_pkg.walker_desc(T=0.4, seed=700 + s, thermalization=6) for s in range(4)]
    b = sse_pkg.Batch([flat], wd, mode=mode)
    b.init()
    b.run(6, measure=False)
    bins = carlo.bin_walkers(b, 10, 5)
    b.close()
    assert carlo.binned_names(flat)[-1] == "WormLengthFraction"
    for k, w in enumerate(wd):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, w)
        o.init()
        rows, wlf = [], []
        for _ in range(6 + 50):
            o.sweep(1)
            sc = o.scalars()
            if sc["sweeps"] > 6:
                rows.append(o.measure())
                wlf.append(sc["last_worm_length_fraction"])
        rows = np.array(rows)
        assert len(rows) == 50
        ref = np.zeros((5, rows.shape[1] + 1))
        for i in range(5):
            acc = np.zeros(rows.shape[1])
            for r in rows[10 * i:10 * i + 10]:
                acc = acc + r
            ref[i, :-1] = acc / 10
            ws, wc = 0.0, 0
            for v in wlf[10 * i:10 * i + 10]:
                if v == v:  # measured in thermalized sweeps with n != 0
                    ws, wc = ws + v, wc + 1
            ref[i, -1] = ws / wc if wc else float("nan")
        assert np.allclose(bins[k], ref, rtol=1e-12, atol=1e-12, equal_nan=True), k
        assert np.isfinite(bins[k][:, -1]).all()


def test_run_tasks_results_json_vs_ed(sse_pkg, carlo, tmp_path):
    """Two tasks of the reference's ED job (test/test_jobs.jl:4-29), 8 runs
    each, merged into results.json: Energy and SpecificHeat within 4 errors
    of exact diagonalisation; WormLengthFraction merged like the others."""
    _say("carlo: run_tasks vs ED")
    Ts = [float(TS[2]), float(TS[4])]
    tasks = [{"name": f"task{i + 1:04d}", "model": job_magnet_square(sse_pkg, T), "T": T,
              "thermalization": 1000, "sweeps": 5000, "binsize": 100} for i, T in enumerate(Ts)]
    entries = carlo.run_tasks(tasks, runs_per_task=8, seed0=4242, progress=_say)
    path = str(tmp_path / "results.json")
    carlo.write_results(path, entries)
    out = json.load(open(path))
    exact = ed.magnet_ed(job_magnet_square(sse_pkg, TS[0]), Ts, sse_pkg.spin_operators)
    for i, e in enumerate(out):
        assert e["task"] == f"task{i + 1:04d}" and e["parameters"]["T"] == Ts[i]
        r = e["results"]
        assert r["OperatorCount"]["internal_bin_len"] == 100
        assert r["OperatorCount"]["rebin_count"] == carlo.calc_rebin_count(8 * 50)
        assert r["WormLengthFraction"]["mean"] > 0 and r["WormLengthFraction"]["error"] > 0
        for name in ("Energy", "SpecificHeat"):
            mean, err = r[name]["mean"][0], r[name]["error"][0]
            z = abs(mean - exact[name][i]) / max(err, 1e-8)
            _say(f"T={Ts[i]:.2f} {name}: {mean:.6f}+-{err:.2g} ED {exact[name][i]:.6f} z={z:.2f}")
            assert z <= 4


def test_run_tasks_parallel_tempering_vs_ed(sse_pkg, carlo, tmp_path):
    """Four temperatures of the reference's ED job as one parallel-tempering
    group (8 ladders, an exchange round after every sweep): each task's
    results, which follow its temperature across walkers, agree with exact
    diagonalisation."""
    _say("carlo: run_tasks with parallel tempering vs ED")
    Ts = [float(t) for t in TS[1:5]]
    model = job_magnet_square(sse_pkg, Ts[0])  # one model object: T is the walkers'
    tasks = [{"name": f"task{i + 1:04d}", "model": model, "T": T, "thermalization": 1000,
              "sweeps": 4000, "binsize": 100} for i, T in enumerate(Ts)]
    entries = carlo.run_tasks(tasks, runs_per_task=8, seed0=777, pt_interval=1, pt_seed=5,
                              progress=_say)
    carlo.write_results(str(tmp_path / "results.json"), entries)
    exact = ed.magnet_ed(job_magnet_square(sse_pkg, TS[0]), Ts, sse_pkg.spin_operators)
    for i, e in enumerate(entries):
        assert e["parameters"]["parallel_tempering_interval"] == 1
        for name in ("Energy", "SpecificHeat"):
            mean, err = e["results"][name]["mean"][0], e["results"][name]["error"][0]
            z = abs(mean - exact[name][i]) / max(err, 1e-8)
            _say(f"PT T={Ts[i]:.2f} {name}: {mean:.6f}+-{err:.2g} ED {exact[name][i]:.6f} z={z:.2f}")
            assert z <= 4
    with pytest.raises(ValueError):
        carlo.run_tasks([dict(tasks[0]), dict(tasks[1], model=job_magnet_square(sse_pkg, Ts[1]))],
                        pt_interval=1)

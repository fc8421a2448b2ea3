"""Statistical end-to-end check of the oracle chain against exact
diagonalisation for the reference's three ED jobs (test/test_ed_compare.jl,
test/test_jobs.jl:4-110): 7 temperatures in [0.04, 4], 40000 sweeps after
5000 thermalization sweeps, |z| <= 4 for every observable."""
import numpy as np
import pytest

import ed
import stats
from fixtures import TS, job_bilayer, job_honeycomb, job_magnet_square

SWEEPS = 40000
THERM = 5000
BINSIZE = 1000


def _run(sse_pkg, oracle_mod, model, T, seed):
    flat = model.flat()
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat,
                                sse_pkg.walker_desc(T=T, seed=seed, thermalization=THERM))
    o.init()
    rows = o.run_series(SWEEPS + THERM + 1)
    bins = stats.bin_rows(rows, BINSIZE)
    return stats.jackknife(flat.observable_names(), bins, flat.normalization_site_count)


@pytest.mark.parametrize("job", ["magnet_square", "honeycomb", "bilayer"])
def test_ed_compare_oracle(sse_pkg, oracle_mod, job):
    make = {"magnet_square": job_magnet_square, "honeycomb": job_honeycomb,
            "bilayer": job_bilayer}[job]
    model0 = make(sse_pkg, TS[0])
    H = None
    if job == "bilayer":
        H = ed.magnet_hamiltonian(model0.inner_model, sse_pkg.spin_operators)
        exact = ed.magnet_ed(model0.inner_model, TS, sse_pkg.spin_operators, hamiltonian=H)
        exact = {k: exact[k] for k in ("Energy", "SpecificHeat")}
    else:
        exact = ed.magnet_ed(model0, TS, sse_pkg.spin_operators)
    worst = 0.0
    for i, T in enumerate(TS):
        res = _run(sse_pkg, oracle_mod, make(sse_pkg, T), T, 124535 + i)
        for name, ev in exact.items():
            mean, err = res[name]
            err = err if err != 0 else 1e-8
            z = abs(mean - ev[i]) / err
            worst = max(worst, z)
            assert z <= 4, f"{job} T={T:.2f} {name}: MC {mean}+-{err} ED {ev[i]} z={z:.2f}"

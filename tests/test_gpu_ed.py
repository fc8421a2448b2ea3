"""Observables of the HIP path against exact diagonalisation for the
reference's three ED jobs (test/test_ed_compare.jl, test/test_jobs.jl:4-110).

Each temperature runs 64 independent walkers in one batch; the error bar is
the jackknife over walkers (independent chains).  Gate: |z| <= 4 as in the
reference; the fraction of observables within 2 sigma is also reported."""
import numpy as np
import pytest

import ed
import stats
from fixtures import TS, job_bilayer, job_honeycomb, job_magnet_square

pytestmark = pytest.mark.gpu

WALKERS = 64
THERM = 1000
SWEEPS = 3000


@pytest.mark.parametrize("job", ["magnet_square", "honeycomb", "bilayer"])
def test_ed_compare_gpu(sse_pkg, job):
    make = {"magnet_square": job_magnet_square, "honeycomb": job_honeycomb,
            "bilayer": job_bilayer}[job]
    model0 = make(sse_pkg, TS[0])
    if job == "bilayer":
        H = ed.magnet_hamiltonian(model0.inner_model, sse_pkg.spin_operators)
        exact = ed.magnet_ed(model0.inner_model, TS, sse_pkg.spin_operators, hamiltonian=H)
        exact = {k: exact[k] for k in ("Energy", "SpecificHeat")}
    else:
        exact = ed.magnet_ed(model0, TS, sse_pkg.spin_operators)
    flat = model0.flat()
    walkers = []
    for i, T in enumerate(TS):
        for k in range(WALKERS):
            walkers.append(sse_pkg.walker_desc(model=i, T=float(T), seed=1000 * i + k,
                                               thermalization=THERM))
    flats = [make(sse_pkg, T).flat() for T in TS]
    print(f"ed gpu: {job} {len(walkers)} walkers", flush=True)
    batch = sse_pkg.Batch(flats, walkers)
    batch.init()
    batch.run(THERM + SWEEPS, measure=True)
    acc = batch.accumulators()
    names = flat.observable_names()
    within2 = total = 0
    for i, T in enumerate(TS):
        means = np.array([acc[i * WALKERS + k][0] / acc[i * WALKERS + k][2]
                          for k in range(WALKERS)])
        res = stats.jackknife(names, means, flat.normalization_site_count)
        for name, ev in exact.items():
            mean, err = res[name]
            err = err if err != 0 else 1e-8
            z = abs(mean - ev[i]) / err
            total += 1
            within2 += z <= 2
            if z > 2:
                print(f"  {job} T={T:.2f} {name}: GPU {mean:.6g}+-{err:.2g} ED {ev[i]:.6g} "
                      f"z={z:.2f}", flush=True)
            assert z <= 4, f"{job} T={T:.2f} {name}: GPU {mean}+-{err} ED {ev[i]} z={z:.2f}"
    print(f"{job}: {within2}/{total} observables within 2 sigma", flush=True)
    assert within2 >= 0.9 * total
    batch.close()

"""sse_run_budgets_noalloc (the entry point that allocates no device memory:
task counter and budgets in a static per-GPU scratch array) against
sse_run_budgets, bit for bit, on two calls in a row, the second on a smaller
batch, so the scratch's leftover contents from the first call must not
matter; and a batch beyond the scratch's maximum is rejected."""
import numpy as np
import pytest

from fixtures import square_heisenberg

pytestmark = pytest.mark.gpu

KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _batch_pair(sse, n, seed):
    rng = np.random.default_rng(seed)
    flats, walkers = [], []
    for k in range(n):
        L = int(rng.choice([4, 6, 8]))
        flats.append(square_heisenberg(sse, L, d=float(rng.uniform(-0.4, 0.4)), hz=0.1).flat())
        beta = float(np.exp(rng.uniform(np.log(0.5), np.log(6.0))))
        walkers.append(sse.walker_desc(model=k, T=1.0 / beta, seed=seed * 1000 + k,
                                       thermalization=4))
    a = sse.Batch(flats, walkers)
    b = sse.Batch(flats, walkers)
    a.init()
    b.init()
    return a, b, rng


def _same(a, b, n, tag):
    for k in range(n):
        sa, oa, pa = a.get_walker(k)
        sb, ob, pb = b.get_walker(k)
        assert np.array_equal(oa, ob) and np.array_equal(pa, pb), f"{tag} walker {k}"
        for key in KEYS:
            assert sa[key] == sb[key], f"{tag} walker {k}: {key}"
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2), tag


def test_noalloc_equals_budgets_two_calls(sse_pkg):
    print("no-alloc: 40 walkers, then 12", flush=True)
    for n, seed, measure in ((40, 5, True), (12, 6, False)):
        a, b, rng = _batch_pair(sse_pkg, n, seed)
        budgets = rng.integers(0, 15, size=n)
        a.run_budgets_noalloc(budgets, measure=measure)
        b.run_budgets(budgets, measure=measure)
        _same(a, b, n, f"{n} walkers")
        a.close()
        b.close()


def test_noalloc_rejects_too_many_walkers(sse_pkg):
    flat = square_heisenberg(sse_pkg, 4).flat()
    n = 65537
    b = sse_pkg.Batch([flat], [sse_pkg.walker_desc(T=1.0, seed=s) for s in range(n)])
    b.init()
    with pytest.raises(sse_pkg.capi.SSEError, match="SSE_NOALLOC_MAX_WALKERS"):
        b.run_budgets_noalloc(np.ones(n, dtype=np.int64))
    b.close()

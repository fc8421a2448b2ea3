"""Bit-exact parity of the HIP path (through the C ABI) with the oracle.

Every walker shares one Philox4x32-10 stream with its oracle twin; after
init and after every sweep the operator strings, spins, cutoff, operator
count, worm controller (FP64 bits), RNG counter and scatter count must be
identical.  Measurements must agree to 1e-12 relative (same expressions;
summation order of the estimator prefix sums may differ)."""
import sys

import numpy as np
import pytest

from fixtures import (isconsistent, job_bilayer, job_honeycomb, ladder_dimer, square_heisenberg,
                      vertex_list_fixture)

pytestmark = pytest.mark.gpu

MODES = [0, 1]  # SSE_MODE_WAVE, SSE_MODE_LANE
KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(batch, oracles, tag):
    for k, o in enumerate(oracles):
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert gs["cutoff"] == os_["cutoff"], f"{tag} walker {k}: cutoff"
        assert np.array_equal(gops, oops), f"{tag} walker {k}: ops differ at " \
            f"{np.nonzero(gops != oops)[0][:8]}"
        assert np.array_equal(gspins, ospins), f"{tag} walker {k}: spins"
        for key in KEYS:
            assert gs[key] == os_[key], f"{tag} walker {k}: {key} {gs[key]} != {os_[key]}"


def _run_parity(sse, oracle_mod, model, T, n_walkers, sweeps, mode, seed0=0, records=None,
                chase=None, generic=None, **kw):
    flat = model.flat()
    walkers = [sse.walker_desc(T=T, seed=seed0 + s, **kw) for s in range(n_walkers)]
    _say(f"parity: mode={mode} N={flat.n_sites} walkers={n_walkers} sweeps={sweeps} "
         f"records={records} chase={chase} generic={generic}")
    batch = sse.Batch([flat], walkers, mode=mode)
    if records is not None:
        batch.set_option(sse.capi.SSE_OPT_WORM_RECORDS, records)
    if chase is not None:
        batch.set_option(sse.capi.SSE_OPT_CHASE, chase)
    if generic is not None:
        batch.set_option(sse.capi.SSE_OPT_GENERIC, generic)
    oracles = [oracle_mod.OracleWalker(sse.capi, flat, w) for w in walkers]
    batch.init()
    for o in oracles:
        o.init()
    _compare(batch, oracles, "init")
    for s in range(sweeps):
        batch.sweep(1)
        for o in oracles:
            o.sweep(1)
        _compare(batch, oracles, f"sweep {s}")
        if s % 5 == 4 or s == sweeps - 1:
            gm = batch.measure()
            for k, o in enumerate(oracles):
                om = o.measure()
                assert np.allclose(gm[k], om, rtol=1e-12, atol=1e-13), f"measure walker {k}"
    batch.close()
    return flat, oracles


@pytest.mark.parametrize("mode", MODES)
def test_parity_square_heisenberg(sse_pkg, oracle_mod, mode):
    _run_parity(sse_pkg, oracle_mod, square_heisenberg(sse_pkg, 4), 0.5, 8, 40, mode,
                thermalization=20)


@pytest.mark.parametrize("mode", MODES)
def test_parity_growth_from_zero_cutoff(sse_pkg, oracle_mod, mode):
    _run_parity(sse_pkg, oracle_mod, square_heisenberg(sse_pkg, 6), 0.1, 4, 30, mode,
                init_opstring_cutoff=0, thermalization=5)


@pytest.mark.parametrize("mode", MODES)
def test_parity_mixed_spin_honeycomb(sse_pkg, oracle_mod, mode):
    _run_parity(sse_pkg, oracle_mod, job_honeycomb(sse_pkg, 0.3), 0.3, 6, 25, mode,
                thermalization=10)


@pytest.mark.parametrize("mode", MODES)
def test_parity_xxz_field(sse_pkg, oracle_mod, mode):
    model = square_heisenberg(sse_pkg, 8, d=-0.3, hz=0.1)
    _run_parity(sse_pkg, oracle_mod, model, 0.25, 6, 25, mode, thermalization=10)


@pytest.mark.parametrize("mode", MODES)
def test_parity_ladder_dimer_basis(sse_pkg, oracle_mod, mode):
    _run_parity(sse_pkg, oracle_mod, ladder_dimer(sse_pkg, 16, 0.6), 0.2, 4, 25, mode,
                thermalization=10)


@pytest.mark.parametrize("mode", MODES)
def test_parity_bilayer_cluster(sse_pkg, oracle_mod, mode):
    _run_parity(sse_pkg, oracle_mod, job_bilayer(sse_pkg, 0.2), 0.2, 4, 25, mode,
                thermalization=10)


@pytest.mark.parametrize("records", [1, 2, 3])
def test_parity_worm_record_layouts(sse_pkg, oracle_mod, records):
    # 10-byte LDS records (1), global-memory records (2) and the hybrid (3: codes
    # in LDS, links in global memory) against the oracle; the default 8-byte
    # layout runs in every other wave-mode test
    _run_parity(sse_pkg, oracle_mod, square_heisenberg(sse_pkg, 4), 0.5, 6, 25, 0,
                thermalization=10, records=records)
    _run_parity(sse_pkg, oracle_mod, job_honeycomb(sse_pkg, 0.4), 0.4, 4, 20, 0, seed0=50,
                thermalization=10, records=records)


@pytest.mark.parametrize("records", [0, 1])
def test_parity_lookahead_chase(sse_pkg, oracle_mod, records):
    # the record-lookahead traversal over 8-byte (0) and 10-byte (1) LDS
    # records: one-worm (S=1/2) and multi-worm (spin-1, dim-4 cluster) tables,
    # operators alone on their site (self-linked through the closure) at 4x4
    _run_parity(sse_pkg, oracle_mod, square_heisenberg(sse_pkg, 4), 0.5, 6, 25, 0,
                thermalization=10, records=records, chase=1)
    _run_parity(sse_pkg, oracle_mod, job_honeycomb(sse_pkg, 0.4), 0.4, 4, 20, 0, seed0=50,
                thermalization=10, records=records, chase=1)
    _run_parity(sse_pkg, oracle_mod, ladder_dimer(sse_pkg, 8, 0.6), 0.2, 4, 20, 0, seed0=80,
                thermalization=10, records=records, chase=1)


@pytest.mark.parametrize("kind", ["square", "honeycomb", "xxz", "ladder", "bilayer"])
def test_parity_generic_kernels_two_site(sse_pkg, oracle_mod, kind):
    """SSE_OPT_GENERIC: two-site models through the generic wave kernels
    (chase codes in LDS, the generic chase table) equal the oracle after
    every sweep: spin-1 exit lists beyond four entries (honeycomb), a field
    (xxz), four-state sites (ladder, bilayer)."""
    model, T = {"square": (square_heisenberg(sse_pkg, 6), 0.4),
                "honeycomb": (job_honeycomb(sse_pkg, 0.4), 0.4),
                "xxz": (square_heisenberg(sse_pkg, 8, d=-0.3, hz=0.1), 0.25),
                "ladder": (ladder_dimer(sse_pkg, 16, 0.6), 0.2),
                "bilayer": (job_bilayer(sse_pkg, 0.2), 0.2)}[kind]
    _run_parity(sse_pkg, oracle_mod, model, T, 4, 20, 0, seed0=90, generic=1, thermalization=6)


def test_generic_switch_mid_chain(sse_pkg):
    """Switching SSE_OPT_GENERIC between calls continues the same chains:
    a batch alternating the kernel families equals one that never switches,
    with and without measurement; the option's errors."""
    hc = job_honeycomb(sse_pkg, 0.5).flat()
    sq = square_heisenberg(sse_pkg, 6).flat()
    wd = [sse_pkg.walker_desc(model=s % 2, T=0.3 + 0.05 * (s % 3), seed=500 + s, thermalization=8)
          for s in range(12)]
    _say("generic kernels switched mid-chain: 12 walkers, 2 models")
    a = sse_pkg.Batch([sq, hc], wd)
    b = sse_pkg.Batch([sq, hc], wd)
    a.init()
    b.init()
    for rep, (k, meas) in enumerate(((15, True), (4, False), (21, True), (3, True))):
        a.set_option(sse_pkg.capi.SSE_OPT_GENERIC, rep % 2)
        a.run(k, measure=meas)
        b.run(k, measure=meas)
        _assert_same_walkers(a, b, len(wd), f"call {rep}")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.allclose(s1, s2, rtol=1e-12) and np.allclose(q1, q2, rtol=1e-12)
    with pytest.raises(RuntimeError):
        a.set_option(sse_pkg.capi.SSE_OPT_GENERIC, 2)
    a.set_option(sse_pkg.capi.SSE_OPT_GENERIC, 1)
    with pytest.raises(RuntimeError):  # no work-queue launch in the generic family
        a.set_option(sse_pkg.capi.SSE_OPT_PERSISTENT, 1)
    a.close()
    b.close()
    lane = sse_pkg.Batch([sq], [wd[0]], mode=1)
    with pytest.raises(RuntimeError):
        lane.set_option(sse_pkg.capi.SSE_OPT_GENERIC, 1)
    lane.close()


@pytest.mark.parametrize("state", [0, 3])
def test_leg_state_out_of_range_rejected(sse_pkg, state):
    """A vertex leg state outside 1..dim (the kernels index per-state tables
    with it) is rejected when the batch is created; the valid model is
    accepted."""
    import copy
    flat = square_heisenberg(sse_pkg, 4).flat()
    bad = copy.copy(flat)
    bad.types = [dict(t) for t in flat.types]
    ls = bad.types[0]["leg_states"].copy()
    ls[1, 2] = state
    bad.types[0]["leg_states"] = ls
    wd = [sse_pkg.walker_desc(T=0.5, seed=1)]
    _say(f"leg state {state} out of range")
    with pytest.raises(sse_pkg.capi.SSEError, match="state out of range"):
        sse_pkg.Batch([bad], wd)
    sse_pkg.Batch([flat], wd).close()


@pytest.mark.parametrize("mode", MODES)
def test_vertex_list_golden_gpu(sse_pkg, mode):
    """Reference golden vertex list, built by the device vertex-list phase."""
    n_sites, bond_sites, ops, links, first, last = vertex_list_fixture()
    _say(f"vertex list golden: mode={mode}")
    # a 4th bond touches site 4 so every site has a dimension; no operator uses it
    bonds = [sse_pkg.SSEBond(1, (1, 2)), sse_pkg.SSEBond(1, (2, 3)), sse_pkg.SSEBond(1, (1, 3)),
             sse_pkg.SSEBond(1, (3, 4))]
    sp, sz = sse_pkg.spin_operators(2)
    H = np.kron(sz, sz) + 0.5 * (np.kron(sp, sp.T) + np.kron(sp.T, sp))
    vd = sse_pkg.VertexData((2, 2), H)
    flat = sse_pkg.FlatModel(sse_pkg.SSEData([vd], bonds))
    batch = sse_pkg.Batch([flat], [sse_pkg.walker_desc(T=1.0, seed=1)], mode=mode)
    batch.init()
    batch.set_walker(0, {"num_operators": int((ops != 0).sum()), "num_worms": 5.0,
                         "avg_worm_length": 1.0}, ops, np.ones(4, dtype=np.uint8))
    batch.sweep_phases(2)
    gl, gf, gla = batch.vertex_list(0)
    assert np.array_equal(gl, links)
    assert np.array_equal(gf, first)
    assert np.array_equal(gla, last)
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_checkpoint_roundtrip_gpu(sse_pkg, oracle_mod, mode):
    model = square_heisenberg(sse_pkg, 6)
    flat = model.flat()
    wd = [sse_pkg.walker_desc(T=0.3, seed=40 + s, thermalization=5) for s in range(3)]
    _say(f"checkpoint: mode={mode}")
    a = sse_pkg.Batch([flat], wd, mode=mode)
    a.init()
    a.sweep(12)
    saved = [a.get_walker(k) for k in range(3)]
    b = sse_pkg.Batch([flat], wd, mode=mode)
    b.init()
    for k, (sc, ops, spins) in enumerate(saved):
        b.set_walker(k, sc, ops, spins)
    a.sweep(15)
    b.sweep(15)
    for k in range(3):
        sa, oa, pa = a.get_walker(k)
        sb, ob, pb = b.get_walker(k)
        assert np.array_equal(oa, ob) and np.array_equal(pa, pb)
        for key in KEYS:
            assert sa[key] == sb[key]
    a.close()
    b.close()


@pytest.mark.parametrize("mode", MODES)
def test_fused_run_accumulators(sse_pkg, oracle_mod, mode):
    """sse_run (sweep + measure once thermalized, accumulated on device) equals
    the oracle's per-sweep measurement series summed."""
    model = square_heisenberg(sse_pkg, 4)
    flat = model.flat()
    wd = [sse_pkg.walker_desc(T=0.4, seed=70 + s, thermalization=10) for s in range(4)]
    _say(f"fused run: mode={mode}")
    batch = sse_pkg.Batch([flat], wd, mode=mode)
    batch.init()
    batch.run(40, measure=True)
    acc = batch.accumulators()
    for k, w in enumerate(wd):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, w)
        o.init()
        rows = o.run_series(40)
        s, s2, cnt = acc[k]
        assert cnt == len(rows) == 40 - 10
        assert np.allclose(s, rows.sum(axis=0), rtol=1e-12, atol=1e-12)
        assert np.allclose(s2, (rows ** 2).sum(axis=0), rtol=1e-12, atol=1e-12)
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_consistency_long_chain(sse_pkg, mode):
    """Worldlines stay consistent over many sweeps of a batch (spin-1 sites too)."""
    model = job_honeycomb(sse_pkg, 0.1)
    flat = model.flat()
    wd = [sse_pkg.walker_desc(T=0.1, seed=s, thermalization=50) for s in range(16)]
    _say(f"long chain: mode={mode}")
    batch = sse_pkg.Batch([flat], wd, mode=mode)
    batch.init()
    for _ in range(10):
        batch.sweep(20)
        for k in range(16):
            sc, ops, spins = batch.get_walker(k)
            assert isconsistent(flat, ops, spins)
            assert sc["num_operators"] == int((ops != 0).sum())
    batch.close()


def test_parity_forced_serial_chunks(sse_pkg, oracle_mod):
    """The slot-by-slot chunk path of the wave kernel (taken on a Lemire
    rejection) is forced for every chunk and must stay bit-exact."""
    model = square_heisenberg(sse_pkg, 6)
    flat = model.flat()
    walkers = [sse_pkg.walker_desc(T=0.3, seed=90 + s, thermalization=8) for s in range(4)]
    _say("forced serial chunks: wave mode")
    batch = sse_pkg.Batch([flat], walkers, mode=0)
    oracles = [oracle_mod.OracleWalker(sse_pkg.capi, flat, w) for w in walkers]
    batch.init()
    for o in oracles:
        o.init()
    for s in range(15):
        batch.sweep_phases(1 | 2 | 4 | 8)
        for o in oracles:
            o.sweep(1)
        _compare(batch, oracles, f"serial sweep {s}")
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_parity_many_walkers_c2_shape(sse_pkg, oracle_mod, mode):
    """16x16 beta=8 Heisenberg (the batched workload's model) for a few sweeps
    from init, 32 walkers with seeds 0..31."""
    model = square_heisenberg(sse_pkg, 16)
    flat = model.flat()
    walkers = [sse_pkg.walker_desc(T=0.125, seed=s, thermalization=2000) for s in range(32)]
    _say(f"c2 shape: mode={mode}")
    batch = sse_pkg.Batch([flat], walkers, mode=mode)
    batch.init()
    batch.sweep(12)
    for k in (0, 7, 31):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, walkers[k])
        o.init()
        o.sweep(12)
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins)
        for key in KEYS:
            assert gs[key] == os_[key]
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_parity_heterogeneous_batch(sse_pkg, oracle_mod, mode):
    """One batch mixing lattices, temperatures and XXZ anisotropies (one model
    per walker); every walker bit-exact against its own oracle."""
    rng = np.random.default_rng(11)
    flats, walkers = [], []
    for k in range(8):
        L = int(rng.choice([4, 6, 8]))
        jz = float(rng.uniform(0.5, 1.5))
        model = square_heisenberg(sse_pkg, L, d=jz - 1.0, hz=0.1)
        flats.append(model.flat())
        beta = float(np.exp(rng.uniform(np.log(0.5), np.log(16.0))))
        walkers.append(sse_pkg.walker_desc(model=k, T=1.0 / beta, seed=300 + k, thermalization=6))
    _say(f"heterogeneous: mode={mode}")
    batch = sse_pkg.Batch(flats, walkers, mode=mode)
    batch.init()
    oracles = []
    for k, w in enumerate(walkers):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flats[k], sse_pkg.walker_desc(
            T=w.T, seed=w.seed, thermalization=6))
        o.init()
        oracles.append(o)
    for s in range(15):
        batch.sweep(1)
        for o in oracles:
            o.sweep(1)
        _compare(batch, oracles, f"hetero sweep {s}")
    gm = batch.measure()
    for k, o in enumerate(oracles):
        assert np.allclose(gm[k], o.measure(), rtol=1e-12, atol=1e-13)
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_graph_replay_equals_eager(sse_pkg, oracle_mod, mode):
    """sse_run replaying a captured HIP graph (several replays with eager work
    and host synchronization in between) gives the eager result bit for bit."""
    model = square_heisenberg(sse_pkg, 6)
    flat = model.flat()
    wd = [sse_pkg.walker_desc(T=0.3, seed=120 + s, thermalization=10) for s in range(6)]
    _say(f"graph replay: mode={mode}")
    a = sse_pkg.Batch([flat], wd, mode=mode)
    b = sse_pkg.Batch([flat], wd, mode=mode)
    for x in (a, b):
        x.set_option(sse_pkg.capi.SSE_OPT_STREAMS, 1)
    a.set_option(sse_pkg.capi.SSE_OPT_TIMING, 0)      # graph path
    b.set_option(sse_pkg.capi.SSE_OPT_GRAPH, 0)       # eager path
    b.set_option(sse_pkg.capi.SSE_OPT_TIMING, 0)
    a.init()
    b.init()
    for rep in range(4):
        a.run(70, measure=True)    # 2 replays of a 32-sweep graph + 6 eager
        b.run(70, measure=True)
        a.measure()                # eager work between replays
        for k in range(6):
            sa, oa, pa = a.get_walker(k)
            sb, ob, pb = b.get_walker(k)
            assert np.array_equal(oa, ob) and np.array_equal(pa, pb), f"replay {rep} walker {k}"
            for key in KEYS:
                assert sa[key] == sb[key]
    acc_a, acc_b = a.accumulators(), b.accumulators()
    for (s1, q1, c1), (s2, q2, c2) in zip(acc_a, acc_b):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    # and against the oracle for one walker
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd[0])
    o.init()
    o.sweep(280)
    sa, oa, pa = a.get_walker(0)
    assert np.array_equal(oa, o.ops()) and sa["rng_counter"] == o.scalars()["rng_counter"]
    a.close()
    b.close()


def _eager(sse, batch):
    """one stream, no graphs, no per-kernel events: plain kernel launches"""
    batch.set_option(sse.capi.SSE_OPT_STREAMS, 1)
    for opt in (sse.capi.SSE_OPT_GRAPH, sse.capi.SSE_OPT_TIMING):
        batch.set_option(opt, 0)


def _assert_same_walkers(a, b, n, tag):
    for k in range(n):
        sa, oa, pa = a.get_walker(k)
        sb, ob, pb = b.get_walker(k)
        assert np.array_equal(oa, ob), f"{tag} walker {k}: ops"
        assert np.array_equal(pa, pb), f"{tag} walker {k}: spins"
        for key in KEYS:
            assert sa[key] == sb[key], f"{tag} walker {k}: {key}"


@pytest.mark.parametrize("graph", [1, 0])
def test_persistent_equals_per_sweep_kernels(sse_pkg, oracle_mod, graph):
    """The persistent work-queue launch (several calls, with and without the
    fused measurement, graph replays or direct launches) gives the per-sweep
    kernels' result bit for bit, on a batch mixing two models."""
    sq = square_heisenberg(sse_pkg, 6).flat()
    hc = job_honeycomb(sse_pkg, 0.5).flat()
    wd = [sse_pkg.walker_desc(model=s % 2, T=0.3 + 0.05 * (s % 3), seed=300 + s, thermalization=12)
          for s in range(10)]
    _say(f"persistent vs per-sweep kernels: 10 walkers, 2 models, graph={graph}")
    a = sse_pkg.Batch([sq, hc], wd)
    b = sse_pkg.Batch([sq, hc], wd)
    a.set_option(sse_pkg.capi.SSE_OPT_PERSISTENT, 1)
    a.set_option(sse_pkg.capi.SSE_OPT_GRAPH, graph)
    _eager(sse_pkg, b)
    a.init()
    b.init()
    for rep, (k, meas) in enumerate(((37, True), (9, False), (70, True), (2, True), (37, True))):
        a.run(k, measure=meas)
        b.run(k, measure=meas)
        ms, sw = a.call_time()
        assert sw == k and ms > 0.0
        _assert_same_walkers(a, b, len(wd), f"call {rep}")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    # one walker of each model against the oracle
    for w in (0, 1):
        o = oracle_mod.OracleWalker(sse_pkg.capi, [sq, hc][w], sse_pkg.walker_desc(
            T=wd[w].T, seed=wd[w].seed, thermalization=12))
        o.init()
        o.sweep(37 + 9 + 70 + 2 + 37)
        gs, gops, gspins = a.get_walker(w)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins)
        for key in KEYS:
            assert gs[key] == os_[key]
    a.close()
    b.close()


@pytest.mark.parametrize("streams,graph", [(1, 0), (1, 1), (3, 0)])
def test_fused_worm_measure_equals_separate(sse_pkg, oracle_mod, streams, graph):
    """SSE_OPT_FUSE_MEASURE: the worm update and the measurement in one
    kernel give the separate kernels' walkers and accumulators bit for bit
    (one stream eager and through graphs, three stream groups), on a batch
    mixing two models, with sweeps before and after thermalization and a
    call without measurement in between."""
    sq = square_heisenberg(sse_pkg, 6).flat()
    hc = job_honeycomb(sse_pkg, 0.5).flat()
    wd = [sse_pkg.walker_desc(model=s % 2, T=0.3 + 0.05 * (s % 3), seed=500 + s, thermalization=12)
          for s in range(12)]
    _say(f"fused worm+measure: streams={streams} graph={graph}")
    a = sse_pkg.Batch([sq, hc], wd)
    b = sse_pkg.Batch([sq, hc], wd)
    for x in (a, b):
        x.set_option(sse_pkg.capi.SSE_OPT_STREAMS, streams)
        x.set_option(sse_pkg.capi.SSE_OPT_GRAPH, graph)
    a.set_option(sse_pkg.capi.SSE_OPT_FUSE_MEASURE, 1)
    a.init()
    b.init()
    for rep, (k, meas) in enumerate(((20, True), (5, False), (33, True))):
        a.run(k, measure=meas)
        b.run(k, measure=meas)
        _assert_same_walkers(a, b, len(wd), f"call {rep}")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    for (s1, q1, c1), (s2, q2, c2) in zip(a.worm_fractions(), b.worm_fractions()):
        assert c1 == c2 and s1 == s2 and q1 == q2
    o = oracle_mod.OracleWalker(sse_pkg.capi, hc, sse_pkg.walker_desc(
        T=wd[1].T, seed=wd[1].seed, thermalization=12))
    o.init()
    rows = o.run_series(58)  # sweeps 13..58; the call without measurement ran 21..25
    kept = np.concatenate([rows[:8], rows[13:]])
    s1, _, c1 = a.accumulators()[1]
    assert c1 == len(kept) == 41
    assert np.allclose(s1, kept.sum(axis=0), rtol=1e-12, atol=1e-12)
    a.close()
    b.close()


def test_persistent_many_walkers(sse_pkg):
    """More walkers than the launch has workgroups: each workgroup takes
    several walkers from the task counter."""
    flat = square_heisenberg(sse_pkg, 4).flat()
    wd = [sse_pkg.walker_desc(T=0.5, seed=900 + s, thermalization=3) for s in range(3000)]
    _say("persistent: 3000 walkers x 12 sweeps vs per-sweep kernels")
    a = sse_pkg.Batch([flat], wd)
    b = sse_pkg.Batch([flat], wd)
    a.set_option(sse_pkg.capi.SSE_OPT_PERSISTENT, 1)
    _eager(sse_pkg, b)
    a.init()
    b.init()
    a.run(12, measure=True)
    b.run(12, measure=True)
    _assert_same_walkers(a, b, len(wd), "3000 walkers")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    a.close()
    b.close()


@pytest.mark.parametrize("records,mode", [(0, 0), (2, 0), (3, 0), (0, 1)])
def test_run_budgets_against_oracle(sse_pkg, oracle_mod, records, mode):
    """sse_run_budgets: every walker of a heterogeneous batch advances by its
    own sweep count in one persistent launch (0 for some), twice; each walker
    equals its oracle twin after the same number of sweeps, in the LDS
    records and in the global-memory layouts (RecGV / RecHV, links read by
    vector loads), and in lane mode (one launch per sweep index, walkers
    past their budget idle)."""
    rng = np.random.default_rng(21)
    flats, walkers = [], []
    for k in range(12):
        L = int(rng.choice([4, 6, 8]))
        model = square_heisenberg(sse_pkg, L, d=float(rng.uniform(-0.5, 0.5)), hz=0.1)
        flats.append(model.flat())
        beta = float(np.exp(rng.uniform(np.log(0.5), np.log(8.0))))
        walkers.append(sse_pkg.walker_desc(model=k, T=1.0 / beta, seed=700 + k, thermalization=5))
    _say(f"run_budgets: 12 walkers, records={records}, mode={mode}")
    batch = sse_pkg.Batch(flats, walkers, mode=mode)
    if mode == 0:
        batch.set_option(sse_pkg.capi.SSE_OPT_WORM_RECORDS, records)
    batch.init()
    b1 = rng.integers(0, 20, size=12)
    b1[3] = 0
    b2 = rng.integers(0, 20, size=12)
    batch.run_budgets(b1, measure=True)
    batch.run_budgets(b2, measure=True)
    acc = batch.accumulators()
    for k, w in enumerate(walkers):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flats[k], sse_pkg.walker_desc(
            T=w.T, seed=w.seed, thermalization=5))
        o.init()
        rows = o.run_series(int(b1[k] + b2[k]))
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins), f"walker {k}"
        for key in KEYS:
            assert gs[key] == os_[key], f"walker {k}: {key}"
        assert gs["sweeps"] == b1[k] + b2[k]
        s1, s2, cnt = acc[k]
        assert cnt == len(rows)
        if cnt:
            assert np.allclose(s1, rows.sum(axis=0), rtol=1e-12, atol=1e-12), f"walker {k}"
    batch.close()


def test_run_budgets_kernel_timing(sse_pkg):
    """SSE_OPT_TIMING 2 on sse_run_budgets: each call that launches reports
    one persistent launch with a positive duration through sse_kernel_times
    (none for a call whose budgets are all 0), and the chains and
    accumulators equal those of the same calls without per-kernel events."""
    flat = square_heisenberg(sse_pkg, 6).flat()
    wd = [sse_pkg.walker_desc(T=0.5, seed=1300 + s, thermalization=4) for s in range(40)]
    _say("run_budgets kernel timing: 40 walkers, 3 calls, timing 2 vs 1")
    a = sse_pkg.Batch([flat], wd)
    b = sse_pkg.Batch([flat], wd)
    a.set_option(sse_pkg.capi.SSE_OPT_TIMING, 2)
    a.init()
    b.init()
    rng = np.random.default_rng(5)
    for budgets in (rng.integers(1, 9, size=40), np.zeros(40, dtype=np.int64),
                    rng.integers(0, 9, size=40)):
        a.run_budgets(budgets, measure=True)
        b.run_budgets(budgets, measure=True)
        ms, n = a.kernel_times()
        if budgets.any():
            assert list(n) == [0, 0, 0, 1] and ms[3] > 0.0, (ms, n)
        else:
            assert list(n) == [0, 0, 0, 0], (ms, n)
    _assert_same_walkers(a, b, len(wd), "timed budgets")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    a.close()
    b.close()


@pytest.mark.parametrize("graph", [1, 2])
def test_walker_groups_equal_single_stream(sse_pkg, oracle_mod, graph):
    """Walker groups on separate streams give the single-stream result bit for
    bit: calls of several lengths with and without the fused measurement, a
    batch mixing two models, and the oracle for two walkers; direct launches
    (graph 1) and replays of the captured multi-stream pipeline (graph 2:
    chunks of 10 sweeps, the rest launched directly)."""
    sq = square_heisenberg(sse_pkg, 6).flat()
    hc = job_honeycomb(sse_pkg, 0.5).flat()
    wd = [sse_pkg.walker_desc(model=s % 2, T=0.3 + 0.05 * (s % 3), seed=300 + s, thermalization=12)
          for s in range(10)]
    _say(f"walker groups vs one stream: 10 walkers, 2 models, graph={graph}")
    a = sse_pkg.Batch([sq, hc], wd)
    b = sse_pkg.Batch([sq, hc], wd)
    a.set_option(sse_pkg.capi.SSE_OPT_STREAMS, 3)
    a.set_option(sse_pkg.capi.SSE_OPT_GRAPH, graph)
    _eager(sse_pkg, b)
    a.init()
    b.init()
    for rep, (k, meas) in enumerate(((37, True), (9, False), (70, True), (1, True))):
        a.run(k, measure=meas)
        b.run(k, measure=meas)
        ms, sw = a.call_time()
        assert sw == k and ms > 0.0
        _assert_same_walkers(a, b, len(wd), f"call {rep}")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    for w in (0, 1):
        o = oracle_mod.OracleWalker(sse_pkg.capi, [sq, hc][w], sse_pkg.walker_desc(
            T=wd[w].T, seed=wd[w].seed, thermalization=12))
        o.init()
        o.sweep(37 + 9 + 70 + 1)
        gs, gops, gspins = a.get_walker(w)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins)
        for key in KEYS:
            assert gs[key] == os_[key]
    a.close()
    b.close()


def test_walker_groups_many_walkers(sse_pkg):
    """3000 walkers in 7 groups against one stream, bit for bit."""
    flat = square_heisenberg(sse_pkg, 4).flat()
    wd = [sse_pkg.walker_desc(T=0.5, seed=900 + s, thermalization=3) for s in range(3000)]
    _say("walker groups: 3000 walkers x 12 sweeps, 7 groups vs one stream")
    a = sse_pkg.Batch([flat], wd)
    b = sse_pkg.Batch([flat], wd)
    a.set_option(sse_pkg.capi.SSE_OPT_STREAMS, 7)
    _eager(sse_pkg, b)
    a.init()
    b.init()
    a.run(12, measure=True)
    b.run(12, measure=True)
    _assert_same_walkers(a, b, len(wd), "3000 walkers")
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    a.close()
    b.close()


def test_full_size_c2_properties(sse_pkg, oracle_mod):
    """The benchmark's full workload (16x16 Heisenberg, beta=8, 4096 walkers in
    two stream groups, 150 sweeps with the fused measurement after 50
    thermalization sweeps), checked through size-independent properties:
    sampled walkers bit-exact against the oracle, operator counts equal to the
    non-identity operators of every walker, every cutoff on the growth
    sequence of its initial cutoff, worldline consistency of a sample, and
    one measurement per walker per thermalized sweep."""
    model = square_heisenberg(sse_pkg, 16)
    flat = model.flat()
    nw, sweeps, therm = 4096, 150, 50
    walkers = [sse_pkg.walker_desc(T=0.125, seed=s, thermalization=therm) for s in range(nw)]
    _say(f"full size: {nw} walkers, {sweeps} sweeps")
    batch = sse_pkg.Batch([flat], walkers)
    batch.init()
    batch.run(sweeps, measure=True)
    acc = batch.accumulators()
    assert all(a[2] == sweeps - therm for a in acc)
    rng = np.random.default_rng(11)
    sample = sorted(set(rng.integers(0, nw, 4).tolist()) | {0, nw - 1})
    for k in sample:
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, walkers[k])
        o.init()
        rows = o.run_series(sweeps)
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins), f"walker {k}"
        for key in KEYS:
            assert gs[key] == os_[key], f"walker {k}: {key}"
        assert np.allclose(acc[k][0], rows.sum(axis=0), rtol=1e-11, atol=1e-9), f"walker {k}"
    L0 = int(round(flat.n_sites * 0.125))
    growth = {L0}
    x = L0
    for _ in range(64):
        x = (3 * x) // 2 + 100
        growth.add(x)
    for k in range(nw):
        sc, ops, spins = batch.get_walker(k, with_spins=(k % 64 == 0))
        assert sc["num_operators"] == int((ops != 0).sum()), f"walker {k}"
        assert sc["cutoff"] in growth, f"walker {k}: cutoff {sc['cutoff']}"
        if k % 64 == 0:
            assert isconsistent(flat, ops, spins), f"walker {k}"
    batch.close()


def test_schedule_order_and_group_lds(sse_pkg):
    """A heterogeneous batch swept with the size-ordered schedule (stream
    groups of similar walkers, each group's worm kernel sized for its own
    largest walker) and with walker index order ends in identical walkers
    and accumulators."""
    rng = np.random.default_rng(5)
    flats, walkers = [], []
    for k in range(96):
        L = int(rng.choice([4, 8, 12, 16]))
        model = square_heisenberg(sse_pkg, L, d=float(rng.uniform(-0.5, 0.5)), hz=0.1)
        flats.append(model.flat())
        beta = float(np.exp(rng.uniform(np.log(0.5), np.log(16.0))))
        walkers.append(sse_pkg.walker_desc(model=k, T=1.0 / beta, seed=900 + k, thermalization=10))
    results = []
    for order in (1, 0):
        _say(f"schedule order={order}")
        batch = sse_pkg.Batch(flats, walkers)
        batch.set_option(sse_pkg.capi.SSE_OPT_ORDER, order)
        batch.set_option(sse_pkg.capi.SSE_OPT_STREAMS, 3)
        batch.init()
        for _ in range(4):
            batch.run(15, measure=True)
        states = [batch.get_walker(k) for k in range(len(walkers))]
        results.append((states, batch.accumulators()))
        batch.close()
    (s1, a1), (s0, a0) = results
    for k in range(len(walkers)):
        g1, g0 = s1[k], s0[k]
        assert np.array_equal(g1[1], g0[1]) and np.array_equal(g1[2], g0[2]), f"walker {k}"
        for key in KEYS:
            assert g1[0][key] == g0[0][key], f"walker {k}: {key}"
        assert np.array_equal(a1[k][0], a0[k][0]) and a1[k][2] == a0[k][2], f"walker {k} sums"

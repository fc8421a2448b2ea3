"""Operator-string estimators on the host (the path for estimators without a
device table), against the oracle's fused measurement on the CPU and the
measurement kernel on the GPU.

The host loop (src/sse.jl:321-376) with the reference's MagnetizationEstimator
restated as a host estimator must give the same observables as the oracle
(which restates the same estimator, pinned against ED) for spin-1/2, mixed
spin-1/2 / spin-1 and cluster-basis models; a user estimator sees every
non-identity operator once, in order, with the state propagated through the
off-diagonal ones, and the propagated state closes on itself."""
import numpy as np
import pytest

from fixtures import job_bilayer, job_honeycomb, job_magnet_square, ladder_dimer, square_heisenberg


def _host_types(sse_pkg, model):
    H = sse_pkg.host_estimators.HostMagnetizationEstimator
    return [H.make(e.q, e.stagger_uc, e.prefix, e.tag) for e in model.get_opstring_estimators()]


def _models(sse_pkg):
    return [("square", square_heisenberg(sse_pkg, 4), 0.5),
            ("job_square", job_magnet_square(sse_pkg, 0.7), 0.7),
            ("honeycomb", job_honeycomb(sse_pkg, 0.6), 0.6),
            ("ladder", ladder_dimer(sse_pkg, 6, 0.6), 0.3),
            ("bilayer", job_bilayer(sse_pkg, 0.8), 0.8)]


@pytest.mark.parametrize("which", range(5))
def test_host_loop_matches_oracle(sse_pkg, oracle_mod, which):
    name, model, T = _models(sse_pkg)[which]
    flat = model.flat()
    sse_data = model.generate_sse_data()
    names = flat.observable_names()
    hm = sse_pkg.host_estimators
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(T=T, seed=11 + which))
    o.init()
    for step in range(4):
        o.sweep(5)
        sc, ops, spins = o.state()
        host = hm.measure_state(model, sse_data, T, sc, sse_pkg.carlo.opcodes_to_reference(ops),
                                spins, _host_types(sse_pkg, model))
        ref = o.measure()
        got = np.array([host[n] for n in names])
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-13), \
            f"{name} step {step}: {[(n, a, b) for n, a, b in zip(names, got, ref) if a != b][:3]}"


class _Recorder:
    """A user estimator: counts off-diagonal operators, checks that the state
    it is handed at each operator is the state after that operator on the
    operator's sites (its upper legs), and records the state after the last
    operator."""

    @classmethod
    def init(cls, model, state):
        e = cls()
        e.start = np.array(state)
        e.seen, e.offdiag, e.last, e.mismatch = [], 0, None, 0
        return e

    def measure(self, op, state, sse_data):
        self.seen.append(op.code)
        self.offdiag += not op.isdiagonal()
        bond = sse_data.bonds[op.get_bond() - 1]
        ls = sse_data.get_vertex_data(op.get_bond()).get_leg_state(op.get_vertex())
        ns = len(bond.sites)
        self.mismatch += sum(int(state[s - 1]) != int(ls[ns + i]) for i, s in enumerate(bond.sites))
        self.last = np.array(state)

    def result(self, T, sign):
        return {"SignOffDiagonal": sign * self.offdiag / max(1, len(self.seen))}


def test_user_estimator_protocol(sse_pkg, oracle_mod):
    model = job_honeycomb(sse_pkg, 0.5)
    flat = model.flat()
    sse_data = model.generate_sse_data()
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(T=0.5, seed=3))
    o.init()
    o.sweep(20)
    sc, ops, spins = o.state()
    codes = sse_pkg.carlo.opcodes_to_reference(ops)
    rec = []

    class R(_Recorder):
        @classmethod
        def init(cls, model, state):
            e = super().init(model, state)
            rec.append(e)
            return e

    out = sse_pkg.host_estimators.measure_state(model, sse_data, 0.5, sc, codes, spins, [R])
    e = rec[0]
    nz = [int(c) for c in codes if c]
    assert e.seen == nz and len(nz) == sc["num_operators"]
    off = int(np.count_nonzero((ops != 0) & ((ops >> 1) & 1 == 0)))
    assert e.offdiag == off and off > 0
    assert e.mismatch == 0
    assert out["SignOffDiagonal"] == out["Sign"] * off / len(nz)
    assert np.array_equal(e.start, spins)
    assert np.array_equal(e.last, spins)  # periodic in imaginary time


@pytest.mark.gpu
def test_host_walker_measure_matches_device(sse_pkg):
    print("host estimators vs measurement kernel", flush=True)
    for name, model, T in _models(sse_pkg):
        flat = model.flat()
        walkers = [sse_pkg.walker_desc(T=T, seed=300 + s) for s in range(3)]
        b = sse_pkg.Batch([flat], walkers)
        b.init()
        b.sweep(15)
        dev = b.measure()
        sse_data = model.generate_sse_data()
        names = flat.observable_names()
        for k in range(3):
            host = sse_pkg.host_estimators.measure_walker(b, k, model, _host_types(sse_pkg, model),
                                                          sse_data)
            got = np.array([host[n] for n in names])
            assert np.allclose(got, dev[k], rtol=1e-12, atol=1e-13), f"{name} walker {k}"
        b.close()

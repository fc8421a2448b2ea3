"""Operator-string estimators on the host: the path for estimators that have
no device table.

The reference lets a model supply any `AbstractOpstringEstimator`
(src/abstract_opstring_estimator.jl:3-69): `init(type, model, state)`, then
`measure(est, op, state, sse_data)` on every non-identity operator with the
state propagated through the off-diagonal ones, then `result(est, ctx, T,
sign)`; src/sse.jl:321-376 fuses them into one loop over the operator string.
The device evaluates the MagnetizationEstimator family in its measurement
kernel; any other estimator runs here, on a walker snapshot taken with
`sse_get_walker` (`measure_walker`), in the same loop order.

A host estimator is a class with
    init(cls, model, state) -> estimator          (classmethod)
    measure(self, op, state, sse_data)            op: OperCode, state: 1-based
    result(self, T, sign) -> {name: value}        what the reference passes
                                                  to measure!(ctx, name, value)
and optionally register_evaluables(cls, names, model) -> [(name, inputs,
function)] for carlo.evaluate_task.
"""
import numpy as np

from .carlo import opcodes_to_reference

_VCODE_BITS = 25        # src/opercode.jl:2
_BOND_SHIFT = 1 + _VCODE_BITS


class OperCode:
    """One operator in the reference encoding (src/opercode.jl:37-71)."""
    __slots__ = ("code",)

    def __init__(self, code):
        self.code = int(code)

    def isidentity(self):
        return self.code == 0

    def get_bond(self):
        """1-based bond index."""
        return self.code >> _BOND_SHIFT

    def get_vertex(self):
        """VertexCode value: diag | vertex_idx << 1 (1-based vertex_idx)."""
        return (self.code & ((1 << _VCODE_BITS) - 1)) >> 1

    def isdiagonal(self):
        return bool(self.get_vertex() & 1)


def measure_sign(codes, sse_data):
    """(-1)^(number of operators with a negative vertex sign) (src/sse.jl:305-314)."""
    neg = 0
    for c in codes:
        if c:
            op = OperCode(c)
            neg += sse_data.get_vertex_data(op.get_bond()).get_sign(op.get_vertex()) < 0
    return -1.0 if neg & 1 else 1.0


def measure_opstring(model, sse_data, codes, state, num_operators, T, sign, estimator_types):
    """The fused loop of src/sse.jl:321-376 over reference-encoded operator
    codes; `state` (1-based local states) is the state at the start of the
    string and is not modified.  Returns {name: value} of every estimator's
    result."""
    if not estimator_types:
        return {}
    state = np.array(state, dtype=np.int64)
    ests = [E.init(model, state.copy()) for E in estimator_types]
    n = 0
    for c in codes:
        if not c:
            continue
        op = OperCode(c)
        if not op.isdiagonal():
            bond = sse_data.bonds[op.get_bond() - 1]
            ls = sse_data.get_vertex_data(op.get_bond()).get_leg_state(op.get_vertex())
            ns = len(bond.sites)
            for i, s in enumerate(bond.sites):
                state[s - 1] = ls[ns + i]
        if n < num_operators:
            for e in ests:
                e.measure(op, state, sse_data)
        n += 1
    out = {}
    for e in ests:
        out.update(e.result(T, sign))
    return out


def measure_state(model, sse_data, T, scalars, codes, state, estimator_types=()):
    """Carlo.measure! (src/sse.jl:70-87) of one walker state: Sign,
    OperatorCount, SignOperatorCount, SignOperatorCount2, SignEnergy and the
    host estimators' results, as {name: value}."""
    sign = measure_sign(codes, sse_data)
    n = int(scalars["num_operators"])
    out = {"Sign": float(sign), "OperatorCount": float(n), "SignOperatorCount": float(sign * n),
           "SignOperatorCount2": sign * float(n) ** 2,
           "SignEnergy": -sign * (n * T + sse_data.energy_offset)
           / model.normalization_site_count()}
    out.update(measure_opstring(model, sse_data, codes, state, n, T, sign, estimator_types))
    return out


def measure_walker(batch, walker, model, estimator_types, sse_data=None):
    """Measure walker `walker` of a Batch on the host from a snapshot
    (sse_get_walker): the base observables and `estimator_types`."""
    sse_data = sse_data or model.generate_sse_data()
    sc, ops, spins = batch.get_walker(walker)
    T = batch.pt_get_parameter(walker)
    return measure_state(model, sse_data, T, sc, opcodes_to_reference(ops), spins,
                         estimator_types)


class HostMagnetizationEstimator:
    """The reference's MagnetizationEstimator
    (src/models/common/magnetization_estimator.jl:96-230) as a host
    estimator, for models or orderings without device tables and as the
    check of this host path against the device kernel.  Subclass with
    q, stagger_uc, prefix (and tag) set, or use `make`."""
    q = ()
    stagger_uc = False
    prefix = ""
    tag = None

    @classmethod
    def make(cls, q, stagger_uc, prefix, tag=None):
        return type(f"HostMagnetizationEstimator{prefix}", (cls,),
                    {"q": tuple(q), "stagger_uc": bool(stagger_uc), "prefix": prefix, "tag": tag})

    @classmethod
    def init(cls, model, state):
        e = cls()
        e.model = model
        site_idx = getattr(model, "magnetization_lattice_site_idx", lambda s: s)
        e._site = site_idx
        tmp = 0.0
        for site in range(1, len(state) + 1):
            tmp += e._sign(site) * model.magnetization_state(cls.tag, site, int(state[site - 1]))
        e.tmpmag = tmp
        e.mag, e.absmag, e.mag2, e.mag4, e.n = tmp, abs(tmp), tmp ** 2, tmp ** 4, 1
        return e

    def _sign(self, site):
        return self.model.staggered_sign(self.q, self.stagger_uc, site)

    def measure(self, op, state, sse_data):
        if not op.isdiagonal():
            bond = sse_data.bonds[op.get_bond() - 1]
            vd = sse_data.get_vertex_data(op.get_bond())
            ls = vd.get_leg_state(op.get_vertex())
            ns = len(vd.dims)
            for k in range(ns):
                site = self._site(bond.sites[k])
                if site is not None:
                    m = self.model.magnetization_state
                    self.tmpmag += self._sign(site) * (m(self.tag, site, int(ls[ns + k]))
                                                       - m(self.tag, site, int(ls[k])))
        t = self.tmpmag
        self.mag += t
        self.absmag += abs(t)
        t2 = t * t
        self.mag2 += t2
        self.mag4 += t2 * t2
        self.n += 1

    def result(self, T, sign):
        p = self.prefix
        N = self.model.normalization_site_count()
        norm = 1 / N
        mag, absmag = self.mag * norm, self.absmag * norm
        mag2, mag4 = self.mag2 * norm ** 2, self.mag4 * norm ** 4
        chi = 1 / T / (self.n + 1) / self.n * (mag ** 2 + mag2) * N
        return {f"Sign{p}Mag": sign * mag / self.n, f"Sign{p}AbsMag": sign * absmag / self.n,
                f"Sign{p}Mag2": sign * mag2 / self.n, f"Sign{p}Mag4": sign * mag4 / self.n,
                f"Sign{p}MagChi": sign * chi}

    @classmethod
    def register_evaluables(cls, names, model):
        """{p}Mag ... {p}BinderRatio (magnetization_estimator.jl:236-258)."""
        from .carlo import evaluables
        return [e for e in evaluables([f"Sign{cls.prefix}Mag4"], model.normalization_site_count())
                if e[0] not in ("Energy", "SpecificHeat")]

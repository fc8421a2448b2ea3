"""VertexData tables, operators and models (test/test_vertex_data.jl,
test/test_common_operators.jl, test/test_sse_data.jl, test/models/*.jl)."""
import itertools

import numpy as np
import pytest


def detailed_balance(sse, vd):
    u = sse.util
    vcode_idx = sse.vertex_data.vcode_idx
    for (leg_in, worm_in, v), (offset, length) in vd.transitions.items():
        weight = vd.weights[v - 1]
        cum = vd.transition_cumprobs[offset - 1:offset + length]
        probs = np.diff(np.concatenate([[0.0], cum]))
        targets = vd.transition_targets[offset - 1:offset + length]
        outs = vd.transition_step_outs[offset - 1:offset + length]
        for prob, target, (leg_out, worm_out) in zip(probs, targets, outs):
            inv = (leg_out, u.worm_inverse(worm_out, vd.dims[u.site_of_leg(leg_out, len(vd.dims)) - 1]),
                   vcode_idx(target))
            t_inv = vd.transitions[inv]
            tg = [vcode_idx(t) for t in vd.transition_targets[t_inv[0] - 1:t_inv[0] + t_inv[1]]]
            assert v in tg
            i_inv = tg.index(v)
            cb = vd.transition_cumprobs[t_inv[0] - 1:t_inv[0] + t_inv[1]]
            pb = np.diff(np.concatenate([[0.0], cb]))[i_inv]
            w_new = vd.weights[vcode_idx(target) - 1]
            assert prob * weight == pytest.approx(pb * w_new, rel=1e-6, abs=1e-9)


def test_vertex_apply_change_inverse(sse_pkg):
    u = sse_pkg.util
    dim, ns = 4, 2
    ls = np.array(list(itertools.product(range(1, dim + 1), repeat=2 * ns)), dtype=np.uint8).T
    rng = np.random.default_rng(1)
    for step_in in itertools.product(range(1, 2 * ns + 1), range(1, u.worm_count(dim) + 1)):
        v = int(rng.integers(1, dim ** (2 * ns) + 1))
        back = (step_in[0], u.worm_inverse(step_in[1], dim))
        assert sse_pkg.vertex_data.vertex_apply_change(ls, (dim,) * (2 * ns), v, step_in, back) == v


def test_heisenberg_closed_form(sse_pkg):
    """epsilon = 0 S=1/2 Heisenberg: deterministic bounce-free loops
    (test_vertex_data.jl:176-195)."""
    sp, sz = sse_pkg.spin_operators(2)
    H = np.kron(sz, sz) + 0.5 * (np.kron(sp, sp.T) + np.kron(sp.T, sp))
    vd = sse_pkg.VertexData((2, 2), H, energy_offset_factor=0.0)
    detailed_balance(sse_pkg, vd)
    assert vd.energy_offset == pytest.approx(-0.25)
    assert sum(v is None for v in vd.diagonal_vertices) == 2
    assert np.allclose(vd.weights, 0.5)
    assert np.allclose(vd.transition_cumprobs, np.ones(16))
    for v in range(1, vd.vertex_count() + 1):
        for leg in range(1, 5):
            off, _ = vd.transitions[(leg, 1, v)]
            assert vd.transition_step_outs[off - 1][0] == ((leg - 1) ^ 1) + 1


def test_heisenberg_offset_factor(sse_pkg):
    """MagnetModel always uses factor 0.25: weights 0.125/0.625/0.5, offset -0.375."""
    m = sse_pkg.MagnetModel({"lattice": {"unitcell": sse_pkg.UnitCells.square, "size": (4, 4)},
                             "J": 1.0})
    vd = m.generate_sse_data().vertex_data[0]
    assert vd.energy_offset == pytest.approx(-0.375)
    assert sorted(np.round(vd.weights, 12)) == [0.125, 0.125, 0.5, 0.5, 0.625, 0.625]
    detailed_balance(sse_pkg, vd)


@pytest.mark.parametrize("dims", [(4, 4), (2, 4), (3, 2)])
def test_random_hamiltonian(sse_pkg, dims):
    rng = np.random.default_rng(sum(dims))
    H = rng.random((int(np.prod(dims)),) * 2)
    vd = sse_pkg.VertexData(dims, H)
    detailed_balance(sse_pkg, vd)
    u = sse_pkg.util
    for v in range(vd.vertex_count()):
        ls = vd.leg_states[:, v]
        assert all(ls <= np.array(dims + dims))
    for state in range(1, int(np.prod(dims)) + 1):
        split = u.split_idx(dims, state)
        diag = vd.get_leg_state(vd.get_diagonal_vertex(state))
        assert list(diag) == list(split) + list(split)
    for (leg, worm, v) in vd.transitions:
        assert worm <= u.worm_count(dims[u.site_of_leg(leg, 2) - 1])
    for (offset, length) in vd.transitions.values():
        assert vd.transition_cumprobs[offset + length - 1] == pytest.approx(1.0, abs=1e-7)


def test_spin_operators(sse_pkg):
    sp, sz = sse_pkg.spin_operators(2)
    assert np.allclose(sp, [[0, 1], [0, 0]]) and np.allclose(sz, [[0.5, 0], [0, -0.5]])
    for d in (2, 3, 5, 17, 64):
        sp, sz = sse_pkg.spin_operators(d)
        s = (d - 1) / 2
        assert np.allclose(sz @ sz + 0.5 * (sp.T @ sp + sp @ sp.T), np.eye(d) * s * (s + 1))
        a = sse_pkg.bosonic_a_operator(d)
        assert np.allclose(a.T @ a, np.diag(np.arange(d)))


def test_sse_data_inconsistent_dims(sse_pkg):
    vd32 = sse_pkg.VertexData((3, 2), np.eye(6))
    vd23 = sse_pkg.VertexData((2, 3), np.eye(6))
    B = sse_pkg.SSEBond
    sites = sse_pkg.sse_data.generate_sites_from_bonds([vd32, vd23], [B(1, (1, 2)), B(2, (2, 3))])
    assert [s.dim for s in sites] == [3, 2, 3]
    with pytest.raises(ValueError):
        sse_pkg.sse_data.generate_sites_from_bonds([vd32, vd23], [B(1, (1, 2)), B(1, (2, 3))])
    with pytest.raises(ValueError):
        sse_pkg.sse_data.generate_sites_from_bonds([vd32, vd23], [B(1, (1, 2)), B(2, (3, 2))])


def test_magnet_params(sse_pkg):
    UC = sse_pkg.UnitCells
    L, hz = 5, 0.3
    m = sse_pkg.MagnetModel({"lattice": {"unitcell": UC.square, "size": (L, L)}, "J": 1.2,
                             "hz": hz, "Dz": 0.32, "Dx": 0.06})
    assert len(m.spin_states) == L * L and len(m.bond_params) == 2 * L * L
    assert np.allclose(m.bond_params[0].hz, (hz / 4, hz / 4))
    m = sse_pkg.MagnetModel({"lattice": {"unitcell": UC.square, "size": (L, 2 * L)},
                             "parameter_map": {"J": ("Jx", "Jy")}, "hz": hz, "Jx": 1.0, "Jy": 2.0})
    assert m.bond_params[0].J == 1.0 and m.bond_params[1].J == 2.0
    with pytest.raises(KeyError):
        sse_pkg.MagnetModel({"lattice": {"unitcell": UC.square, "size": (L, 2 * L)},
                             "parameter_map": {"J": ("Jx", "nonexistent")}, "hz": hz, "Jx": 1.0,
                             "Jy": 2.0})


def test_lift_twobody(sse_pkg):
    A = np.arange(1, 17).reshape(4, 4, order="F").astype(float)
    B = np.arange(1, 10).reshape(3, 3, order="F").astype(float)
    Cm = np.arange(1, 5).reshape(2, 2, order="F").astype(float)
    lift = sse_pkg.lift_twobody_operator
    assert np.allclose(lift(np.kron(A, B), (4, 3, 2), (1, 2)), np.kron(np.kron(A, B), np.eye(2)))
    l31 = lift(np.kron(A, B), (3, 2, 4), (3, 1))
    assert np.allclose(l31 @ np.kron(np.kron(np.eye(3), Cm), np.eye(4)),
                       np.kron(np.kron(B, Cm), A))
    assert np.allclose(l31 @ np.kron(np.kron(B, Cm), np.eye(4)), np.kron(np.kron(B @ B, Cm), A))


def test_dimer_cluster_model(sse_pkg):
    """test/models/test_cluster.jl:156-223."""
    params = {"inner_model": sse_pkg.MagnetModel,
              "lattice": {"unitcell": sse_pkg.UnitCells.fully_frust_square_bilayer, "size": (2, 2)},
              "parameter_map": {"J": [f"J{i}" for i in range(1, 10)], "Dz": ["Dz1", "Dz2"]},
              "cluster_bases": (sse_pkg.ClusterBases.dimer,), "measure_quantum_numbers": []}
    for i in range(1, 10):
        params[f"J{i}"] = 1 + 0.1 * (i - 1)
    model = sse_pkg.ClusterModel(params)
    intra, inter = model.build_cluster_hamiltonians()
    U = model.basis[0].transformation
    assert np.allclose(U.T @ intra[0] @ U, np.diag([-0.75, 0.25, 0.25, 0.25]))
    assert set(inter.keys()) == {(1, 1, (1, 0)), (1, 1, (0, 1))}
    UU = np.kron(U, U)
    expect = np.array([[1, 0, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 1, 0, 0, 0, 0, 0],
                       [0, 0, -1, 0, 1, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0, 0, 0, 0],
                       [0, 0, 1, 0, 0, 0, 1, 0, 0], [0, 0, 0, 0, 0, 0, 0, 1, 0],
                       [0, 0, 0, 0, 1, 0, -1, 0, 0], [0, 0, 0, 0, 0, 1, 0, 0, 0],
                       [0, 0, 0, 0, 0, 0, 0, 0, 1]], dtype=float)
    for h in inter.values():
        H = (UU.T @ h @ UU).reshape(4, 4, 4, 4, order="F")
        H[np.abs(H) < 1e-10] = 0
        Ht = H[1:4, 1:4, 1:4, 1:4].reshape(9, 9, order="F")
        Ht = Ht / Ht[0, 0]
        assert np.allclose(Ht, expect)
        assert not H[0, :, 0, :].any()
        assert np.allclose(np.diag(np.diag(H[:, 0, 0, :])), H[:, 0, 0, :])
        assert not H[:, 0, :, 0].any()
        assert np.allclose(np.diag(np.diag(H[0, :, :, 0])), H[0, :, :, 0])
    sd = model.generate_sse_data()
    assert len(sd.bonds) == 8

"""Exact diagonalisation of small magnets -- the statistical oracle.

Restates the reference's test/ed/ed.jl and test/ed/magnet.jl with numpy:
dense Hamiltonian over all sites (kron, first site slowest), full eigen
decomposition, thermal ensemble, and the magnetization observables incl.
the integrated correlator used for MagChi.
"""
import numpy as np


class Lifter:
    def __init__(self, site_dims):
        self.site_dims = list(site_dims)
        self.cum = [1] + list(np.cumprod(self.site_dims))

    def lift(self, pos, op):
        dl = self.cum[pos - 1]
        dr = self.cum[-1] // self.cum[pos]
        return np.kron(np.kron(np.eye(dl), op), np.eye(dr))

    def spin(self, pos, idx, spin_operators):
        sp, sz = spin_operators(self.site_dims[pos - 1])
        if idx == 1:
            return self.lift(pos, 0.5 * (sp + sp.T))
        if idx == 2:
            return self.lift(pos, 0.5j * (sp - sp.T))
        return self.lift(pos, sz)


def magnet_hamiltonian(mag, spin_operators):
    """test/ed/magnet.jl:10-26."""
    lf = Lifter(mag.spin_states)
    D = lf.cum[-1]
    H = np.zeros((D, D))
    sx = {i: lf.spin(i, 1, spin_operators) for i in range(1, len(mag.spin_states) + 1)}
    sy = {i: lf.spin(i, 2, spin_operators) for i in range(1, len(mag.spin_states) + 1)}
    sz = {i: lf.spin(i, 3, spin_operators) for i in range(1, len(mag.spin_states) + 1)}
    for bond, p in zip(mag.lattice.bonds, mag.bond_params):
        i, j = bond.i, bond.j
        heis = np.real(sx[i] @ sx[j] + sy[i] @ sy[j] + sz[i] @ sz[j])
        H = H + p.J * heis + p.J * p.d * np.real(sz[i] @ sz[j]) + p.hz[0] * np.real(sz[i]) \
            + p.hz[1] * np.real(sz[j]) + p.Dx[0] * np.real(sx[i] @ sx[i]) \
            + p.Dx[1] * np.real(sx[j] @ sx[j]) + p.Dz[0] * np.real(sz[i] @ sz[i]) \
            + p.Dz[1] * np.real(sz[j] @ sz[j])
    return H


class Ensemble:
    def __init__(self, Ts, Es, psi):
        self.Ts = np.asarray(Ts, dtype=float)
        self.Es = Es
        self.psi = psi
        En = Es - Es.min()
        rho = np.exp(-En[:, None] / self.Ts[None, :])
        self.rho = rho / rho.sum(axis=0)

    def mean(self, A):
        a = np.einsum("in,ij,jn->n", self.psi.conj(), A, self.psi).real
        return (a[:, None] * self.rho).sum(axis=0)

    def diag_mean(self, x):
        return (x[:, None] * self.rho).sum(axis=0)

    def integrated_correlator(self, A, B):
        Anm = self.psi.conj().T @ A @ self.psi
        Bnm = self.psi.conj().T @ B @ self.psi
        E = self.Es
        dE = E[None, :] - E[:, None]  # E_m - E_n at [n, m]
        out = []
        for i, T in enumerate(self.Ts):
            den = np.where(np.isclose(E[:, None], E[None, :], atol=1e-6), 2 * T, dE)
            term = 2 * self.rho[:, i][:, None] / den * Anm * Bnm.T
            out.append(np.real(term.sum()))
        return np.array(out)


def magnet_ed(model, Ts, spin_operators, hamiltonian=None):
    """Energy, SpecificHeat and, per estimator of `model`, Mag/AbsMag/Mag2/
    Mag4/BinderRatio/MagChi (test/ed/ed.jl:376-412, magnet.jl:28-97)."""
    H = magnet_hamiltonian(model, spin_operators) if hamiltonian is None else hamiltonian
    Es, psi = np.linalg.eigh(H)
    ens = Ensemble(Ts, Es, psi)
    N = model.normalization_site_count()
    Ts = np.asarray(Ts)
    obs = {"Energy": ens.diag_mean(Es) / N,
           "SpecificHeat": (ens.diag_mean(Es ** 2) - ens.diag_mean(Es) ** 2) / (Ts ** 2 * N)}
    if hasattr(model, "lattice"):
        lf = Lifter(model.spin_states)
        for est in model.get_opstring_estimators():
            M = np.zeros_like(H)
            for i in range(1, len(model.spin_states) + 1):
                d = model.spin_states[i - 1]
                mdiag = np.diag([model.magnetization_state(est.tag, i, s) for s in range(1, d + 1)])
                M = M + model.staggered_sign(est.q, est.stagger_uc, i) * lf.lift(i, mdiag)
            M = M / N
            p = est.prefix
            m2 = ens.mean(M @ M)
            m4 = ens.mean(M @ M @ M @ M)
            obs[f"{p}Mag"] = ens.mean(M)
            obs[f"{p}Mag2"] = m2
            obs[f"{p}Mag4"] = m4
            obs[f"{p}AbsMag"] = ens.mean(np.abs(M))
            obs[f"{p}BinderRatio"] = m2 ** 2 / m4
            obs[f"{p}MagChi"] = ens.integrated_correlator(M, M) * N
    return obs

"""Binning statistics and evaluables (the role Carlo's ResultTools plays).

Observables are per-sweep measurement rows in the order of
FlatModel.observable_names(); evaluables follow sse.jl:124-131 and
magnetization_estimator.jl:243-255.  Errors: bins of `binsize` rows, then
jackknife over bins for ratios.
"""
import numpy as np


def bin_rows(rows, binsize):
    nb = len(rows) // binsize
    return rows[:nb * binsize].reshape(nb, binsize, -1).mean(axis=1)


def evaluables(names, means, n_sites):
    """Evaluables from (possibly jackknifed) means of the signed observables."""
    ix = {n: k for k, n in enumerate(names)}
    s = means[..., ix["Sign"]]
    out = {"Energy": means[..., ix["SignEnergy"]] / s}
    sn2 = means[..., ix["SignOperatorCount2"]]
    sn = means[..., ix["SignOperatorCount"]]
    out["SpecificHeat"] = (sn2 / s - sn * sn / s ** 2 - sn / s) / n_sites
    for n in names:
        if n.startswith("Sign") and n.endswith(("Mag", "AbsMag", "Mag2", "Mag4", "MagChi")) \
                and n not in ("Sign",):
            out[n[4:]] = means[..., ix[n]] / s
    for n in names:
        if n.startswith("Sign") and n.endswith("Mag4"):
            p = n[4:-4]
            m4 = means[..., ix[n]]
            m2 = means[..., ix[f"Sign{p}Mag2"]]
            with np.errstate(divide="ignore", invalid="ignore"):
                out[f"{p}BinderRatio"] = np.where((m2 == 0) & (m4 == 0), 0.0, m2 ** 2 / m4 / s)
    return out


def jackknife(names, samples, n_sites):
    """samples: [n_samples][width] independent means (bins or walkers).
    Returns {evaluable: (mean, error)}."""
    samples = np.asarray(samples)
    n = len(samples)
    tot = samples.mean(axis=0)
    full = evaluables(names, tot, n_sites)
    jk = (tot[None, :] * n - samples) / (n - 1)
    jks = evaluables(names, jk, n_sites)
    res = {}
    for k, v in full.items():
        e = np.sqrt((n - 1) / n * ((jks[k] - jks[k].mean()) ** 2).sum())
        res[k] = (float(v), float(e))
    return res

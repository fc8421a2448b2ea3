"""Extract the binning metadata of every task in the reference's published
results file (docs/src/bani2v2o8.results.json, written by Carlo 0.2.2) into
tests/golden/carlo_results_meta.json.

Data only, per task: sweeps, binsize, and for a few observables the entry
fields Carlo wrote (mean, error, autocorr_time, rebin_len, internal_bin_len,
rebin_count).  The CPU tests check this library's merge rules against them.
Run here, where the reference tree is present:
    python tests/golden/make_carlo_results_meta.py
"""
import json
import os

SRC = "/root/reference/docs/src/bani2v2o8.results.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "carlo_results_meta.json")
OBS = ("Sign", "OperatorCount", "SignOperatorCount", "SignEnergy", "Energy", "SignMagChi",
       "MagChi", "SpecificHeat")


def main():
    tasks = []
    for task in json.load(open(SRC)):
        p = task["parameters"]
        tasks.append({
            "size": p["lattice"]["size"], "T": p["T"], "sweeps": p["sweeps"],
            "thermalization": p["thermalization"], "binsize": p["binsize"],
            "results": {k: task["results"][k] for k in OBS},
            "result_keys": sorted(task["results"].keys()),
            "entry_keys": sorted(task.keys()),
            "version": task["version"],
        })
    with open(OUT, "w") as f:
        f.write('{"source": "docs/src/bani2v2o8.results.json", "tasks": [\n')
        f.write(",\n".join(json.dumps(t, separators=(",", ":")) for t in tasks))
        f.write("\n]}\n")
    print(f"{len(tasks)} tasks -> {OUT}")


if __name__ == "__main__":
    main()

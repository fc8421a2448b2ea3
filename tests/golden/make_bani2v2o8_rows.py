"""Extract the rows of the reference's published results file
(docs/src/bani2v2o8.results.json: spin-1 honeycomb Heisenberg with single-ion
anisotropy Dz, job examples/bani2v2o8.jl) into tests/golden/bani2v2o8_rows.json.

Data only: task parameters and the mean/error of the observables the GPU test
compares.  Run here, where the reference tree is present:
    python tests/golden/make_bani2v2o8_rows.py
"""
import json
import os

SRC = "/root/reference/docs/src/bani2v2o8.results.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bani2v2o8_rows.json")
OBS = ("Energy", "SpecificHeat", "OperatorCount", "Mag2", "AbsMag", "MagChi", "BinderRatio")


def scalar(x):
    return x[0] if isinstance(x, list) else x


def main():
    rows = []
    for task in json.load(open(SRC)):
        p = task["parameters"]
        r = task["results"]
        rows.append({
            "size": p["lattice"]["size"], "T": p["T"], "S": p["S"], "J": p["J"],
            "Dz": p["Dz"], "sweeps": p["sweeps"], "thermalization": p["thermalization"],
            "binsize": p["binsize"],
            "results": {k: [scalar(r[k]["mean"]), scalar(r[k]["error"])] for k in OBS},
        })
    json.dump({"source": "docs/src/bani2v2o8.results.json", "rows": rows}, open(OUT, "w"),
              indent=1)
    print(f"{len(rows)} rows -> {OUT}")


if __name__ == "__main__":
    main()

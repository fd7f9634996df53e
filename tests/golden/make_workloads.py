#!/usr/bin/env python3
"""Writes tests/golden/workloads.json: the instance lists of the batched
workload (512 blast instances, seed 20261202) and the serving workload (32
requests, arrivals at 20/s and 1000/s), as hydro_amd.workloads draws them
(numpy PCG64).  Floats are stored as their repr (exact round trip), so the
test can require the generator to reproduce them bit for bit.

    python tests/golden/make_workloads.py

Then refresh MANIFEST.json's entry for workloads.json (sha256).
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hydro_amd import workloads as W  # noqa: E402


def rows(runs, arrivals=False):
    out = []
    for r in runs:
        row = [r.params.nx, repr(r.params.gamma), r.bx, r.by, repr(r.e)]
        if arrivals:
            row.append(repr(r.arrival))
        out.append(row)
    return out


def main():
    data = {
        "batch_instances": {"n": 512, "seed": W.SEED,
                            "fields": ["side", "gamma", "bx", "by", "e"],
                            "rows": rows(W.batch_instances(512))},
        "serving_requests": {
            str(rate): {"n": 32, "seed": W.SEED + 1,
                        "fields": ["side", "gamma", "bx", "by", "e", "arrival_s"],
                        "rows": rows(W.serving_requests(32, rate), arrivals=True)}
            for rate in (20.0, 1000.0)},
    }
    path = os.path.join(HERE, "workloads.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=0)
    man_path = os.path.join(HERE, "MANIFEST.json")
    man = json.load(open(man_path))
    man["workloads.json"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
    with open(man_path, "w") as f:
        json.dump(dict(sorted(man.items())), f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()

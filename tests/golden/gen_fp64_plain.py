"""Writes tests/golden/fp64_plain_gp.json: plain backtracking global placement
(EplaceNesterovOpt::opt_step_bktrk, nesterov.hpp:310-369, to overflow 0.10,
nesterov.hpp:99-117) run by the CPU oracle in FP64 on the 1M-cell bench
design and the 200K-cell design of the 512x512-bin configuration, from the
same initial placement the GPU starts from (design centres, fillers from
rand() seed 1).  The final HPWL is the reference point that the GPU's plain
and learned runs are held to (within 0.5%).

    python tests/golden/gen_fp64_plain.py [--only NAME]   (CPU, ~10-30 min for 1M)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from eplace_amd import synthetic  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "fp64_plain_gp.json")

# name: (cells, nets, seed, bins; 0 = the reference's bin rule)
DESIGNS = {
    "synth_1000000": (1_000_000, 1_050_000, 2024, 0),
    "cells200k": (200_000, 210_000, 21, 512),
}


def run(name):
    cells, nets, seed, bins = DESIGNS[name]
    d = synthetic(cells, nets, seed=seed, num_io=1000, util=0.7, name=name)
    o = Oracle(d, 1.0, mode="fp64", bins=bins)
    # corners re-derived from centres, as the device holds them
    o.set("positions", o.get("positions"))
    t0 = time.time()
    o.run("total_gradient_update", 0, -1)
    o.run("penalty_init")
    o.run("opt_init")
    k = 0
    traj = []
    while not o.stop_condition(0.10):
        o.run("opt_step", 1, -1)
        k += 1
        if k % 10 == 0:
            i = o.info()
            traj.append([k, i["overflow"], i["last_hpwl"], i["lambda"]])
            print(f"{name} it {k} overflow {i['overflow']:.4f} hpwl {i['last_hpwl']:.6g} "
                  f"({time.time() - t0:.0f}s)", flush=True)
    final_overflow = o.info()["overflow"]
    o.run("wrap_up")
    i = o.info()
    return {"name": name, "cells": cells, "nets": nets, "seed": seed, "bins": bins,
            "bins_used": [i["nx"], i["ny"]], "method": "bktrk", "target_overflow": 0.10,
            "iterations": k, "final_overflow": final_overflow, "final_hpwl": o.hpwl(),
            "backtracks": i["bktrk_count"], "seconds": time.time() - t0,
            "trajectory_every_10": {"columns": ["iteration", "overflow", "hpwl", "lambda"], "rows": traj}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=list(DESIGNS))
    args = ap.parse_args()
    try:
        with open(OUT) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        doc = {"note": "FP64 oracle, plain backtracking to overflow 0.10; written by gen_fp64_plain.py",
               "runs": []}
    for name in ([args.only] if args.only else list(DESIGNS)):
        r = run(name)
        doc["runs"] = [x for x in doc["runs"] if x["name"] != name] + [r]
        with open(OUT, "w") as f:
            json.dump(doc, f, indent=1)
        print(f"{name}: {r['iterations']} iterations, HPWL {r['final_hpwl']:.9g}, overflow "
              f"{r['final_overflow']:.4f} ({r['seconds']:.0f}s)", flush=True)


if __name__ == "__main__":
    main()
